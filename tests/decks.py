"""Seeded synthetic decks shared by the CPU and GPU tests.

A deck is everything one run needs: grid/config, model, sources and
receivers, in the reference's layouts.  Models are padded the way
read_model_a pads them (io_utils.c:1089-1143: edge replication into the
CPML, z then y then x), which is np.pad(..., mode="edge").
"""
from __future__ import annotations

import numpy as np

VX, VY, VZ, TXX, TZZ, TXZ, P = 0, 1, 2, 5, 6, 7, 8
SX, SY, SZ, SXY, SXZ, SYZ, BULK = range(7)


def layered_3d(nx, ny, nz, rng, vmin=1500.0, vmax=4500.0, nlayers=4, perturb=0.05):
    """Layered Vp along z plus a smooth random perturbation; Gardner rho."""
    tops = np.sort(rng.integers(1, max(nz - 1, 2), size=nlayers - 1))
    vals = np.sort(rng.uniform(vmin, vmax, size=nlayers))
    vz = np.empty(nz)
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vz[:] = vals[layer]
    vp = np.broadcast_to(vz, (nx, ny, nz)).copy()
    if perturb:
        noise = rng.standard_normal((nx, ny, nz))
        # cheap smoothing: a few passes of a 3-point box filter on each axis
        for _ in range(3):
            for ax in range(3):
                noise = (np.roll(noise, 1, ax) + noise + np.roll(noise, -1, ax)) / 3.0
        noise /= max(np.abs(noise).max(), 1e-30)
        vp *= 1.0 + perturb * noise
    rho = 310.0 * vp ** 0.25
    return vp, rho


def a3d_deck(nx=10, ny=8, nz=6, np_=3, seed=0, dx=10.0, dy=None, cfl=0.5, f0=25.0,
             nsteps=60, kappa_max=1.0, alpha_max=1, Rc=1e-5, src_at=None, nrec=6):
    rng = np.random.default_rng(seed)
    dy = dx if dy is None else dy
    vp_i, rho_i = layered_3d(nx, ny, nz, rng)
    vp = np.pad(vp_i, np_, mode="edge")
    rho = np.pad(rho_i, np_, mode="edge")
    vpmax = vp_i.max()
    dt = cfl / (vpmax * np.sqrt(2.0 / dx ** 2 + 1.0 / dy ** 2))
    nx2, ny2, nz2 = vp.shape
    cfg = dict(nx=nx, ny=ny, nz=nz, np=np_, dx=dx, dy=dy, dz=dx, dt=dt, f0=f0, vmax=5000.0,
               pml_order=2.0, Rc=Rc, kappa_max=kappa_max, alpha_max=alpha_max)
    if src_at is None:
        src_at = (nx // 2, ny // 2, max(nz // 3, 0))
    si, sj, sk = (c + np_ for c in src_at)
    src_idx = np.array([(si * ny2 + sj) * nz2 + sk], dtype=np.uintp)
    recs = []
    comps = [VX, VY, VZ, P]
    for r in range(nrec):
        # interior of the padded grid: the outermost planes of some fields
        # never change (a_iso_3d.c:605-661)
        i = int(rng.integers(1, nx2 - 1))
        j = int(rng.integers(1, ny2 - 1))
        k = int(rng.integers(1, nz2 - 1))
        recs.append(((i * ny2 + j) * nz2 + k, comps[r % 4]))
    rec_idx = np.array([r[0] for r in recs], dtype=np.uintp)
    rec_comp = np.array([r[1] for r in recs], dtype=np.intc)
    return dict(cfg=cfg, vp=vp, rho=rho, src_idx=src_idx, f=np.array([f0]), A=np.array([1.0]),
                rec_idx=rec_idx, rec_comp=rec_comp, nsteps=nsteps)


def e2d_deck(nx=16, nz=12, npml=3, seed=0, h=5.0, cfl=0.5, f0=40.0, nsteps=80,
             src_type=BULK, q1_compat=True, nrec=8, two_layer=True):
    rng = np.random.default_rng(seed)
    vp = np.empty((nx, nz))
    vs = np.empty((nx, nz))
    rho = np.empty((nx, nz))
    if two_layer:
        zc = nz // 2
        vp[:, :zc], vs[:, :zc], rho[:, :zc] = 2500.0, 1400.0, 2000.0
        vp[:, zc:], vs[:, zc:], rho[:, zc:] = 3500.0, 2000.0, 2400.0
        # lateral variation so the PML sees different edge values
        vp *= 1.0 + 0.05 * rng.uniform(-1, 1, size=(nx, nz))
        vs *= 1.0 + 0.05 * rng.uniform(-1, 1, size=(nx, nz))
        rho *= 1.0 + 0.05 * rng.uniform(-1, 1, size=(nx, nz))
    else:
        vp[:], vs[:], rho[:] = 3000.0, 1700.0, 2200.0
    dt = cfl * h / (np.sqrt(2.0) * vp.max())
    cfg = dict(nx=nx, nz=nz, npml=npml, h=h, dt=dt, q1_compat=q1_compat)
    src_idx = np.array([(nx // 2) * nz + nz // 3], dtype=np.uintp)
    comps = [VX, VZ, TXX, TZZ, TXZ]
    rec_idx = np.array([int(rng.integers(0, nx)) * nz + int(rng.integers(0, nz))
                        for _ in range(nrec)], dtype=np.uintp)
    rec_comp = np.array([comps[r % 5] for r in range(nrec)], dtype=np.intc)
    return dict(cfg=cfg, vp=vp, vs=vs, rho=rho, src_idx=src_idx,
                src_type=np.array([src_type], dtype=np.intc), f=np.array([f0]),
                A=np.array([1.0]), rec_idx=rec_idx, rec_comp=rec_comp, nsteps=nsteps)
