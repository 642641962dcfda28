"""GPU parity of the e_cyl_iso solver (2.5D elastic, cylindrical
coordinates, CPML) against its FP64 oracle, through the C ABI.

FP64 storage runs the reference's expressions in its operation order
(propagate.c:2706-4119), so traces and all nine final fields must equal
the oracle's bit for bit: axisymmetric and azimuthal modes (odd modes
update v_r and tau_rtheta on the axis), force and Kurkjian sources,
divergence and curl records, no absorbing layers, layers thicker than the
model.  FP32 storage is held to a relative-L2 tolerance.
"""
import math

import numpy as np
import pytest
from tolerances import note

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

VR, VT, VZ, TRR, TTT, TZZ, TRZ, TRT, TTZ, DIV, CURL = 13, 14, 2, 15, 18, 6, 17, 16, 19, 11, 12
FR, FT, FZ, BULK, KURK = 9, 10, 11, 6, 12
CYL_FP32_TOL = 4e-6


def deck(nx=30, nz=44, np_=6, n=0, seed=0, h=2.0, src=((FZ, 0, 20),), nsteps=150, f0=30.0,
         kappa_max=1.0, alpha_max=1, fluid=False):
    rng = np.random.default_rng(seed)
    nx2, nz2 = nx + np_, nz + 2 * np_
    zc = nz2 // 2
    vp = np.where(np.arange(nz2)[None, :] < zc, 2500.0, 3500.0) * np.ones((nx2, 1))
    vp = vp * (1 + 0.05 * rng.uniform(-1, 1, (nx2, nz2)))
    vs = vp / rng.uniform(1.6, 1.9, (nx2, nz2))
    if fluid:
        vs[:3, :] = 0.0                                   # a fluid borehole around the axis
    rho = 2200.0 * (1 + 0.05 * rng.uniform(-1, 1, (nx2, nz2)))
    dt = 0.4 * h / (math.sqrt(2) * vp.max())
    cfg = dict(nx=nx, nz=nz, np=np_, dx=h, dz=h, dt=dt, n=n, kappa_max=kappa_max,
               alpha_max=(math.pi * f0 if alpha_max else 0.0))
    types = np.array([s[0] for s in src], np.intc)
    idx = np.array([s[1] * nz2 + np_ + s[2] for s in src], np.uintp)
    comps = [VR, VT, VZ, TRR, TTT, TZZ, TRZ, TRT, TTZ, DIV, CURL]
    ridx = np.array([int(rng.integers(0, nx2)) * nz2 + int(rng.integers(0, nz2))
                     for _ in range(len(comps))] + [0 * nz2 + zc, 1 * nz2 + zc], np.uintp)
    rcomp = np.array(comps + [VR, CURL], np.intc)
    if types[0] == KURK:
        m, _, _ = O.cyl_material(vp, vs, rho)
        lam = m["lij"].reshape(-1)[idx[0]]
        fct = O.kurkjian_fct(lam, h, h, src[0][1] * h, n, 1e-6, f0, dt)
        L = np.array([len(fct)], np.uintp)
    else:
        L, fct = O.source_functions([1.0] * len(src), [f0] * len(src), dt)
    return dict(cfg=cfg, vp=vp, vs=vs, rho=rho, types=types, idx=idx, L=L, fct=fct, ridx=ridx,
                rcomp=rcomp, nsteps=nsteps, f0=f0, alpha_max=alpha_max)


def run_gpu(ctx, d, storage):
    from seismod_amd.solver import CylConfig, Cylindrical
    c = d["cfg"]
    cfg = CylConfig(nx=c["nx"], nz=c["nz"], np=c["np"], dr=c["dx"], dz=c["dz"], dt=c["dt"],
                    f0=d["f0"], mode=c["n"], kappa_max=c["kappa_max"], alpha_max=d["alpha_max"],
                    storage=storage)
    s = Cylindrical(ctx, cfg, d["vp"], d["vs"], d["rho"])
    s.set_sources(d["types"], d["idx"], d["L"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.run(d["nsteps"])
    tr = s.traces(0, d["nsteps"])
    f = np.concatenate([s.field(k).reshape(-1) for k in O.CYL_FIELDS])
    s.close()
    return tr, f


def run_oracle(d):
    return O.cyl_run(d["cfg"], d["vp"], d["vs"], d["rho"], d["types"], d["idx"], d["L"], d["fct"],
                     d["ridx"], d["rcomp"], d["nsteps"])


CASES = [
    dict(n=0, seed=1),                                            # Fz on the axis
    dict(n=0, seed=2, src=((FR, 5, 10), (FT, 9, 30), (BULK, 12, 25)), fluid=True),
    dict(n=1, seed=3, src=((FR, 0, 22),)),                        # odd mode: axis rows
    dict(n=2, seed=4, src=((FZ, 4, 15), (FT, 4, 15))),
    dict(n=3, seed=5, src=((FT, 0, 18),), kappa_max=2.5),
    dict(n=0, seed=6, src=((KURK, 0, 20),)),                      # Kurkjian, axisymmetric
    dict(n=1, seed=7, src=((KURK, 2, 20),)),                      # Kurkjian, dipole
    dict(n=0, seed=8, np_=0),                                     # abs 0
    dict(nx=4, nz=5, np_=9, n=1, seed=9, src=((FZ, 1, 2),), nsteps=80),   # layers > model
    dict(nx=70, nz=130, np_=12, n=0, seed=10, src=((FZ, 0, 60),), nsteps=260, alpha_max=0),
]


def ids(c):
    return f"n{c['n']}_seed{c['seed']}"


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_cyl_fp64_bit_exact(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f64")
    assert np.abs(f_o).max() > 0
    assert np.array_equal(tr, tr_o), np.abs(tr - tr_o).max()
    assert np.array_equal(f, f_o), np.abs(f - f_o).max()


@pytest.mark.parametrize("case", CASES[:4] + CASES[5:6], ids=ids)
def test_cyl_fp32_tolerance(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f32")
    e_tr = np.linalg.norm(tr - tr_o) / np.linalg.norm(tr_o)
    e_f = np.linalg.norm(f - f_o) / np.linalg.norm(f_o)
    note("cyl_fp32_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < CYL_FP32_TOL and e_f < CYL_FP32_TOL


def test_cyl_cfl_and_material_errors(ctx):
    from seismod_amd._lib import SM_ERR_ARG, SM_ERR_CFL, SeismodError
    from seismod_amd.solver import CylConfig, Cylindrical
    d = deck()
    c = d["cfg"]
    cfg = CylConfig(nx=c["nx"], nz=c["nz"], np=c["np"], dr=c["dx"], dz=c["dz"], dt=c["dt"] * 3,
                    f0=30.0)
    with pytest.raises(SeismodError) as e:
        Cylindrical(ctx, cfg, d["vp"], d["vs"], d["rho"])
    assert e.value.status == SM_ERR_CFL
    cfg.dt = c["dt"]
    rho = d["rho"].copy()
    rho[3, 4] = 0.0                                       # 1/rho not finite
    with pytest.raises(SeismodError) as e:
        Cylindrical(ctx, cfg, d["vp"], d["vs"], rho)
    assert e.value.status == SM_ERR_ARG


def test_cyl_timing_counts_steps_and_leaves_results(ctx):
    """Per-step timing brackets every step and changes no value."""
    from seismod_amd.solver import CylConfig, Cylindrical
    d = deck(n=1, seed=11, src=((FR, 3, 20),), nsteps=60)
    tr0, f0 = run_gpu(ctx, d, "f32")
    c = d["cfg"]
    cfg = CylConfig(nx=c["nx"], nz=c["nz"], np=c["np"], dr=c["dx"], dz=c["dz"], dt=c["dt"],
                    f0=d["f0"], mode=c["n"], kappa_max=c["kappa_max"], alpha_max=d["alpha_max"],
                    storage="f32")
    s = Cylindrical(ctx, cfg, d["vp"], d["vs"], d["rho"])
    s.set_sources(d["types"], d["idx"], d["L"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.set_timing(True)
    s.run(25)
    s.run(35)
    ms, n = s.timing()
    assert n == 60 and ms > 0
    assert np.array_equal(s.traces(0, d["nsteps"]), tr0)
    s.set_timing(False)
    assert s.timing() == (0.0, 0)
    s.close()
