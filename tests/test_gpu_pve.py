"""GPU parity of the pve_iso_pml solver (poroviscoelastic P-SV, 2D
isotropic Biot medium, pseudo-spectral derivatives, CPML, RK4 with the
viscous coupling stepped analytically) against its FP64 oracle
(oracle/oracle_pve.py), through the C ABI.

The GPU FFTs (hand-written LDS transforms, hipFFT for other lengths) and
the oracle's (numpy) differ in rounding, so FP64 results agree to FFT
rounding rather than bit for bit: each final field is held to PVE_FP64_TOL
relative L2 and the traces of each recorded component to PVE_FP64_TOL of
its scale (errors()); FP32 storage to PVE_FP32_TOL.  Decks: layered rock with layers on every side, every
receiver component (divergence and curl included), a model wider than deep
with kappa_max > 1 and alpha_max = 0, no layers (abs 0) with odd nx2, an
odd nz2 (x derivatives by real transforms), every source type on one cell
with a layer of zero shear and frame moduli, layers with the lossless
branch (Q >= 1e4) and an inviscid fluid (eta = 0), the median smoothing of
the layers (pml_median), and VTI Biot media
(pve_vti_pml: Thomsen's epsilon and delta, tortuosity and permeability
different along x and z) with layers, kappa_max > 1 and without layers.
"""
import math

import numpy as np
import pytest
from tolerances import note

from oracle import oracle_pve as B
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

PVE_FP64_TOL = 2e-13     # measured 1.2e-15 to 5.9e-15 (two FFT implementations' rounding)
PVE_FP32_TOL = 2e-5      # measured 5.1e-7 to 2.7e-6

COMPS = [B.VX, B.VZ, B.QX, B.QZ, B.TXX, B.TZZ, B.TXZ, B.P, B.DIV, B.CURL]
FIELDS = ("txx", "tzz", "txz", "p", "vx", "vz", "qx", "qz")


def biot_vmax(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T):
    """Fast P velocity of the undrained, unrelaxed medium (m/s)."""
    Km, Ks, Kf, MU = K_m * 1e3, K_s * 1e3, K_f * 1e3, mu * 1e3
    alpha = 1 - Km / Ks
    M = 1. / ((alpha - phi) / Ks + phi / Kf)
    rs, rf = rho_s * 1e-6, rho_f * 1e-6
    rho = (1 - phi) * rs + phi * rf
    mm = T * rf / phi
    v = 0.0
    for n in range(K_m.size):
        C = np.array([[Km.flat[n] + 4. / 3. * MU.flat[n] + alpha.flat[n] ** 2 * M.flat[n], alpha.flat[n] * M.flat[n]],
                      [alpha.flat[n] * M.flat[n], M.flat[n]]])
        R = np.array([[rho.flat[n], rf.flat[n]], [rf.flat[n], mm.flat[n]]])
        v = max(v, float(np.sqrt(np.max(np.linalg.eigvals(np.linalg.solve(R, C)).real))))
    return v


def deck(nx=40, nz=50, np_=8, seed=0, h=10.0, nsteps=120, f0=25.0, kappa_max=1.0, alpha_max=1, abs_=True,
         srcs=None, fluid_layer=False, lossless_layer=False, vti=False, median=False):
    rng = np.random.default_rng(seed)
    npad = np_ if abs_ else 0
    nx2, nz2 = nx + 2 * npad, nz + 2 * npad
    nl = int(rng.integers(3, 5))
    tops = np.sort(rng.choice(np.arange(1, nz2), size=nl - 1, replace=False))
    lay = np.searchsorted(tops, np.arange(nz2), side="right")
    ones = np.ones((nx2, 1))

    def prop(lo, hi, jitter=0.0):
        v = rng.uniform(lo, hi, nl)[lay] * ones
        return v * (1 + jitter * rng.uniform(-1, 1, (nx2, nz2))) if jitter else v
    K_s = prop(30, 40)
    K_m = prop(3, 15, 0.03)
    K_f = prop(1.5, 3.0)
    phi = prop(0.1, 0.35)
    mu = prop(3, 10, 0.03)
    rho_s = prop(2500, 2800)
    rho_f = prop(900, 1100)
    T = prop(1.5, 3.0)
    eta = prop(0.5, 2.0)
    kappa = prop(100, 2000)
    Q = prop(10, 100)
    fq = prop(10, 80)
    if fluid_layer:          # a suspension: no frame, no shear
        K_m[:, lay == 1] = 0.0
        mu[:, lay == 1] = 0.0
    if lossless_layer:
        Q[:, lay == 0] = 2e4
        eta[:, lay == 2] = 0.0
    model = (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa, Q, fq)
    vmax = biot_vmax(*model[:8])
    if vti:
        # tortuosity and permeability along z, Thomsen's epsilon and delta
        T3, kappa3 = prop(1.5, 3.0), prop(100, 2000)
        eps, dl = prop(0.0, 0.25), prop(-0.05, 0.15)
        model = (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, T3, eta, kappa, kappa3, Q, fq, eps, dl)
        vmax *= float(np.sqrt(1 + 2 * eps.max()))
    dt = 0.2 * h / vmax
    if srcs is None:
        srcs = [(B.BULK, nx2 // 2, nz2 // 2, 1.0), (B.SXZ, nx2 // 3, nz2 // 3, 0.6), (B.SF, nx2 // 2 + 3, nz2 // 2 - 5, 0.8)]
    types = np.array([s[0] for s in srcs], np.intc)
    idx = np.array([s[1] * nz2 + s[2] for s in srcs], np.uintp)
    Lf, fct = O.source_functions([s[3] for s in srcs], [f0] * len(srcs), 0.5 * dt)
    fct = fct / (h * h)
    nrec = 2 * len(COMPS)
    # receivers in the middle half of the grid, where every component carries signal
    ridx = np.array([int(rng.integers(nx2 // 4, 3 * nx2 // 4)) * nz2 + int(rng.integers(nz2 // 4, 3 * nz2 // 4))
                     for _ in range(nrec)], np.uintp)
    rcomp = np.array([COMPS[k % len(COMPS)] for k in range(nrec)], np.intc)
    cfg = dict(nx=nx, nz=nz, np=npad, dx=h, dz=h, dt=dt, vmax=vmax, kappa_max=kappa_max,
               alpha_max=(math.pi * f0 if alpha_max else 0.0), iwipe=1 if abs_ else 0)
    return dict(cfg=cfg, model=model, types=types, idx=idx, Lf=Lf, fct=fct, ridx=ridx, rcomp=rcomp, nsteps=nsteps,
                f0=f0, alpha_max=alpha_max, abs=abs_, np=np_, vti=vti, median=median)


def config(d, storage="f64", hipfft=False):
    from seismod_amd.solver import PveConfig
    c = d["cfg"]
    return PveConfig(nx=c["nx"], nz=c["nz"], np=d["np"], dx=c["dx"], dz=c["dz"], dt=c["dt"], f0=d["f0"],
                     vmax=c["vmax"], kappa_max=c["kappa_max"], alpha_max=d["alpha_max"], abs=d["abs"],
                     storage=storage, hipfft=hipfft, pml_median=d.get("median", False))


def solver(ctx, d, storage="f64", hipfft=False):
    from seismod_amd.solver import Poroviscoelastic, PoroviscoelasticVTI
    return (PoroviscoelasticVTI if d["vti"] else Poroviscoelastic)(ctx, config(d, storage, hipfft), *d["model"])


def run_gpu(ctx, d, storage, hipfft=False):
    s = solver(ctx, d, storage, hipfft)
    s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.run(d["nsteps"] // 2)
    s.run(d["nsteps"] - d["nsteps"] // 2)
    tr = s.traces(0, d["nsteps"])
    f = np.stack([s.field(k) for k in FIELDS])
    dc = {k: s.field(k) for k in ("div", "curl")}
    s.close()
    return tr, f, dc


def run_oracle(d):
    off = np.concatenate([[0], np.cumsum(d["Lf"])]).astype(np.int64)
    src = [(int(t), int(i), d["fct"][off[k]:off[k + 1]]) for k, (t, i) in enumerate(zip(d["types"], d["idx"]))]
    if d["vti"]:
        mat = B.pve_vti_material(*d["model"], d["cfg"]["dt"])
    else:
        mat = B.pve_material(*d["model"], d["cfg"]["dt"], median_np=d["cfg"]["np"] if d.get("median") else 0)
    tr, W = B.pve_run(d["cfg"], mat, src, d["ridx"], d["rcomp"], d["nsteps"])
    return tr, W[:8]


CASES = [
    dict(seed=1),
    dict(nx=70, nz=20, np_=6, seed=3, kappa_max=3.0, alpha_max=0),     # nx2 > nz2
    dict(nx=33, nz=28, np_=5, seed=4, abs_=False),                     # no layers, odd nx2
    dict(nx=31, nz=27, np_=6, seed=5, lossless_layer=True),            # odd nz2: real x transforms
    dict(nx=30, nz=36, np_=7, seed=6, fluid_layer=True,
         srcs=[(B.SX, 20, 25, 1.0), (B.SZ, 20, 25, -0.4), (B.BULK, 20, 25, 0.7), (B.SXZ, 20, 25, 0.5),
               (B.SF, 20, 25, 0.3), (B.BULK_S, 20, 25, -0.6), (B.SXZ, 21, 24, 0.2), (3, 22, 22, 1.0)]),
    dict(nx=30, nz=34, np_=6, seed=7, median=True),                    # median smoothing of the layers
    # VTI Biot media (pve_vti_pml)
    dict(seed=11, vti=True),
    dict(nx=70, nz=20, np_=6, seed=13, kappa_max=3.0, alpha_max=0, vti=True),
    dict(nx=33, nz=28, np_=5, seed=14, abs_=False, vti=True, lossless_layer=True),
]


def ids(c):
    return f"seed{c['seed']}" + ("_vti" if c.get("vti") else "") + ("_median" if c.get("median") else "")


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


FIELD_OF = {B.VX: 4, B.VZ: 5, B.QX: 6, B.QZ: 7, B.TXX: 0, B.TZZ: 1, B.TXZ: 2, B.P: 3}


def errors(d, tr, f, tr_o, f_o):
    """Each component on its own (their scales differ by orders of
    magnitude): the final fields by relative L2; the traces by their largest
    difference over the component's scale, the larger of its largest
    recorded value and its final field's (a receiver the slow, diffusive
    fluid motion hardly reaches records values far below the rounding of
    the transforms that carry the whole field); divergence and curl share
    the scale of the velocity derivatives, the larger of their records (the
    curl of a mostly dilatational field is a small difference of two
    derivatives)."""
    per = {}
    dc = np.isin(d["rcomp"], [B.DIV, B.CURL])
    for c in COMPS:
        sel = d["rcomp"] == c
        scale = np.abs(tr_o[:, sel]).max()
        if c in FIELD_OF:
            scale = max(scale, np.abs(f_o[FIELD_OF[c]]).max())
        else:
            scale = np.abs(tr_o[:, dc]).max()
        per[f"tr{c}"] = float(np.abs(tr[:, sel] - tr_o[:, sel]).max() / scale)
    for q, k in enumerate(FIELDS):
        per[k] = rel(f[q], f_o[q])
    errors.last = per
    return max(per[f"tr{c}"] for c in COMPS), max(per[k] for k in FIELDS)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_pve_fp64_matches_oracle(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f, dc = run_gpu(ctx, d, "f64")
    assert all(np.abs(tr_o[:, d["rcomp"] == c]).max() > 0 for c in COMPS)
    e_tr, e_f = errors(d, tr, f, tr_o, f_o)
    note("pve_fp64_vs_oracle", case=ids(case), **errors.last)
    assert e_tr < PVE_FP64_TOL and e_f < PVE_FP64_TOL, errors.last
    # the divergence and curl fields hold the last recorded row
    for c, k in ((B.DIV, "div"), (B.CURL, "curl")):
        sel = d["rcomp"] == c
        assert np.allclose(dc[k].reshape(-1)[d["ridx"][sel].astype(np.int64)], tr[-1, sel], rtol=0, atol=0)


@pytest.mark.parametrize("case", CASES[:2] + CASES[6:7], ids=ids)
def test_pve_fp32_tolerance(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f, _ = run_gpu(ctx, d, "f32")
    e_tr, e_f = errors(d, tr, f, tr_o, f_o)
    note("pve_fp32_vs_oracle", case=ids(case), **errors.last)
    assert e_tr < PVE_FP32_TOL and e_f < PVE_FP32_TOL, errors.last


def test_pve_reset_timing_and_errors(ctx):
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    from seismod_amd.solver import Poroviscoelastic
    d = deck(seed=8, nsteps=40)
    tr0, _, _ = run_gpu(ctx, d, "f64")
    s = Poroviscoelastic(ctx, config(d), *d["model"])
    s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.run(17)
    s.reset()
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.set_timing(True)
    s.run(d["nsteps"])
    ms, n = s.timing()
    assert n == d["nsteps"] and ms > 0
    assert np.array_equal(s.traces(0, d["nsteps"]), tr0)
    with pytest.raises(SeismodError) as e:
        s.set_receivers(d["ridx"][:1], np.array([1], np.intc), 5)      # Vy is not a P-SV field
    assert e.value.status == SM_ERR_ARG
    s.close()
    model = list(d["model"])
    k = model[9].copy()
    k[3, 4] = 0.0                                                      # zero permeability
    model[9] = k
    with pytest.raises(SeismodError) as e:
        Poroviscoelastic(ctx, config(d), *model)
    assert e.value.status == SM_ERR_ARG


@pytest.mark.parametrize("case", [CASES[0], CASES[2]], ids=ids)
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_pve_own_transforms_match_hipfft(ctx, case, storage):
    """The hand-written LDS transforms against hipFFT on every axis
    (SM_PSTD_HIPFFT): the same results to rounding."""
    d = deck(**case)
    tr, f, _ = run_gpu(ctx, d, storage)
    tr_h, f_h, _ = run_gpu(ctx, d, storage, hipfft=True)
    tol = 2e-13 if storage == "f64" else 4e-5
    e_tr, e_f = errors(d, tr, f, tr_h, f_h)
    note("pve_own_vs_hipfft", case=ids(case), storage=storage, **errors.last)
    assert e_tr < tol and e_f < tol, errors.last


def test_pve_energy_of_a_region(ctx):
    """The kinetic energy record (saveEnergy): the quadratic form of v and q
    summed over a region after every step, against the oracle's v and q at
    every region cell; the same bits on a second run (a fixed summation
    order); a region outside the grid is refused."""
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    from seismod_amd.solver import Poroviscoelastic
    d = deck(seed=9, nsteps=60)
    nx2, nz2 = d["model"][0].shape
    i1, i2, j1, j2 = nx2 // 2 - 3, nx2 // 2 + 2, nz2 // 2 - 2, nz2 // 2 + 4
    cells = [i * nz2 + j for i in range(i1, i2 + 1) for j in range(j1, j2 + 1)]
    rng = np.random.default_rng(3)
    r11, r22 = rng.uniform(2000, 2600, len(cells)), rng.uniform(1500, 2500, len(cells))
    r12 = -rng.uniform(100, 400, len(cells))
    d["ridx"] = np.array(cells * 4, np.uintp)
    d["rcomp"] = np.repeat(np.array([B.VX, B.VZ, B.QX, B.QZ], np.intc), len(cells))
    tr_o, _ = run_oracle(d)
    vx, vz, qx, qz = (tr_o[:, k * len(cells):(k + 1) * len(cells)] for k in range(4))
    want = (0.5 * r11 * (vx * vx + vz * vz) + r12 * (vx * qx + vz * qz) + 0.5 * r22 * (qx * qx + qz * qz)).sum(1)
    got = []
    for _ in range(2):
        s = Poroviscoelastic(ctx, config(d), *d["model"])
        s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
        s.set_energy(i1, i2, j1, j2, r11, r12, r22, d["nsteps"])
        s.run(d["nsteps"])
        got.append(s.energy(0, d["nsteps"]))
        with pytest.raises(SeismodError) as e:
            big = np.ones((nx2 - i1 + 1) * (j2 - j1 + 1))
            s.set_energy(i1, nx2, j1, j2, big, big, big, 5)
        assert e.value.status == SM_ERR_ARG
        s.close()
    assert np.abs(want).max() > 0
    err = float(np.abs(got[0] - want).max() / np.abs(want).max())
    note("pve_energy_vs_oracle", err=err)
    assert err < PVE_FP64_TOL, err
    assert np.array_equal(got[0], got[1])


def test_pve_energy_per_axis(ctx):
    """sm_pve_set_energy_axes (the VTI program's energy, eq. 7.171 of
    Carcione 2007: different coefficients along x and z) on a VTI deck
    against the oracle's v and q at every region cell."""
    d = deck(seed=15, nsteps=50, vti=True)
    nx2, nz2 = d["model"][0].shape
    i1, i2, j1, j2 = nx2 // 2 - 2, nx2 // 2 + 3, nz2 // 2 - 3, nz2 // 2 + 2
    cells = [i * nz2 + j for i in range(i1, i2 + 1) for j in range(j1, j2 + 1)]
    rng = np.random.default_rng(4)
    co = [rng.uniform(-500, 2500, len(cells)) for _ in range(6)]
    d["ridx"] = np.array(cells * 4, np.uintp)
    d["rcomp"] = np.repeat(np.array([B.VX, B.VZ, B.QX, B.QZ], np.intc), len(cells))
    tr_o, _ = run_oracle(d)
    vx, vz, qx, qz = (tr_o[:, k * len(cells):(k + 1) * len(cells)] for k in range(4))
    want = (0.5 * co[0] * vx * vx + co[1] * vx * qx + 0.5 * co[2] * qx * qx + 0.5 * co[3] * vz * vz +
            co[4] * vz * qz + 0.5 * co[5] * qz * qz).sum(1)
    s = solver(ctx, d)
    s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
    s.set_energy_axes(i1, i2, j1, j2, *co, d["nsteps"])
    s.run(d["nsteps"])
    got = s.energy(0, d["nsteps"])
    s.close()
    err = float(np.abs(got - want).max() / np.abs(want).max())
    note("pve_energy_axes_vs_oracle", err=err)
    assert err < PVE_FP64_TOL, err


@pytest.mark.parametrize("case,storage", [(CASES[0], "f64"), (CASES[6], "f32")],
                         ids=["seed1-f64", "seed11_vti-f32"])
def test_pve_checkpoint_continues_bit_for_bit(ctx, case, storage):
    """sm_pve_get_state / sm_pve_set_state (the state behind chkpt_inc and
    -c, io_utils.c:2355-2811): a run stopped after k steps, saved, loaded
    into a new solver and continued gives the same traces, energy record
    and fields bit for bit as the run without a stop; a blob cut short or
    saved from another grid is refused."""
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    d = deck(**case)
    nx2, nz2 = d["model"][0].shape
    i1, i2, j1, j2 = nx2 // 2 - 3, nx2 // 2 + 3, nz2 // 2 - 2, nz2 // 2 + 2
    npts = (i2 - i1 + 1) * (j2 - j1 + 1)
    co = [np.random.default_rng(5 + q).uniform(100, 2500, npts) for q in range(6)]

    def make():
        s = solver(ctx, d, storage)
        s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
        s.set_energy_axes(i1, i2, j1, j2, *co, d["nsteps"])
        s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
        return s

    def finish(s):
        out = (s.traces(0, d["nsteps"]), s.energy(0, d["nsteps"]), np.stack([s.field(k) for k in FIELDS]))
        s.close()
        return out

    s = make()
    s.run(d["nsteps"])
    want = finish(s)
    k = d["nsteps"] // 3 + 1
    s = make()
    s.run(k)
    blob = s.get_state()
    s.close()
    s = make()
    s.set_state(blob)
    s.run(d["nsteps"] - k)
    got = finish(s)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert np.abs(want[1]).max() > 0 and np.abs(want[0]).max() > 0
    s = make()
    for bad in (blob[:-8], solver_blob_other_grid(ctx, d, storage)):
        with pytest.raises(SeismodError) as e:
            s.set_state(bad)
        assert e.value.status == SM_ERR_ARG
    s.close()


def solver_blob_other_grid(ctx, d, storage):
    d2 = deck(nx=d["cfg"]["nx"] + 2, nz=d["cfg"]["nz"], np_=d["np"], seed=21, vti=d["vti"])
    s = solver(ctx, d2, storage)
    blob = s.get_state()
    s.close()
    return blob
