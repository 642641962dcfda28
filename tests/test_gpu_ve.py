"""GPU parity of the ve_vti_pml solver (viscoelastic P-SV, 2D VTI,
pseudo-spectral derivatives, CPML, RK4, L + 1 relaxation mechanisms)
against its FP64 oracle (oracle/oracle_ve.py), through the C ABI.

The GPU FFTs (hand-written LDS transforms, hipFFT for other lengths) and
the oracle's (numpy) differ in rounding, so FP64 results
agree to FFT rounding rather than bit for bit: traces and final fields are
held to VE_FP64_TOL relative L2; FP32 storage to VE_FP32_TOL.  Decks:
layered VTI viscoelastic with layers on every side and one or two
dilatational mechanisms (L = 2 exercises the relaxation sum's read of
neighbouring cells), a model wider than deep with kappa_max > 1 and
alpha_max = 0, no layers (abs 0) with odd nx, an odd nz2 (x derivatives by
real transforms), and every source type, several on one cell.
"""
import math

import numpy as np
import pytest
from tolerances import note

from oracle import oracle_ve as V
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

VE_FP64_TOL = 2e-13     # measured 8.2e-16 to 2.1e-14 (two FFT implementations' rounding)
VE_FP32_TOL = 1e-5      # measured 5.3e-7 to 1.1e-6


def deck(nx=40, nz=50, np_=8, seed=0, h=10.0, nsteps=120, f0=25.0, kappa_max=1.0, alpha_max=1,
         abs_=True, L=1, srcs=None):
    rng = np.random.default_rng(seed)
    npad = np_ if abs_ else 0
    nx2, nz2 = nx + 2 * npad, nz + 2 * npad
    nl = int(rng.integers(2, 5))
    tops = np.sort(rng.choice(np.arange(1, nz2), size=nl - 1, replace=False))
    lay = np.searchsorted(tops, np.arange(nz2), side="right")
    ones = np.ones((nx2, 1))
    vp = rng.uniform(2200, 4000, nl)[lay] * (1 + 0.03 * rng.uniform(-1, 1, (nx2, nz2)))
    vs = vp / rng.uniform(1.6, 2.1, nl)[lay]
    rho = rng.uniform(1800, 2600, nl)[lay] * ones
    eps = rng.uniform(0.0, 0.25, nl)[lay] * ones
    dlt = rng.uniform(-0.05, 0.15, nl)[lay] * ones
    Q1 = np.stack([rng.uniform(20, 90, nl)[lay] * ones for _ in range(L)])
    f1 = np.stack([rng.uniform(5, 60, nl)[lay] * ones for _ in range(L)])
    Q2 = rng.uniform(15, 70, nl)[lay] * ones
    f2 = rng.uniform(5, 60, nl)[lay] * ones
    mu = vs * vs * rho * 1e-9
    K = 1e-9 * (vp * vp * rho - 4. / 3. * vs * vs * rho)
    vmax = float((vp * np.sqrt(1 + 2 * eps)).max())
    dt = 0.2 * h / vmax
    if srcs is None:
        srcs = [(V.BULK, nx2 // 2, nz2 // 2, 1.0), (V.SXZ, nx2 // 3, nz2 // 3, 0.6)]
    types = np.array([s[0] for s in srcs], np.intc)
    idx = np.array([s[1] * nz2 + s[2] for s in srcs], np.uintp)
    Lf, fct = O.source_functions([s[3] for s in srcs], [f0] * len(srcs), 0.5 * dt)
    fct = fct / (h * h)
    comps = [V.VX, V.VZ, V.TXX, V.TZZ, V.TXZ]
    ridx = np.array([int(rng.integers(0, nx2)) * nz2 + int(rng.integers(0, nz2)) for _ in range(10)], np.uintp)
    rcomp = np.array([comps[k % 5] for k in range(10)], np.intc)
    cfg = dict(nx=nx, nz=nz, np=npad, dx=h, dz=h, dt=dt, vmax=vmax, kappa_max=kappa_max,
               alpha_max=(math.pi * f0 if alpha_max else 0.0), iwipe=1 if abs_ else 0)
    return dict(cfg=cfg, model=(K, mu, rho, eps, dlt, Q1, f1, Q2, f2), L=L, types=types, idx=idx, Lf=Lf,
                fct=fct, ridx=ridx, rcomp=rcomp, nsteps=nsteps, f0=f0, alpha_max=alpha_max, abs=abs_, np=np_)


def run_gpu(ctx, d, storage, hipfft=False):
    from seismod_amd.solver import VeConfig, ViscoelasticPSV
    c = d["cfg"]
    cfg = VeConfig(nx=c["nx"], nz=c["nz"], np=d["np"], dx=c["dx"], dz=c["dz"], dt=c["dt"], f0=d["f0"], L=d["L"],
                   vmax=c["vmax"], kappa_max=c["kappa_max"], alpha_max=d["alpha_max"], abs=d["abs"],
                   storage=storage, hipfft=hipfft)
    s = ViscoelasticPSV(ctx, cfg, *d["model"])
    s.set_sources(d["types"], d["idx"], d["Lf"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.run(d["nsteps"] // 2)
    s.run(d["nsteps"] - d["nsteps"] // 2)
    tr = s.traces(0, d["nsteps"])
    f = np.stack([s.field(k) for k in ("txx", "tzz", "txz", "vx", "vz")])
    s.close()
    return tr, f


def run_oracle(d):
    off = np.concatenate([[0], np.cumsum(d["Lf"])]).astype(np.int64)
    src = [(int(t), int(i), d["fct"][off[k]:off[k + 1]]) for k, (t, i) in enumerate(zip(d["types"], d["idx"]))]
    mat = V.ve_material(*d["model"])
    tr, W = V.ve_run(d["cfg"], mat, src, d["ridx"], d["rcomp"], d["nsteps"])
    return tr, W[:5]


CASES = [
    dict(seed=1),
    dict(seed=2, L=2),                                                 # relaxation sum over neighbours
    dict(nx=70, nz=20, np_=6, seed=3, kappa_max=3.0, alpha_max=0),     # nx2 > nz2
    dict(nx=33, nz=28, np_=5, seed=4, abs_=False, L=2),                # no layers, odd nx2
    dict(nx=31, nz=27, np_=6, seed=5),                                 # odd nz2: real x transforms
    dict(nx=30, nz=36, np_=7, seed=6, L=3, srcs=[(V.SX, 20, 25, 1.0), (V.SZ, 20, 25, -0.4), (V.BULK, 20, 25, 0.7),
                                                 (V.SXZ, 20, 25, 0.5), (V.SXZ, 21, 24, 0.2), (3, 22, 22, 1.0)]),
]


def ids(c):
    return f"seed{c['seed']}"


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_ve_fp64_matches_oracle(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f64")
    assert np.abs(tr_o).max() > 0
    e_tr, e_f = rel(tr, tr_o), rel(f, f_o)
    note("ve_fp64_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < VE_FP64_TOL and e_f < VE_FP64_TOL, (e_tr, e_f)


@pytest.mark.parametrize("case", CASES[:2], ids=ids)
def test_ve_fp32_tolerance(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f32")
    e_tr, e_f = rel(tr, tr_o), rel(f, f_o)
    note("ve_fp32_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < VE_FP32_TOL and e_f < VE_FP32_TOL, (e_tr, e_f)


def test_ve_reset_timing_and_errors(ctx):
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    from seismod_amd.solver import VeConfig, ViscoelasticPSV
    d = deck(seed=8, nsteps=40)
    tr0, _ = run_gpu(ctx, d, "f64")
    c = d["cfg"]
    cfg = VeConfig(nx=c["nx"], nz=c["nz"], np=d["np"], dx=c["dx"], dz=c["dz"], dt=c["dt"], f0=d["f0"],
                   vmax=c["vmax"])
    s = ViscoelasticPSV(ctx, cfg, *d["model"])
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
    q = model[7].copy()
    q[3, 4] = 0.0
    model[7] = q
    with pytest.raises(SeismodError) as e:
        ViscoelasticPSV(ctx, cfg, *model)
    assert e.value.status == SM_ERR_ARG


@pytest.mark.parametrize("case", [CASES[0], CASES[3]], ids=ids)
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_ve_own_transforms_match_hipfft(ctx, case, storage):
    """The hand-written LDS transforms (axis lengths of small prime factors)
    against hipFFT on every axis (SM_PSTD_HIPFFT): the same results to
    rounding."""
    d = deck(**case)
    tr, f = run_gpu(ctx, d, storage)
    tr_h, f_h = run_gpu(ctx, d, storage, hipfft=True)
    # measured: FP64 3.3e-15 to 1.8e-14, FP32 3.5e-7 to 4.0e-6 (two FP32 paths,
    # each within 2e-6 of the FP64 oracle)
    tol = 2e-13 if storage == "f64" else 4e-5
    e_tr, e_f = rel(tr, tr_h), rel(f, f_h)
    note("ve_own_vs_hipfft", case=ids(case), storage=storage, traces=e_tr, fields=e_f)
    assert e_tr < tol and e_f < tol, (e_tr, e_f)
