"""GPU parity of the ve_vti_sh_pml solver (viscoelastic SH, 2D VTI,
pseudo-spectral derivatives, CPML, RK4) against its FP64 oracle
(oracle/oracle_sh.py), through the C ABI.

The GPU FFTs (hand-written LDS transforms, hipFFT for other lengths) and
the oracle's (numpy) differ in rounding, so FP64 results
agree to FFT rounding rather than bit for bit: traces and final fields are
held to SH_FP64_TOL relative L2.  FP32 storage is held to SH_FP32_TOL.
Decks: layered VTI viscoelastic with layers on every side, a model wider
than deep (nx2 > nz2: the dx_vy layer indexed over nx2 columns),
kappa_max > 1 with alpha_max = 0, no layers (abs 0) with odd nx, several
sources on one cell, and an odd nz2 (x derivatives by real transforms
instead of complex transforms of column pairs).
"""
import math

import numpy as np
import pytest
from tolerances import note

from oracle import oracle_sh as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

SH_FP64_TOL = 1.2e-13   # measured 1.5e-15 to 1.25e-14 (two FFT implementations' rounding)
SH_FP32_TOL = 2e-5      # measured 7.0e-7 to 2.0e-6


def deck(nx=40, nz=50, np_=8, seed=0, h=10.0, nsteps=120, f0=25.0, kappa_max=1.0, alpha_max=1,
         abs_=True, srcs=None):
    rng = np.random.default_rng(seed)
    npad = np_ if abs_ else 0
    nx2, nz2 = nx + 2 * npad, nz + 2 * npad
    nl = int(rng.integers(2, 5))
    tops = np.sort(rng.choice(np.arange(1, nz2), size=nl - 1, replace=False))
    lay = np.searchsorted(tops, np.arange(nz2), side="right")
    vs = rng.uniform(1200, 2500, nl)[lay] * (1 + 0.03 * rng.uniform(-1, 1, (nx2, nz2)))
    rho = rng.uniform(1800, 2600, nl)[lay] * np.ones((nx2, 1))
    gamma = rng.uniform(0.0, 0.3, nl)[lay] * np.ones((nx2, 1))
    Q = rng.uniform(15, 80, nl)[lay] * np.ones((nx2, 1))
    f = rng.uniform(10, 40, nl)[lay] * np.ones((nx2, 1))
    mu = vs * vs * rho * 1e-9
    vmax = float((vs * np.sqrt(1 + 2 * gamma)).max())
    dt = 0.25 * h / vmax
    if srcs is None:
        srcs = [(S.SYZ, nx2 // 2, nz2 // 2, 1.0), (S.SXY, nx2 // 3, nz2 // 3, 0.6)]
    types = np.array([s[0] for s in srcs], np.intc)
    idx = np.array([s[1] * nz2 + s[2] for s in srcs], np.uintp)
    L, fct = O.source_functions([s[3] for s in srcs], [f0] * len(srcs), 0.5 * dt)
    fct = fct / (h * h)
    comps = [S.VY, S.TXY, S.TYZ]
    ridx = np.array([int(rng.integers(0, nx2)) * nz2 + int(rng.integers(0, nz2)) for _ in range(9)],
                    np.uintp)
    rcomp = np.array([comps[k % 3] for k in range(9)], np.intc)
    cfg = dict(nx=nx, nz=nz, np=npad, dx=h, dz=h, dt=dt, vmax=vmax, kappa_max=kappa_max,
               alpha_max=(math.pi * f0 if alpha_max else 0.0), iwipe=1 if abs_ else 0)
    return dict(cfg=cfg, mu=mu, rho=rho, gamma=gamma, Q=Q, f=f, types=types, idx=idx, L=L, fct=fct,
                ridx=ridx, rcomp=rcomp, nsteps=nsteps, f0=f0, alpha_max=alpha_max, abs=abs_, np=np_)


def run_gpu(ctx, d, storage, hipfft=False):
    from seismod_amd.solver import ShConfig, ViscoelasticSH
    c = d["cfg"]
    cfg = ShConfig(nx=c["nx"], nz=c["nz"], np=d["np"], dx=c["dx"], dz=c["dz"], dt=c["dt"], f0=d["f0"],
                   vmax=c["vmax"], kappa_max=c["kappa_max"], alpha_max=d["alpha_max"], abs=d["abs"],
                   storage=storage, hipfft=hipfft)
    s = ViscoelasticSH(ctx, cfg, d["mu"], d["rho"], d["gamma"], d["Q"], d["f"])
    s.set_sources(d["types"], d["idx"], d["L"], d["fct"])
    s.set_receivers(d["ridx"], d["rcomp"], d["nsteps"])
    s.run(d["nsteps"] // 2)
    s.run(d["nsteps"] - d["nsteps"] // 2)
    tr = s.traces(0, d["nsteps"])
    f = np.stack([s.field(k) for k in ("txy", "tyz", "vy")])
    s.close()
    return tr, f


def run_oracle(d):
    off = np.concatenate([[0], np.cumsum(d["L"])]).astype(np.int64)
    src = [(int(t), int(i), d["fct"][off[k]:off[k + 1]]) for k, (t, i) in enumerate(zip(d["types"], d["idx"]))]
    tr, W = S.sh_run(d["cfg"], d["mu"], d["rho"], d["gamma"], d["Q"], d["f"], src, d["ridx"], d["rcomp"],
                     d["nsteps"])
    return tr, W[:3]


CASES = [
    dict(seed=1),
    dict(nx=70, nz=20, np_=6, seed=2, kappa_max=3.0, alpha_max=0),     # nx2 > nz2
    dict(nx=33, nz=28, np_=5, seed=3, abs_=False),                     # no layers, odd nx2
    dict(nx=30, nz=36, np_=7, seed=4, srcs=[(S.SYZ, 20, 25, 1.0), (S.SYZ, 20, 25, -0.4),
                                            (S.SXY, 20, 25, 0.7), (S.SXY, 21, 24, 0.2)]),
    dict(nx=31, nz=27, np_=6, seed=6),                                 # odd nz2: real x transforms
]


def ids(c):
    return f"seed{c['seed']}"


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_sh_fp64_matches_oracle(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f64")
    assert np.abs(tr_o).max() > 0
    e_tr, e_f = rel(tr, tr_o), rel(f, f_o)
    note("sh_fp64_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < SH_FP64_TOL and e_f < SH_FP64_TOL, (e_tr, e_f)


@pytest.mark.parametrize("case", CASES[:2], ids=ids)
def test_sh_fp32_tolerance(ctx, case):
    d = deck(**case)
    tr_o, f_o = run_oracle(d)
    tr, f = run_gpu(ctx, d, "f32")
    e_tr, e_f = rel(tr, tr_o), rel(f, f_o)
    note("sh_fp32_vs_oracle", case=ids(case), traces=e_tr, fields=e_f)
    assert e_tr < SH_FP32_TOL and e_f < SH_FP32_TOL, (e_tr, e_f)


def test_sh_reset_timing_and_errors(ctx):
    from seismod_amd._lib import SM_ERR_ARG, SeismodError
    from seismod_amd.solver import ShConfig, ViscoelasticSH
    d = deck(seed=5, nsteps=40)
    tr0, _ = run_gpu(ctx, d, "f64")
    c = d["cfg"]
    cfg = ShConfig(nx=c["nx"], nz=c["nz"], np=d["np"], dx=c["dx"], dz=c["dz"], dt=c["dt"], f0=d["f0"],
                   vmax=c["vmax"])
    s = ViscoelasticSH(ctx, cfg, d["mu"], d["rho"], d["gamma"], d["Q"], d["f"])
    s.set_sources(d["types"], d["idx"], d["L"], d["fct"])
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
        s.set_receivers(d["ridx"][:1], np.array([2], np.intc), 5)      # Vz is not an SH field
    assert e.value.status == SM_ERR_ARG
    s.close()
    q = d["Q"].copy()
    q[3, 4] = 0.0
    with pytest.raises(SeismodError) as e:
        ViscoelasticSH(ctx, cfg, d["mu"], d["rho"], d["gamma"], q, d["f"])
    assert e.value.status == SM_ERR_ARG


@pytest.mark.parametrize("case", [CASES[0], CASES[3]], ids=ids)
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_sh_own_transforms_match_hipfft(ctx, case, storage):
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
    note("sh_own_vs_hipfft", case=ids(case), storage=storage, traces=e_tr, fields=e_f)
    assert e_tr < tol and e_f < tol, (e_tr, e_f)
