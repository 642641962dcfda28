"""GPU tests of 16-bit wavefield storage (BF16 and FP16 with power-of-two
per-field scales) in the fused 3D kernel.

Arithmetic stays FP32; only the stored state (P, v, CPML memory variables)
is rounded to 16 bits once per step.  Every storage mode is compared with
the FP64 oracle on the same deck: a 48^3 layered model with a 10-cell CPML,
a 15 Hz Ricker and a receiver plane of Vz plus a few P traces, run long
enough for the wavelet to cross the grid and reach the CPML.

Tolerances are relative L2 over all traces together (the trace error the
receiver plane sees), and over the pressure field at mid-run, while the
wavelet is inside the grid (at the end only the small late-time residue is
left, whose relative error says little).
"""
import numpy as np
import pytest
from tolerances import FP32_REL_L2

from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

VZ, P = 2, 8

# relative L2 vs the FP64 oracle
TOL = {
    "f32": dict(traces=FP32_REL_L2, field=FP32_REL_L2),
    "f16": dict(traces=5e-3, field=1e-2),
    "bf16": dict(traces=5e-2, field=1e-1),
    "f16-sr": dict(traces=5e-3, field=1e-2),
    "bf16-sr": dict(traces=5e-2, field=1e-1),
}
MID = 200


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def deck(n=48, np_=10, nsteps=400, seed=3, amp=1.0):
    rng = np.random.default_rng(seed)
    from decks import layered_3d
    vp_i, rho_i = layered_3d(n, n, n, rng, vmin=1500.0, vmax=5000.0, nlayers=5)
    vp = np.pad(vp_i, np_, mode="edge")
    rho = np.pad(rho_i, np_, mode="edge")
    dx = 10.0
    dt = 0.47 / (vp_i.max() * np.sqrt(3.0) / dx)
    n2 = n + 2 * np_
    cfg = dict(nx=n, ny=n, nz=n, np=np_, dx=dx, dy=dx, dz=dx, dt=dt, f0=15.0, vmax=5000.0,
               pml_order=2.0, Rc=1e-5, kappa_max=1.0, alpha_max=1)
    si = sj = n2 // 2
    sk = np_ + 5
    src_idx = np.array([(si * n2 + sj) * n2 + sk], dtype=np.uintp)
    rec = []
    for a in range(np_ + 2, n2 - np_ - 2, 6):
        for b in range(np_ + 2, n2 - np_ - 2, 6):
            rec.append(((a * n2 + b) * n2 + np_ + 3, VZ))
    for k in range(np_ + 8, n2 - np_, 8):
        rec.append(((si * n2 + sj) * n2 + k, P))
    rec_idx = np.array([r[0] for r in rec], dtype=np.uintp)
    rec_comp = np.array([r[1] for r in rec], dtype=np.intc)
    return dict(cfg=cfg, vp=vp, rho=rho, src_idx=src_idx, A=np.array([amp]),
                f=np.array([15.0]), rec_idx=rec_idx, rec_comp=rec_comp, nsteps=nsteps)


def gpu_run(ctx, d, storage):
    """storage: f32, bf16, f16, or bf16-sr / f16-sr (stochastic rounding)"""
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    c = d["cfg"]
    cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                    dz=c["dz"], dt=c["dt"], f0=c["f0"], vmax=c["vmax"], pml_order=c["pml_order"],
                    Rc=c["Rc"], kappa_max=c["kappa_max"], alpha_max=c["alpha_max"],
                    storage=storage.split("-")[0], sround=storage.endswith("-sr"))
    s = Acoustic3D(ctx, cfg, d["vp"], d["rho"])
    lengths, fct = source_functions(d["A"], d["f"], c["dt"])
    s.set_sources(d["src_idx"], lengths, fct)
    s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
    s.run(MID)
    mid = s.field(P)
    s.run(d["nsteps"] - MID)
    out = s.traces(0, d["nsteps"]), mid
    s.close()
    return out


def oracle_run(d):
    c = d["cfg"]
    lengths, fct = O.source_functions(d["A"], d["f"], c["dt"])
    tr, _ = O.a3d_run(c, d["vp"], d["rho"], d["src_idx"], lengths, fct, d["rec_idx"],
                      d["rec_comp"], d["nsteps"], nthreads=3, want_fields=False)
    _, fields = O.a3d_run(c, d["vp"], d["rho"], d["src_idx"], lengths, fct, d["rec_idx"],
                          d["rec_comp"], MID, nthreads=3)
    return tr, fields[:d["vp"].size]    # fields are P, vx, vy, vz


@pytest.fixture(scope="module")
def ref():
    d = deck()
    return d, oracle_run(d)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16", "f16-sr", "bf16-sr"])
def test_storage_error_vs_fp64_oracle(ctx, ref, storage):
    d, (tr_ref, p_ref) = ref
    tr, p = gpu_run(ctx, d, storage)
    assert np.isfinite(tr).all() and np.isfinite(p).all()
    e_tr = rel_l2(tr, tr_ref)
    e_p = rel_l2(p.reshape(-1), p_ref)
    print(f"\n{storage}: traces rel-L2 {e_tr:.3e}  P rel-L2 at step {MID} {e_p:.3e}")
    assert e_tr <= TOL[storage]["traces"], e_tr
    assert e_p <= TOL[storage]["field"], e_p


@pytest.mark.parametrize("amp", [1e-6, 1e4])
def test_f16_scaling_follows_source_amplitude(ctx, amp):
    """FP16 has 5 exponent bits: without the per-field scales a weak source
    underflows to zero and a strong one overflows to inf.  The scales come
    from the injected amplitude, so the error relative to the signal stays
    that of a unit source."""
    d = deck(n=32, np_=8, nsteps=250, amp=amp)
    tr32, p32 = gpu_run(ctx, d, "f32")
    tr16, p16 = gpu_run(ctx, d, "f16")
    assert np.isfinite(tr16).all() and np.isfinite(p16).all()
    assert np.abs(tr32).max() > 0
    e = rel_l2(tr16, tr32)
    print(f"\namp {amp:g}: f16 vs f32 traces rel-L2 {e:.3e}")
    assert e <= TOL["f16"]["traces"], e


@pytest.mark.parametrize("storage", ["bf16-sr", "f16-sr"])
def test_stochastic_rounding_is_reproducible_and_launch_invariant(ctx, storage):
    """The rounding bits depend on element index and step only."""
    from seismod_amd.solver import A3dConfig, Acoustic3D, source_functions
    d = deck(n=32, np_=8, nsteps=120)
    c = d["cfg"]
    outs = []
    for launch in ((4, 2), (2, 5), (8, 1)):
        cfg = A3dConfig(nx=c["nx"], ny=c["ny"], nz=c["nz"], np=c["np"], dx=c["dx"], dy=c["dy"],
                        dz=c["dz"], dt=c["dt"], f0=c["f0"], storage=storage.split("-")[0],
                        sround=True)
        s = Acoustic3D(ctx, cfg, d["vp"], d["rho"])
        s.set_launch(*launch)
        L, fct = source_functions(d["A"], d["f"], c["dt"])
        s.set_sources(d["src_idx"], L, fct)
        s.set_receivers(d["rec_idx"], d["rec_comp"], d["nsteps"])
        s.run(50)
        s.run(70)
        outs.append((s.traces(0, d["nsteps"]), s.field(P)))
        s.close()
    for tr, p in outs[1:]:
        assert np.array_equal(tr, outs[0][0]) and np.array_equal(p, outs[0][1])
