"""CPU tests that pin the FP64 oracle of ve_vti_sh_pml (oracle/oracle_sh.py).

The reference cannot be built here (structs.h includes fftw3.h) and ships
no tests.  The restatement is pinned by
  - the derivative operators: on band-limited periodic signals the shifted
    spectral derivatives D^+ and D^- are exact, at the half-cell positions
    (propagate.c:566-577, 2487-2510);
  - the CPML stage coefficients (pml.c:1959-2106): their semi-implicit
    update has the exact fixed point of the memory-variable ODE and
    approximates its decay to third order in the step;
  - the material (ve_vti_sh_pml.c:475-585): a standard linear solid with
    its relaxation peak at f and quality Q (omega^2 tau_e tau_s = 1,
    omega (tau_e - tau_s) = 2/Q), c66 = (1 + 2 gamma) c44;
  - invariants: zero source, linearity;
  - physics: a Syz source in a homogeneous, nearly elastic medium against
    the 2D Green's function (v_y = (1/rho) d/dz of the scalar solution),
    and in a VTI medium against the same solution with x stretched by
    sqrt(1 + 2 gamma); the CPML absorbs.
The finite-Q time response is pinned only through the material relations
("parity unpinned" for attenuation itself).
"""
import math

import numpy as np
import pytest

from oracle import oracle_sh as S


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [32, 33, 48])
@pytest.mark.parametrize("axis", [0, 1])
def test_shifted_spectral_derivatives_are_exact(n, axis):
    d = 2.5
    L = n * d
    i = np.arange(n)
    for m in (1, 3, n // 2 - 1):
        k = 2 * math.pi * m / L
        f = np.sin(k * i * d + 0.3)
        a = np.stack([f] * 5, axis=1 - axis)
        for fwd, shift in ((True, 0.5), (False, -0.5)):
            got = S.deriv(a, axis, S.k_multipliers(n, d, fwd))
            want1 = k * np.cos(k * (i + shift) * d + 0.3)
            want = np.stack([want1] * 5, axis=1 - axis)
            assert np.abs(got - want).max() < 1e-11 * k, (m, fwd)


@pytest.mark.parametrize("np_,kmax,amax", [(10, 1.0, 31.4), (20, 3.0, 0.0), (8, 7.0, 62.8)])
def test_cpml_stage_coefficients(np_, kmax, amax):
    dx, dz, dt, vmax, order, Rc = 5.0, 8.0, 4e-4, 4500.0, 2.0, 1e-5
    t = S.sh_cpml(np_, dx, dz, dt, vmax, order, Rc, kmax, amax, 1)
    for ax, d in (("x", dx), ("z", dz)):
        L = np_ * d
        d0 = -(order + 1.) * vmax * math.log(Rc) / (2. * L)
        for half, off, kname in (("", 0.0, "k"), ("h", -0.5, "kh"), ("H", 0.5, "kH")):
            for m in range(np_ + 1 if half != "h" else np_):
                dist = (np_ - m + off) * d / L
                k = 1. + (kmax - 1.) * dist ** order
                dd = d0 * dist ** order
                al = (m + {"": 0.0, "h": 0.5, "H": -0.5}[half]) * amax / np_
                assert t[1][kname + "_" + ax][m] == pytest.approx(k, rel=1e-14)
                assert t[1]["b" + half + "_" + ax][m] == 1.0 and t[1]["c" + half + "_" + ax][m] == 0.0
                lam = dd / k + al
                for s, h in ((23, 0.5 * dt), (4, dt)):
                    b = t[s]["b" + half + "_" + ax][m]
                    c = t[s]["c" + half + "_" + ax][m]
                    # decay factor of psi' = -lam psi - (d/k^2) g, third order
                    assert abs(b - math.exp(-lam * h)) <= (lam * h) ** 3 / 12 + 1e-15
                    # the update's fixed point is the ODE's steady state
                    if lam > 0:
                        assert c / (1. - b) == pytest.approx(-dd / (k * k * lam), rel=1e-10, abs=1e-300)
                    else:
                        assert c == 0.0


def test_material_is_a_standard_linear_solid():
    rng = np.random.default_rng(1)
    shape = (9, 11)
    vs = rng.uniform(800, 2500, shape)
    rho = rng.uniform(1500, 2700, shape)
    gamma = rng.uniform(0.0, 0.3, shape)
    Q = rng.uniform(10, 200, shape)
    f = rng.uniform(5, 60, shape)
    mu = vs * vs * rho * 1e-9
    m = S.sh_material(mu, rho, gamma, Q, f)
    # (i + 1/2, j): c66, tau_s_2, eta2; (i, j + 1/2): c44, tau_s_1, eta1
    for ts_, eta_, Qa, fa in ((m["ts2"][:-1], m["eta2"][:-1], 0.5 * (Q[:-1] + Q[1:]), 0.5 * (f[:-1] + f[1:])),
                              (m["ts1"][:, :-1], m["eta1"][:, :-1], 0.5 * (Q[:, :-1] + Q[:, 1:]),
                               0.5 * (f[:, :-1] + f[:, 1:]))):
        tau_s = 1.0 / ts_
        eta = 1.0 / (1.0 - eta_)
        tau_e = tau_s / eta
        w = 2 * math.pi * fa
        assert np.allclose(w * w * tau_e * tau_s, 1.0, rtol=1e-12)
        assert np.allclose(w * (tau_e - tau_s), 2.0 / Qa, rtol=1e-9)
    c44 = 0.5 * (mu[:, :-1] + mu[:, 1:]) * 1000.0
    assert np.allclose(m["c44"][:, :-1], c44, rtol=1e-15)
    g = 0.5 * (gamma[:-1] + gamma[1:])
    assert np.allclose(m["c66"][:-1], (1 + 2 * g) * 0.5 * (mu[:-1] + mu[1:]) * 1000.0, rtol=1e-14)
    # c44_0 multiplies the x-average eta2 (as written), c66_0 its own eta2
    eta2 = 1.0 / (1.0 - m["eta2"])
    assert np.allclose(m["c44_0"][:, :-1], m["c44"][:, :-1] * eta2[:, :-1], rtol=1e-14)
    assert np.allclose(m["c66_0"][:-1], m["c66"][:-1] * eta2[:-1], rtol=1e-14)
    assert np.allclose(m["irho"], 1.0 / (rho * 1e-6), rtol=1e-15)


def _ricker(f0, dt_half, n):
    tcut = 1.0 / f0
    t = np.arange(n) * dt_half - tcut
    a2 = (math.pi * f0) ** 2
    return 2 * a2 * (1 - 2 * a2 * t * t) * np.exp(-a2 * t * t)


def _homogeneous(n, h, vs0, rho0, gamma0, Q0, f0, dt, nsteps, recs, np_=0, src_type=S.SYZ, A=1.0):
    nx2 = nz2 = n + 2 * np_
    shape = (nx2, nz2)
    mu = np.full(shape, vs0 * vs0 * rho0 * 1e-9)
    rho = np.full(shape, rho0)
    gamma = np.full(shape, gamma0)
    Q = np.full(shape, Q0)
    fr = np.full(shape, f0)
    L = int(math.ceil((1.0 / f0) * 2.0 / (0.5 * dt)))
    fct = A * _ricker(f0, 0.5 * dt, L) / (h * h)
    c = nx2 // 2
    src = [(src_type, c * nz2 + c, fct)]
    ridx = [(c + a) * nz2 + c + b for a, b, _ in recs]
    rcomp = [cp for _, _, cp in recs]
    cfg = dict(nx=n, nz=n, np=np_, dx=h, dz=h, dt=dt, vmax=vs0 * math.sqrt(1 + 2 * gamma0),
               kappa_max=1.0, alpha_max=math.pi * f0, iwipe=1 if np_ else 0)
    tr, W = S.sh_run(cfg, mu, rho, gamma, Q, fr, src, ridx, rcomp, nsteps)
    return tr, W, fct


def test_zero_source_and_linearity():
    recs = [(5, 3, S.VY), (-4, 6, S.TXY), (2, -7, S.TYZ)]
    tr1, W1, _ = _homogeneous(40, 10.0, 2000.0, 2200.0, 0.1, 30.0, 20.0, 1e-3, 60, recs, np_=6)
    tr3, W3, _ = _homogeneous(40, 10.0, 2000.0, 2200.0, 0.1, 30.0, 20.0, 1e-3, 60, recs, np_=6, A=3.0)
    tr0, W0, _ = _homogeneous(40, 10.0, 2000.0, 2200.0, 0.1, 30.0, 20.0, 1e-3, 60, recs, np_=6, A=0.0)
    assert np.abs(tr1).max() > 0 and not np.any(W0)
    assert np.allclose(tr3, 3.0 * tr1, rtol=1e-12, atol=1e-14 * np.abs(tr1).max())


def _green_vy(x, z, t, f0, c, rho_s, stretch=1.0):
    """v_y at (x, z) from s(t) = Ricker(t - 1/f0) injected into tau_yz at the
    origin: rho v_tt = mu lap(v) + d/dz(s delta), so v = d/dz psi with
    psi = 1/(2 pi rho c^2) int_0^acosh(ct/r) s(t - (r/c) cosh u) du."""
    a2 = (math.pi * f0) ** 2

    def s(tt):
        tau = tt - 1.0 / f0
        return 2 * a2 * (1 - 2 * a2 * tau * tau) * np.exp(-a2 * tau * tau) * (np.abs(tau) <= 1.0 / f0)

    def psi(xx, zz):
        r = math.hypot(xx / stretch, zz)
        out = np.zeros_like(t)
        for k, tt in enumerate(t):
            if c * tt > r:
                u = np.linspace(0.0, math.acosh(c * tt / r), 4001)
                out[k] = np.trapezoid(s(tt - (r / c) * np.cosh(u)), u)
        return out / (2 * math.pi * rho_s * c * c) / stretch

    e = 1e-3 * math.hypot(x, z)
    return (psi(x, z + e) - psi(x, z - e)) / (2 * e)


@pytest.mark.parametrize("gamma0", [0.0, 0.25])
def test_syz_source_against_green_function(gamma0):
    """Homogeneous medium, Q = 1e7 (elastic limit), no layers, 128^2 cells of
    10 m, 10 Hz: receivers 20-35 cells from the source (v_y at nodes, the
    source at the tau_yz node half a cell below its index), compared until
    just before the wrapped-around wave could arrive.  PSTD has no grid
    dispersion at 20 cells per wavelength: measured 0.05-0.15%; held to 1%,
    where a half-cell shift of the source gives 7-19% and a missing
    sqrt(1 + 2 gamma) stretch 18-184%."""
    n, h, vs0, rho0, f0, dt = 128, 10.0, 2000.0, 2200.0, 10.0, 1e-3
    recs = [(0, 25, S.VY), (0, -30, S.VY), (22, 22, S.VY), (30, 10, S.VY), (-35, -5, S.VY)]
    nsteps = 330
    tr, _, _ = _homogeneous(n, h, vs0, rho0, gamma0, 1e7, f0, dt, nsteps, recs)
    rho_s = rho0 * 1e-6
    t = (np.arange(nsteps) + 1) * dt
    for k, (a, b, _) in enumerate(recs):
        want = _green_vy(a * h, (b - 0.5) * h, t, f0, vs0, rho_s, math.sqrt(1 + 2 * gamma0))
        assert rel_l2(tr[:, k], want) < 0.01, (k, rel_l2(tr[:, k], want))


def test_cpml_absorbs():
    """With layers, the energy left in the interior after the wavelet has
    crossed it is a small fraction of what remains without them (periodic
    wrap-around)."""
    n, h, f0, dt = 60, 10.0, 15.0, 1e-3
    e = []
    for np_ in (0, 15):
        _, W, _ = _homogeneous(n, h, 2000.0, 2200.0, 0.0, 1e6, f0, dt, 500, [(0, 1, S.VY)], np_=np_)
        inner = W[2][np_:np_ + n, np_:np_ + n]
        e.append(float(np.sum(inner * inner)))
    assert e[1] < 0.02 * e[0], e
