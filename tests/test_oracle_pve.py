"""CPU tests that pin the FP64 oracle of pve_iso_pml (oracle/oracle_pve.py).

The reference cannot be built here (structs.h includes fftw3.h) and ships
no tests.  The restatement shares its derivative operators, CPML tables
and layer loops with the SH and P-SV oracles (pinned in their own tests)
and is pinned by
  - the material (pve_iso_pml.c:795-1145) against textbook Biot theory in
    an independent form: M = [(alpha - phi)/K_s + phi/K_f]^-1, alpha =
    1 - K_m/K_s, Gassmann's undrained modulus K_m + alpha^2 M, a standard
    linear solid (omega^2 tau_e tau_s = 1, omega (tau_e - tau_s) = 2/Q at
    f0), m = T rho_f / phi;
  - the analytic step of the viscous coupling (eq. 76-77 of Carcione 1996):
    it must conserve the momentum rho v + rho_f q and decay q at the rate
    of rho v' + rho_f q' = 0, rho_f v' + m q' = -(eta/kappa) q;
  - invariants: zero source, linearity;
  - physics, lossless limit (eta = 0, Q >= 1e4): a line of Bulk sources is
    a 1D problem whose solution is the sum of Biot's fast and slow P waves,
    v = -sgn(x) sum_k g_k phi_k s(t - |x|/c_k) / (2 c_k^2) from the
    eigenvectors phi_k of rho^-1 C (C the undrained stiffness matrix,
    rho the density matrix), g the source projected on them; the CPML
    absorbs.
The finite-Q and viscous responses beyond these relations are "parity
unpinned".
"""
import math

import numpy as np
import pytest

from oracle import oracle_pve as B


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rock(shape, K_m=8.0, K_s=36.0, K_f=2.25, phi=0.25, mu=6.0, rho_s=2650., rho_f=1000., T=2.0, eta=0.0,
         kappa=1000.0, Q=1e5, f0=30.0):
    one = np.ones(shape)
    return tuple(v * one for v in (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa, Q, f0))


def test_material_against_biot_theory():
    rng = np.random.default_rng(5)
    shape = (7, 9)
    K_s = rng.uniform(30, 40, shape); K_m = rng.uniform(2, 15, shape); K_f = rng.uniform(1.5, 3.0, shape)
    phi = rng.uniform(0.1, 0.35, shape); mu = rng.uniform(3, 10, shape)
    rho_s = rng.uniform(2500, 2800, shape); rho_f = rng.uniform(900, 1100, shape); T = rng.uniform(1.5, 3, shape)
    eta = rng.uniform(0.5, 2, shape); kappa = rng.uniform(100, 2000, shape)
    Q = rng.uniform(10, 100, shape); f0 = rng.uniform(10, 80, shape)
    dt = 1e-4
    m = B.pve_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa, Q, f0, dt)
    Km, Ks, Kf, MU = K_m * 1e3, K_s * 1e3, K_f * 1e3, mu * 1e3
    alpha = 1 - Km / Ks
    M = 1. / ((alpha - phi) / Ks + phi / Kf)
    assert np.allclose(m["alpha"], alpha, rtol=1e-14)
    assert np.allclose(m["M"], M, rtol=1e-12)
    assert np.allclose(m["E"], Km + 4. / 3. * MU, rtol=1e-14)
    assert np.allclose(m["Emu"], Km - 2. / 3. * MU, rtol=1e-12)
    # Gassmann
    Kg = Km + (1 - Km / Ks) ** 2 / (phi / Kf + (1 - phi) / Ks - Km / Ks ** 2)
    assert np.allclose(Km + m["alpha"] * m["alpha"] * m["M"], Kg, rtol=1e-12)
    # standard linear solid: tau_s stored inverted, varphi stored as M varphi/(1+varphi)
    tau_s = 1. / m["ts"]
    vphi = m["varphi"] / (m["M"] - m["varphi"])
    tau_e = tau_s * (1 + vphi)
    w = 2 * math.pi * f0
    assert np.allclose(w * w * tau_e * tau_s, 1.0, rtol=1e-12)
    assert np.allclose(w * (tau_e - tau_s), 2.0 / Q, rtol=1e-9)
    # the analytic step: momentum conserved, q decays at the coupled rate
    rho = (1 - phi) * rho_s * 1e-6 + phi * rho_f * 1e-6
    rf = rho_f * 1e-6
    mm = T * rf / phi
    nk = eta * B.ETA_SCALE / kappa
    ri, rfi, mi, nki = (0.5 * (a[:-1] + a[1:]) for a in (rho, rf, mm, nk))
    lam = -nki * ri / (ri * mi - rfi * rfi)
    assert np.allclose(m["cq_i"][:-1], np.exp(lam * dt), rtol=1e-13)
    # v1 = v0 + cv q0, q1 = cq q0: rho v1 + rho_f q1 = rho v0 + rho_f q0
    assert np.allclose(ri * m["cv_i"][:-1] + rfi * m["cq_i"][:-1], rfi, rtol=1e-12)
    assert np.allclose(1. / m["irho_i"][:-1], ri, rtol=1e-14)
    assert np.allclose(m["rf_i"][:-1], rfi / ri, rtol=1e-14)
    assert np.allclose(m["m_i"][:-1], rfi * rfi / ri - mi, rtol=1e-12)
    h = 4.0 / (1 / MU[:-1, :-1] + 1 / MU[1:, :-1] + 1 / MU[:-1, 1:] + 1 / MU[1:, 1:])
    assert np.allclose(m["mu_ij"][:-1, :-1], h, rtol=1e-14)


def _ricker(f0, dt_half, n):
    tcut = 1.0 / f0
    t = np.arange(n) * dt_half - tcut
    a2 = (math.pi * f0) ** 2
    return 2 * a2 * (1 - 2 * a2 * t * t) * np.exp(-a2 * t * t)


def _ricker_at(t, f0):
    tau = t - 1.0 / f0
    a2 = (math.pi * f0) ** 2
    return 2 * a2 * (1 - 2 * a2 * tau * tau) * np.exp(-a2 * tau * tau) * (np.abs(tau) <= 1.0 / f0)


def _run(model, h, dt, f0, nsteps, cells, recs, np_=0, typ=B.BULK, A=1.0):
    mat = B.pve_material(*model, dt)
    nx2, nz2 = mat["E"].shape
    Lf = int(math.ceil((1.0 / f0) * 2.0 / (0.5 * dt)))
    fct = A * _ricker(f0, 0.5 * dt, Lf) / (h * h)
    src = [(typ, i * nz2 + j, fct) for i, j in cells]
    cfg = dict(np=np_, dx=h, dz=h, dt=dt, vmax=4000.0, kappa_max=1.0, alpha_max=math.pi * f0,
               iwipe=1 if np_ else 0)
    return B.pve_run(cfg, mat, src, [i * nz2 + j for i, j, _ in recs], [c for _, _, c in recs], nsteps)


def test_zero_source_and_linearity():
    model = rock((40, 36), eta=1.0, kappa=500.0, Q=40.0)
    recs = [(5, 7, B.VX), (30, 20, B.VZ), (12, 25, B.TXX), (22, 9, B.P), (17, 17, B.QX), (8, 30, B.DIV)]
    tr1, _ = _run(model, 5., 2e-4, 60., 60, [(20, 18)], recs)
    tr3, _ = _run(model, 5., 2e-4, 60., 60, [(20, 18)], recs, A=3.0)
    tr0, W0 = _run(model, 5., 2e-4, 60., 60, [(20, 18)], recs, A=0.0)
    assert np.abs(tr1).max() > 0 and not np.any(W0)
    assert np.allclose(tr3, 3.0 * tr1, rtol=1e-12, atol=1e-14 * np.abs(tr1).max())


def test_line_source_fast_and_slow_waves():
    """Lossless Biot medium, Bulk sources on every cell of one column (a 1D
    problem along x; the column is long enough for the velocities the
    reference never updates on the outer rows to stay out of the window):
    v_x and q_x against the two-wave closed form (module docstring).  The
    slow wave (about 900 m/s) is resolved by 3 or more nodes per wavelength
    up to 2.5 f0.  Measured 6.3e-4 to 9.7e-4, held to 0.5%; the slow wave
    alone is 44-100% of the expectation."""
    h, f0, dt = 2.0, 50.0, 1e-4
    nx, nz = 128, 200
    model = rock((nx, nz))
    K_m, K_s, K_f, phi, mu, rho_s, rho_f, T = (a[0, 0] for a in model[:8])
    Km, Ks, Kf, MU = K_m * 1e3, K_s * 1e3, K_f * 1e3, mu * 1e3
    alpha = 1 - Km / Ks
    M = 1. / ((alpha - phi) / Ks + phi / Kf)
    E = Km + 4. / 3. * MU
    rs, rf = rho_s * 1e-6, rho_f * 1e-6
    rho = (1 - phi) * rs + phi * rf
    mm = T * rf / phi
    C = np.array([[E + alpha * alpha * M, alpha * M], [alpha * M, M]])
    R = np.array([[rho, rf], [rf, mm]])
    lam, vec = np.linalg.eig(np.linalg.solve(R, C))
    # the source adds s to tau_xx and to p: S = (tau_xx, -p) gains (1, -1)
    g = np.linalg.solve(vec, np.linalg.solve(R, np.array([1.0, -1.0])))
    c = np.sqrt(lam.real)
    c0 = nx // 2
    offs = [-15, 12]
    recs = [(c0 + o, nz // 2, comp) for o in offs for comp in (B.VX, B.QX)]
    nsteps = 500
    tr, _ = _run(model, h, dt, f0, nsteps, [(c0, j) for j in range(nz)], recs)
    t = (np.arange(nsteps) + 1) * dt
    for k, o in enumerate(offs):
        x = (o + 0.5) * h
        want = np.zeros((2, nsteps))
        slow = np.zeros((2, nsteps))
        for q in range(2):
            term = -np.sign(x) * g[q] * _ricker_at(t - abs(x) / c[q], f0) / h / (2 * c[q] ** 2)
            want += np.outer(vec[:, q].real, term)
            if c[q] < c.max():
                slow += np.outer(vec[:, q].real, term)
        for comp in range(2):
            got = tr[:, 2 * k + comp]
            assert rel_l2(got, want[comp]) < 5e-3, (o, comp, rel_l2(got, want[comp]))
            assert rel_l2(want[comp] - slow[comp], want[comp]) > 0.4


def test_cpml_absorbs():
    """Viscous fluid (eta 1 cP, kappa 1000 mD: the slow wave diffuses away):
    the energy left in the interior after the fast wave has crossed it is a
    small fraction of what remains without layers (measured 1e-5)."""
    n, h, f0, dt = 60, 5.0, 30.0, 2e-4
    e = []
    for np_ in (0, 15):
        model = rock((n + 2 * np_, n + 2 * np_), eta=1.0)
        c = n // 2 + np_
        _, W = _run(model, h, dt, f0, 700, [(c, c)], [(c, c + 1, B.VZ)], np_=np_)
        inner = W[4:6, np_:np_ + n, np_:np_ + n]
        e.append(float(np.sum(inner * inner)))
    assert e[1] < 1e-3 * e[0], e


# ---------------------------------------------------------------- VTI ------
def vti_rock(shape, eps=0.15, delta=0.08, T1=1.8, T3=2.6, kappa1=1500.0, kappa3=400.0, eta=0.0, Q=1e5, f0=30.0,
             **iso):
    K_m, K_s, K_f, phi, mu, rho_s, rho_f = rock(shape, **iso)[:7]
    one = np.ones(shape)
    return (K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1 * one, T3 * one, eta * one, kappa1 * one, kappa3 * one,
            Q * one, f0 * one, eps * one, delta * one)


def test_vti_material_against_anisotropic_biot_theory():
    """pve_vti_pml.c:787-1145 against relations written independently:
    Thomsen's c11 = (1 + 2 eps) c33 and (c13 + c55)^2 = (c33 - c55)^2 +
    2 delta c33 (c33 - c55); the effective-stress coefficients alpha_i =
    1 - sum_j C_ij / (3 K_s) and M = K_s / (1 - K*/K_s - phi (1 - K_s/K_f)),
    K* = sum_ij C_ij / 9 over the upper-left 3 x 3 block of the VTI frame
    (Carcione 1996, eq. 9-11); directional mass m_i = T_i rho_f / phi;
    and the isotropic limit (eps = delta = 0, T1 = T3, kappa1 = kappa3)
    equal to the isotropic material."""
    rng = np.random.default_rng(9)
    shape = (6, 7)
    K_s = rng.uniform(30, 40, shape); K_m = rng.uniform(2, 15, shape); K_f = rng.uniform(1.5, 3.0, shape)
    phi = rng.uniform(0.1, 0.35, shape); mu = rng.uniform(3, 10, shape)
    rho_s = rng.uniform(2500, 2800, shape); rho_f = rng.uniform(900, 1100, shape)
    T1, T3 = rng.uniform(1.5, 3, shape), rng.uniform(1.5, 3, shape)
    eta = rng.uniform(0.5, 2, shape); k1, k3 = rng.uniform(100, 2000, shape), rng.uniform(100, 2000, shape)
    Q = rng.uniform(10, 100, shape); f0 = rng.uniform(10, 80, shape)
    eps, dl = rng.uniform(0, 0.3, shape), rng.uniform(-0.05, 0.2, shape)
    dt = 1e-4
    m = B.pve_vti_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3, eta, k1, k3, Q, f0, eps, dl, dt)
    Km, Ks, Kf, MU = K_m * 1e3, K_s * 1e3, K_f * 1e3, mu * 1e3
    c33 = Km + 4. / 3. * MU
    assert np.allclose(m["c33"], c33, rtol=1e-14)
    assert np.allclose(m["c11"], (1 + 2 * eps) * c33, rtol=1e-14)
    assert np.allclose((m["c13"] + MU) ** 2, (c33 - MU) ** 2 + 2 * dl * c33 * (c33 - MU), rtol=1e-12)
    c12 = m["c11"] - 2 * MU
    C = np.array([[m["c11"], c12, m["c13"]], [c12, m["c11"], m["c13"]], [m["c13"], m["c13"], c33]])
    assert np.allclose(m["alpha1"], 1 - C[0].sum(0) / (3 * Ks), rtol=1e-13)
    assert np.allclose(m["alpha3"], 1 - C[2].sum(0) / (3 * Ks), rtol=1e-13)
    Kstar = C.sum((0, 1)) / 9
    assert np.allclose(m["M"], Ks / (1 - Kstar / Ks - phi * (1 - Ks / Kf)), rtol=1e-12)
    rf = rho_f * 1e-6
    rho = (1 - phi) * rho_s * 1e-6 + phi * rf
    ri, rfi = 0.5 * (rho[:-1] + rho[1:]), 0.5 * (rf[:-1] + rf[1:])
    m1i = 0.5 * (T1 * rf / phi)[:-1] + 0.5 * (T1 * rf / phi)[1:]
    assert np.allclose(m["m_i"][:-1], rfi * rfi / ri - m1i, rtol=1e-12)
    rj, rfj = 0.5 * (rho[:, :-1] + rho[:, 1:]), 0.5 * (rf[:, :-1] + rf[:, 1:])
    m3j = 0.5 * (T3 * rf / phi)[:, :-1] + 0.5 * (T3 * rf / phi)[:, 1:]
    assert np.allclose(m["m_j"][:, :-1], rfj * rfj / rj - m3j, rtol=1e-12)
    nk3 = eta * B.ETA_SCALE / k3
    lam = -0.5 * (nk3[:, :-1] + nk3[:, 1:]) * rj / (rj * m3j - rfj * rfj)
    assert np.allclose(m["cq_j"][:, :-1], np.exp(lam * dt), rtol=1e-13)
    # isotropic limit
    zero = np.zeros(shape)
    mv = B.pve_vti_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T1, eta, k1, k1, Q, f0, zero, zero, dt)
    mi = B.pve_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, eta, k1, Q, f0, dt)
    assert np.allclose(mv["c11"], mi["E"], rtol=1e-14) and np.allclose(mv["c33"], mi["E"], rtol=1e-14)
    assert np.allclose(mv["c13"], mi["Emu"], rtol=1e-12)
    for k in ("M", "ts", "varphi", "mu_ij", "cq_i", "cv_i", "cq_j", "cv_j", "m_i", "m_j", "rf_i", "irho_j"):
        assert np.allclose(mv[k], mi[k], rtol=1e-11), k
    assert np.allclose(mv["alpha1"], mi["alpha"], rtol=1e-12) and np.allclose(mv["alpha3"], mi["alpha"], rtol=1e-12)


def _run_vti(model, h, dt, f0, nsteps, cells, recs):
    mat = B.pve_vti_material(*model, dt)
    nx2, nz2 = mat["M"].shape
    Lf = int(math.ceil((1.0 / f0) * 2.0 / (0.5 * dt)))
    fct = _ricker(f0, 0.5 * dt, Lf) / (h * h)
    src = [(B.BULK, i * nz2 + j, fct) for i, j in cells]
    cfg = dict(np=0, dx=h, dz=h, dt=dt, vmax=4000.0, kappa_max=1.0, alpha_max=math.pi * f0, iwipe=0)
    return B.pve_run(cfg, mat, src, [i * nz2 + j for i, j, _ in recs], [c for _, _, c in recs], nsteps)


@pytest.mark.parametrize("axis", ["x", "z"])
def test_vti_line_source_fast_and_slow_waves(axis):
    """Lossless VTI Biot medium (eps 0.15, delta 0.08, T1 1.8, T3 2.6), a
    line of Bulk sources: waves along x see C = [[c11 + a1^2 M, a1 M],
    [a1 M, M]] and R = [[rho, rho_f], [rho_f, T1 rho_f / phi]], waves along
    z the c33, alpha3, T3 counterparts; v and q against the two-wave closed
    form, and the cross stress (tau_zz for waves along x, tau_xx along z)
    against sum_k g_k [(c13 + a1 a3 M) phi_v + a_other M phi_q]_k
    s(t - |x|/c_k) / (2 c_k^3 h), which pins c13.  Held to 0.5%."""
    h, f0, dt = 2.0, 50.0, 1e-4
    n_along, n_across = 128, 200
    shape = (n_along, n_across) if axis == "x" else (n_across, n_along)
    model = vti_rock(shape)
    mat = B.pve_vti_material(*model, dt)
    K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3 = (a[0, 0] for a in model[:9])
    eps, dl = model[14][0, 0], model[15][0, 0]
    Km, Ks, Kf, MU = K_m * 1e3, K_s * 1e3, K_f * 1e3, mu * 1e3
    c33 = Km + 4. / 3. * MU
    c11 = (1 + 2 * eps) * c33
    c13 = math.sqrt((c33 - MU) ** 2 + 2 * dl * c33 * (c33 - MU)) - MU
    c12 = c11 - 2 * MU
    a1, a3 = 1 - (c11 + c12 + c13) / (3 * Ks), 1 - (2 * c13 + c33) / (3 * Ks)
    M = Ks / (1 - (2 * c11 + c33 + 2 * c12 + 4 * c13) / 9 / Ks - phi * (1 - Ks / Kf))
    rs, rf = rho_s * 1e-6, rho_f * 1e-6
    rho = (1 - phi) * rs + phi * rf
    cl, al, ao, Tl = (c11, a1, a3, T1) if axis == "x" else (c33, a3, a1, T3)
    C = np.array([[cl + al * al * M, al * M], [al * M, M]])
    R = np.array([[rho, rf], [rf, Tl * rf / phi]])
    lam, vec = np.linalg.eig(np.linalg.solve(R, C))
    g = np.linalg.solve(vec, np.linalg.solve(R, np.array([1.0, -1.0])))
    c = np.sqrt(lam.real)
    c0, mid = n_along // 2, n_across // 2
    offs = [-15, 12]
    if axis == "x":
        comps = (B.VX, B.QX, B.TZZ)
        recs = [(c0 + o, mid, comp) for o in offs for comp in comps]
        cells = [(c0, j) for j in range(n_across)]
    else:
        comps = (B.VZ, B.QZ, B.TXX)
        recs = [(mid, c0 + o, comp) for o in offs for comp in comps]
        cells = [(i, c0) for i in range(n_across)]
    nsteps = 500
    tr, _ = _run_vti(model, h, dt, f0, nsteps, cells, recs)
    t = (np.arange(nsteps) + 1) * dt
    for k, o in enumerate(offs):
        x = (o + 0.5) * h
        want = np.zeros((3, nsteps))
        for q in range(2):
            s = _ricker_at(t - abs(x) / c[q], f0)
            want[:2] += np.outer(vec[:, q].real, -np.sign(x) * g[q] * s / h / (2 * c[q] ** 2))
            # the cross stress sits on the node, half a cell nearer the source
            xs = o * h
            s2 = _ricker_at(t - abs(xs) / c[q], f0)
            want[2] += g[q] * ((c13 + a1 * a3 * M) * vec[0, q].real + ao * M * vec[1, q].real) * s2 / (
                2 * c[q] ** 3 * h)
        for comp in range(3):
            got = tr[:, 3 * k + comp]
            assert rel_l2(got, want[comp]) < 5e-3, (axis, o, comp, rel_l2(got, want[comp]))


def test_pml_median_smoothing():
    """pml_median_average (pml.c:2254-2329) as the oracle restates it: a
    constant field stays constant; nodes outside the rings (the model, and
    the ring columns below nz2 - np - n - 1) keep their values; a ring node
    takes the lower median of its 3 x 3 window of the previous ring's
    result (for the four values of a grid corner the second smallest);
    an isolated spike in the layers is removed."""
    np_ = 3
    nx2, nz2 = 14, 12
    assert np.array_equal(B.pml_median(np.full((nx2, nz2), 2.5), np_), np.full((nx2, nz2), 2.5))
    rng = np.random.default_rng(2)
    a = rng.uniform(0, 1, (nx2, nz2))
    b = B.pml_median(a, np_)
    touched = np.zeros((nx2, nz2), bool)
    for n in range(1, np_ + 1):
        d = np_ - n
        touched[d:nx2 - d, d] = touched[d:nx2 - d, nz2 - d - 1] = True
        touched[d, d + 1:nz2 - np_ - n - 1] = touched[nx2 - d - 1, d + 1:nz2 - np_ - n - 1] = True
    assert np.array_equal(b[~touched], a[~touched])
    spike = np.ones((nx2, nz2))
    spike[1, 5] = 100.0
    assert B.pml_median(spike, np_)[1, 5] == 1.0
    # the lower median of an even window
    z = np.zeros((nx2, nz2))
    z[0, 0], z[0, 1], z[1, 0], z[1, 1] = 4.0, 3.0, 2.0, 1.0
    z[1:, 1:] = np.where(z[1:, 1:] == 0, 9.0, z[1:, 1:])
    assert B.pml_median(z, 1)[0, 0] == 2.0
