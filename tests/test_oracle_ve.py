"""CPU tests that pin the FP64 oracle of ve_vti_pml (oracle/oracle_ve.py).

The reference cannot be built here (structs.h includes fftw3.h) and ships
no tests.  The restatement shares its derivative operators and CPML tables
with the SH oracle (pinned in test_oracle_sh.py) and is pinned by
  - the material (ve_vti_pml.c:518-683): Thomsen's definitions of epsilon
    and delta (c11 = (1 + 2 eps) c33, (c13 + c55)^2 = (c33 - c55)^2 +
    2 delta c33 (c33 - c55)), one standard linear solid per mechanism with
    its peak at f and quality Q, the closed form K0 = (D - c55) sum tau_s/tau_e,
    the harmonic mean of mu at the cell corners;
  - invariants: zero source, linearity;
  - physics, elastic limit (Q = 1e7):
      * a line source (Bulk on every cell of one column or row, no layers)
        is a 1D problem with the closed-form solution
        v = -sgn(x) s(t - |x|/c) / (2 rho c^2), tau_xx = c11 s(.)/(2 rho c^3),
        tau_zz = c13 s(.)/(2 rho c^3): pins c11, c33 (hence epsilon) and c13
        (hence delta) separately, in a VTI medium;
      * a point Bulk source in an isotropic medium against the 2D Green's
        function of the P potential (v = grad phi);
      * the CPML absorbs.
The finite-Q response is pinned only through the material relations
("parity unpinned" for attenuation itself).
"""
import math

import numpy as np
import pytest

from oracle import oracle_ve as V


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _model(shape, vp, vs, rho, eps, dlt, Q1, f1, Q2, f2):
    """read_model_ve units: K, mu in GPa from Vp, Vs (m/s), rho (kg/m^3)."""
    L = len(Q1)
    mu = np.full(shape, vs * vs * rho) * 1e-9
    K = np.full(shape, 1e-9 * (vp * vp * rho - 4. / 3. * vs * vs * rho))
    one = np.ones(shape)
    return V.ve_material(K, mu, rho * one, eps * one, dlt * one,
                         np.stack([q * one for q in Q1]), np.stack([f * one for f in f1]), Q2 * one, f2 * one), L


def test_material_relations():
    rng = np.random.default_rng(3)
    shape = (7, 9)
    L = 3
    vp = rng.uniform(2000, 4500, shape); vs = vp / rng.uniform(1.6, 2.2, shape)
    rho = rng.uniform(1800, 2700, shape)
    eps = rng.uniform(0, 0.3, shape); dlt = rng.uniform(-0.1, 0.2, shape)
    Q1 = rng.uniform(10, 120, (L,) + shape); f1 = rng.uniform(5, 80, (L,) + shape)
    Q2 = rng.uniform(10, 120, shape); f2 = rng.uniform(5, 80, shape)
    mu = vs * vs * rho * 1e-9
    K = 1e-9 * (vp * vp * rho - 4. / 3. * vs * vs * rho)
    m = V.ve_material(K, mu, rho, eps, dlt, Q1, f1, Q2, f2)
    c33 = (K + 4. / 3. * mu) * 1000.
    c55 = mu * 1000.
    assert np.allclose(m["c33"], c33, rtol=1e-14)
    assert np.allclose(m["c11"], (1 + 2 * eps) * c33, rtol=1e-14)
    assert np.allclose((m["c13"] + c55) ** 2, (c33 - c55) ** 2 + 2 * dlt * c33 * (c33 - c55), rtol=1e-12)
    # one standard linear solid per mechanism (stored 1/tau_s and (1 - tau_e/tau_s)/L)
    tau_s1 = 1. / m["ts1"]
    tau_e1 = tau_s1 * (1. - L * m["eta1"])
    w1 = 2 * math.pi * f1
    assert np.allclose(w1 * w1 * tau_e1 * tau_s1, 1.0, rtol=1e-12)
    assert np.allclose(w1 * (tau_e1 - tau_s1), 2.0 / Q1, rtol=1e-9)
    tau_s2 = 1. / m["ts2"]
    tau_e2 = tau_s2 * (1. - m["eta2"])
    w2 = 2 * math.pi * f2
    assert np.allclose(w2 * w2 * tau_e2 * tau_s2, 1.0, rtol=1e-12)
    assert np.allclose(w2 * (tau_e2 - tau_s2), 2.0 / Q2, rtol=1e-9)
    # K0 = (D - c55) sum tau_s/tau_e, c55_0 = c55 tau_s2/tau_e2 at the nodes
    D = 0.5 * (m["c11"] + m["c33"])
    assert np.allclose(m["K0"], (D - c55) * np.sum(tau_s1 / tau_e1, axis=0), rtol=1e-10)
    assert np.allclose(m["c55_0"], c55 * tau_s2 / tau_e2, rtol=1e-13)
    # corners: harmonic mean of mu, averages of tau_s2 and tau_s2/tau_e2
    h = 4.0 / (1 / c55[:-1, :-1] + 1 / c55[1:, :-1] + 1 / c55[:-1, 1:] + 1 / c55[1:, 1:])
    assert np.allclose(m["c55"][:-1, :-1], h, rtol=1e-14)
    assert np.array_equal(m["c55"][-1, :-1], m["c55"][-2, :-1])
    assert np.array_equal(m["c55"][:, -1], m["c55"][:, -2])
    avg = lambda a: 0.25 * (a[:-1, :-1] + a[1:, :-1] + a[:-1, 1:] + a[1:, 1:])
    assert np.allclose(1. / m["ts2_ij"][:-1, :-1], avg(tau_s2), rtol=1e-13)
    assert np.allclose(1. / (1. - m["eta2_ij"][:-1, :-1]), avg(tau_s2 / tau_e2), rtol=1e-12)
    assert np.allclose(m["irho_i"][:-1], 1. / (0.5 * (rho[:-1] + rho[1:]) * 1e-6), rtol=1e-14)
    assert np.allclose(m["irho_j"][:, :-1], 1. / (0.5 * (rho[:, :-1] + rho[:, 1:]) * 1e-6), rtol=1e-14)
    assert not np.any(m["c55_0_ij"])


def _ricker(f0, dt_half, n):
    tcut = 1.0 / f0
    t = np.arange(n) * dt_half - tcut
    a2 = (math.pi * f0) ** 2
    return 2 * a2 * (1 - 2 * a2 * t * t) * np.exp(-a2 * t * t)


def _ricker_at(t, f0):
    tau = t - 1.0 / f0
    a2 = (math.pi * f0) ** 2
    return 2 * a2 * (1 - 2 * a2 * tau * tau) * np.exp(-a2 * tau * tau) * (np.abs(tau) <= 1.0 / f0)


def _run(mat, h, dt, f0, nsteps, cells, recs, np_=0, typ=V.BULK, A=1.0, kappa_max=1.0):
    nx2, nz2 = mat["c11"].shape
    Lf = int(math.ceil((1.0 / f0) * 2.0 / (0.5 * dt)))
    fct = A * _ricker(f0, 0.5 * dt, Lf) / (h * h)
    src = [(typ, i * nz2 + j, fct) for i, j in cells]
    ridx = [i * nz2 + j for i, j, _ in recs]
    rcomp = [c for _, _, c in recs]
    vmax = float(np.sqrt(np.max(mat["c11"]) / 1000.0 / 1e-9 * 1e-3))
    cfg = dict(np=np_, dx=h, dz=h, dt=dt, vmax=3000.0 if not np_ else vmax, kappa_max=kappa_max,
               alpha_max=math.pi * f0, iwipe=1 if np_ else 0)
    return V.ve_run(cfg, mat, src, ridx, rcomp, nsteps)


def test_zero_source_and_linearity():
    shape = (40, 36)
    mat, _ = _model(shape, 3000., 1700., 2200., 0.1, 0.05, [30., 50.], [20., 40.], 25., 30.)
    recs = [(5, 7, V.VX), (30, 20, V.VZ), (12, 25, V.TXX), (22, 9, V.TZZ), (17, 17, V.TXZ)]
    tr1, _ = _run(mat, 10., 1e-3, 20., 60, [(20, 18)], recs, typ=V.SX)
    tr3, _ = _run(mat, 10., 1e-3, 20., 60, [(20, 18)], recs, typ=V.SX, A=3.0)
    tr0, W0 = _run(mat, 10., 1e-3, 20., 60, [(20, 18)], recs, typ=V.SX, A=0.0)
    assert np.abs(tr1).max() > 0 and not np.any(W0)
    assert np.allclose(tr3, 3.0 * tr1, rtol=1e-12, atol=1e-14 * np.abs(tr1).max())


@pytest.mark.parametrize("axis", [0, 1])
def test_line_source_1d_closed_form(axis):
    """Bulk source on every cell of one column (axis 0: waves along x) or
    row (axis 1: along z) of a VTI medium without layers: the exact 1D
    solution.  v sits half a cell past its node (v_x at i + 1/2, v_z at
    j + 1/2), the stresses on the nodes.  The reference never updates the
    velocities of the grid's outer rows and columns, so the line is long
    enough (128 cells, receivers in its middle) for that edge to stay out of
    the compared window, which also ends before the periodic images
    arrive.  Measured 3.8e-4 to 6.3e-4, held to 0.5%; a c11/c33 swap in
    the material fails it, and so does c13 off by 10%."""
    h, f0, dt = 10.0, 25.0, 5e-4
    vp, vs, rho, eps, dlt = 3000., 1500., 2200., 0.2, 0.08
    n_long, n_short = 256, 128
    shape = (n_long, n_short) if axis == 0 else (n_short, n_long)
    mat, _ = _model(shape, vp, vs, rho, eps, dlt, [1e7], [20.], 1e7, 20.)
    c0, m = n_long // 2, n_short // 2
    cells = [(c0, j) for j in range(n_short)] if axis == 0 else [(i, c0) for i in range(n_short)]
    offs = [-40, 25, 45]
    if axis == 0:
        recs = [(c0 + o, m, V.VX) for o in offs] + [(c0 + o, m, c) for o in offs for c in (V.TXX, V.TZZ)]
    else:
        recs = [(m, c0 + o, V.VZ) for o in offs] + [(m, c0 + o, c) for o in offs for c in (V.TZZ, V.TXX)]
    nsteps = 360
    tr, _ = _run(mat, h, dt, f0, nsteps, cells, recs)
    # moduli from the model parameters (MPa, rho in Gg/m^3), not from the oracle
    rho_s = rho * 1e-6
    c33, c55 = vp * vp * rho_s, vs * vs * rho_s
    c_along = (1 + 2 * eps) * c33 if axis == 0 else c33
    c_cross = math.sqrt(2 * dlt * c33 * (c33 - c55) + (c33 - c55) ** 2) - c55
    vel = math.sqrt(c_along / rho_s)
    t = (np.arange(nsteps) + 1) * dt
    s = lambda tt: _ricker_at(tt, f0) / h          # line source per unit length
    for k, o in enumerate(offs):
        x = (o + 0.5) * h
        want = -np.sign(x) * s(t - abs(x) / vel) / (2 * rho_s * vel * vel)
        assert rel_l2(tr[:, k], want) < 5e-3, (axis, o, rel_l2(tr[:, k], want))
        x = o * h
        for q, c in ((0, c_along), (1, c_cross)):
            want = c * s(t - abs(x) / vel) / (2 * rho_s * vel ** 3)
            got = tr[:, len(offs) + 2 * k + q]
            assert rel_l2(got, want) < 5e-3, (axis, o, q, rel_l2(got, want))


def _green_grad(x, z, t, f0, c, rho_s):
    """grad phi at (x, z), phi = 1/(2 pi rho c^2) int_0^acosh(ct/r) s(t - (r/c) cosh u) du,
    the P potential of a Bulk point source s(t) = Ricker(t - 1/f0)."""
    def phi(xx, zz):
        r = math.hypot(xx, zz)
        out = np.zeros_like(t)
        for k, tt in enumerate(t):
            if c * tt > r:
                u = np.linspace(0.0, math.acosh(c * tt / r), 4001)
                out[k] = np.trapezoid(_ricker_at(tt - (r / c) * np.cosh(u), f0), u)
        return out / (2 * math.pi * rho_s * c * c)
    e = 1e-3 * math.hypot(x, z)
    return (phi(x + e, z) - phi(x - e, z)) / (2 * e), (phi(x, z + e) - phi(x, z - e)) / (2 * e)


def test_bulk_point_source_against_green_function():
    """Isotropic, Q = 1e7, no layers, 128^2 cells of 10 m, 10 Hz: receivers
    20-35 cells away, until the periodic images could arrive.  Measured
    0.04-0.18%; held to 1%, where v taken at the nodes instead of half a
    cell further gives 12%."""
    n, h, f0, dt = 128, 10.0, 10.0, 5e-4
    vp, vs, rho = 3000., 1700., 2200.
    mat, _ = _model((n, n), vp, vs, rho, 0.0, 0.0, [1e7], [20.], 1e7, 20.)
    c = n // 2
    pts = [(0, 25), (22, 22), (-30, 10), (28, -12)]
    recs = [(c + a, c + b, comp) for a, b in pts for comp in (V.VX, V.VZ)]
    nsteps = 440
    tr, _ = _run(mat, h, dt, f0, nsteps, [(c, c)], recs)
    t = (np.arange(nsteps) + 1) * dt
    for k, (a, b) in enumerate(pts):
        gx, _ = _green_grad((a + 0.5) * h, b * h, t, f0, vp, rho * 1e-6)
        _, gz = _green_grad(a * h, (b + 0.5) * h, t, f0, vp, rho * 1e-6)
        for q, want in ((0, gx), (1, gz)):
            if np.abs(want).max() < 0.05 * max(np.abs(gx).max(), np.abs(gz).max()):
                continue
            got = tr[:, 2 * k + q]
            assert rel_l2(got, want) < 0.01, (a, b, q, rel_l2(got, want))


def test_cpml_absorbs():
    n, h, f0, dt = 60, 10.0, 15.0, 5e-4
    e = []
    for np_ in (0, 15):
        mat, _ = _model((n + 2 * np_, n + 2 * np_), 3000., 1700., 2200., 0.1, 0.05, [1e6], [20.], 1e6, 20.)
        c = n // 2 + np_
        _, W = _run(mat, h, dt, f0, 700, [(c, c)], [(c, c + 1, V.VZ)], np_=np_)
        inner = W[3:5, np_:np_ + n, np_:np_ + n]
        e.append(float(np.sum(inner * inner)))
    assert e[1] < 0.02 * e[0], e
