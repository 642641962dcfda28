"""FP64 restatement of pve_iso_pml (poroviscoelastic waves in 2D isotropic
media after Carcione and Helle (1999): Biot solid and fluid motion, one
relaxation mechanism for the fluid pressure, pseudo-spectral derivatives,
CPML, fourth-order Runge-Kutta with the stiff viscous coupling solved
analytically at the start of every step).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline
leg as the checker.  The product path never calls it.

Follows the reference step by step:
  - model units and material: pve_iso_pml.c:795-830 (E, M, alpha, tau_s,
    varphi, rho and m at the nodes), 1009-1110 (staggered averages, the
    analytic-solution coefficients and the fill of the last row and column),
    1123-1145 (stored as 1/tau_s, M varphi/(1 + varphi), E - 2 mu,
    rho_f^2/rho - m, rho_f/rho and 1/rho);
  - CPML coefficients (pml.c:1959-2106): the tables of oracle_sh.sh_cpml;
  - derivative operators and layer loops: oracle_sh / oracle_ve;
  - propagateCPML: propagate.c:884-1329, as written (the viscous term of the
    fluid equation is not in it: it lives in the analytic step);
  - the step (pve_iso_pml.c:1249-1328): vs = v + coeff_v q and qs = coeff_q q
    (eq. 76-77 of Carcione 1996), then the Runge-Kutta stages;
  - source injection: src.c:77-112 (add_src_rk4);
  - divergence and curl records (propagate.c:2672-2704): spectral
    derivatives without layer corrections of the fields the FFT plans are
    bound to, the last stage's input.

The FFTs are numpy's (pocketfft), not FFTW: results agree with an FFTW
build to rounding, so parity with the GPU path is stated as a relative
tolerance, not bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .oracle_sh import deriv, k_multipliers
from .oracle_sh import sh_cpml as cpml_tables
from .oracle_ve import _x_layer, _z_layer

SX, SZ, SXZ, BULK, SF, BULK_S = 0, 2, 4, 6, 7, 8
VX, VZ, QX, QZ, TXX, TZZ, TXZ, P, DIV, CURL = 0, 2, 3, 4, 5, 6, 7, 8, 11, 12
# state: tau_xx, tau_zz, tau_xz, p, v_x, v_z, q_x, q_z, e
I_TXX, I_TZZ, I_TXZ, I_P, I_VX, I_VZ, I_QX, I_QZ, I_E = range(9)
ETA_SCALE = 1.013249965828145e6


def pml_median(a, np_):
    """pml_median_average (pml.c:2254-2329): ring by ring from the layer
    next to the model outwards, each ring's nodes replaced by the lower
    median ((n - 1)/2-th smallest) of their 3 x 3 neighbourhood inside the
    grid, computed from the previous ring's result; the window stays 3 x 3
    (the size update of pml.c:2324 assigns nmed its own old value).  The
    columns of a ring stop at nz2 - np - n - 1 (pml.c:2297), short of its
    bottom row."""
    a = np.array(a, np.float64)
    nx2, nz2 = a.shape

    def med(src, i, j):
        w = src[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2].reshape(-1)
        return np.sort(w)[(w.size - 1) // 2]
    for n in range(1, np_ + 1):
        d = np_ - n
        out = a.copy()
        for i in range(d, nx2 - d):
            out[i, d] = med(a, i, d)
            out[i, nz2 - d - 1] = med(a, i, nz2 - d - 1)
        for j in range(d + 1, nz2 - np_ - n - 1):
            out[d, j] = med(a, d, j)
            out[nx2 - d - 1, j] = med(a, nx2 - d - 1, j)
        a = out
    return a


def pve_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T, eta, kappa, Q, f0, dt, median_np=0):
    """Material of pve_iso_pml.c:795-1145 from the padded model as read_model
    leaves it (moduli in GPa, densities in kg/m^3, eta in cP, kappa in mD),
    with the median smoothing of the layers (833-845) when median_np > 0.
    Returns the stored arrays and the analytic-step coefficients."""
    pi = 4.0 * math.atan(1.0)
    a = lambda x: np.asarray(x, np.float64)
    K_m, K_s, K_f, mu = a(K_m) * 1000., a(K_s) * 1000., a(K_f) * 1000., a(mu) * 1000.
    eta = a(eta) * ETA_SCALE
    rho_s, rho_f = a(rho_s) * 1.e-6, a(rho_f) * 1.e-6
    phi, T, kappa, Q, f0 = a(phi), a(T), a(kappa), a(Q), a(f0)
    # nodes, 809-829
    tau0 = 1. / (2. * pi * f0)
    E = K_m + 4. / 3. * mu
    D = K_s * (1. + phi * (K_s / K_f - 1.))
    M = K_s * K_s / (D - K_m)
    alpha = 1. - K_m / K_s
    m = T * rho_f / phi
    lowQ = Q < 1.e4
    Qs = np.where(lowQ, Q, 1.0)
    tau_e = tau0 / Qs * (np.sqrt(Qs * Qs + 1.) + 1.)
    tau_s = np.where(lowQ, tau0 / Qs * (np.sqrt(Qs * Qs + 1.) - 1.), tau0)
    varphi = np.where(lowQ, tau_e / np.where(lowQ, tau_s, 1.0) - 1., 0.0)
    rho = (1. - phi) * rho_s + phi * rho_f
    if median_np:
        mu, E, M, alpha, varphi, tau_s, rho, rho_f, eta, kappa, m = (
            pml_median(x, median_np) for x in (mu, E, M, alpha, varphi, tau_s, rho, rho_f, eta, kappa, m))
    # staggered, 1009-1110
    z = np.zeros_like
    rho_i, rho_f_i, nk_i, m_i, cq_i, cv_i = z(E), z(E), z(E), z(E), z(E), z(E)
    rho_j, rho_f_j, nk_j, m_j, cq_j, cv_j = z(E), z(E), z(E), z(E), z(E), z(E)
    mu_ij = z(E)
    nk = eta / kappa

    def coeffs(r, rf, mm, nkk):
        tmp = 1. / (rf * rf - r * mm)
        beta12 = tmp * rf
        beta22 = tmp * -r
        lam = -nkk * beta22
        cq = np.exp(lam * dt)
        return cq, beta12 / beta22 * (cq - 1.)
    rho_i[:-1] = 0.5 * (rho[:-1] + rho[1:]); rho_f_i[:-1] = 0.5 * (rho_f[:-1] + rho_f[1:])
    nk_i[:-1] = 0.5 * (nk[:-1] + nk[1:]); m_i[:-1] = 0.5 * (m[:-1] + m[1:])
    cq_i[:-1], cv_i[:-1] = coeffs(rho_i[:-1], rho_f_i[:-1], m_i[:-1], nk_i[:-1])
    rho_j[:, :-1] = 0.5 * (rho[:, :-1] + rho[:, 1:]); rho_f_j[:, :-1] = 0.5 * (rho_f[:, :-1] + rho_f[:, 1:])
    nk_j[:, :-1] = 0.5 * (nk[:, :-1] + nk[:, 1:]); m_j[:, :-1] = 0.5 * (m[:, :-1] + m[:, 1:])
    cq_j[:, :-1], cv_j[:, :-1] = coeffs(rho_j[:, :-1], rho_f_j[:, :-1], m_j[:, :-1], nk_j[:, :-1])
    with np.errstate(divide="ignore"):      # mu = 0: a zero corner modulus, as in IEEE arithmetic
        mu_ij[:-1, :-1] = 4.0 / (1. / mu[:-1, :-1] + 1. / mu[1:, :-1] + 1. / mu[:-1, 1:] + 1. / mu[1:, 1:])
    # last row, then last column (1076-1105)
    for x in (rho_i, rho_f_i, nk_i, m_i, cq_i, cv_i, mu_ij):
        x[-1, :] = x[-2, :]
    for x in (rho_j, rho_f_j, nk_j, m_j, cq_j, cv_j, mu_ij):
        x[:, -1] = x[:, -2]
    # 1131-1145
    out = dict(E=E, Emu=E - 2. * mu, M=M, alpha=alpha, ts=1. / tau_s, varphi=M * varphi / (1. + varphi),
               mu_ij=mu_ij, cq_i=cq_i, cv_i=cv_i, cq_j=cq_j, cv_j=cv_j)
    out["m_i"] = rho_f_i * rho_f_i / rho_i - m_i
    out["m_j"] = rho_f_j * rho_f_j / rho_j - m_j
    out["rf_i"] = rho_f_i / rho_i
    out["rf_j"] = rho_f_j / rho_j
    out["irho_i"] = 1. / rho_i
    out["irho_j"] = 1. / rho_j
    return out


def pve_vti_material(K_m, K_s, K_f, phi, mu, rho_s, rho_f, T1, T3, eta, kappa1, kappa3, Q, f0, epsilon, delta,
                     dt):
    """Material of pve_vti_pml.c:787-1145 (VTI Biot medium after Carcione
    1996): c33 = K_m + 4/3 mu, c55 = mu, Thomsen's c11 and c13, the Biot
    modulus and effective-stress coefficients alpha1, alpha3 of the
    anisotropic frame, tortuosities T1, T3 and permeabilities kappa1,
    kappa3 along x and z (the x ones at (i+1/2, j), the z ones at
    (i, j+1/2)).  Inputs as read_modelVTI leaves them.  Returns the stored
    arrays under the keys the propagation reads."""
    pi = 4.0 * math.atan(1.0)
    a = lambda x: np.asarray(x, np.float64)
    K_m, K_s, K_f, mu = a(K_m) * 1000., a(K_s) * 1000., a(K_f) * 1000., a(mu) * 1000.
    eta = a(eta) * ETA_SCALE
    rho_s, rho_f = a(rho_s) * 1.e-6, a(rho_f) * 1.e-6
    phi, T1, T3, kappa1, kappa3, Q, f0 = a(phi), a(T1), a(T3), a(kappa1), a(kappa3), a(Q), a(f0)
    epsilon, delta = a(epsilon), a(delta)
    tau0 = 1. / (2. * pi * f0)
    D = K_s * (1. + phi * (K_s / K_f - 1.))
    c33 = K_m + 4. / 3. * mu
    c55 = mu
    c11 = 2. * epsilon * c33 + c33
    c12 = c11 - 2. * c55
    c13 = np.sqrt(2. * delta * c33 * (c33 - c55) + (c33 - c55) * (c33 - c55)) - c55
    M = K_s * K_s / (D - (2. * c11 + c33 + 2. * c12 + 4. * c13) / 9.)
    alpha1 = 1. - 1. / 3. * (c11 + c12 + c13) / K_s
    alpha3 = 1. - 1. / 3. * (2. * c13 + c33) / K_s
    m1 = T1 * rho_f / phi
    m3 = T3 * rho_f / phi
    lowQ = Q < 1.e4
    Qs = np.where(lowQ, Q, 1.0)
    tau_e = tau0 / Qs * (np.sqrt(Qs * Qs + 1.) + 1.)
    tau_s = np.where(lowQ, tau0 / Qs * (np.sqrt(Qs * Qs + 1.) - 1.), tau0)
    varphi = np.where(lowQ, tau_e / np.where(lowQ, tau_s, 1.0) - 1., 0.0)
    rho = (1. - phi) * rho_s + phi * rho_f
    z = np.zeros_like
    rho_i, rho_f_i, nk_i, m_i, cq_i, cv_i = z(M), z(M), z(M), z(M), z(M), z(M)
    rho_j, rho_f_j, nk_j, m_j, cq_j, cv_j = z(M), z(M), z(M), z(M), z(M), z(M)
    c55_ij = z(M)

    def coeffs(r, rf, mm, nkk):
        tmp = 1. / (rf * rf - r * mm)
        beta12 = tmp * rf
        beta22 = tmp * -r
        cq = np.exp(-nkk * beta22 * dt)
        return cq, beta12 / beta22 * (cq - 1.)
    nk1, nk3 = eta / kappa1, eta / kappa3
    rho_i[:-1] = 0.5 * (rho[:-1] + rho[1:]); rho_f_i[:-1] = 0.5 * (rho_f[:-1] + rho_f[1:])
    nk_i[:-1] = 0.5 * (nk1[:-1] + nk1[1:]); m_i[:-1] = 0.5 * (m1[:-1] + m1[1:])
    cq_i[:-1], cv_i[:-1] = coeffs(rho_i[:-1], rho_f_i[:-1], m_i[:-1], nk_i[:-1])
    rho_j[:, :-1] = 0.5 * (rho[:, :-1] + rho[:, 1:]); rho_f_j[:, :-1] = 0.5 * (rho_f[:, :-1] + rho_f[:, 1:])
    nk_j[:, :-1] = 0.5 * (nk3[:, :-1] + nk3[:, 1:]); m_j[:, :-1] = 0.5 * (m3[:, :-1] + m3[:, 1:])
    cq_j[:, :-1], cv_j[:, :-1] = coeffs(rho_j[:, :-1], rho_f_j[:, :-1], m_j[:, :-1], nk_j[:, :-1])
    with np.errstate(divide="ignore"):
        c55_ij[:-1, :-1] = 4.0 / (1. / mu[:-1, :-1] + 1. / mu[1:, :-1] + 1. / mu[:-1, 1:] + 1. / mu[1:, 1:])
    for x in (rho_i, rho_f_i, nk_i, m_i, cq_i, cv_i, c55_ij):
        x[-1, :] = x[-2, :]
    for x in (rho_j, rho_f_j, nk_j, m_j, cq_j, cv_j, c55_ij):
        x[:, -1] = x[:, -2]
    return dict(c11=c11, c13=c13, c33=c33, alpha1=alpha1, alpha3=alpha3, M=M, ts=1. / tau_s,
                varphi=M * varphi / (1. + varphi), mu_ij=c55_ij, cq_i=cq_i, cv_i=cv_i, cq_j=cq_j, cv_j=cv_j,
                m_i=rho_f_i * rho_f_i / rho_i - m_i, m_j=rho_f_j * rho_f_j / rho_j - m_j, rf_i=rho_f_i / rho_i,
                rf_j=rho_f_j / rho_j, irho_i=1. / rho_i, irho_j=1. / rho_j)


class PveState:
    """W (9 components) and the CPML memory variables mem[0..2]
    (alloc_cpml, pml.c:1409-1491)."""

    DX = ("dx_txx", "dx_p", "dx_txz", "dx_vx", "dx_qx", "dx_vz")
    DZ = ("dz_txz", "dz_tzz", "dz_p", "dz_vz", "dz_qz", "dz_vx")

    def __init__(self, nx2, nz2, np_):
        self.W = np.zeros((9, nx2, nz2))
        self.delta = np.zeros((9, nx2, nz2))
        self.mem = []
        for _ in range(3):
            m = {nm: np.zeros((2 * np_ + 1) * nz2) for nm in self.DX}
            m.update({nm: np.zeros((2 * np_ + 1) * nx2) for nm in self.DZ})
            self.mem.append(m)


def propagate(Wt, delta, mat, fp, mem, np_, mx_f, mx_b, mz_f, mz_b):
    """propagateCPML (propagate.c:884-1329) on Wt -> delta."""
    txx, tzz, txz, p, vx, vz, qx, qz, e = Wt
    inner = (slice(1, -1), slice(1, -1))
    H = (fp["bh_x"], fp["ch_x"], fp["kh_x"], fp["bH_x"], fp["cH_x"], fp["kH_x"])
    Hz = (fp["bh_z"], fp["ch_z"], fp["kh_z"], fp["bH_z"], fp["cH_z"], fp["kH_z"])
    Bx = (fp["b_x"], fp["c_x"], fp["k_x"]) * 2
    Bz = (fp["b_z"], fp["c_z"], fp["k_z"]) * 2
    # v_x, q_x (909-1005)
    t1 = deriv(txx, 0, mx_f)
    _x_layer(t1, "dx_txx", mem, np_, np_, *H)
    t2 = deriv(txz, 1, mz_b)
    _z_layer(t2, "dz_txz", mem, np_, np_ - 1, *Bz)
    t1 = t1 + t2
    t2 = deriv(p, 0, mx_f)
    _x_layer(t2, "dx_p", mem, np_, np_, *H)
    q2 = (t2 + mat["rf_i"] * t1) / mat["m_i"]
    delta[I_QX][inner] = q2[inner]
    delta[I_VX][inner] = (mat["irho_i"] * t1 - mat["rf_i"] * q2)[inner]
    # v_z, q_z (1007-1098)
    t1 = deriv(txz, 0, mx_b)
    _x_layer(t1, "dx_txz", mem, np_, np_ - 1, *Bx)
    t2 = deriv(tzz, 1, mz_f)
    _z_layer(t2, "dz_tzz", mem, np_, np_, *Hz)
    t1 = t1 + t2
    t2 = deriv(p, 1, mz_f)
    _z_layer(t2, "dz_p", mem, np_, np_, *Hz)
    q2 = (t2 + mat["rf_j"] * t1) / mat["m_j"]
    delta[I_QZ][inner] = q2[inner]
    delta[I_VZ][inner] = (mat["irho_j"] * t1 - mat["rf_j"] * q2)[inner]
    # epsilon = alpha (D_x^- v_x + D_z^- v_z) + D_x^- q_x + D_z^- q_z (1100-1211)
    t1 = deriv(vx, 0, mx_b)
    _x_layer(t1, "dx_vx", mem, np_, np_ - 1, *Bx)
    t2 = deriv(vz, 1, mz_b)
    _z_layer(t2, "dz_vz", mem, np_, np_ - 1, *Bz)
    if "c11" in mat:         # propagateVTI_CPML (propagate.c:1786)
        eps = mat["alpha1"] * t1 + mat["alpha3"] * t2
    else:
        eps = mat["alpha"] * (t1 + t2)
    t3 = deriv(qx, 0, mx_b)
    _x_layer(t3, "dx_qx", mem, np_, np_ - 1, *Bx)
    eps = eps + t3
    t3 = deriv(qz, 1, mz_b)
    _z_layer(t3, "dz_qz", mem, np_, np_ - 1, *Bz)
    eps = eps + t3
    # e, p, tau_xx, tau_zz (1213-1251)
    delta[I_E] = -mat["ts"] * (mat["varphi"] * eps + e)
    eps = mat["M"] * eps + e
    delta[I_P] = -eps
    if "c11" in mat:         # propagate.c:1871-1880
        delta[I_TXX] = mat["c11"] * t1 + mat["c13"] * t2 + mat["alpha1"] * eps
        delta[I_TZZ] = mat["c13"] * t1 + mat["c33"] * t2 + mat["alpha3"] * eps
    else:
        delta[I_TXX] = mat["E"] * t1 + mat["Emu"] * t2 + mat["alpha"] * eps
        delta[I_TZZ] = mat["Emu"] * t1 + mat["E"] * t2 + mat["alpha"] * eps
    # tau_xz (1253-1311)
    t1 = deriv(vx, 1, mz_f)
    _z_layer(t1, "dz_vx", mem, np_, np_, *Hz)
    t2 = deriv(vz, 0, mx_f)
    _x_layer(t2, "dx_vz", mem, np_, np_, *H)
    delta[I_TXZ] = mat["mu_ij"] * (t1 + t2)
    # 1312-1328
    nx2, nz2 = Wt.shape[1:]
    m1, m2 = mem[1], mem[2]
    n1, n2 = (2 * np_ + 1) * nz2, (2 * np_ + 1) * nx2
    for nm in PveState.DX:
        m2[nm][:n1] = m1[nm][:n1]
    for nm in PveState.DZ:
        m2[nm][:n2] = m1[nm][:n2]


def add_src(W, Wt, src, dt_w, dt_t, it):
    """add_src_rk4 (src.c:77-112): p (Sf, Bulk), tau_xx (Sx, Bulk, Bulk_s),
    tau_zz (Sz, Bulk, Bulk_s), tau_xz (Sxz), in that order per source."""
    for typ, idx, fct in src:
        k = 2 * it + 1
        if k >= len(fct):
            continue
        i, j = divmod(int(idx), W.shape[2])
        qs = []
        if typ in (SF, BULK):
            qs.append(I_P)
        if typ in (SX, BULK, BULK_S):
            qs.append(I_TXX)
        if typ in (SZ, BULK, BULK_S):
            qs.append(I_TZZ)
        if typ == SXZ:
            qs.append(I_TXZ)
        for q in qs:
            if dt_t != 0.0:
                Wt[q, i, j] += dt_t * fct[k]
            W[q, i, j] += dt_w * fct[k]


COMP_Q = {VX: I_VX, VZ: I_VZ, QX: I_QX, QZ: I_QZ, TXX: I_TXX, TZZ: I_TZZ, TXZ: I_TXZ, P: I_P}


def pve_run(cfg, mat, src, rec_idx, rec_comp, nsteps):
    """Run the restatement.  cfg: np, dx, dz, dt, vmax, pml_order, Rc,
    kappa_max, alpha_max (value, 0 for none), iwipe.  Returns traces
    (nsteps x nrec; DIV and CURL from the last stage's input) and the final
    W."""
    np_ = cfg["np"]
    nx2, nz2 = mat["M"].shape
    tabs = cpml_tables(np_, cfg["dx"], cfg["dz"], cfg["dt"], cfg.get("vmax", 5000.0), cfg.get("pml_order", 2.0),
                       cfg.get("Rc", 1e-5), cfg.get("kappa_max", 1.0), cfg.get("alpha_max", 0.0),
                       cfg.get("iwipe", 1))
    mx_f = k_multipliers(nx2, cfg["dx"], True); mx_b = k_multipliers(nx2, cfg["dx"], False)
    mz_f = k_multipliers(nz2, cfg["dz"], True); mz_b = k_multipliers(nz2, cfg["dz"], False)
    st = PveState(nx2, nz2, np_)
    W, delta, mem = st.W, st.delta, st.mem
    dt = cfg["dt"]
    traces = np.zeros((nsteps, len(rec_idx)))
    rest = (mem, np_, mx_f, mx_b, mz_f, mz_b)
    for it in range(nsteps):
        Ws = W.copy()
        # analytic step of the viscous coupling (1255-1262)
        Ws[I_VX] = W[I_VX] + mat["cv_i"] * W[I_QX]
        Ws[I_VZ] = W[I_VZ] + mat["cv_j"] * W[I_QZ]
        Ws[I_QX] = mat["cq_i"] * W[I_QX]
        Ws[I_QZ] = mat["cq_j"] * W[I_QZ]
        # 1270-1285: over 2np + 1 rows here (2np in the viscoelastic programs
        # and pve_vti_pml.c:1226-1241)
        rows = 2 * np_ if "c11" in mat else 2 * np_ + 1
        n1 = rows * nz2
        for nm in PveState.DX:
            mem[0][nm][:n1] = mem[1][nm][:n1]
        n2 = rows * nx2
        for nm in PveState.DZ:
            mem[0][nm][:n2] = mem[1][nm][:n2]
        Wt = Ws.copy()
        propagate(Wt, delta, mat, tabs[1], *rest)
        Wt = Ws + dt / 2. * delta
        W = Ws + dt / 6. * delta
        add_src(W, Wt, src, dt / 6., dt / 2., it)
        propagate(Wt, delta, mat, tabs[23], *rest)
        Wt = Ws + dt / 2. * delta
        W = W + dt / 3. * delta
        add_src(W, Wt, src, dt / 3., dt / 2., it)
        propagate(Wt, delta, mat, tabs[23], *rest)
        Wt = Ws + dt * delta
        W = W + dt / 3. * delta
        add_src(W, Wt, src, dt / 3., dt, it)
        propagate(Wt, delta, mat, tabs[4], *rest)
        W = W + dt / 6. * delta
        add_src(W, Wt, src, dt / 6., 0.0, it)
        div = curl = None
        for r, (idx, c) in enumerate(zip(rec_idx, rec_comp)):
            i, j = divmod(int(idx), nz2)
            if c == DIV:
                if div is None:
                    div = deriv(Wt[I_VX], 0, mx_b) + deriv(Wt[I_VZ], 1, mz_b)
                traces[it, r] = div[i, j]
            elif c == CURL:
                if curl is None:
                    curl = deriv(Wt[I_VZ], 0, mx_f) - deriv(Wt[I_VX], 1, mz_f)
                traces[it, r] = curl[i, j]
            else:
                traces[it, r] = W[COMP_Q[c], i, j]
    return traces, W
