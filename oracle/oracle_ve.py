"""FP64 restatement of ve_vti_pml (viscoelastic P-SV waves in a 2D VTI
medium, pseudo-spectral derivatives, CPML, fourth-order Runge-Kutta, L
relaxation mechanisms for the quasi-dilatational mode and one for shear).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline
leg as the checker.  The product path never calls it.

Follows the reference step by step:
  - model units and material: ve_vti_pml.c:518-557 (c11, c13, c33, K0,
    c55_0 and the relaxation times at the nodes), 559-648 (staggered
    averages and the fill of the last row and column), 650-683 (stored as
    1/rho, 1/tau_s, (1 - 1/eta) and (1 - 1/eta1)/L);
  - CPML coefficients (pml.c:1959-2106): the tables of oracle_sh.sh_cpml;
  - derivative operators: propagate.c:325-375 (multipliers),
    2463-2486 (partial_*), plans 386-497;
  - propagateVE_CPML: propagate.c:1959-2278, its region loops and
    memory-variable copies as written;
  - Runge-Kutta loop and source injection: ve_vti_pml.c:720-784,
    src.c:114-142 (add_src_ve_rk4).

Two reads of the reference are kept as written although they look like
slips, because they define its results:
  - the relaxation sum of tau_xx and tau_zz reads e1[m + nnodes + n]
    (propagate.c:2190-2195), i.e. the state at flat offset 6 nnodes + n + m
    rather than e1 of mechanism m: with L = 1 that is e2;
  - c55_0_ij, which multiplies e3 in tau_xz (propagate.c:2263), is
    allocated (ve_vti_pml.c:368) and never set.  Freshly allocated memory
    is zero-filled on the reference's platform (glibc serves it from new
    zeroed pages), so it is 0 here: e3 evolves but does not feed back.

The FFTs are numpy's (pocketfft), not FFTW: results agree with an FFTW
build to rounding, so parity with the GPU path is stated as a relative
tolerance, not bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .oracle_sh import deriv, k_multipliers
from .oracle_sh import sh_cpml as cpml_tables

SX, SZ, SXZ, BULK = 0, 2, 4, 6
VX, VZ, TXX, TZZ, TXZ = 0, 2, 5, 6, 7
# state components: tau_xx, tau_zz, tau_xz, v_x, v_z, e1 (L), e2, e3
I_TXX, I_TZZ, I_TXZ, I_VX, I_VZ = 0, 1, 2, 3, 4


def ve_material(K_gpa, mu_gpa, rho, eps, delta, Q1, f1, Q2, f2):
    """Material of ve_vti_pml.c:518-683 from the padded model as
    read_model_ve leaves it (K, mu in GPa, rho in kg/m^3; Q1, f1 of shape
    (L, nx2, nz2)).  Returns a dict of the stored arrays (see the module
    docstring for c55_0_ij)."""
    pi = 4.0 * math.atan(1.0)
    K = np.asarray(K_gpa, np.float64) * 1000.
    mu = np.asarray(mu_gpa, np.float64) * 1000.
    rho = np.asarray(rho, np.float64) * 1.e-6
    eps = np.asarray(eps, np.float64)
    delta = np.asarray(delta, np.float64)
    Q1 = np.asarray(Q1, np.float64)
    f1 = np.asarray(f1, np.float64)
    Q2 = np.asarray(Q2, np.float64)
    f2 = np.asarray(f2, np.float64)
    L = Q1.shape[0]
    nx2, nz2 = K.shape
    # nodes, 527-557
    c33 = K + 4. / 3. * mu
    c55 = mu
    c11 = 2. * eps * c33 + c33
    c13 = np.sqrt(2. * delta * c33 * (c33 - c55) + (c33 - c55) * (c33 - c55)) - c55
    ts1 = np.empty_like(Q1); eta1 = np.empty_like(Q1)
    for m in range(L):
        tau_e = (np.sqrt(Q1[m] * Q1[m] + 1) + 1.) / (2. * pi * f1[m] * Q1[m])
        ts1[m] = (np.sqrt(Q1[m] * Q1[m] + 1) - 1.) / (2. * pi * f1[m] * Q1[m])
        eta1[m] = ts1[m] / tau_e
    tau_e = (np.sqrt(Q2 * Q2 + 1) + 1.) / (2. * pi * f2 * Q2)
    ts2 = (np.sqrt(Q2 * Q2 + 1) - 1.) / (2. * pi * f2 * Q2)
    eta2 = ts2 / tau_e
    c55_0 = c55 * eta2
    D = 0.5 * (c11 + c33)
    Kk = D - c55
    e1s = np.zeros_like(K)
    for m in range(L):
        e1s = e1s + eta1[m]
    c11_0 = c11 - D + Kk * e1s + c55 * eta2
    c33_0 = c33 - D + Kk * e1s + c55 * eta2
    K0 = 0.5 * (c11_0 + c33_0) - c55_0
    # staggered, 559-608
    rho_i = np.zeros_like(K); rho_j = np.zeros_like(K)
    c55s = np.zeros_like(K); ts2_ij = np.zeros_like(K); eta2_ij = np.zeros_like(K)
    rho_i[:-1, :] = 0.5 * (rho[:-1, :] + rho[1:, :])
    rho_j[:, :-1] = 0.5 * (rho[:, :-1] + rho[:, 1:])
    c55s[:-1, :-1] = 4.0 / (1. / mu[:-1, :-1] + 1. / mu[1:, :-1] + 1. / mu[:-1, 1:] + 1. / mu[1:, 1:])
    ts2_ij[:-1, :-1] = 0.25 * (ts2[:-1, :-1] + ts2[1:, :-1] + ts2[:-1, 1:] + ts2[1:, 1:])
    eta2_ij[:-1, :-1] = 0.25 * (eta2[:-1, :-1] + eta2[1:, :-1] + eta2[:-1, 1:] + eta2[1:, 1:])
    # fill of the last row, then of the last column (610-635): the row
    # fill reads the corner before the column fill sets it, and the column
    # fill then overwrites that corner with a defined value
    for a in (rho_i, c55s, ts2_ij, eta2_ij):
        a[-1, :] = a[-2, :]
    for a in (rho_j, c55s, ts2_ij, eta2_ij):
        a[:, -1] = a[:, -2]
    # 650-683
    return dict(c11=c11, c13=c13, c33=c33, c55=c55s, K0=K0, c55_0=c55_0, c55_0_ij=np.zeros_like(K),
                irho_i=1. / rho_i, irho_j=1. / rho_j, eta2=1. - 1. / eta2, eta2_ij=1. - 1. / eta2_ij,
                ts2=1. / ts2, ts2_ij=1. / ts2_ij, eta1=(1. - 1. / eta1) / L, ts1=1. / ts1, L=L)


class VeState:
    """W = (tau_xx, tau_zz, tau_xz, v_x, v_z, e1[L], e2, e3) and the CPML
    memory variables mem[0..2] (alloc_cpml_ve, pml.c:1553-1631)."""

    DX = ("dx_txx", "dx_txz", "dx_vx", "dx_vz")
    DZ = ("dz_txz", "dz_tzz", "dz_vz", "dz_vx")

    def __init__(self, nx2, nz2, np_, L):
        self.W = np.zeros((7 + L, nx2, nz2))
        self.delta = np.zeros((7 + L, nx2, nz2))
        self.mem = []
        for _ in range(3):
            m = {nm: np.zeros((2 * np_ + 1) * nz2) for nm in self.DX}
            m.update({nm: np.zeros((2 * np_ + 1) * nx2) for nm in self.DZ})
            self.mem.append(m)


def _x_layer(t, nm, mem, np_, last, b1, c1, k1, b2, c2, k2):
    """x layers: left rows nl < np with tables 1, right rows nx2-1-nl for
    nl <= last with tables 2, one slot per (row, column) in loop order."""
    m0, m1, m2 = mem
    nx2, nz2 = t.shape
    tf = t.reshape(-1)
    ipml = 0
    for nl in range(np_):
        ind = nl * nz2 + np.arange(nz2)
        sl = slice(ipml, ipml + nz2)
        m1[nm][sl] = b1[nl] * m0[nm][sl] + c1[nl] * tf[ind]
        tf[ind] = tf[ind] / k1[nl] + m2[nm][sl]
        ipml += nz2
    for nl in range(last + 1):
        ind = (nx2 - 1 - nl) * nz2 + np.arange(nz2)
        sl = slice(ipml, ipml + nz2)
        m1[nm][sl] = b2[nl] * m0[nm][sl] + c2[nl] * tf[ind]
        tf[ind] = tf[ind] / k2[nl] + m2[nm][sl]
        ipml += nz2


def _z_layer(t, nm, mem, np_, last, b1, c1, k1, b2, c2, k2):
    """z layers: top columns nl < np, bottom columns nz2-1-nl for nl <= last."""
    m0, m1, m2 = mem
    nx2, nz2 = t.shape
    tf = t.reshape(-1)
    ipml = 0
    for nl in range(np_):
        ind = np.arange(nx2) * nz2 + nl
        sl = slice(ipml, ipml + nx2)
        m1[nm][sl] = b1[nl] * m0[nm][sl] + c1[nl] * tf[ind]
        tf[ind] = tf[ind] / k1[nl] + m2[nm][sl]
        ipml += nx2
    for nl in range(last + 1):
        ind = np.arange(nx2) * nz2 + nz2 - 1 - nl
        sl = slice(ipml, ipml + nx2)
        m1[nm][sl] = b2[nl] * m0[nm][sl] + c2[nl] * tf[ind]
        tf[ind] = tf[ind] / k2[nl] + m2[nm][sl]
        ipml += nx2


def propagate(Wt, delta, mat, fp, mem, np_, mx_f, mx_b, mz_f, mz_b):
    """propagateVE_CPML (propagate.c:1959-2278) on Wt -> delta."""
    L = mat["L"]
    nx2, nz2 = Wt.shape[1:]
    N = nx2 * nz2
    txx, tzz, txz, vx, vz = Wt[0], Wt[1], Wt[2], Wt[3], Wt[4]
    e1 = Wt[5:5 + L]
    e2 = Wt[5 + L]
    e3 = Wt[6 + L]
    Wf = Wt.reshape(-1)
    inner = (slice(1, -1), slice(1, -1))
    # rho v_x' = D_x^+ tau_xx + D_z^- tau_xz (1981-2047)
    t1 = deriv(txx, 0, mx_f)
    t2 = deriv(txz, 1, mz_b)
    _x_layer(t1, "dx_txx", mem, np_, np_, fp["bh_x"], fp["ch_x"], fp["kh_x"], fp["bH_x"], fp["cH_x"], fp["kH_x"])
    _z_layer(t2, "dz_txz", mem, np_, np_ - 1, fp["b_z"], fp["c_z"], fp["k_z"], fp["b_z"], fp["c_z"], fp["k_z"])
    t1 = t1 + t2
    delta[I_VX][inner] = mat["irho_i"][inner] * t1[inner]
    # rho v_z' = D_x^- tau_xz + D_z^+ tau_zz (2049-2112)
    t1 = deriv(txz, 0, mx_b)
    t2 = deriv(tzz, 1, mz_f)
    _x_layer(t1, "dx_txz", mem, np_, np_ - 1, fp["b_x"], fp["c_x"], fp["k_x"], fp["b_x"], fp["c_x"], fp["k_x"])
    _z_layer(t2, "dz_tzz", mem, np_, np_, fp["bh_z"], fp["ch_z"], fp["kh_z"], fp["bH_z"], fp["cH_z"], fp["kH_z"])
    t1 = t1 + t2
    delta[I_VZ][inner] = mat["irho_j"][inner] * t1[inner]
    # relaxation of the dilatation (2115-2180)
    t1 = deriv(vx, 0, mx_b)
    _x_layer(t1, "dx_vx", mem, np_, np_ - 1, fp["b_x"], fp["c_x"], fp["k_x"], fp["b_x"], fp["c_x"], fp["k_x"])
    t2 = deriv(vz, 1, mz_b)
    _z_layer(t2, "dz_vz", mem, np_, np_ - 1, fp["b_z"], fp["c_z"], fp["k_z"], fp["b_z"], fp["c_z"], fp["k_z"])
    for m in range(L):
        delta[5 + m] = mat["ts1"][m] * (mat["eta1"][m] * (t1 + t2) - e1[m])
    delta[5 + L] = 0.5 * mat["ts2"] * (mat["eta2"] * (t1 + t2) - 2. * e2)
    # tau_xx', tau_zz' (2183-2199): the relaxation sum reads the state at
    # flat offset 6 N + n + m, as written
    s = np.zeros(N)
    for m in range(L):
        s = s + Wf[6 * N + m:7 * N + m]
    s = s.reshape(nx2, nz2)
    delta[I_TXX] = mat["c11"] * t1 + mat["c13"] * t2 + mat["K0"] * s + 2. * mat["c55_0"] * e2
    delta[I_TZZ] = mat["c13"] * t1 + mat["c33"] * t2 + mat["K0"] * s - 2. * mat["c55_0"] * e2
    # e3', tau_xz' (2201-2263)
    t1 = deriv(vx, 1, mz_f)
    t2 = deriv(vz, 0, mx_f)
    _z_layer(t1, "dz_vx", mem, np_, np_, fp["bh_z"], fp["ch_z"], fp["kh_z"], fp["bH_z"], fp["cH_z"], fp["kH_z"])
    _x_layer(t2, "dx_vz", mem, np_, np_, fp["bh_x"], fp["ch_x"], fp["kh_x"], fp["bH_x"], fp["cH_x"], fp["kH_x"])
    delta[6 + L] = mat["ts2_ij"] * (mat["eta2_ij"] * (t1 + t2) - e3)
    delta[I_TXZ] = mat["c55"] * (t1 + t2) + mat["c55_0_ij"] * e3
    # 2265-2277
    m1, m2 = mem[1], mem[2]
    n1 = (2 * np_ + 1) * nz2
    for nm in VeState.DX:
        m2[nm][:n1] = m1[nm][:n1]
    n2 = (2 * np_ + 1) * nx2
    for nm in VeState.DZ:
        m2[nm][:n2] = m1[nm][:n2]


def add_src(W, Wt, src, dt_w, dt_t, it):
    """add_src_ve_rk4 (src.c:114-142): sample 2 it + 1 of each source
    function into tau_xx (Sx, Bulk), tau_zz (Sz, Bulk) or tau_xz (Sxz)."""
    for typ, idx, fct in src:
        k = 2 * it + 1
        if k >= len(fct):
            continue
        i, j = divmod(int(idx), W.shape[2])
        qs = []
        if typ in (SX, BULK):
            qs.append(I_TXX)
        if typ in (SZ, BULK):
            qs.append(I_TZZ)
        if typ == SXZ:
            qs.append(I_TXZ)
        for q in qs:
            if dt_t != 0.0:
                Wt[q, i, j] += dt_t * fct[k]
            W[q, i, j] += dt_w * fct[k]


COMP_Q = {TXX: I_TXX, TZZ: I_TZZ, TXZ: I_TXZ, VX: I_VX, VZ: I_VZ}


def ve_run(cfg, mat, src, rec_idx, rec_comp, nsteps):
    """Run the restatement.  cfg: np, dx, dz, dt, vmax, pml_order, Rc,
    kappa_max, alpha_max (value, 0 for none), iwipe.  mat: ve_material's
    dict.  src: list of (type, padded index, fct at dt/2 already divided by
    dx dz).  Returns traces (nsteps x nrec) and the final W."""
    np_ = cfg["np"]
    nx2, nz2 = mat["c11"].shape
    L = mat["L"]
    tabs = cpml_tables(np_, cfg["dx"], cfg["dz"], cfg["dt"], cfg.get("vmax", 5000.0), cfg.get("pml_order", 2.0),
                       cfg.get("Rc", 1e-5), cfg.get("kappa_max", 1.0), cfg.get("alpha_max", 0.0),
                       cfg.get("iwipe", 1))
    mx_f = k_multipliers(nx2, cfg["dx"], True); mx_b = k_multipliers(nx2, cfg["dx"], False)
    mz_f = k_multipliers(nz2, cfg["dz"], True); mz_b = k_multipliers(nz2, cfg["dz"], False)
    st = VeState(nx2, nz2, np_, L)
    W, delta, mem = st.W, st.delta, st.mem
    dt = cfg["dt"]
    traces = np.zeros((nsteps, len(rec_idx)))
    rest = (mem, np_, mx_f, mx_b, mz_f, mz_b)
    for it in range(nsteps):
        Ws = W.copy()
        n1 = 2 * np_ * nz2
        for nm in VeState.DX:
            mem[0][nm][:n1] = mem[1][nm][:n1]
        n2 = 2 * np_ * nx2
        for nm in VeState.DZ:
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
        for r, (idx, c) in enumerate(zip(rec_idx, rec_comp)):
            i, j = divmod(int(idx), nz2)
            traces[it, r] = W[COMP_Q[c], i, j]
    return traces, W
