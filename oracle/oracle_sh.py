"""FP64 restatement of ve_vti_sh_pml (viscoelastic SH waves in a 2D VTI
medium, pseudo-spectral derivatives, CPML, fourth-order Runge-Kutta).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline
leg as the checker.  The product path never calls it.

Follows the reference step by step:
  - material at staggered nodes: ve_vti_sh_pml.c:475-585;
  - CPML coefficients of the three Runge-Kutta stage sets fp1, fp23, fp4:
    pml.c:1959-2106 (compute_cpml), shared kappa tables pml.c:1747-1752;
  - derivative operators: propagate.c:540-591 (k multipliers with the
    half-cell shift), 2487-2510 (partial_*_sh), FFT layout 593-669;
  - propagateVE_SH_CPML: propagate.c:2280-2462, including its region
    loops and memory-variable bookkeeping as written;
  - Runge-Kutta loop and source injection: ve_vti_sh_pml.c:613-700,
    src.c:145-168 (add_src_ve_sh_rk4).

The FFTs are numpy's (pocketfft), not FFTW: results agree with an FFTW
build to rounding, so parity with the GPU path is stated as a relative
tolerance, not bit for bit.

Three reads of the reference are undefined (uninitialised or out-of-bounds
memory) and are given their natural values here:
  - c44_0 of the last x row multiplies eta2 before the reference fills that
    row (ve_vti_sh_pml.c:520 reads it, 531-540 fills it): the filled value
    (a copy of row nx2-2) is used;
  - the right-side loop of dx_txy and the bottom loop of dz_vy/dx_vy run to
    nl = np, one past the np-entry b/c/k tables (propagate.c:2327-2334): the
    tables' own formulas at n = np give k = 1 and c = 0 (no damping at the
    layer's inner edge), which leaves the derivative unchanged;
  - dx_vy is indexed like a z layer over nx2 columns in an array of
    (2np+1)*nz2 entries (propagate.c:2410-2425): the array is sized for the
    larger of the two here, so nx2 > nz2 does not overrun it.
"""
from __future__ import annotations

import math

import numpy as np

SXY, SYZ = 3, 5
VY, TXY, TYZ = 1, 9, 10


def sh_material(mu_gpa, rho, gamma, Q, f):
    """Staggered material of ve_vti_sh_pml.c:475-585 from the padded model
    as read_model_ve_sh leaves it (mu = Vs0^2 rho 1e-9, in GPa).  Returns
    c44, c66, c44_0, c66_0 (MPa), tau_s_1, tau_s_2 (stored as 1/tau_s),
    eta1, eta2 (stored as 1 - 1/eta) and irho (1/rho, rho in Gg/m^3)."""
    pi = 4.0 * math.atan(1.0)
    mu = np.asarray(mu_gpa, dtype=np.float64) * 1000.0
    rho = np.asarray(rho, dtype=np.float64) * 1.e-6
    gamma = np.asarray(gamma, dtype=np.float64)
    Q = np.asarray(Q, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    nx2, nz2 = mu.shape
    c44 = np.zeros_like(mu); c66 = np.zeros_like(mu)
    c44_0 = np.zeros_like(mu); c66_0 = np.zeros_like(mu)
    ts1 = np.zeros_like(mu); ts2 = np.zeros_like(mu)
    eta1 = np.zeros_like(mu); eta2 = np.zeros_like(mu)
    # (i+1/2, j), ve_vti_sh_pml.c:488-507
    q = 0.5 * (Q[:-1, :] + Q[1:, :])
    fr = 0.5 * (f[:-1, :] + f[1:, :])
    tau_e = (np.sqrt(q * q + 1) + 1.) / (2. * pi * fr * q)
    ts2[:-1, :] = (np.sqrt(q * q + 1) - 1.) / (2. * pi * fr * q)
    eta2[:-1, :] = ts2[:-1, :] / tau_e
    c = 0.5 * (mu[:-1, :] + mu[1:, :])
    g = 0.5 * (gamma[:-1, :] + gamma[1:, :])
    c66[:-1, :] = 2. * c * g + c
    c66_0[:-1, :] = c66[:-1, :] * eta2[:-1, :]
    # last row filled from the one before (531-540); filled before it is
    # read below (see the module docstring)
    c66[-1, :] = c66[-2, :]; c66_0[-1, :] = c66_0[-2, :]
    ts2[-1, :] = ts2[-2, :]; eta2[-1, :] = eta2[-2, :]
    # (i, j+1/2), 513-528; c44_0 multiplies eta2, as written
    q = 0.5 * (Q[:, :-1] + Q[:, 1:])
    fr = 0.5 * (f[:, :-1] + f[:, 1:])
    tau_e = (np.sqrt(q * q + 1) + 1.) / (2. * pi * fr * q)
    ts1[:, :-1] = (np.sqrt(q * q + 1) - 1.) / (2. * pi * fr * q)
    eta1[:, :-1] = ts1[:, :-1] / tau_e
    c44[:, :-1] = 0.5 * (mu[:, :-1] + mu[:, 1:])
    c44_0[:, :-1] = c44[:, :-1] * eta2[:, :-1]
    c44[:, -1] = c44[:, -2]; c44_0[:, -1] = c44_0[:, -2]
    ts1[:, -1] = ts1[:, -2]; eta1[:, -1] = eta1[:, -2]
    # 574-582
    return dict(c44=c44, c66=c66, c44_0=c44_0, c66_0=c66_0, ts1=1. / ts1, ts2=1. / ts2,
                eta1=1. - 1. / eta1, eta2=1. - 1. / eta2, irho=1. / rho)


def sh_cpml(np_, dx, dz, dt, vmax, order, Rc, kappa_max, alpha_max, iwipe):
    """pml.c:1959-2106.  Returns {stage: {name: array}} for stages 1, 23, 4
    with arrays b_x, c_x, bh_x, ch_x, b_z, ... of np + 1 entries (entry np is
    the tables' own formula there, see the module docstring) and bH_*, cH_*
    of np + 1 entries; k_x, kh_x, kH_x (and z) are shared."""
    n = np_
    out = {s: {} for s in (1, 23, 4)}
    ks = {}
    if iwipe and n > 0:
        Lx, Lz = n * dx, n * dz
        d0x = -(order + 1.) * vmax * math.log(Rc) / (2. * Lx)
        d0z = -(order + 1.) * vmax * math.log(Rc) / (2. * Lz)
        da = alpha_max / n
        theta = 0.5
        for ax, d, L, d0 in (("x", dx, Lx, d0x), ("z", dz, Lz, d0z)):
            k = np.empty(n + 1); kh = np.empty(n + 1); kH = np.empty(n + 1)
            tabs = {s: {nm: np.empty(n + 1) for nm in ("b", "c", "bh", "ch", "bH", "cH")}
                    for s in (1, 23, 4)}
            for m in range(n + 1):
                k[m] = 1. + (kappa_max - 1.) * ((n - m) * d / L) ** order
                kh[m] = 1. + (kappa_max - 1.) * ((n - m - 0.5) * d / L) ** order if m < n else 1.
                dd = d0 * ((n - m) * d / L) ** order
                al = m * da
                ddh = d0 * ((n - m - 0.5) * d / L) ** order if m < n else 0.
                alh = (m + 0.5) * da
                kH[m] = 1. + (kappa_max - 1.) * ((n - m + 0.5) * d / L) ** order
                ddH = d0 * ((n - m + 0.5) * d / L) ** order
                alH = (m - 0.5) * da
                tabs[1]["b"][m] = tabs[1]["bh"][m] = tabs[1]["bH"][m] = 1.0
                tabs[1]["c"][m] = tabs[1]["ch"][m] = tabs[1]["cH"][m] = 0.0
                for s, h in ((23, dt * 0.5), (4, dt)):
                    def bc(kk, dv, a):
                        b = (kk - (1. - theta) * h * (dv + kk * a)) / (kk + theta * h * (dv + kk * a))
                        c = -dv * h / (kk * (kk + theta * h * (dv + kk * a)))
                        return b, c
                    tabs[s]["b"][m], tabs[s]["c"][m] = bc(k[m], dd, al)
                    tabs[s]["bh"][m], tabs[s]["ch"][m] = bc(kh[m], ddh, alh)
                    tabs[s]["bH"][m], tabs[s]["cH"][m] = bc(kH[m], ddH, alH)
            ks["k_" + ax], ks["kh_" + ax], ks["kH_" + ax] = k, kh, kH
            for s in (1, 23, 4):
                for nm, v in tabs[s].items():
                    out[s][nm + "_" + ax] = v
    else:
        for ax in ("x", "z"):
            ks["k_" + ax] = np.ones(n + 1); ks["kh_" + ax] = np.ones(n + 1)
            ks["kH_" + ax] = np.ones(n + 1)
            for s in (1, 23, 4):
                for nm in ("b", "c", "bh", "ch", "bH", "cH"):
                    out[s][nm + "_" + ax] = np.zeros(n + 1)
    for s in (1, 23, 4):
        out[s].update(ks)
    return out


def k_multipliers(n, d, forward):
    """i k e^{+-i k d / 2} for the rfft bins of n samples of spacing d
    (propagate.c:566-577, the reference folds the 1/n of its unnormalised
    inverse FFT into k; numpy's irfft normalises instead)."""
    kk = 2. * math.pi * np.arange(n // 2 + 1) / (d * n)
    return 1j * kk * np.exp((0.5j if forward else -0.5j) * kk * d)


def deriv(a, axis, mult):
    n = a.shape[axis]
    shape = [1, 1]
    shape[axis] = -1
    return np.fft.irfft(np.fft.rfft(a, axis=axis) * mult.reshape(shape), n=n, axis=axis)


class ShState:
    """Fields W = (tau_xy, tau_yz, v_y, e1, e2) and memory variables."""

    def __init__(self, nx2, nz2, np_):
        self.W = np.zeros((5, nx2, nz2))
        self.delta = np.zeros((5, nx2, nz2))
        m = max(nx2, nz2)
        # mem[0..2] of propagate.c: dx_txy, dx_vy over (2np+1) rows of nz2
        # (dx_vy indexed with nx2, sized for the larger), dz_* over (2np+1) x nx2
        self.mem = [dict(dx_txy=np.zeros((2 * np_ + 1) * nz2), dx_vy=np.zeros((2 * np_ + 1) * m),
                         dz_tyz=np.zeros((2 * np_ + 1) * nx2), dz_vy=np.zeros((2 * np_ + 1) * nx2))
                    for _ in range(3)]


def propagate(Wt, delta, mat, fp, mem, np_, mx_f, mx_b, mz_f, mz_b):
    """propagateVE_SH_CPML (propagate.c:2280-2462) on Wt -> delta."""
    nx2, nz2 = Wt.shape[1:]
    npml = np_
    txy, tyz, vy, e1, e2 = Wt
    t1 = deriv(txy, 0, mx_b)             # D_x^- tau_xy
    t2 = deriv(tyz, 1, mz_b)             # D_z^- tau_yz
    m0, m1, m2 = mem
    t1f = t1.reshape(-1); t2f = t2.reshape(-1)
    # dx_txy: left rows nl < np, right rows nx2-1-nl for nl <= np
    ipml = 0
    for nl in range(npml):
        ind = nl * nz2 + np.arange(nz2)
        sl = slice(ipml, ipml + nz2)
        m1["dx_txy"][sl] = fp["b_x"][nl] * m0["dx_txy"][sl] + fp["c_x"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["k_x"][nl] + m2["dx_txy"][sl]
        ipml += nz2
    for nl in range(npml + 1):
        ind = (nx2 - 1 - nl) * nz2 + np.arange(nz2)
        sl = slice(ipml, ipml + nz2)
        m1["dx_txy"][sl] = fp["b_x"][nl] * m0["dx_txy"][sl] + fp["c_x"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["k_x"][nl] + m2["dx_txy"][sl]
        ipml += nz2
    # dz_tyz: top j = nl, bottom j = nz2-1-nl, nl < np each
    ipml = 0
    for nl in range(npml):
        ind = np.arange(nx2) * nz2 + nl
        sl = slice(ipml, ipml + nx2)
        m1["dz_tyz"][sl] = fp["b_z"][nl] * m0["dz_tyz"][sl] + fp["c_z"][nl] * t2f[ind]
        t2f[ind] = t2f[ind] / fp["k_z"][nl] + m2["dz_tyz"][sl]
        ipml += nx2
    for nl in range(npml):
        ind = np.arange(nx2) * nz2 + nz2 - 1 - nl
        sl = slice(ipml, ipml + nx2)
        m1["dz_tyz"][sl] = fp["b_z"][nl] * m0["dz_tyz"][sl] + fp["c_z"][nl] * t2f[ind]
        t2f[ind] = t2f[ind] / fp["k_z"][nl] + m2["dz_tyz"][sl]
        ipml += nx2
    t1 = t1 + t2
    delta[2][1:-1, 1:-1] = mat["irho"][1:-1, 1:-1] * t1[1:-1, 1:-1]
    # D_z^+ v_y and its layers (half nodes on top, end half nodes below)
    t1 = deriv(vy, 1, mz_f)
    t1f = t1.reshape(-1)
    ipml = 0
    for nl in range(npml):
        ind = np.arange(nx2) * nz2 + nl
        sl = slice(ipml, ipml + nx2)
        m1["dz_vy"][sl] = fp["bh_z"][nl] * m0["dz_vy"][sl] + fp["ch_z"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["kh_z"][nl] + m2["dz_vy"][sl]
        ipml += nx2
    for nl in range(npml + 1):
        ind = np.arange(nx2) * nz2 + nz2 - 1 - nl
        sl = slice(ipml, ipml + nx2)
        m1["dz_vy"][sl] = fp["bH_z"][nl] * m0["dz_vy"][sl] + fp["cH_z"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["kH_z"][nl] + m2["dz_vy"][sl]
        ipml += nx2
    delta[3] = mat["ts1"] * (mat["eta1"] * t1 - e1)
    delta[1] = mat["c44"] * t1 + mat["c44_0"] * e1
    # D_x^+ v_y; its layers indexed like z layers, as written
    t1 = deriv(vy, 0, mx_f)
    t1f = t1.reshape(-1)
    ipml = 0
    for nl in range(npml):
        ind = np.arange(nx2) * nz2 + nl
        sl = slice(ipml, ipml + nx2)
        m1["dx_vy"][sl] = fp["bh_x"][nl] * m0["dx_vy"][sl] + fp["ch_x"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["kh_x"][nl] + m2["dx_vy"][sl]
        ipml += nx2
    for nl in range(npml + 1):
        ind = np.arange(nx2) * nz2 + nz2 - 1 - nl
        sl = slice(ipml, ipml + nx2)
        m1["dx_vy"][sl] = fp["bH_z"][nl] * m0["dx_vy"][sl] + fp["cH_z"][nl] * t1f[ind]
        t1f[ind] = t1f[ind] / fp["kH_z"][nl] + m2["dx_vy"][sl]
        ipml += nx2
    delta[4] = mat["ts2"] * (mat["eta2"] * t1 - e2)
    delta[0] = mat["c66"] * t1 + mat["c66_0"] * e2
    # 2451-2461
    n1 = (2 * npml + 1) * nz2
    m2["dx_txy"][:n1] = m1["dx_txy"][:n1]
    m2["dx_vy"][:n1] = m1["dx_vy"][:n1]
    n2 = (2 * npml + 1) * nx2
    m2["dz_tyz"][:n2] = m1["dz_tyz"][:n2]
    m2["dz_vy"][:n2] = m1["dz_vy"][:n2]


def add_src(W, Wt, src, dt_w, dt_t, it):
    """add_src_ve_sh_rk4 (src.c:145-168): sample 2 it + 1 of each source
    function (computed at dt/2) into tau_xy (Sxy) or tau_yz (Syz)."""
    for typ, idx, fct in src:
        k = 2 * it + 1
        if k >= len(fct):
            continue
        q = 0 if typ == SXY else 1 if typ == SYZ else None
        if q is None:
            continue
        i, j = divmod(int(idx), W.shape[2])
        if dt_t != 0.0:
            Wt[q, i, j] += dt_t * fct[k]
        W[q, i, j] += dt_w * fct[k]


def sh_run(cfg, mu, rho, gamma, Q, f, src, rec_idx, rec_comp, nsteps):
    """Run the restatement.  cfg: nx, nz, np, dx, dz, dt, vmax, pml_order,
    Rc, kappa_max, alpha_max (value of alpha_max, 0 for none), iwipe.
    src: list of (type, padded index, fct at dt/2 already normalised by
    1/(dx dz)).  Returns traces (nsteps x nrec) and the final W."""
    np_ = cfg["np"]
    nx2, nz2 = mu.shape
    mat = sh_material(mu, rho, gamma, Q, f)
    tabs = sh_cpml(np_, cfg["dx"], cfg["dz"], cfg["dt"], cfg.get("vmax", 5000.0),
                   cfg.get("pml_order", 2.0), cfg.get("Rc", 1e-5), cfg.get("kappa_max", 1.0),
                   cfg.get("alpha_max", 0.0), cfg.get("iwipe", 1))
    mx_f = k_multipliers(nx2, cfg["dx"], True); mx_b = k_multipliers(nx2, cfg["dx"], False)
    mz_f = k_multipliers(nz2, cfg["dz"], True); mz_b = k_multipliers(nz2, cfg["dz"], False)
    st = ShState(nx2, nz2, np_)
    W, delta, mem = st.W, st.delta, st.mem
    dt = cfg["dt"]
    traces = np.zeros((nsteps, len(rec_idx)))
    comp_q = {TXY: 0, TYZ: 1, VY: 2}
    for it in range(nsteps):
        Ws = W.copy()
        n1 = 2 * np_ * nz2
        for nm in ("dx_txy", "dx_vy"):
            mem[0][nm][:n1] = mem[1][nm][:n1]
        n2 = 2 * np_ * nx2
        for nm in ("dz_tyz", "dz_vy"):
            mem[0][nm][:n2] = mem[1][nm][:n2]
        Wt = Ws.copy()
        propagate(Wt, delta, mat, tabs[1], mem, np_, mx_f, mx_b, mz_f, mz_b)
        Wt = Ws + dt / 2. * delta
        W = Ws + dt / 6. * delta
        add_src(W, Wt, src, dt / 6., dt / 2., it)
        propagate(Wt, delta, mat, tabs[23], mem, np_, mx_f, mx_b, mz_f, mz_b)
        Wt = Ws + dt / 2. * delta
        W = W + dt / 3. * delta
        add_src(W, Wt, src, dt / 3., dt / 2., it)
        propagate(Wt, delta, mat, tabs[23], mem, np_, mx_f, mx_b, mz_f, mz_b)
        Wt = Ws + dt * delta
        W = W + dt / 3. * delta
        add_src(W, Wt, src, dt / 3., dt, it)
        propagate(Wt, delta, mat, tabs[4], mem, np_, mx_f, mx_b, mz_f, mz_b)
        W = W + dt / 6. * delta
        add_src(W, Wt, src, dt / 6., 0.0, it)
        for r, (idx, c) in enumerate(zip(rec_idx, rec_comp)):
            i, j = divmod(int(idx), nz2)
            traces[it, r] = W[comp_q[c], i, j]
    return traces, W
