"""Independent numpy formulations used to pin the oracle.

These are written from the physical description of the two schemes, not
from the oracle's loops:
  - every field lives on one padded grid, values off the grid are zero;
  - absorbing coefficients are computed from the distance of each
    (staggered) node position to the interior, in metres;
  - the 2D split PML is expressed as "split where the axis is inside the
    absorbing layer", so the eight regions of pml.c never appear.
A mistake in the oracle's region tables, loop bounds, mirror indices or
table assignment shows up as a mismatch against these.
"""
from __future__ import annotations

import numpy as np


def _shift(a, axis, step):
    """a[x+step] along axis with zeros off the grid."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    if step > 0:
        src[axis] = slice(step, n)
        dst[axis] = slice(0, n - step)
    else:
        src[axis] = slice(0, n + step)
        dst[axis] = slice(-step, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


# ---------------------------------------------------------------------------
# 3D acoustic CPML
# ---------------------------------------------------------------------------
def cpml_profile(n2, np_, d, dt, vmax=5000.0, order=2.0, Rc=1e-5, kappa_max=1.0,
                 alpha_max=0.0, half=False):
    """(a, b, kappa) of one axis from the distance to the interior nodes.

    Interior node positions are np .. n2-1-np (in cells); a node (or half
    node at n+1/2) at distance dist > 0 outside gets
      kappa = 1 + (kappa_max-1)(dist/L)^m,  d = d0 (dist/L)^m,
      alpha = alpha_max (1 - dist/L),  b = exp(-(d/kappa + alpha) dt),
      a = d (b-1) / (kappa (d + kappa alpha)),
    with L = np*d and d0 = -(m+1) vmax ln(Rc) / (2L) (Komatitsch & Martin
    2007); elsewhere a = 0, b = kappa = 1.
    """
    count = n2 - 1 if half else n2
    pos = np.arange(count) + (0.5 if half else 0.0)
    lo, hi = np_, n2 - 1 - np_
    dist = np.maximum(np.maximum(lo - pos, pos - hi), 0.0) * d
    a = np.zeros(count)
    b = np.ones(count)
    k = np.ones(count)
    if np_ == 0:
        return a, b, k
    L = np_ * d
    d0 = -(order + 1.0) * vmax * np.log(Rc) / (2.0 * L)
    m = dist > 0
    x = dist[m] / L
    k[m] = 1.0 + (kappa_max - 1.0) * x ** order
    dd = d0 * x ** order
    al = alpha_max * (1.0 - x)
    b[m] = np.exp(-(dd / k[m] + al) * dt)
    a[m] = dd * (b[m] - 1.0) / (k[m] * (dd + k[m] * al))
    return a, b, k


def a3d_numpy(cfg, vp, rho, src_idx, src_len, src_fct, rec_idx, rec_comp, nsteps):
    nx2, ny2, nz2 = vp.shape
    dx, dy, dz, dt = cfg["dx"], cfg["dy"], cfg["dz"], cfg["dt"]
    amax = np.pi * cfg["f0"] if cfg.get("alpha_max", 1) else 0.0
    kw = dict(vmax=cfg.get("vmax", 5000.0), order=cfg.get("pml_order", 2.0),
              Rc=cfg.get("Rc", 1e-5), kappa_max=cfg.get("kappa_max", 1.0), alpha_max=amax)
    tab = []
    for n2, d in ((nx2, dx), (ny2, dy), (nz2, dz)):
        tab.append((cpml_profile(n2, cfg["np"], d, dt, **kw),
                    cpml_profile(n2, cfg["np"], d, dt, half=True, **kw)))
    K = rho * vp * vp * dt
    P = np.zeros_like(vp)
    v = [np.zeros_like(vp) for _ in range(3)]
    mP = [np.zeros_like(vp) for _ in range(3)]
    mv = [np.zeros_like(vp) for _ in range(3)]
    h = (dx, dy, dz)
    flat_src = [np.unravel_index(int(s), vp.shape) for s in src_idx]
    offs = np.concatenate([[0], np.cumsum(src_len)[:-1]]).astype(int)
    traces = np.zeros((nsteps, len(rec_idx)))
    for it in range(nsteps):
        for ax in range(3):
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax] = slice(0, vp.shape[ax] - 1)
            hi[ax] = slice(1, vp.shape[ax])
            lo, hi = tuple(lo), tuple(hi)
            a_h, b_h, k_h = tab[ax][1]
            shape = [1, 1, 1]
            shape[ax] = -1
            a_h, b_h, k_h = (t.reshape(shape) for t in (a_h, b_h, k_h))
            dP = (P[hi] - P[lo]) / h[ax]
            mP[ax][lo] = b_h * mP[ax][lo] + a_h * dP
            dP = dP / k_h + mP[ax][lo]
            v[ax][lo] -= dt * 2.0 / (rho[lo] + rho[hi]) * dP
        inner = (slice(1, None),) * 3
        div = []
        for ax in range(3):
            prev = [slice(1, None)] * 3
            prev[ax] = slice(0, -1)
            a_n, b_n, k_n = tab[ax][0]
            shape = [1, 1, 1]
            shape[ax] = -1
            a_n, b_n, k_n = (t[1:].reshape(shape) for t in (a_n, b_n, k_n))
            dv = (v[ax][inner] - v[ax][tuple(prev)]) / h[ax]
            mv[ax][inner] = b_n * mv[ax][inner] + a_n * dv
            div.append(dv / k_n + mv[ax][inner])
        P[inner] -= K[inner] * (div[0] + div[1] + div[2])
        for s, (i, j, k) in enumerate(flat_src):
            if it < src_len[s]:
                P[i, j, k] += dt * src_fct[offs[s] + it] * 1.0e6
        fields = {0: v[0], 1: v[1], 2: v[2], 8: P}
        for r, (idx, comp) in enumerate(zip(rec_idx, rec_comp)):
            traces[it, r] = fields[int(comp)].reshape(-1)[int(idx)]
    return traces, np.concatenate([P.reshape(-1)] + [x.reshape(-1) for x in v])


# ---------------------------------------------------------------------------
# 2D elastic split PML
# ---------------------------------------------------------------------------
def e2d_numpy(cfg, vp, vs, rho, src_type, src_idx, src_len, src_fct, rec_idx, rec_comp,
              nsteps, damping_scale=1.0):
    """Padded-grid elastic scheme; split PML where an axis is absorbing.

    Damping of a component differentiated along an axis is
    d0 ((distance of its staggered position to the interior)/delta)^2 with
    d0 = 1.5 ln(1/R) mean(Vp)/delta, R = 1e-3, delta = Np h; positions of
    vx and txz are integer along x, of vz, txx, tzz half-integer; along z
    vx, txx, tzz are integer and vz, txz half-integer.  PML cells use the
    material of the nearest interior cell (the lower-right corner of the
    normal stresses may use cell nx*nx-1 instead).
    """
    nx, nz, Np, h, dt = cfg["nx"], cfg["nz"], cfg["npml"], cfg["h"], cfg["dt"]
    NX, NZ = nx + 2 * Np, nz + 2 * Np
    # Lame constants and staggered averages
    lam = rho * (vp * vp - 2.0 * vs * vs)
    mu0 = rho * vs * vs
    l2m0 = lam + 2.0 * mu0
    b0 = 1.0 / rho
    mu = mu0.copy()
    mu[:, :-1] = (mu0[:, :-1] + mu0[:, 1:]) / 2.0
    la = lam.copy()
    l2m = l2m0.copy()
    la[:-1, :] = (lam[:-1, :] + lam[1:, :]) / 2.0
    l2m[:-1, :] = (l2m0[:-1, :] + l2m0[1:, :]) / 2.0
    bm = b0.copy()
    bm[:-1, :-1] = (b0[:-1, :-1] + b0[1:, :-1] + b0[:-1, 1:] + b0[1:, 1:]) / 4.0
    bm[:-1, -1] = (b0[:-1, -1] + b0[1:, -1]) / 2.0
    bm[-1, :-1] = (b0[-1, :-1] + b0[-1, 1:]) / 2.0
    ci = np.clip(np.arange(NX) - Np, 0, nx - 1)
    cj = np.clip(np.arange(NZ) - Np, 0, nz - 1)
    pad = lambda m: m[np.ix_(ci, cj)]
    B, BM, LA, L2M, MU = pad(b0), pad(bm), pad(la), pad(l2m), pad(mu)
    if cfg.get("q1_compat", True) and nx * nx <= nx * nz:
        q = divmod(nx * nx - 1, nz)
        LA[Np + nx:, Np + nz:] = la[q]
        L2M[Np + nx:, Np + nz:] = l2m[q]
    # damping from distances
    delta = Np * h
    d0 = np.log(1.0 / 1e-3) * 1.5 * vp.mean() / delta * damping_scale

    def damp(n, count, half):
        pos = np.arange(count) - Np + (0.5 if half else 0.0)
        dist = np.maximum(np.maximum(-pos, pos - (n - 1)), 0.0) * h
        return d0 * (dist / delta) ** 2

    in_x = (np.arange(NX) < Np) | (np.arange(NX) >= Np + nx)
    in_z = (np.arange(NZ) < Np) | (np.arange(NZ) >= Np + nz)
    dxi, dxh = damp(nx, NX, False)[:, None], damp(nx, NX, True)[:, None]
    dzi, dzh = damp(nz, NZ, False)[None, :], damp(nz, NZ, True)[None, :]
    px = np.broadcast_to(in_x[:, None], (NX, NZ))
    pz = np.broadcast_to(in_z[None, :], (NX, NZ))
    pml = px | pz

    def split(u, diff, coef, dd, absorbing):
        c1 = (2.0 * dt) / (2.0 + dt * dd)
        c2 = (2.0 - dt * dd) / (2.0 * dt)
        damped = (u * c2 + coef * diff / h) * c1
        plain = u + dt * coef * diff / h
        return np.where(absorbing, damped, plain)

    F = {k: np.zeros((NX, NZ)) for k in ("vx", "vz", "txx", "tzz", "txz")}
    S = {k: np.zeros((NX, NZ)) for k in ("vx_x", "vx_z", "vz_x", "vz_z", "txx_x", "txx_z",
                                          "tzz_x", "tzz_z", "txz_x", "txz_z")}
    offs = np.concatenate([[0], np.cumsum(src_len)[:-1]]).astype(int)
    traces = np.zeros((nsteps, len(rec_idx)))
    for it in range(nsteps):
        txx, tzz, txz = F["txx"], F["tzz"], F["txz"]
        d_xx = txx - _shift(txx, 0, -1)
        d_xz = txz - _shift(txz, 1, -1)
        d_zx = _shift(txz, 0, 1) - txz
        d_zz = _shift(tzz, 1, 1) - tzz
        vx_int = F["vx"] + dt * B * (d_xx / h + d_xz / h)
        vz_int = F["vz"] + dt * BM * (d_zx / h + d_zz / h)
        S["vx_x"] = split(S["vx_x"], d_xx, B, dxi, px)
        S["vx_z"] = split(S["vx_z"], d_xz, B, dzi, pz)
        S["vz_x"] = split(S["vz_x"], d_zx, BM, dxh, px)
        S["vz_z"] = split(S["vz_z"], d_zz, BM, dzh, pz)
        F["vx"] = np.where(pml, S["vx_x"] + S["vx_z"], vx_int)
        F["vz"] = np.where(pml, S["vz_x"] + S["vz_z"], vz_int)
        vx, vz = F["vx"], F["vz"]
        e_xx = _shift(vx, 0, 1) - vx
        e_zz = vz - _shift(vz, 1, -1)
        e_xz = _shift(vx, 1, 1) - vx
        e_zx = vz - _shift(vz, 0, -1)
        ddx = dt * e_xx / h
        ddz = dt * e_zz / h
        txx_int = F["txx"] + (L2M * ddx + LA * ddz)
        tzz_int = F["tzz"] + (L2M * ddz + LA * ddx)
        txz_int = F["txz"] + MU * (dt * e_xz / h + dt * e_zx / h)
        S["txx_x"] = split(S["txx_x"], e_xx, L2M, dxh, px)
        S["tzz_x"] = split(S["tzz_x"], e_xx, LA, dxh, px)
        S["txx_z"] = split(S["txx_z"], e_zz, LA, dzi, pz)
        S["tzz_z"] = split(S["tzz_z"], e_zz, L2M, dzi, pz)
        S["txz_x"] = split(S["txz_x"], e_zx, MU, dxi, px)
        S["txz_z"] = split(S["txz_z"], e_xz, MU, dzh, pz)
        F["txx"] = np.where(pml, S["txx_x"] + S["txx_z"], txx_int)
        F["tzz"] = np.where(pml, S["tzz_x"] + S["tzz_z"], tzz_int)
        F["txz"] = np.where(pml, S["txz_x"] + S["txz_z"], txz_int)
        for s in range(len(src_idx)):
            if it < src_len[s]:
                i, j = divmod(int(src_idx[s]), nz)
                inc = dt * src_fct[offs[s] + it]
                if src_type[s] in (0, 6):
                    F["txx"][i + Np, j + Np] += inc
                if src_type[s] in (2, 6):
                    F["tzz"][i + Np, j + Np] += inc
                if src_type[s] == 4:
                    F["txz"][i + Np, j + Np] += inc
        names = {0: "vx", 2: "vz", 5: "txx", 6: "tzz", 7: "txz"}
        for r, (idx, comp) in enumerate(zip(rec_idx, rec_comp)):
            i, j = divmod(int(idx), nz)
            traces[it, r] = F[names[int(comp)]][i + Np, j + Np]
    fields = np.concatenate([F[k].reshape(-1) for k in ("vx", "vz", "txx", "tzz", "txz")])
    return traces, fields
