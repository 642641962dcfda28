// cyl.hip -- 2.5D elastic FDTD in cylindrical coordinates (e_cyl_iso):
// staggered velocity / stress in (r, z) for one azimuthal mode n, CPML on
// the outer r side and at both z ends, force and Kurkjian sources,
// receivers including divergence and curl.
//
// The reference (e_cyl_iso.c:601-866, propagate.c:2706-4119) runs seven
// update routines per step, each a set of region loops.  Here one thread
// owns one (r, z) cell of the padded nx2 x nz2 grid (nx2 = nx + np: no
// layer at the axis, nz2 = nz + 2 np) and runs every update that the
// reference applies to that cell, with the reference's expression and
// operand order for that region (FP64 storage reproduces it bit for bit;
// -ffp-contract=off keeps multiply and add separate).  Two launches per
// step: velocities, then force sources at their cells; stresses, then the
// Kurkjian source.  Receivers are gathered by a third small kernel.
//
// CPML memory variables exist only where the reference keeps them:
//   r layers: rows i >= nx-1, (np+1) rows x nz2, index (i-(nx-1))*nz2 + j;
//   z layers: 2np+1 slots per row, index i*(2np+1) + s, s = j for the top
//   rows j < np and s = np + (j - (nz+np-1)) for the bottom ones.
// Each slot belongs to one cell of one update, as in the reference's
// compact arrays (pml.c:1811-1912), so values are identical.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "sm_internal.hpp"

namespace sm {
namespace {

enum { C_VR, C_VT, C_VZ, C_TRR, C_TTT, C_TZZ, C_TRZ, C_TRT, C_TTZ, C_NF };
// r-layer memory variables
enum {
    R_DTRR_DR, R_TRR_R1, R_TRT_R1, R_TTT_R1, R_TRT_R2, R_DTRT_DR, R_TTT_R2, R_DTRZ_DR, R_TRZ_R,
    R_TTZ_R, R_DVR_DR, R_VR_R1, R_VT_R1, R_DVZ_DR, R_DVT_DR, R_VR_R2, R_VT_R2, R_VZ_R, R_NM
};
// z-layer memory variables
enum { Z_DTRZ_DZ, Z_DTTZ_DZ, Z_DTZZ_DZ, Z_DVZ_DZ, Z_DVR_DZ, Z_DVT_DZ, Z_NM };
// material arrays
enum { M_LIJ, M_L2MIJ, M_MU, M_MUI, M_MUJ, M_BI, M_BJ, M_BIJ, M_NMAT };
// CPML tables: 9 of np entries, then 6 of np+1 (pml.c:2107-2164)
enum { T_IK_R, T_IK_Z, T_IKH_Z, T_B_R, T_C_R, T_B_Z, T_C_Z, T_BH_Z, T_CH_Z };
enum { T_IKH_R = 0, T_IKHZ_END, T_BHR, T_CHR, T_BHZ, T_CHZ };

template <typename T>
struct CylDev {
    int nx, nz, np, nx2, nz2, n;
    T dx, dz, dt;
    T *f[C_NF];
    T *mr, *mz;            // R_NM x (np+1) x nz2, Z_NM x nx2 x (2np+1)
    const T *m;            // M_NMAT x N
    const T *tab;          // 9 np + 6 (np+1)
    // sources: cell, length, offset (int64 each), type (int), fct (T)
    int nsrc;
    const int64_t *src;
    const int *stype;
    const T *sfct;
    int64_t it;
};

template <typename T>
struct Coef {
    bool on;
    T b, c, ik;
};

template <typename T>
__device__ __forceinline__ const T *tabs(const CylDev<T> &d, int q)
{
    return d.tab + q * d.np;
}
template <typename T>
__device__ __forceinline__ const T *tabH(const CylDev<T> &d, int q)
{
    return d.tab + 9 * d.np + q * (d.np + 1);
}

// z coefficients at half nodes: bh_z[j] on top, bH_z[j - jbot] from jbot
template <typename T>
__device__ __forceinline__ Coef<T> zc_half(const CylDev<T> &d, int j, int jbot)
{
    if (j < d.np) return {true, tabs(d, T_BH_Z)[j], tabs(d, T_CH_Z)[j], tabs(d, T_IKH_Z)[j]};
    if (j >= jbot) {
        const int jj = j - jbot;
        return {true, tabH(d, T_BHZ)[jj], tabH(d, T_CHZ)[jj], tabH(d, T_IKHZ_END)[jj]};
    }
    return {false, T(0), T(0), T(1)};
}
// z coefficients at nodes: b_z[j] on top, b_z[np-1-(j-(nz+np))] at the bottom
template <typename T>
__device__ __forceinline__ Coef<T> zc_node(const CylDev<T> &d, int j)
{
    if (j < d.np) return {true, tabs(d, T_B_Z)[j], tabs(d, T_C_Z)[j], tabs(d, T_IK_Z)[j]};
    if (j >= d.nz + d.np) {
        const int q = d.np - 1 - (j - (d.nz + d.np));
        return {true, tabs(d, T_B_Z)[q], tabs(d, T_C_Z)[q], tabs(d, T_IK_Z)[q]};
    }
    return {false, T(0), T(0), T(1)};
}
// r coefficients of rows at r = i dr: b_r[np-1-(i-nx)] from i = nx
template <typename T>
__device__ __forceinline__ Coef<T> rc_node(const CylDev<T> &d, int i)
{
    if (i >= d.nx) {
        const int q = d.np - 1 - (i - d.nx);
        return {true, tabs(d, T_B_R)[q], tabs(d, T_C_R)[q], tabs(d, T_IK_R)[q]};
    }
    return {false, T(0), T(0), T(1)};
}
// r coefficients bH_r[i - i0] from row i0
template <typename T>
__device__ __forceinline__ Coef<T> rc_half(const CylDev<T> &d, int i, int i0)
{
    if (i >= i0) {
        const int ii = i - i0;
        return {true, tabH(d, T_BHR)[ii], tabH(d, T_CHR)[ii], tabH(d, T_IKH_R)[ii]};
    }
    return {false, T(0), T(0), T(1)};
}

// memory variable update and stretched derivative (pml.c / propagate.c form)
template <typename T>
__device__ __forceinline__ T cpml(T *m, const Coef<T> &k, T dv)
{
    const T v = k.b * *m + k.c * dv;
    *m = v;
    return k.ik * dv + v;
}

template <typename T>
__device__ __forceinline__ T *rmem(const CylDev<T> &d, int q, int i, int j)
{
    return d.mr + ((size_t)q * (d.np + 1) + (size_t)(i - (d.nx - 1))) * d.nz2 + j;
}
template <typename T>
__device__ __forceinline__ T *zmem(const CylDev<T> &d, int q, int i, int j, int jbot)
{
    const int s = j < d.np ? j : d.np + (j - jbot);
    return d.mz + ((size_t)q * d.nx2 + i) * (2 * d.np + 1) + s;
}

#define FLD(q, i, j) d.f[q][(size_t)(i) * d.nz2 + (j)]
#define MAT(q, c) d.m[(size_t)(q) * N + (c)]

template <typename T>
__global__ void __launch_bounds__(256) cyl_velocity(CylDev<T> d)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= d.nx2 || j >= d.nz2) return;
    const size_t N = (size_t)d.nx2 * d.nz2;
    const size_t c = (size_t)i * d.nz2 + j;
    const T dx = d.dx, dz = d.dz, dt = d.dt;
    const int n = d.n;
    const T dn = (T)n;
    const int jb_half = d.nz + d.np - 1;

    // ---- v_r (propagate.c:2706-2813, 3312-3477): rows 1.. (0 for odd n)
    if ((i >= 1 || (n % 2)) && j + 1 < d.nz2) {
        const Coef<T> kz = zc_half(d, j, jb_half);
        T ddz = (FLD(C_TRZ, i, j + 1) - FLD(C_TRZ, i, j)) / dz;
        if (kz.on) ddz = cpml(zmem(d, Z_DTRZ_DZ, i, j, jb_half), kz, ddz);
        if (i == 0) {
            FLD(C_VR, i, j) += dt * MAT(M_BJ, c) * ((FLD(C_TRR, i, j) - -FLD(C_TRR, i, j)) / dx + ddz);
        } else {
            const Coef<T> kr = rc_node(d, i);
            const T twoidx = T(2.) * (T)i * dx, idx_ = (T)i * dx;
            T ddr1 = (FLD(C_TRR, i, j) - FLD(C_TRR, i - 1, j)) / dx;
            T dr1 = (FLD(C_TRR, i, j) + FLD(C_TRR, i - 1, j)) / twoidx;
            T dr3 = (FLD(C_TTT, i, j) + FLD(C_TTT, i - 1, j)) / twoidx;
            if (n == 0) {
                if (kr.on) {
                    ddr1 = cpml(rmem(d, R_DTRR_DR, i, j), kr, ddr1);
                    dr1 = cpml(rmem(d, R_TRR_R1, i, j), kr, dr1);
                    dr3 = cpml(rmem(d, R_TTT_R1, i, j), kr, dr3);
                }
                FLD(C_VR, i, j) += dt * MAT(M_BJ, c) * (ddr1 + dr1 - dr3 + ddz);
            } else if (kr.on) {
                ddr1 = cpml(rmem(d, R_DTRR_DR, i, j), kr, ddr1);
                dr1 = cpml(rmem(d, R_TRR_R1, i, j), kr, dr1);
                T dr2 = FLD(C_TRT, i, j) / idx_;
                dr2 = cpml(rmem(d, R_TRT_R1, i, j), kr, dr2);
                dr3 = cpml(rmem(d, R_TTT_R1, i, j), kr, dr3);
                FLD(C_VR, i, j) += dt * MAT(M_BJ, c) * (ddr1 + dr1 + dn * dr2 - dr3 + ddz);
            } else {
                FLD(C_VR, i, j) +=
                    dt * MAT(M_BJ, c) * (ddr1 + dr1 + dn * FLD(C_TRT, i, j) / idx_ - dr3 + ddz);
            }
        }
    }
    if (i + 1 < d.nx2) {
        const T ihdx = ((T)i + T(0.5)) * dx;
        // ---- v_theta (propagate.c:2815-2906, 3479-3585)
        if (j + 1 < d.nz2) {
            const Coef<T> kr = rc_half(d, i, d.nx - 1);
            const Coef<T> kz = zc_half(d, j, jb_half);
            T ddz = (FLD(C_TTZ, i, j + 1) - FLD(C_TTZ, i, j)) / dz;
            if (kz.on) ddz = cpml(zmem(d, Z_DTTZ_DZ, i, j, jb_half), kz, ddz);
            T dr1 = (FLD(C_TRT, i + 1, j) + FLD(C_TRT, i, j)) / ihdx;
            T ddr = (FLD(C_TRT, i + 1, j) - FLD(C_TRT, i, j)) / dx;
            if (kr.on) {
                dr1 = cpml(rmem(d, R_TRT_R2, i, j), kr, dr1);
                ddr = cpml(rmem(d, R_DTRT_DR, i, j), kr, ddr);
            }
            if (n == 0) {
                FLD(C_VT, i, j) += dt * MAT(M_BIJ, c) * (dr1 + ddr + ddz);
            } else if (kr.on) {
                T dr2 = FLD(C_TTT, i, j) / ihdx;
                dr2 = cpml(rmem(d, R_TTT_R2, i, j), kr, dr2);
                FLD(C_VT, i, j) += dt * MAT(M_BIJ, c) * (dr1 + ddr - dn * dr2 + ddz);
            } else {
                FLD(C_VT, i, j) += dt * MAT(M_BIJ, c) * (dr1 + ddr - dn * FLD(C_TTT, i, j) / ihdx + ddz);
            }
        }
        // ---- v_z (propagate.c:2908-2997, 3587-3693)
        if (j >= 1) {
            const Coef<T> kr = rc_half(d, i, d.nx - 1);
            const Coef<T> kz = zc_node(d, j);
            const T two_ihdx = T(2.) * ((T)i + T(0.5)) * dx;
            T ddz = (FLD(C_TZZ, i, j) - FLD(C_TZZ, i, j - 1)) / dz;
            if (kz.on) ddz = cpml(zmem(d, Z_DTZZ_DZ, i, j, d.nz + d.np), kz, ddz);
            T ddr = (FLD(C_TRZ, i + 1, j) - FLD(C_TRZ, i, j)) / dx;
            T dr1 = (FLD(C_TRZ, i + 1, j) + FLD(C_TRZ, i, j)) / two_ihdx;
            if (kr.on) {
                ddr = cpml(rmem(d, R_DTRZ_DR, i, j), kr, ddr);
                dr1 = cpml(rmem(d, R_TRZ_R, i, j), kr, dr1);
            }
            if (n == 0) {
                FLD(C_VZ, i, j) += dt * MAT(M_BI, c) * (ddz + ddr + dr1);
            } else if (kr.on) {
                T dr2 = FLD(C_TTZ, i, j) / ihdx;
                dr2 = cpml(rmem(d, R_TTZ_R, i, j), kr, dr2);
                FLD(C_VZ, i, j) += dt * MAT(M_BI, c) * (ddz + ddr + dr1 + dn * dr2);
            } else {
                FLD(C_VZ, i, j) += dt * MAT(M_BI, c) * (ddz + ddr + dr1 + dn * FLD(C_TTZ, i, j) / ihdx);
            }
        }
    }
    // ---- force sources at this cell, in source order (src.c:186-200)
    for (int s = 0; s < d.nsrc; ++s) {
        if ((size_t)d.src[s] != c || d.it >= d.src[d.nsrc + s]) continue;
        const T a = d.sfct[d.src[2 * d.nsrc + s] + d.it];
        const int ty = d.stype[s];
        if (ty == SM_FR || ty == SM_BULK) FLD(C_VR, i, j) += dt * MAT(M_BJ, c) * a;
        if (ty == SM_FT) FLD(C_VT, i, j) += dt * MAT(M_BIJ, c) * a;
        if (ty == SM_FZ || ty == SM_BULK) FLD(C_VZ, i, j) += dt * MAT(M_BI, c) * a;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) cyl_stress(CylDev<T> d)
{
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int i = blockIdx.y * 4 + threadIdx.y;
    if (i >= d.nx2 || j >= d.nz2) return;
    const size_t N = (size_t)d.nx2 * d.nz2;
    const size_t c = (size_t)i * d.nz2 + j;
    const T dx = d.dx, dz = d.dz, dt = d.dt;
    const int n = d.n;
    const T dn = (T)n;
    const int jb_half = d.nz + d.np - 1;

    if (i + 1 < d.nx2) {
        const T ihdx = ((T)i + T(0.5)) * dx, two_ihdx = T(2.) * ((T)i + T(0.5)) * dx;
        const Coef<T> kr = rc_half(d, i, d.nx - 1);
        // ---- tau_rr, tau_thetatheta, tau_zz (propagate.c:2999-3134, 3695-3845)
        if (j + 1 < d.nz2) {
            const Coef<T> kz = zc_half(d, j, jb_half);
            T dvrdr = (FLD(C_VR, i + 1, j) - FLD(C_VR, i, j)) / dx;
            T vrr = (FLD(C_VR, i + 1, j) + FLD(C_VR, i, j)) / two_ihdx;
            T vtr = T(0);
            if (n != 0) vtr = FLD(C_VT, i, j) / ihdx;
            T dvzdz = (FLD(C_VZ, i, j + 1) - FLD(C_VZ, i, j)) / dz;
            if (kr.on) {
                dvrdr = cpml(rmem(d, R_DVR_DR, i, j), kr, dvrdr);
                vrr = cpml(rmem(d, R_VR_R1, i, j), kr, vrr);
                if (n != 0) vtr = cpml(rmem(d, R_VT_R1, i, j), kr, vtr);
            }
            if (kz.on) dvzdz = cpml(zmem(d, Z_DVZ_DZ, i, j, jb_half), kz, dvzdz);
            const T l2m = dt * MAT(M_L2MIJ, c), la = dt * MAT(M_LIJ, c);
            if (n == 0) {
                FLD(C_TRR, i, j) += l2m * dvrdr + la * (vrr + dvzdz);
                FLD(C_TTT, i, j) += l2m * vrr + la * (dvrdr + dvzdz);
                FLD(C_TZZ, i, j) += l2m * dvzdz + la * (dvrdr + vrr);
            } else {
                FLD(C_TRR, i, j) += l2m * dvrdr + la * (vrr + dn * vtr + dvzdz);
                FLD(C_TTT, i, j) += l2m * (vrr + dn * vtr) + la * (dvrdr + dvzdz);
                FLD(C_TZZ, i, j) += l2m * dvzdz + la * (dvrdr + vrr + dn * vtr);
            }
        }
        // ---- tau_thetaz (propagate.c:3249-3310, 4010-4088)
        if (j >= 1) {
            const Coef<T> kz = zc_node(d, j);
            const T dvt = FLD(C_VT, i, j) - FLD(C_VT, i, j - 1);
            if (n == 0) {
                if (kz.on) {
                    const T ddz = cpml(zmem(d, Z_DVT_DZ, i, j, d.nz + d.np), kz, dvt / dz);
                    FLD(C_TTZ, i, j) += dt * MAT(M_MUI, c) * ddz;
                } else {
                    FLD(C_TTZ, i, j) += dt * MAT(M_MUI, c) * dvt / dz;
                }
            } else {
                T ddz = dvt / dz;
                if (kz.on) ddz = cpml(zmem(d, Z_DVT_DZ, i, j, d.nz + d.np), kz, ddz);
                if (kr.on) {
                    T vzr = FLD(C_VZ, i, j) / ihdx;
                    vzr = cpml(rmem(d, R_VZ_R, i, j), kr, vzr);
                    FLD(C_TTZ, i, j) += dt * MAT(M_MUI, c) * (ddz - dn * vzr);
                } else {
                    FLD(C_TTZ, i, j) += dt * MAT(M_MUI, c) * (ddz - dn * FLD(C_VZ, i, j) / ihdx);
                }
            }
        }
    }
    // ---- tau_rz, both modes (propagate.c:3136-3213)
    if (i >= 1 && j >= 1) {
        const Coef<T> kr = rc_half(d, i, d.nx);
        const Coef<T> kz = zc_half(d, j, d.nz + d.np);
        T ddz = (FLD(C_VR, i, j) - FLD(C_VR, i, j - 1)) / dz;
        T ddr = (FLD(C_VZ, i, j) - FLD(C_VZ, i - 1, j)) / dx;
        if (kz.on) ddz = cpml(zmem(d, Z_DVR_DZ, i, j, d.nz + d.np), kz, ddz);
        if (kr.on) ddr = cpml(rmem(d, R_DVZ_DR, i, j), kr, ddr);
        FLD(C_TRZ, i, j) += dt * MAT(M_MU, c) * (ddz + ddr);
    }
    // ---- tau_rtheta (propagate.c:3215-3247, 3958-4008)
    if (i >= 1 || (n % 2)) {
        if (i == 0) {
            FLD(C_TRT, i, j) += dt * MAT(M_MUJ, c) * ((FLD(C_VT, i, j) - -FLD(C_VT, i, j)) / dx);
        } else {
            const Coef<T> kr = rc_node(d, i);
            const T twoidx = T(2.) * (T)i * dx, idx_ = (T)i * dx;
            T ddr = (FLD(C_VT, i, j) - FLD(C_VT, i - 1, j)) / dx;
            T vtr = (FLD(C_VT, i, j) + FLD(C_VT, i - 1, j)) / twoidx;
            if (n == 0) {
                if (kr.on) {
                    ddr = cpml(rmem(d, R_DVT_DR, i, j), kr, ddr);
                    vtr = cpml(rmem(d, R_VT_R2, i, j), kr, vtr);
                }
                FLD(C_TRT, i, j) += dt * MAT(M_MUJ, c) * (ddr - vtr);
            } else if (kr.on) {
                ddr = cpml(rmem(d, R_DVT_DR, i, j), kr, ddr);
                T vrr = FLD(C_VR, i, j) / idx_;
                vrr = cpml(rmem(d, R_VR_R2, i, j), kr, vrr);
                vtr = cpml(rmem(d, R_VT_R2, i, j), kr, vtr);
                FLD(C_TRT, i, j) += dt * MAT(M_MUJ, c) * (ddr - dn * vrr - vtr);
            } else {
                FLD(C_TRT, i, j) += dt * MAT(M_MUJ, c) * (ddr - dn * FLD(C_VR, i, j) / idx_ - vtr);
            }
        }
    }
    // ---- Kurkjian source, entry 0 (src.c:231-238)
    if (d.nsrc && d.stype[0] == SM_KURKJIAN && (size_t)d.src[0] == c && d.it < d.src[d.nsrc]) {
        const T a = d.sfct[d.src[2 * d.nsrc] + d.it];
        FLD(C_TRR, i, j) += dt * a;
        FLD(C_TZZ, i, j) += dt * a;
    }
}

// receiver values after step it (divergence and curl evaluated at the
// receiver's cell with propagate.c:4090-4119's expressions; cells outside
// their ranges read 0 as the reference's zero-initialised arrays do)
template <typename T>
__global__ void cyl_gather(CylDev<T> d, int nrec, const int64_t *ridx, const int *rcomp, T *trace,
                           int64_t row)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    const int64_t c = ridx[r];
    const int i = (int)(c / d.nz2), j = (int)(c % d.nz2);
    T v = T(0);
    switch (rcomp[r]) {
    case SM_VR: v = d.f[C_VR][c]; break;
    case SM_VT: v = d.f[C_VT][c]; break;
    case SM_VZ: v = d.f[C_VZ][c]; break;
    case SM_TRR: v = d.f[C_TRR][c]; break;
    case SM_TRT: v = d.f[C_TRT][c]; break;
    case SM_TRZ: v = d.f[C_TRZ][c]; break;
    case SM_TTT: v = d.f[C_TTT][c]; break;
    case SM_TTZ: v = d.f[C_TTZ][c]; break;
    case SM_TZZ: v = d.f[C_TZZ][c]; break;
    case SM_DIV:
        if (i + 1 < d.nx2 && j + 1 < d.nz2)
            v = (T(i + 1) * FLD(C_VR, i + 1, j) - T(i) * FLD(C_VR, i, j)) / ((T)i + T(0.5)) +
                T(d.n) * FLD(C_VT, i, j) / (((T)i + T(0.5)) * d.dx) +
                (FLD(C_VZ, i, j + 1) - FLD(C_VZ, i, j)) / d.dz;
        break;
    case SM_CURL:
        if (j >= 1 && (i >= 1 || (d.n % 2)))
            v = (FLD(C_VR, i, j) - FLD(C_VR, i, j - 1)) / d.dz -
                (i == 0 ? (FLD(C_VZ, i, j) - -FLD(C_VZ, i, j)) : (FLD(C_VZ, i, j) - FLD(C_VZ, i - 1, j))) /
                    d.dx;
        break;
    default: break;
    }
    trace[(size_t)row * nrec + r] = v;
}

#undef FLD
#undef MAT

}  // namespace
}  // namespace sm

using namespace sm;

struct sm_cyl {
    sm_ctx *ctx = nullptr;
    sm_cyl_params p{};
    size_t nx2 = 0, nz2 = 0, N = 0, elt = 8;
    sm_dbuf fields, mem, mat, tab, src, rcv, trace;
    size_t nsrc = 0, nrec = 0, max_steps = 0, rcv_step0 = 0, step = 0;
    bool timing = false;            // HIP events around each step's two launches
    double timed_ms = 0.0;
    size_t timed_steps = 0;
};

namespace {

void release(sm_cyl *s)
{
    for (sm_dbuf *b : {&s->fields, &s->mem, &s->mat, &s->tab, &s->src, &s->rcv, &s->trace}) dfree(*b);
}

size_t mem_count(size_t np, size_t nx2, size_t nz2)
{
    return np ? (size_t)R_NM * (np + 1) * nz2 + (size_t)Z_NM * nx2 * (2 * np + 1) : 0;
}

template <typename T>
CylDev<T> view(sm_cyl *s, int64_t it)
{
    CylDev<T> d;
    d.nx = (int)s->p.nx; d.nz = (int)s->p.nz; d.np = (int)s->p.np;
    d.nx2 = (int)s->nx2; d.nz2 = (int)s->nz2; d.n = s->p.mode;
    d.dx = (T)s->p.dr; d.dz = (T)s->p.dz; d.dt = (T)s->p.dt;
    T *f = (T *)s->fields.ptr;
    for (int q = 0; q < C_NF; ++q) d.f[q] = f + (size_t)q * s->N;
    d.mr = (T *)s->mem.ptr;
    d.mz = d.mr ? d.mr + (size_t)R_NM * (s->p.np + 1) * s->nz2 : nullptr;
    d.m = (const T *)s->mat.ptr;
    d.tab = (const T *)s->tab.ptr;
    d.nsrc = (int)s->nsrc;
    d.src = (const int64_t *)s->src.ptr;
    d.stype = d.src ? (const int *)(d.src + 3 * s->nsrc) : nullptr;
    d.sfct = d.src ? (const T *)(d.src + 4 * s->nsrc) : nullptr;
    d.it = it;
    return d;
}

template <typename T>
int upload(void *dst, const double *src, size_t n, hipStream_t st)
{
    std::vector<T> tmp(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
    SM_HIP(hipMemcpyAsync(dst, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    return SM_OK;
}

// events of a timed run, destroyed on every exit path
struct StepEvents {
    std::vector<hipEvent_t> ev;
    ~StepEvents()
    {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

template <typename T>
int run(sm_cyl *s, size_t nsteps)
{
    hipStream_t st = s->ctx->stream;
    const dim3 blk(64, 4, 1);
    const dim3 grd((unsigned)((s->nz2 + 63) / 64), (unsigned)((s->nx2 + 3) / 4), 1);
    const int64_t *ridx = (const int64_t *)s->rcv.ptr;
    const int *rcomp = ridx ? (const int *)(ridx + s->nrec) : nullptr;
    StepEvents te;
    if (s->timing) {
        te.ev.assign(2 * nsteps, nullptr);
        for (auto &e : te.ev) SM_HIP(hipEventCreate(&e));
    }
    for (size_t k = 0; k < nsteps; ++k) {
        const size_t it = s->step + k;
        CylDev<T> d = view<T>(s, (int64_t)it);
        if (s->timing) SM_HIP(hipEventRecord(te.ev[2 * k], st));
        cyl_velocity<T><<<grd, blk, 0, st>>>(d);
        cyl_stress<T><<<grd, blk, 0, st>>>(d);
        if (s->timing) SM_HIP(hipEventRecord(te.ev[2 * k + 1], st));
        if (s->nrec && it >= s->rcv_step0 && it - s->rcv_step0 < s->max_steps)
            cyl_gather<T><<<(unsigned)((s->nrec + 255) / 256), 256, 0, st>>>(
                d, (int)s->nrec, ridx, rcomp, (T *)s->trace.ptr, (int64_t)(it - s->rcv_step0));
    }
    SM_HIP(hipGetLastError());
    s->step += nsteps;
    if (s->timing) {
        SM_HIP(hipStreamSynchronize(st));
        for (size_t k = 0; k < nsteps; ++k) {
            float ms = 0.f;
            SM_HIP(hipEventElapsedTime(&ms, te.ev[2 * k], te.ev[2 * k + 1]));
            s->timed_ms += ms;
            s->timed_steps += 1;
        }
    }
    return SM_OK;
}

template <typename T>
int read_back(sm_cyl *s, const void *dev, size_t n, double *out)
{
    std::vector<T> tmp(n);
    SM_HIP(hipMemcpyAsync(tmp.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    for (size_t i = 0; i < n; ++i) out[i] = (double)tmp[i];
    return SM_OK;
}

int field_slot(int comp)
{
    switch (comp) {
    case SM_VR: return C_VR;
    case SM_VT: return C_VT;
    case SM_VZ: return C_VZ;
    case SM_TRR: return C_TRR;
    case SM_TTT: return C_TTT;
    case SM_TZZ: return C_TZZ;
    case SM_TRZ: return C_TRZ;
    case SM_TRT: return C_TRT;
    case SM_TTZ: return C_TTZ;
    default: return -1;
    }
}

bool recordable(int comp) { return field_slot(comp) >= 0 || comp == SM_DIV || comp == SM_CURL; }

}  // namespace

extern "C" {

size_t sm_cyl_bytes_needed(const sm_cyl_params *p, size_t nrec, size_t max_steps)
{
    if (!p) return 0;
    const size_t e = p->storage == SM_STORE_F64 ? 8 : 4;
    const size_t nx2 = p->nx + p->np, nz2 = p->nz + 2 * p->np, N = nx2 * nz2;
    return e * ((C_NF + M_NMAT) * N + mem_count(p->np, nx2, nz2) + 15 * (p->np + 1) +
                nrec * max_steps) + 12 * nrec;
}

int sm_cyl_create(sm_ctx *ctx, const sm_cyl_params *pp, const double *vp, const double *vs,
                  const double *rho, sm_cyl **out)
{
    SM_CHECK(ctx && pp && vp && vs && rho && out, SM_ERR_ARG, "sm_cyl_create: NULL argument");
    const sm_cyl_params &p = *pp;
    SM_CHECK(p.nx >= 2 && p.nz >= 2, SM_ERR_ARG, "sm_cyl_create: grid must be at least 2x2");
    SM_CHECK(p.dr > 0 && p.dz > 0 && p.dt > 0, SM_ERR_ARG, "sm_cyl_create: dr, dz and dt must be > 0");
    SM_CHECK(p.mode >= 0 && p.mode <= 1000, SM_ERR_ARG, "sm_cyl_create: azimuthal mode %d", p.mode);
    SM_CHECK(p.storage == SM_STORE_F32 || p.storage == SM_STORE_F64, SM_ERR_ARG,
             "sm_cyl_create: storage %d not supported", p.storage);
    SM_CHECK(p.flags == 0, SM_ERR_ARG, "sm_cyl_create: unknown flags %#x", p.flags);
    const size_t nx2 = p.nx + p.np, nz2 = p.nz + 2 * p.np, N = nx2 * nz2;
    SM_CHECK(nx2 < (1u << 20) && nz2 < (1u << 20) && N < (1ull << 31), SM_ERR_ARG,
             "sm_cyl_create: grid too large");
    // ---- material (e_cyl_iso.c:342-415), FP64 on the host
    std::vector<double> m((size_t)M_NMAT * N), mu_c(N);
    double *lij = &m[M_LIJ * N], *l2m = &m[M_L2MIJ * N], *mu = &m[M_MU * N], *mui = &m[M_MUI * N],
           *muj = &m[M_MUJ * N], *bi = &m[M_BI * N], *bj = &m[M_BJ * N], *bij = &m[M_BIJ * N];
    double vpmax = 0.0;
    for (size_t c = 0; c < N; ++c) {
        vpmax = vp[c] > vpmax ? vp[c] : vpmax;
        lij[c] = rho[c] * (vp[c] * vp[c] - 2. * vs[c] * vs[c]);
        mu_c[c] = rho[c] * vs[c] * vs[c];
        l2m[c] = lij[c] + 2.0 * mu_c[c];
        bij[c] = 1. / rho[c];
    }
    auto hm2 = [](double a, double b) { return (a == 0.0 && b == 0.0) ? 0.0 : 2. * a * b / (a + b); };
    for (size_t i = 0; i < nx2; ++i) {
        const size_t r = i * nz2;
        muj[r] = mu_c[r];
        bj[r] = bij[r];
        for (size_t j = 1; j < nz2; ++j) {
            muj[r + j] = hm2(mu_c[r + j], mu_c[r + j - 1]);
            bj[r + j] = 1. / ((rho[r + j] + rho[r + j - 1]) / 2.0);
        }
    }
    for (size_t j = 0; j < nz2; ++j) {
        mui[j] = mu_c[j];
        bi[j] = 1. / rho[j];
        for (size_t i = 1; i < nx2; ++i) {
            mui[i * nz2 + j] = hm2(mu_c[i * nz2 + j], mu_c[(i - 1) * nz2 + j]);
            bi[i * nz2 + j] = 1. / ((rho[i * nz2 + j] + rho[(i - 1) * nz2 + j]) / 2.0);
        }
    }
    mu[0] = mu_c[0];
    for (size_t j = 1; j < nz2; ++j) mu[j] = hm2(mu_c[j], mu_c[j - 1]);
    for (size_t i = 1; i < nx2; ++i) {
        mu[i * nz2] = hm2(mu_c[i * nz2], mu_c[(i - 1) * nz2]);
        for (size_t j = 1; j < nz2; ++j) {
            const double a = mu_c[i * nz2 + j], b = mu_c[i * nz2 + j - 1], c = mu_c[(i - 1) * nz2 + j],
                         e = mu_c[(i - 1) * nz2 + j - 1];
            if (b * c * e == 0.0 && a * c * e == 0.0 && a * b * e == 0.0 && a * b * c == 0.0)
                mu[i * nz2 + j] = 0.0;
            else
                mu[i * nz2 + j] = 4. * a * b * c * e / (b * c * e + a * c * e + a * b * e + a * b * c);
        }
    }
    for (size_t c = 0; c < N; ++c)
        for (int q = 0; q < M_NMAT; ++q)
            SM_CHECK(std::isfinite(m[(size_t)q * N + c]), SM_ERR_ARG,
                     "sm_cyl_create: material value %d not finite at node %zu (e_cyl_iso.c:417-452)",
                     q, c);
    // ---- CFL, e_cyl_iso.c:454-458
    SM_CHECK(!(p.dt > 0.5 * (p.dr + p.dz) / (std::sqrt(2.0) * vpmax)), SM_ERR_CFL,
             "sm_cyl_create: time step %g s too large (Vp max %g m/s)", p.dt, vpmax);
    // ---- CPML tables, pml.c:2107-2164
    const size_t np = p.np;
    std::vector<double> t(9 * np + 6 * (np + 1), 0.0);
    if (np) {
        const double pi = 4.0 * std::atan(1.0);
        const double amax = p.alpha_max == 0 ? 0.0 : pi * p.f0;   // e_cyl_iso.c:532-533
        double *ik_r = &t[0], *ik_z = &t[np], *ikh_z = &t[2 * np], *b_r = &t[3 * np], *c_r = &t[4 * np],
               *b_z = &t[5 * np], *c_z = &t[6 * np], *bh_z = &t[7 * np], *ch_z = &t[8 * np];
        double *H = &t[9 * np];
        double *ikH_r = H, *ikH_z = H + (np + 1), *bH_r = H + 2 * (np + 1), *cH_r = H + 3 * (np + 1),
               *bH_z = H + 4 * (np + 1), *cH_z = H + 5 * (np + 1);
        const double Lx = np * p.dr, Lz = np * p.dz, o = p.pml_order;
        const double d0x = -(o + 1.) * p.vmax * std::log(p.Rc) / (2. * Lx);
        const double d0z = -(o + 1.) * p.vmax * std::log(p.Rc) / (2. * Lz);
        const double da = amax / np, da2 = amax / (np + 1);
        for (size_t k = 0; k < np; ++k) {
            const double k_r = 1. + (p.kappa_max - 1.) * std::pow(((np - k) * p.dr / Lx), o);
            const double k_z = 1. + (p.kappa_max - 1.) * std::pow(((np - k) * p.dz / Lz), o);
            ik_r[k] = 1. / k_r;
            ik_z[k] = 1. / k_z;
            const double kh_z = 1. + (p.kappa_max - 1.) * std::pow(((np - k - 0.5) * p.dr / Lx), o);
            ikh_z[k] = 1. / kh_z;
            const double d_x = d0x * std::pow(((np - k) * p.dr / Lx), o), al_x = k * da;
            b_r[k] = std::exp(-(d_x / k_r + al_x) * p.dt);
            c_r[k] = d_x * (b_r[k] - 1.) / (k_r * (d_x + k_r * al_x));
            const double d_z = d0z * std::pow(((np - k) * p.dz / Lz), o), al_z = k * da;
            const double dh_z = d0z * std::pow(((np - k - 0.5) * p.dz / Lz), o), alh_z = (k + 0.5) * da;
            b_z[k] = std::exp(-(d_z / k_z + al_z) * p.dt);
            c_z[k] = d_z * (b_z[k] - 1.) / (k_z * (d_z + k_z * al_z));
            bh_z[k] = std::exp(-(dh_z / kh_z + alh_z) * p.dt);
            ch_z[k] = dh_z * (bh_z[k] - 1.) / (kh_z * (dh_z + kh_z * alh_z));
        }
        for (size_t k = 0; k < np + 1; ++k) {
            const double kH_r = 1. + (p.kappa_max - 1.) * std::pow(((k + 0.5) * p.dr / Lx), o);
            const double kH_z = 1. + (p.kappa_max - 1.) * std::pow(((k + 0.5) * p.dz / Lz), o);
            ikH_r[k] = 1. / kH_r;
            ikH_z[k] = 1. / kH_z;
            const double dH_x = d0x * std::pow(((k + 0.5) * p.dr / Lx), o), aH_x = (np - k - 0.5) * da2;
            bH_r[k] = std::exp(-(dH_x / kH_r + aH_x) * p.dt);
            cH_r[k] = dH_x * (bH_r[k] - 1.) / (kH_r * (dH_x + kH_r * aH_x));
            const double dH_z = d0z * std::pow(((k + 0.5) * p.dz / Lz), o), aH_z = (np - k - 0.5) * da2;
            bH_z[k] = std::exp(-(dH_z / kH_z + aH_z) * p.dt);
            cH_z[k] = dH_z * (bH_z[k] - 1.) / (kH_z * (dH_z + kH_z * aH_z));
        }
    }
    SM_CHECK(set_device(ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    sm_cyl *s = new sm_cyl;
    s->ctx = ctx;
    s->p = p;
    s->nx2 = nx2; s->nz2 = nz2; s->N = N;
    s->elt = p.storage == SM_STORE_F64 ? 8 : 4;
    int rc = dalloc(s->fields, (size_t)C_NF * N * s->elt);
    if (rc == SM_OK) rc = dalloc(s->mem, mem_count(np, nx2, nz2) * s->elt);
    if (rc == SM_OK) rc = dalloc(s->mat, (size_t)M_NMAT * N * s->elt);
    if (rc == SM_OK) rc = dalloc(s->tab, std::max<size_t>(t.size(), 1) * s->elt);
    if (rc == SM_OK)
        rc = s->elt == 8 ? upload<double>(s->mat.ptr, m.data(), m.size(), ctx->stream)
                         : upload<float>(s->mat.ptr, m.data(), m.size(), ctx->stream);
    if (rc == SM_OK && !t.empty())
        rc = s->elt == 8 ? upload<double>(s->tab.ptr, t.data(), t.size(), ctx->stream)
                         : upload<float>(s->tab.ptr, t.data(), t.size(), ctx->stream);
    if (rc == SM_OK) rc = sm_cyl_reset(s);
    if (rc != SM_OK) {
        release(s);
        delete s;
        return rc;
    }
    *out = s;
    return SM_OK;
}

int sm_cyl_destroy(sm_cyl *s)
{
    if (!s) return SM_OK;
    (void)set_device(s->ctx);
    (void)hipStreamSynchronize(s->ctx->stream);
    release(s);
    delete s;
    return SM_OK;
}

int sm_cyl_set_sources(sm_cyl *s, size_t nsrc, const int *type, const size_t *idx,
                       const size_t *length, const double *fct)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_cyl_set_sources: NULL solver");
    SM_CHECK(nsrc == 0 || (type && idx && length && fct), SM_ERR_ARG,
             "sm_cyl_set_sources: NULL array");
    SM_CHECK(nsrc < (1u << 16), SM_ERR_ARG, "sm_cyl_set_sources: too many sources");
    size_t total = 0;
    for (size_t k = 0; k < nsrc; ++k) {
        SM_CHECK(idx[k] < s->N, SM_ERR_ARG, "source %zu: index %zu outside the %zux%zu padded grid", k,
                 idx[k], s->nx2, s->nz2);
        SM_CHECK(type[k] >= 0 && type[k] <= SM_KURKJIAN, SM_ERR_ARG,
                 "source %zu: type %d is not a source type", k, type[k]);
        SM_CHECK(type[k] != SM_KURKJIAN || k == 0, SM_ERR_ARG,
                 "source %zu: a Kurkjian source must be the first (and is the only one injected, "
                 "src.c:231-238)", k);
        total += length[k];
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->src);
    s->nsrc = 0;
    if (nsrc == 0) return SM_OK;
    const size_t bytes = 4 * nsrc * sizeof(int64_t) + total * s->elt;
    int rc = dalloc(s->src, bytes);
    if (rc != SM_OK) return rc;
    std::vector<unsigned char> hb(bytes, 0);
    int64_t *hi = (int64_t *)hb.data();
    int *ht = (int *)(hi + 3 * nsrc);
    size_t off = 0;
    for (size_t k = 0; k < nsrc; ++k) {
        hi[k] = (int64_t)idx[k];
        hi[nsrc + k] = (int64_t)length[k];
        hi[2 * nsrc + k] = (int64_t)off;
        ht[k] = type[k];
        off += length[k];
    }
    for (size_t q = 0; q < total; ++q) {
        if (s->elt == 8) ((double *)(hi + 4 * nsrc))[q] = fct[q];
        else ((float *)(hi + 4 * nsrc))[q] = (float)fct[q];
    }
    SM_HIP(hipMemcpyAsync(s->src.ptr, hb.data(), bytes, hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nsrc = nsrc;
    return SM_OK;
}

int sm_cyl_set_receivers(sm_cyl *s, size_t nrec, const size_t *idx, const int *comp,
                         size_t max_steps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_cyl_set_receivers: NULL solver");
    SM_CHECK(nrec == 0 || (idx && comp), SM_ERR_ARG, "sm_cyl_set_receivers: NULL array");
    SM_CHECK(nrec < (1u << 24), SM_ERR_ARG, "sm_cyl_set_receivers: too many receivers");
    for (size_t r = 0; r < nrec; ++r) {
        SM_CHECK(idx[r] < s->N, SM_ERR_ARG, "receiver %zu: index %zu outside the padded grid", r,
                 idx[r]);
        SM_CHECK(recordable(comp[r]), SM_ERR_ARG, "receiver %zu: component %d not recordable in "
                 "e_cyl_iso", r, comp[r]);
    }
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    dfree(s->rcv);
    dfree(s->trace);
    s->nrec = 0;
    s->max_steps = 0;
    if (nrec == 0 || max_steps == 0) return SM_OK;
    int rc = dalloc(s->rcv, nrec * (sizeof(int64_t) + sizeof(int)));
    if (rc == SM_OK) rc = dalloc(s->trace, nrec * max_steps * s->elt);
    if (rc != SM_OK) { dfree(s->rcv); dfree(s->trace); return rc; }
    std::vector<unsigned char> hb(nrec * (sizeof(int64_t) + sizeof(int)));
    int64_t *hi = (int64_t *)hb.data();
    int *hc = (int *)(hi + nrec);
    for (size_t r = 0; r < nrec; ++r) { hi[r] = (int64_t)idx[r]; hc[r] = comp[r]; }
    SM_HIP(hipMemcpyAsync(s->rcv.ptr, hb.data(), hb.size(), hipMemcpyHostToDevice, s->ctx->stream));
    SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, s->ctx->stream));
    SM_HIP(hipStreamSynchronize(s->ctx->stream));
    s->nrec = nrec;
    s->max_steps = max_steps;
    s->rcv_step0 = s->step;
    return SM_OK;
}

int sm_cyl_run(sm_cyl *s, size_t nsteps)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_cyl_run: NULL solver");
    SM_CHECK(s->step + nsteps < (1ull << 62), SM_ERR_ARG, "sm_cyl_run: step count too large");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    return s->elt == 8 ? run<double>(s, nsteps) : run<float>(s, nsteps);
}

int sm_cyl_set_timing(sm_cyl *s, int on)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_cyl_set_timing: NULL solver");
    s->timing = on != 0;
    s->timed_ms = 0.0;
    s->timed_steps = 0;
    return SM_OK;
}

int sm_cyl_get_timing(const sm_cyl *s, double *ms_total, size_t *steps)
{
    SM_CHECK(s && ms_total && steps, SM_ERR_ARG, "sm_cyl_get_timing: NULL argument");
    *ms_total = s->timed_ms;
    *steps = s->timed_steps;
    return SM_OK;
}

int sm_cyl_current_step(const sm_cyl *s, size_t *it)
{
    SM_CHECK(s && it, SM_ERR_ARG, "sm_cyl_current_step: NULL argument");
    *it = s->step;
    return SM_OK;
}

int sm_cyl_get_traces(sm_cyl *s, size_t it0, size_t count, double *out)
{
    SM_CHECK(s && (count == 0 || out), SM_ERR_ARG, "sm_cyl_get_traces: NULL argument");
    SM_CHECK(it0 + count <= s->max_steps, SM_ERR_ARG, "sm_cyl_get_traces: rows [%zu, %zu) beyond %zu",
             it0, it0 + count, s->max_steps);
    if (count == 0 || s->nrec == 0) return SM_OK;
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->trace.ptr + it0 * s->nrec * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, count * s->nrec, out)
                       : read_back<float>(s, base, count * s->nrec, out);
}

int sm_cyl_get_field(sm_cyl *s, int comp, double *out)
{
    SM_CHECK(s && out, SM_ERR_ARG, "sm_cyl_get_field: NULL argument");
    const int q = field_slot(comp);
    SM_CHECK(q >= 0, SM_ERR_ARG, "sm_cyl_get_field: component %d is not a field", comp);
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    const char *base = (const char *)s->fields.ptr + (size_t)q * s->N * s->elt;
    return s->elt == 8 ? read_back<double>(s, base, s->N, out) : read_back<float>(s, base, s->N, out);
}

int sm_cyl_reset(sm_cyl *s)
{
    SM_CHECK(s, SM_ERR_ARG, "sm_cyl_reset: NULL solver");
    SM_CHECK(set_device(s->ctx) == SM_OK, SM_ERR_HIP, "hipSetDevice failed");
    hipStream_t st = s->ctx->stream;
    SM_HIP(hipMemsetAsync(s->fields.ptr, 0, s->fields.bytes, st));
    if (s->mem.ptr) SM_HIP(hipMemsetAsync(s->mem.ptr, 0, s->mem.bytes, st));
    if (s->trace.ptr) SM_HIP(hipMemsetAsync(s->trace.ptr, 0, s->trace.bytes, st));
    SM_HIP(hipStreamSynchronize(st));
    s->step = 0;
    s->rcv_step0 = 0;
    return SM_OK;
}

int sm_cyl_device_bytes(const sm_cyl *s, size_t *bytes)
{
    SM_CHECK(s && bytes, SM_ERR_ARG, "sm_cyl_device_bytes: NULL argument");
    *bytes = s->fields.bytes + s->mem.bytes + s->mat.bytes + s->tab.bytes + s->src.bytes +
             s->rcv.bytes + s->trace.bytes;
    return SM_OK;
}

int sm_kurkjian_fct(double lambda, double dr, double dz, double r, int mode, double A, double f,
                    double dt, size_t length, double *fct)
{
    SM_CHECK(fct || length == 0, SM_ERR_ARG, "sm_kurkjian_fct: NULL output");
    SM_CHECK(mode >= 0 && f > 0 && dt > 0 && dr > 0 && dz > 0, SM_ERR_ARG, "sm_kurkjian_fct: bad argument");
    // src.c:202-229: Ricker of unit amplitude normalised to peak 1, times A
    // and lambda/(2 pi) eps_m 2/(dr^2 dz) (r/dr)^n, eps_m = 1 (n = 0) or 2
    const double pi = 4.0 * std::atan(1.0);
    const double pi2 = pi * pi;
    const double eps = mode > 0 ? 2.0 : 1.0;
    const double term = lambda / (2. * pi) * eps * 2. / (dr * dr * dz) * std::pow(r / dr, mode);
    const double tcut = 1.0 / f;
    double amax = 0.0;
    for (size_t i = 0; i < length; ++i) {
        const double time = i * dt - tcut;
        const double K = pi2 * f * f * time * time;
        fct[i] = 2. * pi2 * f * f * (1.0 - 2.0 * K) * std::exp(-K);
        amax = amax > std::fabs(fct[i]) ? amax : std::fabs(fct[i]);
    }
    for (size_t i = 0; i < length; ++i) fct[i] *= A / amax * term;
    return SM_OK;
}

}  // extern "C"
