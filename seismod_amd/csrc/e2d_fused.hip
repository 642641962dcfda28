// e2d_fused.hip -- one-pass FP32 time step of the 2D elastic split-PML
// solver over one shot or a batch of shots (replaces, per step,
// e_iso.c:589-669: update_vx/update_vz + pml_v, update_txxzz/update_txz +
// pml_t, add_src, receiver sampling).
//
// Arithmetic (all FP32, dt/h folded into the material arrays):
//   interior   vx += cb*((txx - txx[i-1]) + (txz - txz[j-1]))
//              vz += cbm*((txz[i+1] - txz) + (tzz[j+1] - tzz))
//              txx += cl2m*dvx_x + cla*dvz_z ; tzz += cl2m*dvz_z + cla*dvx_x
//              txz += cmu*(dvx_z + dvz_x)
//   PML split  u = A*u + B*(c*delta) per component, with
//              A = (2 - dt d)/(2 + dt d), B = 2/(2 + dt d) where the axis is
//              damped (pml.c:59-63 folded: c1*c2 and c1*h/dt) and A = B = 1
//              where it is not (pml.c:123); the total is x + z.
// Neighbours off the padded grid read 0 (pml.c:58, 79).
#include <algorithm>
#include <cstring>
#include <type_traits>

#include <mutex>

#include "e2d_fused.hpp"
#include "sm_internal.hpp"

namespace sm {
namespace {

// Buffer-resource access: one 128-bit resource per array group (uniform,
// in SGPRs), the array's start as the scalar offset and the cell's byte
// offset as the per-lane offset.  A lane that must not touch memory passes
// OFF_NONE: the hardware returns 0 for such loads and drops such stores, so
// masking costs no branch.  Every resource covers less than 2^31 bytes
// (sm_e2d_create refuses larger grids), so OFF_NONE plus any scalar offset
// stays out of range.
constexpr unsigned OFF_NONE = 0x80000000u;
constexpr int RSRC_WORD3 = 0x00020000;   // raw 32-bit buffer on gfx9

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, RSRC_WORD3);
}
// (the b32 builtins move raw 32-bit words: bit casts, not conversions)
__device__ __forceinline__ float bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}
// uniform read-only data through the scalar cache
__device__ __forceinline__ int cld(const int *p)
{
    return *(const __attribute__((address_space(4))) int *)p;
}
__device__ __forceinline__ float cld(const float *p)
{
    return *(const __attribute__((address_space(4))) float *)p;
}
__device__ __forceinline__ float2 cld(const float2 *p)
{
    const float *f = (const float *)p;
    return make_float2(cld(f), cld(f + 1));
}
__device__ __forceinline__ int4 cld(const int4 *p)
{
    const int *i = (const int *)p;
    return make_int4(cld(i), cld(i + 1), cld(i + 2), cld(i + 3));
}

// value held by lane-1 / lane+1 (lanes 0 and 63 receive a value they do not use)
__device__ __forceinline__ float from_left(float x) { return __shfl_up(x, 1, 64); }
__device__ __forceinline__ float from_right(float x) { return __shfl_down(x, 1, 64); }

struct VRow { float t, a, c, coef; };   // total or split pair, and its coefficient
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// Staged loads: the rows every iteration needs (field-rows 0..4 from the
// state buffer, 5..9 from the material buffer) copied in 16-byte chunks
// through the wave's LDS.  Field-row fr: buffer field, row offset from the
// iteration (+1 or +2), first chunk in the wave's LDS area.
__host__ __device__ constexpr int sg_field(int fr)
{
    return fr == 0 ? E2_TXX : fr == 1 ? E2_TZZ : fr == 2 ? E2_TXZ : fr == 3 ? E2_VX : fr == 4 ? E2_VZ
         : fr == 5 ? E2_CB : fr == 6 ? E2_CBM : fr == 7 ? E2_CLA : fr == 8 ? E2_CL2M : E2_CMU;
}
__host__ __device__ constexpr int sg_drow(int fr) { return (fr <= 3 || fr == 5) ? 2 : 1; }
constexpr int SG_CHUNKS = 17;   // 68 cells: 64 plus up to 3 before for 16-byte alignment
__host__ __device__ constexpr int sg_lds(int fr) { return fr < 5 ? fr * SG_CHUNKS : 128 + (fr - 5) * SG_CHUNKS; }
struct TRow { float la, l2m, mu, x[6]; }; // stress coefficients and split pairs

// STAGED: the rows every iteration needs arrive as 16-byte chunk loads
// through LDS (4 loads per lane per iteration instead of 10); the
// arithmetic is the same, so results are identical.
template <bool STAGED>
__global__ void __launch_bounds__(256)
e2d_fused_step(const E2dShot *__restrict__ shots, const int *__restrict__ bshot, int nitems,
               const int *__restrict__ base, int k)
{
    __shared__ u32x4 s_stage[STAGED ? E2D_WAVES * 256 : 1];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // consecutive blocks go to different XCDs: give each XCD a contiguous
    // range of work items (the grid is a multiple of 8 blocks)
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int logical = (b & 7) * (nb >> 3) + (b >> 3);
    const int g = logical * E2D_WAVES + w;
    if (g >= nitems) return;
    const E2dShot *S = shots + bshot[logical];
    const int s = (base ? *base : 0) + k;
    if (s >= S->nsteps) return;
    const int q = g - S->item0;
    const int ntz = S->ntz;
    if (q >= ntz * S->nch) return;   // padding wave of the shot's last block
    const int it = S->it0 + s;

    const int NX = S->NX, NZ = S->NZ, np = S->np, nx = S->nx, nz = S->nz;
    const int zt = q % ntz, ch = q / ntz;
    const int I0 = ch * S->chunk;
    const int I1 = min(I0 + S->chunk, NX);
    const int J = zt * E2D_OWN - 1 + lane;
    const bool jv = J >= 0 && J < NZ;
    const bool own = jv && lane >= 1 && lane <= E2D_OWN;
    const bool pzl = J < np || J >= np + nz;
    // does any lane of this wave lie in the z PML (uniform)
    const bool zpml = __any(jv && pzl);
    const unsigned N = S->N, N4 = N * 4u;
    const auto Rin = rsrc(S->st[it & 1], E2_NSTATE * N4);
    const auto Rout = rsrc(S->st[(it + 1) & 1], E2_NSTATE * N4);
    const auto Rss = rsrc(S->ss, E2_NSPLIT * N4);
    const auto Rmat = rsrc(S->mat, E2_NMAT * N4);
    const auto Rdmp = rsrc(S->dmp, (unsigned)(2 * NX + 2 * NZ) * 8u);
    const float2 *xi = (const float2 *)S->dmp;
    const float2 *xh = xi + NX;
    // z damping of this lane: (A, B) at integer and half stagger
    const unsigned zo = jv ? (unsigned)(2 * NX + J) * 8u : OFF_NONE;
    const unsigned zho = jv ? (unsigned)(2 * NX + NZ + J) * 8u : OFF_NONE;
    const float2 zi = make_float2(bld(Rdmp, zo, 0), bld(Rdmp, zo, 4));
    const float2 zh = make_float2(bld(Rdmp, zho, 0), bld(Rdmp, zho, 4));

    auto rv = [&](int R) { return R >= 0 && R < NX; };
    auto px = [&](int R) { return R < np || R >= np + nx; };
    // byte offset of row R in this lane, or OFF_NONE
    auto voff = [&](int R, bool ok) { return ok ? ((unsigned)R * NZ + J) * 4u : OFF_NONE; };

    // ---- staged loads (STAGED) ----
    // The wave's 64 cells [J0m, J0m + 64) of a field-row lie in a window of
    // 17 chunks starting at the 16-byte boundary at or before J0m, counted
    // from the start of the state allocation (= st[0]) or of the material
    // allocation (padded by 16 floats), both 256-byte aligned.  Chunks of
    // rows outside the grid are not loaded (0); cells outside [0, NZ) are
    // masked by the reader (jv), as the per-lane loads mask them.
    const int J0m = zt * E2D_OWN - 1;
    const int par_off = (int)(S->st[it & 1] - S->st[0]);   // floats: 0 or E2_NSTATE*N
    const auto Rall = rsrc(S->st[0], (2 * E2_NSTATE + E2_NSPLIT) * N4);
    const auto Rmat16 = rsrc(S->mat, E2_NMAT * N4 + 64);
    u32x4 *const stage = s_stage + (STAGED ? w * 256 : 0);
    // float index (within its allocation) of field-row fr's first cell
    auto fr_base = [&](int fr) { return (fr < 5 ? par_off : 0) + sg_field(fr) * (int)N; };
    // issue the chunk loads of iteration I (rows I+1, I+2); the lane's
    // chunk of round r is c = (r & 1) * 64 + lane of the state (r < 2) or
    // material (r >= 2) field-rows, 17 chunks each
    auto stage_issue = [&](int I, u32x4 (&v)[4]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int c = (r & 1) * 64 + lane;
            const int fr = (r < 2 ? 0 : 5) + c / SG_CHUNKS;
            const int R = I + ((fr <= 3 || fr == 5) ? 2 : 1);
            const int a = (fr_base(fr) + R * NZ + J0m) & ~3;
            const bool ok = c < 5 * SG_CHUNKS && rv(R);
            const unsigned off = ok ? (unsigned)(a + 4 * (c % SG_CHUNKS)) * 4u : OFF_NONE;
            v[r] = __builtin_amdgcn_raw_buffer_load_b128(r < 2 ? Rall : Rmat16, off, 0, 0);
        }
    };
    auto stage_write = [&](const u32x4 (&v)[4]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) stage[r * 64 + lane] = v[r];
    };
    // this lane's value of field-row fr for iteration I
    auto stage_read = [&](int I, int fr) {
        const int R = I + sg_drow(fr);
        const int a = (fr_base(fr) + R * NZ + J0m) & 3;
        const float x = ((const float *)stage)[sg_lds(fr) * 4 + lane + a];
        return jv ? x : 0.f;
    };

    const int nsrc = S->nsrc, nrec = S->nrec;
    const int4 *src = S->src;
    const int4 *rcv = S->rcv;
    const int row = it - S->rcv0;
    const bool record = nrec > 0 && row >= 0 && row < S->rmax;
    int sp = nsrc ? cld(S->srow + I0) : 0;
    int rp = record ? cld(S->rrow + I0) : 0;

    // The split fields are touched only in rows [I0-1, I1] of the x PML or
    // by lanes in the z PML; a wave clear of both runs the loop compiled
    // without them (PML false).  In both forms every per-lane load and store
    // of the row loop is issued unconditionally, lanes and rows with nothing
    // to move passing OFF_NONE: the compiler then knows how many accesses
    // are in flight and waits only for the ones whose values it needs, where
    // accesses behind branches made it wait for loads of the next row just
    // issued.
    const bool pmlw = zpml || px(I0 - 1) || px(I1);
    auto body = [&](auto pml_tag) {
        constexpr bool PML = decltype(pml_tag)::value;

        // ---- loads, issued a row ahead of the arithmetic that uses them;
        // ok false: not needed (0, no memory access) ----
        auto fetch_s = [&](int R, bool ok, float &sxx, float &szz, float &sxz) {
            const unsigned o = voff(R, ok && rv(R) && jv);
            sxx = bld(Rin, o, E2_TXX * N4);
            szz = bld(Rin, o, E2_TZZ * N4);
            sxz = bld(Rin, o, E2_TXZ * N4);
        };
        // velocity component at row R: total, or split pair where the
        // lane's cell is in the PML
        auto fetch_v = [&](int R, bool ok, int tot, int sx, int sz, int cf) {
            VRow r{0.f, 0.f, 0.f, 0.f};
            const unsigned o = voff(R, ok && rv(R) && jv);
            r.coef = bld(Rmat, o, cf * N4);
            if constexpr (PML) {
                const bool pm = px(R) || pzl;
                r.a = bld(Rin, pm ? o : OFF_NONE, sx * N4);
                r.c = bld(Rin, pm ? o : OFF_NONE, sz * N4);
                r.t = bld(Rin, pm ? OFF_NONE : o, tot * N4);
            } else {
                r.t = bld(Rin, o, tot * N4);
            }
            return r;
        };
        auto fetch_t = [&](int R, bool ok) {
            TRow r{};
            const unsigned o = voff(R, ok && rv(R) && jv);
            r.la = bld(Rmat, o, E2_CLA * N4);
            r.l2m = bld(Rmat, o, E2_CL2M * N4);
            r.mu = bld(Rmat, o, E2_CMU * N4);
            if constexpr (PML) {
                const unsigned op = (px(R) || pzl) && own ? o : OFF_NONE;
#pragma unroll
                for (int c = 0; c < 6; ++c) r.x[c] = bld(Rss, op, c * N4);
            }
            return r;
        };

        // new vx of row R from old txx of rows R-1, R and old txz of row R
        auto vx_new = [&](int R, const VRow &r, float sxx_m, float sxx_r, float sxz_r, bool store) {
            const float dxs = sxx_r - sxx_m;
            const float dzs = sxz_r - from_left(sxz_r);
            const unsigned o = voff(R, store && own);
            float vx = r.t + r.coef * (dxs + dzs);
            if constexpr (PML) {
                // (R lies in [0, NX]: the damping table read stays in the
                // allocation; off the grid nothing is kept or stored)
                const bool pm = rv(R) && (px(R) || pzl);
                const float2 ax = cld(xi + R);
                const float a = ax.x * r.a + ax.y * (r.coef * dxs);
                const float c = zi.x * r.c + zi.y * (r.coef * dzs);
                if (pm) vx = a + c;
                bst(Rout, pm ? o : OFF_NONE, E2_VXX * N4, a);
                bst(Rout, pm ? o : OFF_NONE, E2_VXZ * N4, c);
            }
            bst(Rout, o, E2_VX * N4, vx);
            return vx;
        };
        // new vz of row R from old txz of rows R, R+1 and old tzz of row R
        auto vz_new = [&](int R, const VRow &r, float sxz_r, float szz_r, float sxz_p, bool store) {
            const float dxs = sxz_p - sxz_r;
            const float dzs = from_right(szz_r) - szz_r;
            const unsigned o = voff(R, store && own);
            float vz = r.t + r.coef * (dxs + dzs);
            if constexpr (PML) {
                // (R lies in [-1, NX): the table read stays in the allocation)
                const bool pm = rv(R) && (px(R) || pzl);
                const float2 ax = cld(xh + R);
                const float a = ax.x * r.a + ax.y * (r.coef * dxs);
                const float c = zh.x * r.c + zh.y * (r.coef * dzs);
                if (pm) vz = a + c;
                bst(Rout, pm ? o : OFF_NONE, E2_VZX * N4, a);
                bst(Rout, pm ? o : OFF_NONE, E2_VZZ * N4, c);
            }
            bst(Rout, o, E2_VZ * N4, vz);
            return vz;
        };

        // Inputs of row r beyond the stress of row r itself: the stress of
        // row r+1, the old vx (and coefficient) of row r+1, the old vz of row
        // r, and row r's stress coefficients (and split stresses)
        struct RowIn {
            float sxx, szz, sxz;
            VRow vr, zr;
            TRow tr;
        };
        // per-lane loads of the inputs of row r (none where ok is false)
        auto load_in = [&](int r, bool ok, RowIn &n) {
            fetch_s(r + 1, ok, n.sxx, n.szz, n.sxz);
            n.vr = fetch_v(r + 1, ok, E2_VX, E2_VXX, E2_VXZ, E2_CB);
            n.zr = fetch_v(r, ok, E2_VZ, E2_VZX, E2_VZZ, E2_CBM);
            n.tr = fetch_t(r, ok);
        };

        // preamble: vx of row I0 and vz of row I0-1, from rows I0-1 and I0;
        // the loads of the first row go out with them.  A: inputs of row I0;
        // C.sxx..: stress of row I0
        float sxx_m, szz_m, sxz_m;
        RowIn A, C;
        fetch_s(I0 - 1, true, sxx_m, szz_m, sxz_m);
        fetch_s(I0, true, C.sxx, C.szz, C.sxz);
        const VRow v_pre = fetch_v(I0, true, E2_VX, E2_VXX, E2_VXZ, E2_CB);
        const VRow z_pre = fetch_v(I0 - 1, true, E2_VZ, E2_VZX, E2_VZZ, E2_CBM);
        load_in(I0, true, A);
        u32x4 sv[4];
        if constexpr (STAGED)
            if (I0 + 1 < I1) stage_issue(I0, sv);
        float vx0 = vx_new(I0, v_pre, sxx_m, C.sxx, C.sxz, true);
        float vz_m = vz_new(I0 - 1, z_pre, sxz_m, szz_m, C.sxz, false);

        // row I: new vx of row I+1, vz of row I, stress of row I, sources
        // and receivers; sxx0..: stress of row I, c: the rest of its inputs
        auto do_row = [&](int I, bool more, float sxx0, float szz0, float sxz0, const RowIn &c) {
            const float vx1 = vx_new(I + 1, c.vr, sxx0, c.sxx, c.sxz, more);
            const float vz0 = vz_new(I, c.zr, sxz0, szz0, c.sxz, true);
            const TRow &tr = c.tr;

            // stress of row I
            const float dvxx = vx1 - vx0;
            const float dvzz = vz0 - from_left(vz0);
            const float dvxz = from_right(vx0) - vx0;
            const float dvzx = vz0 - vz_m;
            const unsigned o = voff(I, own);
            float nxx = sxx0 + (tr.l2m * dvxx + tr.la * dvzz);
            float nzz = szz0 + (tr.l2m * dvzz + tr.la * dvxx);
            float nxz = sxz0 + tr.mu * (dvxz + dvzx);
            if constexpr (PML) {
                const bool pm = px(I) || pzl;
                const float2 ah = cld(xh + I), ai = cld(xi + I);
                const float txxx = ah.x * tr.x[E2_TXXX] + ah.y * (tr.l2m * dvxx);
                const float txxz = zi.x * tr.x[E2_TXXZ] + zi.y * (tr.la * dvzz);
                const float tzzx = ah.x * tr.x[E2_TZZX] + ah.y * (tr.la * dvxx);
                const float tzzz = zi.x * tr.x[E2_TZZZ] + zi.y * (tr.l2m * dvzz);
                const float txzx = ai.x * tr.x[E2_TXZX] + ai.y * (tr.mu * dvzx);
                const float txzz = zh.x * tr.x[E2_TXZZ] + zh.y * (tr.mu * dvxz);
                if (pm) {
                    nxx = txxx + txxz;
                    nzz = tzzx + tzzz;
                    nxz = txzx + txzz;
                }
                const unsigned op = pm ? o : OFF_NONE;
                bst(Rss, op, E2_TXXX * N4, txxx); bst(Rss, op, E2_TXXZ * N4, txxz);
                bst(Rss, op, E2_TZZX * N4, tzzx); bst(Rss, op, E2_TZZZ * N4, tzzz);
                bst(Rss, op, E2_TXZX * N4, txzx); bst(Rss, op, E2_TXZZ * N4, txzz);
            }

            // add_src (src.c:170-184) on this row, after both phases
            const int rowend = (I + 1) * NZ;
            for (; sp < nsrc; ++sp) {
                const int4 e = cld(src + sp);
                if (e.x >= rowend) break;
                if (it < e.z && own && J == e.x - I * NZ) {
                    const float inc = cld(S->sfct + e.w + it);
                    if (e.y == SM_SX || e.y == SM_BULK) nxx += inc;
                    if (e.y == SM_SZ || e.y == SM_BULK) nzz += inc;
                    if (e.y == SM_SXZ) nxz += inc;
                }
            }
            bst(Rout, o, E2_TXX * N4, nxx);
            bst(Rout, o, E2_TZZ * N4, nzz);
            bst(Rout, o, E2_TXZ * N4, nxz);
            // receivers on this row (e_iso.c:641-669 samples after add_src)
            if (record) {
                for (; rp < nrec; ++rp) {
                    const int4 e = cld(rcv + rp);
                    if (e.x >= rowend) break;
                    if (own && J == e.x - I * NZ) {
                        const float val = e.y == 0 ? vx0 : e.y == 1 ? vz0 : e.y == 2 ? nxx
                                        : e.y == 3 ? nzz : nxz;
                        *(__attribute__((address_space(1))) float *)(S->trace + (size_t)row * nrec + e.z) = val;
                    }
                }
            }
            vx0 = vx1;
            vz_m = vz0;
        };

        if constexpr (STAGED) {
            // (two alternating sets of chunk registers, loads issued before
            // the wait for the previous row's, took 103 instead of 85
            // registers, 4 instead of 5 waves per SIMD, and were slower)
            RowIn cur = A;
            float sxx0 = C.sxx, szz0 = C.szz, sxz0 = C.sxz;
            for (int I = I0; I < I1; ++I) {
                // loads of the next iteration, staged (none after the
                // chunk's last row)
                const bool more = I + 1 < I1;
                RowIn nxt{0.f, 0.f, 0.f, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, TRow{}};
                if (more) {
                    stage_write(sv);
                    if (I + 2 < I1) stage_issue(I + 1, sv);
                    nxt.sxx = stage_read(I, 0);
                    nxt.szz = stage_read(I, 1);
                    nxt.sxz = stage_read(I, 2);
                    nxt.vr.t = stage_read(I, 3);
                    nxt.vr.coef = stage_read(I, 5);
                    nxt.zr.t = stage_read(I, 4);
                    nxt.zr.coef = stage_read(I, 6);
                    nxt.tr.la = stage_read(I, 7);
                    nxt.tr.l2m = stage_read(I, 8);
                    nxt.tr.mu = stage_read(I, 9);
                    if constexpr (PML) {
                        // split velocities and stresses, per lane (totals of
                        // split lanes are staged too and not used)
                        auto vsplit = [&](int R, int sx, int sz, VRow &r) {
                            const bool pm = rv(R) && (px(R) || pzl);
                            const unsigned o = voff(R, pm && jv);
                            r.a = bld(Rin, o, sx * N4);
                            r.c = bld(Rin, o, sz * N4);
                        };
                        vsplit(I + 2, E2_VXX, E2_VXZ, nxt.vr);
                        vsplit(I + 1, E2_VZX, E2_VZZ, nxt.zr);
                        const unsigned op = (px(I + 1) || pzl) && own ? voff(I + 1, jv) : OFF_NONE;
#pragma unroll
                        for (int c = 0; c < 6; ++c) nxt.tr.x[c] = bld(Rss, op, c * N4);
                    }
                }
                do_row(I, more, sxx0, szz0, sxz0, cur);
                sxx0 = cur.sxx; szz0 = cur.szz; sxz0 = cur.sxz;
                cur = nxt;
            }
        } else {
            // Two rows per trip with two register sets: the loads of one
            // row land in the set the other row has finished with, so no
            // value still in flight is ever copied (a copy would have to
            // wait for it).  The stress of row I is taken out of that set
            // before it is reloaded; it arrived a row ago.  (Loading two rows
            // ahead from three sets took 84 instead of 62 registers, 5
            // instead of 7 waves per SIMD, and made single shots slower.)
            for (int I = I0; I < I1; I += 2) {
                {
                    const float x0 = C.sxx, z0 = C.szz, xz0 = C.sxz;
                    load_in(I + 1, I + 1 < I1, C);
                    do_row(I, I + 1 < I1, x0, z0, xz0, A);
                }
                if (I + 1 >= I1) break;
                {
                    const float x0 = A.sxx, z0 = A.szz, xz0 = A.sxz;
                    load_in(I + 2, I + 2 < I1, A);
                    do_row(I + 1, I + 2 < I1, x0, z0, xz0, C);
                }
            }
        }
    };
    if (pmlw) body(std::true_type{});
    else body(std::false_type{});
}

__global__ void e2d_base_set(int *base, int value, int add)
{
    if (threadIdx.x == 0) *base = add ? *base + value : value;
}

int grid_blocks(int nitems)
{
    const int nb = (nitems + E2D_WAVES - 1) / E2D_WAVES;
    return (nb + 7) & ~7;
}

}  // namespace

E2dRunner::~E2dRunner()
{
    drop_graphs();
    if (table_) (void)hipFree(table_);
    if (base_) (void)hipFree(base_);
}

void E2dRunner::drop_graphs()
{
    for (auto &kv : graphs_) (void)hipGraphExecDestroy(kv.second.exec);
    graphs_.clear();
}

size_t E2dRunner::device_bytes() const { return table_cap_ + (base_ ? 256 : 0); }

int E2dRunner::ensure(size_t bytes)
{
    if (!base_) SM_HIP(hipMalloc((void **)&base_, 256));
    if (bytes <= table_cap_) return SM_OK;
    drop_graphs();
    if (table_) SM_HIP(hipFree(table_));
    table_ = nullptr;
    table_cap_ = 0;
    const size_t cap = std::max<size_t>(bytes, 4096);
    SM_HIP(hipMalloc(&table_, cap));
    table_cap_ = cap;
    return SM_OK;
}

void E2dRunner::launch_step(int nb, hipStream_t st, const E2dShot *tab, const int *bshot, int nitems,
                            const int *base, int k)
{
    if (staged) e2d_fused_step<true><<<nb, 256, 0, st>>>(tab, bshot, nitems, base, k);
    else e2d_fused_step<false><<<nb, 256, 0, st>>>(tab, bshot, nitems, base, k);
}

int E2dRunner::graph_for(int nact, int nitems, int g, hipStream_t st, hipGraphExec_t *out)
{
    const auto key = std::make_tuple(nact, nitems, g);
    auto f = graphs_.find(key);
    if (f != graphs_.end()) {
        f->second.used = ++tick_;
        *out = f->second.exec;
        return SM_OK;
    }
    if (graphs_.size() >= MAX_GRAPHS) {
        auto lru = graphs_.begin();
        for (auto it = graphs_.begin(); it != graphs_.end(); ++it)
            if (it->second.used < lru->second.used) lru = it;
        // replays of it may still be queued on this stream
        SM_HIP(hipStreamSynchronize(st));
        (void)hipGraphExecDestroy(lru->second.exec);
        graphs_.erase(lru);
    }
    const E2dShot *tab = (const E2dShot *)table_;
    const int *bshot = (const int *)((const char *)table_ + tab_bytes_);
    const int nb = grid_blocks(nitems);
    hipGraph_t graph = nullptr;
    SM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < g; ++k)
        launch_step(nb, st, tab, bshot, nitems, base_, k);
    e2d_base_set<<<1, 64, 0, st>>>(base_, g, 1);
    SM_HIP(hipStreamEndCapture(st, &graph));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(SM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    graphs_[key] = Cached{ex, ++tick_};
    ++graphs_built;
    *out = ex;
    return SM_OK;
}

// Sorts the shots by step count (the active set at any step is then a
// prefix), gives each its first work item, and lays out the table the
// kernel reads: the shot descriptors, then at tab_bytes the block -> shot
// map (every shot's items are whole blocks).  pref[m] is the number of work
// items of the first m shots.
int E2dRunner::plan(std::vector<E2dShot> &shots, std::vector<int> &pref,
                    std::vector<unsigned char> &host, size_t &tab_bytes)
{
    shots.erase(std::remove_if(shots.begin(), shots.end(),
                               [](const E2dShot &d) { return d.nsteps <= 0; }),
                shots.end());
    if (shots.empty()) return SM_OK;
    std::stable_sort(shots.begin(), shots.end(),
                     [](const E2dShot &a, const E2dShot &b) { return a.nsteps > b.nsteps; });
    const size_t n = shots.size();
    pref.assign(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        shots[i].item0 = pref[i];
        int ntz, nch;
        const long long next =
            (long long)pref[i] + e2d_items(shots[i].NX, shots[i].NZ, shots[i].chunk, &ntz, &nch);
        SM_CHECK(next < (1ll << 30), SM_ERR_ARG, "batch too large: %lld work items", next);
        pref[i + 1] = (int)next;
    }
    const int nblk = pref[n] / E2D_WAVES;
    tab_bytes = (n * sizeof(E2dShot) + 255) & ~(size_t)255;
    host.assign(tab_bytes + (size_t)nblk * sizeof(int), 0);
    std::memcpy(host.data(), shots.data(), n * sizeof(E2dShot));
    int *bs = (int *)(host.data() + tab_bytes);
    for (size_t i = 0; i < n; ++i)
        for (int q = pref[i] / E2D_WAVES; q < pref[i + 1] / E2D_WAVES; ++q) bs[q] = (int)i;
    return SM_OK;
}

// One launch per step over the shots still running (a prefix of the sorted
// list); with graph_steps > 1, stretches of at least two graphs' length are
// replayed from captured graphs.  Returns the number of step launches.
int E2dRunner::launch_all(const std::vector<E2dShot> &shots, const std::vector<int> &pref,
                          const void *table, size_t tab_bytes, hipStream_t st, int graph_steps,
                          size_t *count)
{
    const E2dShot *tab = (const E2dShot *)table;
    const int *bshot = (const int *)((const char *)table + tab_bytes);
    const int n = (int)shots.size();
    int s = 0;
    const int total = shots[0].nsteps;
    while (s < total) {
        int m = n;
        while (m > 0 && shots[m - 1].nsteps <= s) --m;
        const int e = shots[m - 1].nsteps;
        const int nitems = pref[m];
        int len = e - s;
        if (graph_steps > 1 && len >= 2 * graph_steps) {
            hipGraphExec_t ex = nullptr;
            int rc = graph_for(m, nitems, graph_steps, st, &ex);
            if (rc != SM_OK) return rc;
            e2d_base_set<<<1, 64, 0, st>>>(base_, s, 0);
            const int reps = len / graph_steps;
            for (int r = 0; r < reps; ++r) SM_HIP(hipGraphLaunch(ex, st));
            s += reps * graph_steps;
            *count += (size_t)reps * graph_steps;
            len = e - s;
        }
        const int nb = grid_blocks(nitems);
        for (int k = 0; k < len; ++k)
            launch_step(nb, st, tab, bshot, nitems, nullptr, s + k);
        s += len;
        *count += (size_t)len;
    }
    SM_HIP(hipGetLastError());
    return SM_OK;
}

int E2dRunner::run(std::vector<E2dShot> shots, hipStream_t st, int graph_steps)
{
    std::vector<int> pref;
    std::vector<unsigned char> host;
    size_t tab_bytes = 0;
    int rc = plan(shots, pref, host, tab_bytes);
    if (rc != SM_OK || shots.empty()) return rc;
    tab_bytes_ = tab_bytes;
    if (host.size() > table_cap_ || tab_bytes_ != last_tab_bytes_ || staged != last_staged_)
        drop_graphs();
    last_tab_bytes_ = tab_bytes_;
    last_staged_ = staged;
    rc = ensure(host.size());
    if (rc != SM_OK) return rc;
    // the table lives in device memory read by the kernels of this run;
    // waiting for the copy keeps the host buffer's lifetime simple
    SM_HIP(hipMemcpyAsync(table_, host.data(), host.size(), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));

    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        SM_HIP(hipEventCreate(&t0));
        SM_HIP(hipEventCreate(&t1));
        SM_HIP(hipEventRecord(t0, st));
    }
    size_t count = 0;
    rc = launch_all(shots, pref, table_, tab_bytes_, st, graph_steps, &count);
    if (rc != SM_OK) return rc;
    if (timing) {
        SM_HIP(hipEventRecord(t1, st));
        SM_HIP(hipEventSynchronize(t1));
        float ms = 0.f;
        SM_HIP(hipEventElapsedTime(&ms, t0, t1));
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        ms_total += ms;
        launches += count;
    }
    return SM_OK;
}

// ---------------------------------------------------------------------
// Allocation-free batches.  The shot table and block -> shot map of a run
// live in this file-scope device array instead of memory the runner
// allocates.  Each device holds its own copy; host code reaches it only
// through the address hipGetSymbolAddress returns on that device (the
// array's name is not a device address in host code).  Every run on a
// device overwrites the whole table it reads, so nothing needs clearing;
// two runs that overlap on one device (from two host threads) would
// overwrite each other's table.
constexpr size_t E2D_SCRATCH_BYTES = 1u << 20;
static __device__ __align__(256) unsigned char e2d_scratch[E2D_SCRATCH_BYTES];

namespace {
constexpr int E2D_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[E2D_MAX_DEVICES] = {};
}  // namespace

int e2d_scratch_init(int device)
{
    SM_CHECK(device >= 0 && device < E2D_MAX_DEVICES, SM_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return SM_OK;
    int cur = -1;
    SM_HIP(hipGetDevice(&cur));
    SM_CHECK(cur == device, SM_ERR_STATE, "e2d_scratch_init: device %d is not current (%d)", device, cur);
    void *p = nullptr;
    SM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(e2d_scratch)));
    SM_CHECK(p != nullptr, SM_ERR_HIP, "hipGetSymbolAddress returned NULL");
    g_scratch[device] = p;
    return SM_OK;
}

size_t e2d_scratch_bytes() { return E2D_SCRATCH_BYTES; }

int E2dRunner::run_static(std::vector<E2dShot> shots, int device, hipStream_t st)
{
    void *scratch = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (device >= 0 && device < E2D_MAX_DEVICES) scratch = g_scratch[device];
    }
    SM_CHECK(scratch, SM_ERR_STATE, "allocation-free batch: no scratch address for device %d "
             "(the context was not created by sm_ctx_create)", device);
    std::vector<int> pref;
    std::vector<unsigned char> host;
    size_t tab_bytes = 0;
    int rc = plan(shots, pref, host, tab_bytes);
    if (rc != SM_OK || shots.empty()) return rc;
    SM_CHECK(host.size() <= E2D_SCRATCH_BYTES, SM_ERR_ARG,
             "allocation-free batch: the shot table needs %zu bytes, the device scratch holds %zu "
             "(fewer shots or more rows per chunk)", host.size(), (size_t)E2D_SCRATCH_BYTES);
    // on one stream the copy is ordered after the kernels of an earlier run
    // that read the table; runs on other streams of the device must not
    // overlap this one (see above)
    SM_HIP(hipMemcpyAsync(scratch, host.data(), host.size(), hipMemcpyHostToDevice, st));
    SM_HIP(hipStreamSynchronize(st));
    size_t count = 0;
    return launch_all(shots, pref, scratch, tab_bytes, st, 0, &count);
}

}  // namespace sm
