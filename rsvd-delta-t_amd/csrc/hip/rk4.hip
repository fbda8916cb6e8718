// Fused CSR SpMM + classical-RK4 stage kernel for gfx950.
//
// Replaces the per-step sequence of SourceCode/TSRK4.c:18-47 (four
// MatMultAdd, three VecWAXPY, four VecAXPY on one test vector at a time) by
// four launches that each apply the operator to the whole N x k block of
// test vectors and finish the stage's vector updates in the epilogue:
//
//   stage 1:  s = A y  + F0 ;  u' = y + (h/2) s ;  acc = y + (h/6) s
//   stage 2:  s = A u' + Fh ;  u''= y + (h/2) s ;  acc = acc + (h/6*2) s
//   stage 3:  s = A u''+ Fh ;  u' = y + h s     ;  acc = acc + (h/6*2) s
//   stage 4:  s = A u' + F1 ;  y  = acc + (h/6) s  [+ FP64 snapshot copy]
//
// which is the reference's rounding order (y += h/6 k1; += h/6*2 k2; ...).
// The accumulator starts from the forcing, as PETSc's MatMultAdd does.
//
// Layout: a block is N x k row-major, so a row gather reads k contiguous
// complex values; LPR lanes (power of two >= k, capped at 64) cooperate on
// one row and each lane owns CPL columns.  With xcd_remap the workgroups of
// one XCD (dispatch is round-robin over the 8 XCDs) get a contiguous row
// range (or interleaved chunks of row blocks, see xcd_block), so the stencil
// neighbours a row gathers sit in that XCD's L2.
#include "common.h"

// LDS budget of the fused small-operator stepper (CSR matrix, y and two
// stage vectors of one column)
#define RSVDT_FUSED_LDS (160 * 1024)

namespace rsvdt {

// Matrix values are FP64 for both state precisions (real or complex); the
// SpMM accumulates in FP64 and only the stored state vectors are FP32 in the
// FP32 form (rounding A itself to FP32 would perturb the operator).
template <typename T, bool AREAL> struct aval { using type = double2; };
template <typename T> struct aval<T, true> { using type = double; };

__device__ __forceinline__ double shfl_val(double v, int src, int w) { return __shfl(v, src, w); }
__device__ __forceinline__ float shfl_val(float v, int src, int w) { return __shfl(v, src, w); }
__device__ __forceinline__ double2 shfl_val(double2 v, int src, int w) {
    return make_double2(__shfl(v.x, src, w), __shfl(v.y, src, w));
}
__device__ __forceinline__ float2 shfl_val(float2 v, int src, int w) {
    return make_float2(__shfl(v.x, src, w), __shfl(v.y, src, w));
}

// Memory-level parallelism: the LPR lanes of a row first fetch up to 16 of
// its (column, value) pairs with one coalesced load each, broadcast them with
// cross-lane shuffles, then issue all the row's gathers back to back; the
// row-local operands (forcing, y, acc) are loaded before the gathers.
typedef double ev2d __attribute__((ext_vector_type(2)));
typedef float ev2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_stream(const double2 *p, bool nt) {
    if (nt) { ev2d v = __builtin_nontemporal_load((const ev2d *)p); return make_double2(v.x, v.y); }
    return *p;
}
__device__ __forceinline__ float2 ld_stream(const float2 *p, bool nt) {
    if (nt) { ev2f v = __builtin_nontemporal_load((const ev2f *)p); return make_float2(v.x, v.y); }
    return *p;
}
__device__ __forceinline__ void st_stream(double2 *p, double2 v, bool nt) {
    if (nt) { ev2d w = {v.x, v.y}; __builtin_nontemporal_store(w, (ev2d *)p); }
    else *p = v;
}
__device__ __forceinline__ void st_stream(float2 *p, float2 v, bool nt) {
    if (nt) { ev2f w = {v.x, v.y}; __builtin_nontemporal_store(w, (ev2f *)p); }
    else *p = v;
}

// s[t] += sum over the row's nonzeros e of A(row, col_e) * x(col_e, lane + t*LPR).
// One batch of a row's nonzeros eb .. eb+min(NB, e1-eb)-1: column indices
// and values are fetched cooperatively (one coalesced load per lane) and
// broadcast with __shfl; then all NB gathers are issued before any FMA, so
// NB independent requests are in flight per lane.  Gathers past the row's
// end read row 0 (valid) and are not used; lanes beyond k read column k-1
// and their sums are never stored.
template <typename T, bool AREAL, int LPR, int CPL, int NB>
__device__ __forceinline__ void gather_batch(int eb, int e1, int lane, int k, const int32_t *__restrict__ colidx,
                                             const typename aval<T, AREAL>::type *__restrict__ vals,
                                             const typename vec2<T>::type *__restrict__ ub, double (&sr)[CPL],
                                             double (&si)[CPL])
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    constexpr int EPL = LPR >= NB ? 1 : NB / LPR;   // entries each lane fetches per batch
    int mc[EPL];
    AV mv[EPL];
#pragma unroll
    for (int u = 0; u < EPL; ++u) {
        const int e = eb + lane + u * LPR;
        const bool ok = (LPR <= NB || lane < NB) && e < e1;
        mc[u] = ok ? colidx[e] : 0;
        if (ok) mv[u] = vals[e];
        else mv[u] = AV{};
    }
    const int cnt = (e1 - eb) < NB ? (e1 - eb) : NB;
    int cs[NB];
    AV as[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        cs[u] = __shfl(mc[u / LPR < EPL ? u / LPR : 0], u % LPR, LPR);
        as[u] = shfl_val(mv[u / LPR < EPL ? u / LPR : 0], u % LPR, LPR);
    }
    V xs[NB][CPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int j = lane + t * LPR;
            xs[u][t] = ub[(int64_t)cs[u] * k + (j < k ? j : k - 1)];
        }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if (u < cnt) {
#pragma unroll
            for (int t = 0; t < CPL; ++t) {
                const double xr = (double)xs[u][t].x, xi = (double)xs[u][t].y;
                if constexpr (AREAL) {
                    sr[t] += as[u] * xr;
                    si[t] += as[u] * xi;
                } else {
                    sr[t] += as[u].x * xr - as[u].y * xi;
                    si[t] += as[u].x * xi + as[u].y * xr;
                }
            }
        }
    }
}

// s += A(row, :) * ub over the row's nonzeros e0 .. e1-1: full batches of NB,
// then the remainder in one batch of NB, NB/2 or NB/4 -- the smallest that
// holds it once it is at most half a batch -- so short tails issue few
// padding gathers (11 nonzeros with NB = 8: 8 + 4 gathers, not 16).
template <typename T, bool AREAL, int LPR, int CPL, int NB>
__device__ __forceinline__ void row_gather(int e0, int e1, int lane, int k, const int32_t *__restrict__ colidx,
                                           const typename aval<T, AREAL>::type *__restrict__ vals,
                                           const typename vec2<T>::type *__restrict__ ub, double (&sr)[CPL],
                                           double (&si)[CPL])
{
    constexpr int NB2 = NB / 2 > 0 ? NB / 2 : 1, NB4 = NB / 4 > 0 ? NB / 4 : 1;
    int eb = e0;
    for (; e1 - eb >= NB; eb += NB) gather_batch<T, AREAL, LPR, CPL, NB>(eb, e1, lane, k, colidx, vals, ub, sr, si);
    const int rem = e1 - eb;
    if (rem > NB2) gather_batch<T, AREAL, LPR, CPL, NB>(eb, e1, lane, k, colidx, vals, ub, sr, si);
    else if (rem > NB4) gather_batch<T, AREAL, LPR, CPL, NB2>(eb, e1, lane, k, colidx, vals, ub, sr, si);
    else if (rem > 0) gather_batch<T, AREAL, LPR, CPL, NB4>(eb, e1, lane, k, colidx, vals, ub, sr, si);
}

// NB: nonzeros gathered per batch (16, or 8 for fewer registers and more
// waves); NT: nontemporal hints on the row-local streams, which keeps the
// L2 for the gathered stage vector.
template <typename T, bool AREAL, int LPR, int CPL, int NB, bool NT>
__global__ __launch_bounds__(256) void k_rk4_stage(
    int nrows, int k, int stage, int xcd_remap, int64_t bstride,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, const void *__restrict__ vals_,
    const typename vec2<T>::type *__restrict__ uin, const typename vec2<T>::type *yv,
    const typename vec2<T>::type *__restrict__ F, typename vec2<T>::type *__restrict__ acc,
    typename vec2<T>::type *out, double cnext, double cacc, double2 *__restrict__ snap)
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    constexpr int RPB = 256 / LPR;                  // rows per workgroup
    const int bx = xcd_block(blockIdx.x, gridDim.x, xcd_remap);
    const int lane = threadIdx.x & (LPR - 1);
    const int row = bx * RPB + (int)(threadIdx.x / LPR);
    if (row >= nrows) return;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const V *ub = uin + boff;
    const AV *vals = (const AV *)vals_;
    const int e0 = rowptr[row], e1 = rowptr[row + 1];

    // row-local operands first
    V fr[CPL], yr[CPL], ar[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = lane + t * LPR;
        const int64_t idx = boff + (int64_t)row * k + j;
        fr[t].x = 0; fr[t].y = 0; yr[t] = fr[t]; ar[t] = fr[t];
        if (j < k) {
            if (F != nullptr) fr[t] = ld_stream(F + idx, NT);
            if (stage != 4) yr[t] = ld_stream(yv + idx, NT);
            if (stage != 1) ar[t] = ld_stream(acc + idx, NT);
        }
    }
    double sr[CPL], si[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) { sr[t] = (double)fr[t].x; si[t] = (double)fr[t].y; }

    row_gather<T, AREAL, LPR, CPL, NB>(e0, e1, lane, k, colidx, vals, ub, sr, si);
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = lane + t * LPR;
        if (j >= k) continue;
        const int64_t idx = boff + (int64_t)row * k + j;
        if (stage == 4) {
            V yn;
            yn.x = (T)((double)ar[t].x + cacc * sr[t]);
            yn.y = (T)((double)ar[t].y + cacc * si[t]);
            out[idx] = yn;
            if (snap != nullptr) st_stream(snap + idx, make_double2((double)yn.x, (double)yn.y), NT);
        } else {
            V u;
            u.x = (T)((double)yr[t].x + cnext * sr[t]);
            u.y = (T)((double)yr[t].y + cnext * si[t]);
            out[idx] = u;                     // gathered by the next stage: keep it cacheable
            V a;
            if (stage == 1) {
                a.x = (T)((double)yr[t].x + cacc * sr[t]);
                a.y = (T)((double)yr[t].y + cacc * si[t]);
            } else {
                a.x = (T)((double)ar[t].x + cacc * sr[t]);
                a.y = (T)((double)ar[t].y + cacc * si[t]);
            }
            st_stream(acc + idx, a, NT);
        }
    }
}

template <typename T, bool AREAL, int LPR, int CPL, int NB, bool NT>
static int launch_rk4(const rsvdt_spmm_desc *d, int stage, const void *uin, const void *y, const void *F,
                      void *acc, void *out, double cn, double ca, double *snap, hipStream_t s)
{
    using V = typename vec2<T>::type;
    constexpr int RPB = 256 / LPR;
    const unsigned gx = ceil_div(d->nrows, RPB);
    const unsigned gy = (unsigned)d->nb;
    const int remap = (d->xcd_remap > 0 && d->xcd_remap <= 3 && gy == 1 && gx >= 8) ? d->xcd_remap : 0;
    hipLaunchKernelGGL((k_rk4_stage<T, AREAL, LPR, CPL, NB, NT>), dim3(gx, gy), dim3(256), 0, s,
                       (int)d->nrows, (int)d->k, stage, remap, d->nrows * d->k, d->rowptr, d->col, d->val,
                       (const V *)uin, (const V *)y, (const V *)F, (V *)acc, (V *)out, cn, ca,
                       (double2 *)snap);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_rk4_stage");
    return RSVDT_HIP_OK;
}

template <typename T, bool AREAL>
static int dispatch_lanes(const rsvdt_spmm_desc *d, int stage, const void *uin, const void *y, const void *F,
                          void *acc, void *out, double cn, double ca, double *snap, hipStream_t s)
{
    int lpr = d->lanes;
    const int k = (int)d->k;
    if (lpr == 0) {
        lpr = 4;
        while (lpr < k && lpr < 64) lpr <<= 1;
    }
    const int cpl = (k + lpr - 1) / lpr;
    const int var = d->variant;
#define RSVDT_CASE(L, C)                                                                                        \
    if (lpr == L && cpl <= C) {                                                                                 \
        if (C == 1 && var == 1) return launch_rk4<T, AREAL, L, C, 16, true>(d, stage, uin, y, F, acc, out, cn, ca, snap, s);  \
        if (C == 1 && var == 2) return launch_rk4<T, AREAL, L, C, 8, false>(d, stage, uin, y, F, acc, out, cn, ca, snap, s);  \
        if (C == 1 && var == 3) return launch_rk4<T, AREAL, L, C, 8, true>(d, stage, uin, y, F, acc, out, cn, ca, snap, s);   \
        return launch_rk4<T, AREAL, L, C, 16, false>(d, stage, uin, y, F, acc, out, cn, ca, snap, s);                         \
    }
    RSVDT_CASE(4, 1) RSVDT_CASE(8, 1) RSVDT_CASE(16, 1) RSVDT_CASE(32, 1) RSVDT_CASE(64, 1)
    RSVDT_CASE(4, 2) RSVDT_CASE(8, 2) RSVDT_CASE(16, 2) RSVDT_CASE(32, 2) RSVDT_CASE(64, 2)
    RSVDT_CASE(4, 4) RSVDT_CASE(8, 4) RSVDT_CASE(16, 4) RSVDT_CASE(32, 4) RSVDT_CASE(64, 4)
#undef RSVDT_CASE
    set_error_msg("rsvdt_hip_rk4_stage: unsupported k / lanes combination (k <= 256)");
    return RSVDT_HIP_ERR_ARG;
}

// ---------------------------------------------------------------------------
// Horner stage.  For a linear operator the classical RK4 step is the degree-4
// Taylor polynomial of hA applied to y plus a forcing term that is linear in
// the three stage forcings:
//   y_{i+1} = P(hA) y_i + g_i,  P(z) = 1 + z(1 + z/2(1 + z/3(1 + z/4))),
// evaluated as four SpMM stages that each read y and one gathered vector:
//   v1 = y + (h/4) A y ;  v2 = y + (h/3) A v1 ;  v3 = y + (h/2) A v2 ;
//   y' = y + h A v3 + g_i
// One stage: out_b = ys_b * y_b + c * (A vin_b) (+ g_b), ys_b a complex scale
// per block (1 when ys is null; the forcing precompute uses it), optional
// FP64 copy of out into snap.
template <typename T, bool AREAL, int LPR, int CPL, int NB, bool NT, int BS = 256>
__global__ __launch_bounds__(BS) void k_horner_stage(
    int nrows, int k, int xcd_remap, int64_t bstride,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, const void *__restrict__ vals_,
    const typename vec2<T>::type *__restrict__ vin, const typename vec2<T>::type *yv,
    const double2 *__restrict__ ys, const typename vec2<T>::type *__restrict__ g,
    typename vec2<T>::type *out, double c, double2 *__restrict__ snap, const int *__restrict__ snap_slot,
    int64_t snap_stride)
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    constexpr int RPB = BS / LPR;
    const int bx = xcd_block(blockIdx.x, gridDim.x, xcd_remap);
    const int lane = threadIdx.x & (LPR - 1);
    const int row = bx * RPB + (int)(threadIdx.x / LPR);
    if (row >= nrows) return;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const AV *vals = (const AV *)vals_;
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
    if (snap_slot != nullptr) {                 // snapshot slot read at run time (graph replays)
        const int sl = *snap_slot;
        snap = sl >= 0 ? snap + (int64_t)sl * snap_stride : nullptr;
    }

    V yr[CPL], gr[CPL];
    // when y is also the gathered vector (first stage) its lines are reused
    // by the gathers: no nontemporal hint on it then
    const bool nty = NT && (const void *)yv != (const void *)vin;
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = lane + t * LPR;
        const int64_t idx = boff + (int64_t)row * k + j;
        yr[t].x = 0; yr[t].y = 0; gr[t] = yr[t];
        if (j < k) {
            yr[t] = ld_stream(yv + idx, nty);
            if (g != nullptr) gr[t] = ld_stream(g + idx, NT);
        }
    }
    double ysr = 1.0, ysi = 0.0;
    if (ys != nullptr) { ysr = ys[blockIdx.y].x; ysi = ys[blockIdx.y].y; }
    double sr[CPL], si[CPL];
#pragma unroll
    for (int t = 0; t < CPL; ++t) { sr[t] = 0; si[t] = 0; }
    row_gather<T, AREAL, LPR, CPL, NB>(e0, e1, lane, k, colidx, vals, vin + boff, sr, si);
#pragma unroll
    for (int t = 0; t < CPL; ++t) {
        const int j = lane + t * LPR;
        if (j >= k) continue;
        const int64_t idx = boff + (int64_t)row * k + j;
        double yx = (double)yr[t].x, yy = (double)yr[t].y;
        if (ys != nullptr) {
            const double tx = ysr * yx - ysi * yy;
            yy = ysr * yy + ysi * yx;
            yx = tx;
        }
        double ox = yx + c * sr[t], oy = yy + c * si[t];
        if (g != nullptr) { ox += (double)gr[t].x; oy += (double)gr[t].y; }
        V o;
        o.x = (T)ox;
        o.y = (T)oy;
        out[idx] = o;                         // gathered by the next stage: keep it cacheable
        if (snap != nullptr) st_stream(snap + idx, make_double2((double)o.x, (double)o.y), NT);
    }
}

template <typename T, bool AREAL, int LPR, int CPL, int NB, bool NT, int BS = 256>
static int launch_horner(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g,
                         double c, void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    using V = typename vec2<T>::type;
    constexpr int RPB = BS / LPR;
    const unsigned gx = ceil_div(d->nrows, RPB);
    const unsigned gy = (unsigned)d->nb;
    const int remap = (d->xcd_remap > 0 && d->xcd_remap <= 3 && gy == 1 && gx >= 8) ? d->xcd_remap : 0;
    hipLaunchKernelGGL((k_horner_stage<T, AREAL, LPR, CPL, NB, NT, BS>), dim3(gx, gy), dim3(BS), 0, s, (int)d->nrows,
                       (int)d->k, remap, d->nrows * d->k, d->rowptr, d->col, d->val, (const V *)vin, (const V *)y,
                       (const double2 *)ys, (const V *)g, (V *)out, c, (double2 *)snap, slot, sstride);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_stage");
    return RSVDT_HIP_OK;
}

// ---------------------------------------------------------------- grouped gathers
// The Horner stage on a row-grouped form of the operator: G consecutive rows
// form a group and the group's distinct columns are listed once, each as a
// word (column << 8 | mask of the group rows that use it), with the values in
// (column, row) order starting at the group's first CSR offset.  Rows of
// multi-field discretisations share the columns of their grid point, so a
// group gathers each shared row of the stage vector once instead of once per
// row (the 5-field 3D operator: 7 instead of 11 gathers per row).  LPR lanes
// (one column each, k <= LPR) own a group; a batch of NB distinct columns is
// fetched cooperatively, its values follow one coalesced load per lane, and
// all NB gathers are in flight before the FMAs.
template <typename T, bool AREAL, int LPR, int G, int NB>
__device__ __forceinline__ void group_batch(int qb, int q1, int lane, int k, int jc, const uint32_t *__restrict__ gword,
                                            const typename aval<T, AREAL>::type *__restrict__ gval, int &vbase,
                                            const typename vec2<T>::type *__restrict__ ub, double (&sr)[G],
                                            double (&si)[G])
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    static_assert(LPR >= NB, "one word per lane");
    constexpr int VPL = (NB * G + LPR - 1) / LPR;     // values each lane fetches per batch
    const int cnt = (q1 - qb) < NB ? (q1 - qb) : NB;
    const uint32_t w = lane < cnt ? gword[qb + lane] : 0u;
    const int pc = __popc(w & 0xffu);
    int incl = pc;                                     // inclusive scan of the row counts
#pragma unroll
    for (int o = 1; o < NB; o <<= 1) {
        const int t = __shfl_up(incl, o, LPR);
        if (lane >= o) incl += t;
    }
    const int excl = incl - pc;
    const int tot = __shfl(incl, NB - 1, LPR);
    AV vv[VPL];
#pragma unroll
    for (int t = 0; t < VPL; ++t) {
        const int e = lane + t * LPR;
        if (e < tot) vv[t] = gval[vbase + e];
        else vv[t] = AV{};
    }
    int cs[NB], os[NB];
    uint32_t ms[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const uint32_t wu = __shfl(w, u, LPR);
        cs[u] = (int)(wu >> 8);
        ms[u] = wu & 0xffu;
        os[u] = __shfl(excl, u, LPR);
    }
    V xs[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) xs[u] = ub[(int64_t)cs[u] * k + jc];   // padding words gather row 0
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if (u >= cnt) continue;
        const double xr = (double)xs[u].x, xi = (double)xs[u].y;
        int idx = os[u];
#pragma unroll
        for (int r = 0; r < G; ++r) {
            if (ms[u] & (1u << r)) {          // uniform over the group's lanes
                AV src = vv[0];
#pragma unroll
                for (int t = 1; t < VPL; ++t)
                    if (idx / LPR == t) src = vv[t];
                const AV a = shfl_val(src, idx % LPR, LPR);
                ++idx;
                if constexpr (AREAL) {
                    sr[r] += a * xr;
                    si[r] += a * xi;
                } else {
                    sr[r] += a.x * xr - a.y * xi;
                    si[r] += a.x * xi + a.y * xr;
                }
            }
        }
    }
    vbase += tot;
}

template <typename T, bool AREAL, int LPR, int G>
__global__ __launch_bounds__(256) void k_horner_group(
    int nrows, int ngroups, int k, int xcd_remap, int64_t bstride, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ gptr, const uint32_t *__restrict__ gword, const void *__restrict__ gval_,
    const typename vec2<T>::type *__restrict__ vin, const typename vec2<T>::type *yv,
    const double2 *__restrict__ ys, const typename vec2<T>::type *__restrict__ g,
    typename vec2<T>::type *out, double c, double2 *__restrict__ snap, const int *__restrict__ snap_slot,
    int64_t snap_stride)
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    constexpr int GPB = 256 / LPR, NB = 8;
    const int bx = xcd_block(blockIdx.x, gridDim.x, xcd_remap);
    const int lane = threadIdx.x & (LPR - 1);
    const int grp = bx * GPB + (int)(threadIdx.x / LPR);
    if (grp >= ngroups) return;
    const int r0 = grp * G;
    const int jc = lane < k ? lane : k - 1;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const AV *gval = (const AV *)gval_;
    if (snap_slot != nullptr) {
        const int sl = *snap_slot;
        snap = sl >= 0 ? snap + (int64_t)sl * snap_stride : nullptr;
    }
    const bool nty = (const void *)yv != (const void *)vin;
    V yr[G], gr[G];
#pragma unroll
    for (int r = 0; r < G; ++r) {
        yr[r].x = 0; yr[r].y = 0; gr[r] = yr[r];
        const int row = r0 + r;
        if (row < nrows && lane < k) {
            const int64_t idx = boff + (int64_t)row * k + lane;
            yr[r] = ld_stream(yv + idx, nty);
            if (g != nullptr) gr[r] = ld_stream(g + idx, true);
        }
    }
    double sr[G], si[G];
#pragma unroll
    for (int r = 0; r < G; ++r) { sr[r] = 0; si[r] = 0; }
    const int q0 = gptr[grp], q1 = gptr[grp + 1];
    int vbase = rowptr[r0];
    const V *ub = vin + boff;
    int qb = q0;
    for (; q1 - qb >= NB; qb += NB) group_batch<T, AREAL, LPR, G, NB>(qb, q1, lane, k, jc, gword, gval, vbase, ub, sr, si);
    const int rem = q1 - qb;
    if (rem > NB / 2) group_batch<T, AREAL, LPR, G, NB>(qb, q1, lane, k, jc, gword, gval, vbase, ub, sr, si);
    else if (rem > NB / 4) group_batch<T, AREAL, LPR, G, NB / 2>(qb, q1, lane, k, jc, gword, gval, vbase, ub, sr, si);
    else if (rem > 0) group_batch<T, AREAL, LPR, G, NB / 4>(qb, q1, lane, k, jc, gword, gval, vbase, ub, sr, si);
    double ysr = 1.0, ysi = 0.0;
    if (ys != nullptr) { ysr = ys[blockIdx.y].x; ysi = ys[blockIdx.y].y; }
#pragma unroll
    for (int r = 0; r < G; ++r) {
        const int row = r0 + r;
        if (row >= nrows || lane >= k) continue;
        const int64_t idx = boff + (int64_t)row * k + lane;
        double yx = (double)yr[r].x, yy = (double)yr[r].y;
        if (ys != nullptr) {
            const double tx = ysr * yx - ysi * yy;
            yy = ysr * yy + ysi * yx;
            yx = tx;
        }
        double ox = yx + c * sr[r], oy = yy + c * si[r];
        if (g != nullptr) { ox += (double)gr[r].x; oy += (double)gr[r].y; }
        V o;
        o.x = (T)ox;
        o.y = (T)oy;
        out[idx] = o;
        if (snap != nullptr) st_stream(snap + idx, make_double2((double)o.x, (double)o.y), true);
    }
}

template <typename T, bool AREAL, int LPR, int G>
static int launch_horner_group(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys,
                               const void *g, double c, void *out, double *snap, const int *slot, int64_t sstride,
                               hipStream_t s)
{
    using V = typename vec2<T>::type;
    constexpr int GPB = 256 / LPR;
    const int64_t ngroups = (d->nrows + G - 1) / G;
    const unsigned gx = ceil_div(ngroups, GPB);
    const unsigned gy = (unsigned)d->nb;
    const int remap = (d->xcd_remap > 0 && d->xcd_remap <= 3 && gy == 1 && gx >= 8) ? d->xcd_remap : 0;
    hipLaunchKernelGGL((k_horner_group<T, AREAL, LPR, G>), dim3(gx, gy), dim3(256), 0, s, (int)d->nrows,
                       (int)ngroups, (int)d->k, remap, d->nrows * d->k, d->rowptr, d->gptr,
                       (const uint32_t *)d->gword, d->gval, (const V *)vin, (const V *)y, (const double2 *)ys,
                       (const V *)g, (V *)out, c, (double2 *)snap, slot, sstride);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_stage");
    return RSVDT_HIP_OK;
}

template <typename T, bool AREAL>
static int dispatch_horner_group(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys,
                                 const void *g, double c, void *out, double *snap, const int *slot, int64_t sstride,
                                 hipStream_t s)
{
    int lpr = d->lanes;
    if (lpr == 0) {
        lpr = 16;
        while (lpr < d->k && lpr < 64) lpr <<= 1;
    }
#define RSVDT_GCASE(L, GG)                                                                                       \
    if (lpr == L && d->group == GG) return launch_horner_group<T, AREAL, L, GG>(d, vin, y, ys, g, c, out, snap,   \
                                                                               slot, sstride, s);
#define RSVDT_GCASES(L) RSVDT_GCASE(L, 2) RSVDT_GCASE(L, 3) RSVDT_GCASE(L, 4) RSVDT_GCASE(L, 5) RSVDT_GCASE(L, 8)
    if (d->k <= lpr) {
        RSVDT_GCASES(16) RSVDT_GCASES(32) RSVDT_GCASES(64)
    }
#undef RSVDT_GCASES
#undef RSVDT_GCASE
    set_error_msg("rsvdt_hip_horner_stage: grouped variant needs group in {2,3,4,5,8}, 16..64 lanes and k <= lanes");
    return RSVDT_HIP_ERR_ARG;
}

// ---------------------------------------------------------------- FP32 column pairs
// The FP32 Horner stage with each lane owning two adjacent columns, so that
// every gather is one 16-byte load (a float4: two complex FP32 values)
// instead of two 8-byte ones: half the load instructions per gathered byte.
// LPR = k/2 lanes per row (k even, k/2 a power of two), 8-nonzero batches,
// nontemporal row-local streams.
template <bool AREAL, int LPR, int NB, int PPL>
__device__ __forceinline__ void gather_batch_pair(int eb, int e1, int lane, int k, const int32_t *__restrict__ colidx,
                                                  const typename aval<float, AREAL>::type *__restrict__ vals,
                                                  const float4 *__restrict__ ub4, double (&sr)[2 * PPL],
                                                  double (&si)[2 * PPL])
{
    using AV = typename aval<float, AREAL>::type;
    constexpr int EPL = LPR >= NB ? 1 : NB / LPR;
    int mc[EPL];
    AV mv[EPL];
#pragma unroll
    for (int u = 0; u < EPL; ++u) {
        const int e = eb + lane + u * LPR;
        const bool ok = (LPR <= NB || lane < NB) && e < e1;
        mc[u] = ok ? colidx[e] : 0;
        if (ok) mv[u] = vals[e];
        else mv[u] = AV{};
    }
    const int cnt = (e1 - eb) < NB ? (e1 - eb) : NB;
    int cs[NB];
    AV as[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        cs[u] = __shfl(mc[u / LPR < EPL ? u / LPR : 0], u % LPR, LPR);
        as[u] = shfl_val(mv[u / LPR < EPL ? u / LPR : 0], u % LPR, LPR);
    }
    const int k2 = k >> 1;
    float4 xs[NB][PPL];
#pragma unroll
    for (int u = 0; u < NB; ++u)
#pragma unroll
        for (int p = 0; p < PPL; ++p) xs[u][p] = ub4[(int64_t)cs[u] * k2 + lane + p * LPR];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        if (u < cnt) {
#pragma unroll
            for (int p = 0; p < PPL; ++p) {
                const double x0r = xs[u][p].x, x0i = xs[u][p].y, x1r = xs[u][p].z, x1i = xs[u][p].w;
                if constexpr (AREAL) {
                    sr[2 * p] += as[u] * x0r; si[2 * p] += as[u] * x0i;
                    sr[2 * p + 1] += as[u] * x1r; si[2 * p + 1] += as[u] * x1i;
                } else {
                    sr[2 * p] += as[u].x * x0r - as[u].y * x0i; si[2 * p] += as[u].x * x0i + as[u].y * x0r;
                    sr[2 * p + 1] += as[u].x * x1r - as[u].y * x1i; si[2 * p + 1] += as[u].x * x1i + as[u].y * x1r;
                }
            }
        }
    }
}

template <bool AREAL, int LPR, int PPL>
__global__ __launch_bounds__(256) void k_horner_pair(
    int nrows, int k, int xcd_remap, int64_t bstride, const int32_t *__restrict__ rowptr,
    const int32_t *__restrict__ colidx, const void *__restrict__ vals_, const float2 *__restrict__ vin,
    const float2 *yv, const double2 *__restrict__ ys, const float2 *__restrict__ g, float2 *out, double c,
    double2 *__restrict__ snap, const int *__restrict__ snap_slot, int64_t snap_stride)
{
    using AV = typename aval<float, AREAL>::type;
    constexpr int RPB = 256 / LPR, NB = 8;
    const int bx = xcd_block(blockIdx.x, gridDim.x, xcd_remap);
    const int lane = threadIdx.x & (LPR - 1);
    const int row = bx * RPB + (int)(threadIdx.x / LPR);
    if (row >= nrows) return;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const AV *vals = (const AV *)vals_;
    const int e0 = rowptr[row], e1 = rowptr[row + 1];
    if (snap_slot != nullptr) {
        const int sl = *snap_slot;
        snap = sl >= 0 ? snap + (int64_t)sl * snap_stride : nullptr;
    }
    // pair p of the lane: columns 2 (lane + p LPR) and 2 (lane + p LPR) + 1
    const int64_t idx0 = boff + (int64_t)row * k + 2 * lane;
    const bool nty = (const void *)yv != (const void *)vin;
    typedef float ev4f __attribute__((ext_vector_type(4)));
    float4 yr[PPL], gr[PPL];
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
        const int64_t idx = idx0 + 2 * p * LPR;
        const ev4f yv4 = nty ? __builtin_nontemporal_load((const ev4f *)(yv + idx)) : *(const ev4f *)(yv + idx);
        ev4f gv4 = {0.f, 0.f, 0.f, 0.f};
        if (g != nullptr) gv4 = __builtin_nontemporal_load((const ev4f *)(g + idx));
        yr[p] = make_float4(yv4.x, yv4.y, yv4.z, yv4.w);
        gr[p] = make_float4(gv4.x, gv4.y, gv4.z, gv4.w);
    }
    double sr[2 * PPL], si[2 * PPL];
#pragma unroll
    for (int t = 0; t < 2 * PPL; ++t) { sr[t] = 0; si[t] = 0; }
    const float4 *ub4 = (const float4 *)(vin + boff);
    int eb = e0;
    for (; e1 - eb >= NB; eb += NB) gather_batch_pair<AREAL, LPR, NB, PPL>(eb, e1, lane, k, colidx, vals, ub4, sr, si);
    const int rem = e1 - eb;
    if (rem > NB / 2) gather_batch_pair<AREAL, LPR, NB, PPL>(eb, e1, lane, k, colidx, vals, ub4, sr, si);
    else if (rem > NB / 4) gather_batch_pair<AREAL, LPR, NB / 2, PPL>(eb, e1, lane, k, colidx, vals, ub4, sr, si);
    else if (rem > 0) gather_batch_pair<AREAL, LPR, NB / 4, PPL>(eb, e1, lane, k, colidx, vals, ub4, sr, si);
#pragma unroll
    for (int p = 0; p < PPL; ++p) {
        const int64_t idx = idx0 + 2 * p * LPR;
        double y0r = yr[p].x, y0i = yr[p].y, y1r = yr[p].z, y1i = yr[p].w;
        if (ys != nullptr) {
            const double a = ys[blockIdx.y].x, b = ys[blockIdx.y].y;
            double t = a * y0r - b * y0i; y0i = a * y0i + b * y0r; y0r = t;
            t = a * y1r - b * y1i; y1i = a * y1i + b * y1r; y1r = t;
        }
        float4 o;
        o.x = (float)(y0r + c * sr[2 * p] + (double)gr[p].x);
        o.y = (float)(y0i + c * si[2 * p] + (double)gr[p].y);
        o.z = (float)(y1r + c * sr[2 * p + 1] + (double)gr[p].z);
        o.w = (float)(y1i + c * si[2 * p + 1] + (double)gr[p].w);
        *(float4 *)(out + idx) = o;
        if (snap != nullptr) {
            st_stream(snap + idx, make_double2((double)o.x, (double)o.y), true);
            st_stream(snap + idx + 1, make_double2((double)o.z, (double)o.w), true);
        }
    }
}

template <bool AREAL, int L, int PPL = 1>
static int launch_horner_pair(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g,
                              double c, void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    constexpr int RPB = 256 / L;
    const unsigned gx = ceil_div(d->nrows, RPB);
    const unsigned gy = (unsigned)d->nb;
    const int remap = (d->xcd_remap > 0 && d->xcd_remap <= 3 && gy == 1 && gx >= 8) ? d->xcd_remap : 0;
    hipLaunchKernelGGL((k_horner_pair<AREAL, L, PPL>), dim3(gx, gy), dim3(256), 0, s, (int)d->nrows, (int)d->k, remap,
                       d->nrows * d->k, d->rowptr, d->col, d->val, (const float2 *)vin, (const float2 *)y,
                       (const double2 *)ys, (const float2 *)g, (float2 *)out, c, (double2 *)snap, slot, sstride);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_stage");
    return RSVDT_HIP_OK;
}

template <typename T, bool AREAL>
static int dispatch_horner(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g,
                           double c, void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    if (d->variant == 6) {
        if (!d->group || !d->gptr || !d->gword || !d->gval) {
            set_error_msg("rsvdt_hip_horner_stage: variant 6 needs the grouped operator arrays");
            return RSVDT_HIP_ERR_ARG;
        }
        return dispatch_horner_group<T, AREAL>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
    }
    if (d->variant == 7) {
        if constexpr (sizeof(T) == 4) {
            const int64_t l = d->k / 2;
            if (d->k % 2 == 0) {
                if (l == 4) return launch_horner_pair<AREAL, 4>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
                if (l == 8) return launch_horner_pair<AREAL, 8>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
                if (l == 16) return launch_horner_pair<AREAL, 16>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
                if (l == 32) return launch_horner_pair<AREAL, 32>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
            }
        }
        set_error_msg("rsvdt_hip_horner_stage: variant 7 needs the FP32 state and k/2 a power of two in 4..32");
        return RSVDT_HIP_ERR_ARG;
    }
    if (d->variant == 8) {
        if constexpr (sizeof(T) == 4) {
            const int64_t l = d->k / 4;
            if (d->k % 4 == 0) {
                if (l == 4) return launch_horner_pair<AREAL, 4, 2>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
                if (l == 8) return launch_horner_pair<AREAL, 8, 2>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
                if (l == 16) return launch_horner_pair<AREAL, 16, 2>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
            }
        }
        set_error_msg("rsvdt_hip_horner_stage: variant 8 needs the FP32 state and k/4 a power of two in 4..16");
        return RSVDT_HIP_ERR_ARG;
    }
    int lpr = d->lanes;
    const int k = (int)d->k;
    if (lpr == 0) {
        lpr = 4;
        while (lpr < k && lpr < 64) lpr <<= 1;
    }
    const int cpl = (k + lpr - 1) / lpr;
    const int var = d->variant;
#define RSVDT_HCASE(L, C)                                                                                         \
    if (lpr == L && cpl <= C) {                                                                                   \
        if constexpr (C <= 2 && L >= 16) {                                                                        \
            if (var == 4) return launch_horner<T, AREAL, L, C, 12, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);     \
            if (var == 5) return launch_horner<T, AREAL, L, C, 12, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);      \
        }                                                                                                         \
        if (var == 1) return launch_horner<T, AREAL, L, C, 16, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);          \
        if (var == 2) return launch_horner<T, AREAL, L, C, 8, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);          \
        if (var == 3) return launch_horner<T, AREAL, L, C, 8, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);           \
        return launch_horner<T, AREAL, L, C, 16, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);                        \
    }
    RSVDT_HCASE(4, 1) RSVDT_HCASE(8, 1) RSVDT_HCASE(16, 1) RSVDT_HCASE(32, 1) RSVDT_HCASE(64, 1)
    RSVDT_HCASE(4, 2) RSVDT_HCASE(8, 2) RSVDT_HCASE(16, 2) RSVDT_HCASE(32, 2) RSVDT_HCASE(64, 2)
    RSVDT_HCASE(4, 4) RSVDT_HCASE(8, 4) RSVDT_HCASE(16, 4) RSVDT_HCASE(32, 4) RSVDT_HCASE(64, 4)
#undef RSVDT_HCASE
    set_error_msg("rsvdt_hip_horner_stage: unsupported k / lanes combination (k <= 256)");
    return RSVDT_HIP_ERR_ARG;
}

}  // namespace rsvdt

using namespace rsvdt;

extern "C" int rsvdt_hip_rk4_stage(const rsvdt_spmm_desc *d, int stage, const void *uin, const void *y,
                                   const void *F, void *acc, void *out, double c_next, double c_acc,
                                   double *snap, rsvdt_stream_t s)
{
    if (!d || d->nrows <= 0 || d->k <= 0 || d->nb <= 0 || stage < 1 || stage > 4) {
        set_error_msg("rsvdt_hip_rk4_stage: bad descriptor or stage");
        return RSVDT_HIP_ERR_ARG;
    }
    if (d->nrows > 0x7fffffff || d->nb > 65535) {
        set_error_msg("rsvdt_hip_rk4_stage: N or nb too large");
        return RSVDT_HIP_ERR_ARG;
    }
    if (d->lanes && (d->lanes & (d->lanes - 1))) {
        set_error_msg("rsvdt_hip_rk4_stage: lanes must be a power of two");
        return RSVDT_HIP_ERR_ARG;
    }
    hipStream_t st = as_stream(s);
    if (d->prec == 64) {
        if (d->a_real) return dispatch_lanes<double, true>(d, stage, uin, y, F, acc, out, c_next, c_acc, snap, st);
        return dispatch_lanes<double, false>(d, stage, uin, y, F, acc, out, c_next, c_acc, snap, st);
    }
    if (d->prec == 32) {
        if (d->a_real) return dispatch_lanes<float, true>(d, stage, uin, y, F, acc, out, c_next, c_acc, snap, st);
        return dispatch_lanes<float, false>(d, stage, uin, y, F, acc, out, c_next, c_acc, snap, st);
    }
    set_error_msg("rsvdt_hip_rk4_stage: prec must be 64 or 32");
    return RSVDT_HIP_ERR_ARG;
}

// ---------------------------------------------------------------- fused small-operator stepping
// For operators that fit in LDS together with their state, whole runs of
// Horner steps in one launch: each workgroup owns one column of one block
// (columns never interact in A*Y), stages the CSR matrix and y and the two
// stage vectors of its column in LDS, and a stage costs one barrier instead
// of one kernel launch.  The last stage of step s adds g + s*gstride (the
// folded forcing sample; may be null) and, when step i0+s stores a snapshot
// (SourceCode/SaveSnapshots.c:17-29), writes y to snap[pos] in FP64.  Stage
// coefficients c1..c4 = h/4, h/3, h/2, h.
struct fused_lds {
    size_t val, state, rp, ci, total;
};
__host__ __device__ inline size_t rnd16(size_t b) { return (b + 15) & ~(size_t)15; }
__host__ __device__ inline fused_lds fused_layout(int64_t nrows, int64_t nnz, int a_real, int prec)
{
    fused_lds l;
    l.val = 0;
    l.state = rnd16((size_t)nnz * (a_real ? 8 : 16));
    l.rp = l.state + rnd16((size_t)3 * (size_t)nrows * (prec == 64 ? 16 : 8));
    l.ci = l.rp + rnd16((size_t)(nrows + 1) * 4);
    l.total = l.ci + rnd16((size_t)nnz * 4);
    return l;
}

template <typename T, bool AREAL>
__global__ __launch_bounds__(256) void k_horner_fused(
    int nrows, int nnz, int k, int nsteps, int64_t bstride,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, const void *__restrict__ vals_,
    typename vec2<T>::type *y, const typename vec2<T>::type *__restrict__ g, int64_t gstride,
    double c1, double c2, double c3, double c4,
    double2 *__restrict__ snap, int64_t snap_stride, int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir)
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    extern __shared__ double2 fsm[];
    const fused_lds lay = fused_layout(nrows, nnz, AREAL ? 1 : 0, sizeof(T) == 8 ? 64 : 32);
    char *base = (char *)fsm;
    AV *sv = (AV *)(base + lay.val);
    V *sy = (V *)(base + lay.state);
    V *sa = sy + nrows;
    V *sb = sa + nrows;
    int *srp = (int *)(base + lay.rp);
    int *sci = (int *)(base + lay.ci);
    const int j = blockIdx.x;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const AV *vals = (const AV *)vals_;
    for (int e = threadIdx.x; e < nnz; e += blockDim.x) { sv[e] = vals[e]; sci[e] = colidx[e]; }
    for (int r = threadIdx.x; r <= nrows; r += blockDim.x) srp[r] = rowptr[r];
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) sy[r] = y[boff + (int64_t)r * k + j];
    // snapshot schedule by counters (as in k_horner_fused_reg below)
    int64_t snap_st = nsteps, spos = 0;
    if (snap != nullptr) {
        const int64_t rel0 = i0 - Nt;
        const int64_t first = rel0 >= 1 ? (rel0 + RR - 1) / RR * RR : RR;
        snap_st = first - rel0;
        spos = (first / RR - 1) % Nstore;
    }
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
        int64_t pos = -1;
        if (st == snap_st) {
            pos = dir ? spos : Nstore - 1 - spos;
            spos = spos + 1 == Nstore ? 0 : spos + 1;
            snap_st += RR;
        }
        const V *gs = g != nullptr ? g + (int64_t)st * gstride + boff : nullptr;
#pragma unroll 1
        for (int stage = 0; stage < 4; ++stage) {
            const V *vin = stage == 0 ? sy : (stage == 2 ? sb : sa);
            V *out = stage == 3 ? sy : (stage == 1 ? sb : sa);
            const double cs = stage == 0 ? c1 : (stage == 1 ? c2 : (stage == 2 ? c3 : c4));
            for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
                double sr = 0, si = 0;
                const int e1 = srp[r + 1];
                for (int p = srp[r]; p < e1; ++p) {
                    const V x = vin[sci[p]];
                    const double xr = (double)x.x, xi = (double)x.y;
                    if constexpr (AREAL) {
                        const double a = sv[p];
                        sr += a * xr;
                        si += a * xi;
                    } else {
                        const double2 a = sv[p];
                        sr += a.x * xr - a.y * xi;
                        si += a.x * xi + a.y * xr;
                    }
                }
                const V yv = sy[r];
                double ox = (double)yv.x + cs * sr, oy = (double)yv.y + cs * si;
                if (stage == 3 && gs != nullptr) {
                    const V gv = gs[(int64_t)r * k + j];
                    ox += (double)gv.x;
                    oy += (double)gv.y;
                }
                V o;
                o.x = (T)ox;
                o.y = (T)oy;
                out[r] = o;       // stage 3 writes sy in place: each entry is read and written by one thread
                if (stage == 3 && pos >= 0)
                    snap[pos * snap_stride + (int64_t)r * k + j] = make_double2((double)o.x, (double)o.y);
            }
            __syncthreads();
        }
    }
    for (int r = threadIdx.x; r < nrows; r += blockDim.x) y[boff + (int64_t)r * k + j] = sy[r];
}

// The same with each thread's rows (at most RPT, 512 threads) and their
// nonzeros (at most MAXNZ per row) held in registers for the whole launch:
// a stage is then MAXNZ LDS reads and FMAs per row and one barrier.
template <typename T, bool AREAL, int MAXNZ, int RPT, bool HASG>
__global__ __launch_bounds__(512) void k_horner_fused_reg(
    int nrows, int k, int nsteps, int64_t bstride,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ colidx, const void *__restrict__ vals_,
    typename vec2<T>::type *y, const typename vec2<T>::type *__restrict__ g, int64_t gstride,
    double c1, double c2, double c3, double c4,
    double2 *__restrict__ snap, int64_t snap_stride, int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir)
{
    using V = typename vec2<T>::type;
    using AV = typename aval<T, AREAL>::type;
    extern __shared__ double2 fsm[];
    V *sy = (V *)fsm;
    V *sa = sy + nrows;
    V *sb = sa + nrows;
    const int j = blockIdx.x;
    const int64_t boff = (int64_t)blockIdx.y * bstride;
    const AV *vals = (const AV *)vals_;
    int cl[RPT][MAXNZ];
    AV av[RPT][MAXNZ];
    bool live[RPT];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int r = threadIdx.x + t * blockDim.x;
        live[t] = r < nrows;
        const int e0 = live[t] ? rowptr[r] : 0, n = live[t] ? rowptr[r + 1] - e0 : 0;
#pragma unroll
        for (int u = 0; u < MAXNZ; ++u) {
            cl[t][u] = u < n ? colidx[e0 + u] : 0;      // padding: column 0 with a zero value
            if (u < n) av[t][u] = vals[e0 + u];
            else av[t][u] = AV{};
        }
        if (live[t]) sy[r] = y[boff + (int64_t)r * k + j];
    }
    // Snapshot schedule, advanced by counters: step i = i0 + st with i > Nt and
    // (i - Nt) % RR == 0 stores snapshot (i - Nt)/RR - 1 (mod Nstore), slot
    // reversed for the adjoint.  The divisions run once, not once per step.
    int64_t snap_st = nsteps, spos = 0;
    if (snap != nullptr) {
        const int64_t rel0 = i0 - Nt;
        const int64_t first = rel0 >= 1 ? (rel0 + RR - 1) / RR * RR : RR;
        snap_st = first - rel0;
        spos = (first / RR - 1) % Nstore;
    }
    // The forcing term (HASG: g is not null) of step st + 1 is loaded as soon
    // as step st's is used: its latency overlaps a whole step (barriers wait
    // on LDS only).  The load is unconditional (the last step reloads its own
    // sample) so that nothing has to wait for it before the next step.
    V gcur[RPT];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int r = threadIdx.x + t * blockDim.x;
        gcur[t] = V{};
        if (HASG && live[t]) gcur[t] = g[boff + (int64_t)r * k + j];
    }
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
        int64_t pos = -1;
        if (st == snap_st) {
            pos = dir ? spos : Nstore - 1 - spos;
            spos = spos + 1 == Nstore ? 0 : spos + 1;
            snap_st += RR;
        }
        const V *gn = HASG ? g + (int64_t)(st + 1 < nsteps ? st + 1 : st) * gstride + boff : nullptr;
#pragma unroll
        for (int stage = 0; stage < 4; ++stage) {
            const V *vin = stage == 0 ? sy : (stage == 2 ? sb : sa);
            V *out = stage == 3 ? sy : (stage == 1 ? sb : sa);
            const double cs = stage == 0 ? c1 : (stage == 1 ? c2 : (stage == 2 ? c3 : c4));
#pragma unroll
            for (int t = 0; t < RPT; ++t) {
                if (!live[t]) continue;
                const int r = threadIdx.x + t * blockDim.x;
                double sr = 0, si = 0;
#pragma unroll
                for (int u = 0; u < MAXNZ; ++u) {
                    const V x = vin[cl[t][u]];
                    const double xr = (double)x.x, xi = (double)x.y;
                    if constexpr (AREAL) {
                        sr += av[t][u] * xr;
                        si += av[t][u] * xi;
                    } else {
                        sr += av[t][u].x * xr - av[t][u].y * xi;
                        si += av[t][u].x * xi + av[t][u].y * xr;
                    }
                }
                const V yv = sy[r];
                double ox = (double)yv.x + cs * sr, oy = (double)yv.y + cs * si;
                if (HASG && stage == 3) {
                    ox += (double)gcur[t].x;
                    oy += (double)gcur[t].y;
                    gcur[t] = gn[(int64_t)r * k + j];
                }
                V o;
                o.x = (T)ox;
                o.y = (T)oy;
                out[r] = o;
                if (stage == 3 && pos >= 0)
                    snap[pos * snap_stride + (int64_t)r * k + j] = make_double2((double)o.x, (double)o.y);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int r = threadIdx.x + t * blockDim.x;
        if (live[t]) y[boff + (int64_t)r * k + j] = sy[r];
    }
}

extern "C" int rsvdt_hip_horner_fused_fits(int64_t nrows, int64_t nnz, int a_real, int prec)
{
    if (nrows <= 0 || nnz < 0 || nrows > 0x7fffffff || nnz > 0x7fffffff) return 0;
    return fused_layout(nrows, nnz, a_real, prec).total <= (size_t)RSVDT_FUSED_LDS;
}

extern "C" int rsvdt_hip_horner_fused(const rsvdt_spmm_desc *d, int64_t nnz, int64_t maxrow, void *y, const void *g,
                                      int64_t gstride,
                                      const double *coef, int64_t nsteps, double *snap, int64_t snap_stride,
                                      int64_t i0, int64_t Nt, int64_t RR, int64_t Nstore, int dir, rsvdt_stream_t s)
{
    if (!d || d->nrows <= 0 || d->k <= 0 || d->k > 65535 || d->nb <= 0 || d->nb > 65535 || !y || !coef ||
        (snap && (RR <= 0 || Nstore <= 0)) || (d->prec != 64 && d->prec != 32)) {
        set_error_msg("rsvdt_hip_horner_fused: bad arguments");
        return RSVDT_HIP_ERR_ARG;
    }
    if (!rsvdt_hip_horner_fused_fits(d->nrows, nnz, d->a_real, d->prec)) {
        set_error_msg("rsvdt_hip_horner_fused: the operator and its state do not fit in LDS");
        return RSVDT_HIP_ERR_ARG;
    }
    if (nsteps <= 0) return RSVDT_HIP_OK;
    const dim3 grid((unsigned)d->k, (unsigned)d->nb);
    hipStream_t st0 = as_stream(s);
    if (maxrow >= 0 && maxrow <= 4 && d->nrows <= 1024) {    // rows and nonzeros in registers
        const size_t lds3 = (size_t)3 * (size_t)d->nrows * (d->prec == 64 ? 16 : 8);
#define RSVDT_FUSED_REG_G(T, AR, HG)                                                                              \
        hipLaunchKernelGGL((k_horner_fused_reg<T, AR, 4, 2, HG>), grid, dim3(512), lds3, st0, (int)d->nrows,       \
                           (int)d->k, (int)nsteps, d->nrows * d->k, d->rowptr, d->col, d->val,                      \
                           (typename vec2<T>::type *)y, (const typename vec2<T>::type *)g, gstride, coef[0],       \
                           coef[1], coef[2], coef[3], (double2 *)snap, snap_stride, i0, Nt, RR, Nstore, dir)
#define RSVDT_FUSED_REG(T, AR)                                                                                    \
        do {                                                                                                      \
            if (g != nullptr) RSVDT_FUSED_REG_G(T, AR, true);                                                     \
            else RSVDT_FUSED_REG_G(T, AR, false);                                                                 \
        } while (0)
        if (d->prec == 64) {
            if (d->a_real) RSVDT_FUSED_REG(double, true);
            else RSVDT_FUSED_REG(double, false);
        } else {
            if (d->a_real) RSVDT_FUSED_REG(float, true);
            else RSVDT_FUSED_REG(float, false);
        }
#undef RSVDT_FUSED_REG
#undef RSVDT_FUSED_REG_G
        RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_fused");
        return RSVDT_HIP_OK;
    }
    const size_t lds = fused_layout(d->nrows, nnz, d->a_real, d->prec).total;
    const int64_t bs = d->nrows * d->k;
    hipStream_t st = as_stream(s);
#define RSVDT_FUSED_LAUNCH(T, AR)                                                                                 \
    hipLaunchKernelGGL((k_horner_fused<T, AR>), grid, dim3(256), lds, st, (int)d->nrows, (int)nnz, (int)d->k,      \
                       (int)nsteps, bs, d->rowptr, d->col, d->val, (typename vec2<T>::type *)y,                    \
                       (const typename vec2<T>::type *)g, gstride, coef[0], coef[1], coef[2], coef[3],             \
                       (double2 *)snap, snap_stride, i0, Nt, RR, Nstore, dir)
    if (d->prec == 64) {
        if (d->a_real) RSVDT_FUSED_LAUNCH(double, true);
        else RSVDT_FUSED_LAUNCH(double, false);
    } else {
        if (d->a_real) RSVDT_FUSED_LAUNCH(float, true);
        else RSVDT_FUSED_LAUNCH(float, false);
    }
#undef RSVDT_FUSED_LAUNCH
    RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_fused");
    return RSVDT_HIP_OK;
}

static int check_desc(const rsvdt_spmm_desc *d, const char *who)
{
    if (!d || d->nrows <= 0 || d->k <= 0 || d->nb <= 0 || d->nrows > 0x7fffffff || d->nb > 65535 ||
        (d->lanes && (d->lanes & (d->lanes - 1)))) {
        char msg[160];
        snprintf(msg, sizeof(msg), "%s: bad descriptor (nrows, k, nb, lanes)", who);
        set_error_msg(msg);
        return RSVDT_HIP_ERR_ARG;
    }
    return RSVDT_HIP_OK;
}

extern "C" int rsvdt_hip_horner_stage_ex(const rsvdt_spmm_desc *d, const void *vin, const void *y,
                                         const double *yscale, const void *g, double c, void *out, double *snap,
                                         const int *snap_slot, int64_t snap_stride, rsvdt_stream_t s)
{
    int rc = check_desc(d, "rsvdt_hip_horner_stage");
    if (rc) return rc;
    if (!vin || !y || !out || (snap_slot && !snap)) {
        set_error_msg("rsvdt_hip_horner_stage: vin, y and out (and snap with a slot) are required");
        return RSVDT_HIP_ERR_ARG;
    }
    hipStream_t st = as_stream(s);
    if (d->variant == 9 || d->variant == 10) return horner_tile(d, vin, y, yscale, g, c, out, snap, snap_slot, snap_stride, st);
    if (d->prec == 64) {
        if (d->a_real) return dispatch_horner<double, true>(d, vin, y, yscale, g, c, out, snap, snap_slot, snap_stride, st);
        return dispatch_horner<double, false>(d, vin, y, yscale, g, c, out, snap, snap_slot, snap_stride, st);
    }
    if (d->prec == 32) {
        if (d->a_real) return dispatch_horner<float, true>(d, vin, y, yscale, g, c, out, snap, snap_slot, snap_stride, st);
        return dispatch_horner<float, false>(d, vin, y, yscale, g, c, out, snap, snap_slot, snap_stride, st);
    }
    set_error_msg("rsvdt_hip_horner_stage: prec must be 64 or 32");
    return RSVDT_HIP_ERR_ARG;
}

extern "C" int rsvdt_hip_horner_stage(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *yscale,
                                      const void *g, double c, void *out, double *snap, rsvdt_stream_t s)
{
    return rsvdt_hip_horner_stage_ex(d, vin, y, yscale, g, c, out, snap, nullptr, 0, s);
}

// ---------------------------------------------------------------------------
// Plain block SpMM Out_b = M * In_b for nb blocks (the weight and input/output
// matrices of SourceCode/ApplyWeightMats.c:16-80 and SVDAllFreqs.c:63-64
// applied to all frequencies at once).  M is rows x cols CSR, complex FP64;
// In blocks are cols x k, Out blocks rows x k.
namespace rsvdt {
__global__ __launch_bounds__(256) void k_spmm_blocks(int nrows, int k, int lpr, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ colidx, const double2 *__restrict__ val,
                                                     const double2 *__restrict__ in, int64_t ibs,
                                                     double2 *__restrict__ out, int64_t obs)
{
    const int rpb = 256 / lpr;
    const int lane = threadIdx.x % lpr;
    const int row = blockIdx.x * rpb + threadIdx.x / lpr;
    if (row >= nrows) return;
    const double2 *ib = in + (int64_t)blockIdx.y * ibs;
    double2 *ob = out + (int64_t)blockIdx.y * obs;
    for (int j = lane; j < k; j += lpr) {
        double sr = 0, si = 0;
        for (int e = rowptr[row]; e < rowptr[row + 1]; ++e) {
            const double2 a = val[e], x = ib[(int64_t)colidx[e] * k + j];
            sr += a.x * x.x - a.y * x.y;
            si += a.x * x.y + a.y * x.x;
        }
        ob[(int64_t)row * k + j] = make_double2(sr, si);
    }
}
}  // namespace rsvdt

extern "C" int rsvdt_hip_spmm_blocks(int64_t nrows, int64_t k, int64_t nb, const int32_t *rowptr, const int32_t *col,
                                     const double *val_complex, const double *in, int64_t in_bstride, double *out,
                                     int64_t out_bstride, rsvdt_stream_t s)
{
    if (nrows <= 0 || k <= 0 || nb <= 0) return RSVDT_HIP_OK;
    if (nb > 65535 || nrows > 0x7fffffffLL) { set_error_msg("rsvdt_hip_spmm_blocks: bad size"); return RSVDT_HIP_ERR_ARG; }
    int lpr = 1;
    while (lpr < k && lpr < 64) lpr <<= 1;
    const int rpb = 256 / lpr;
    hipLaunchKernelGGL(k_spmm_blocks, dim3(ceil_div(nrows, rpb), (unsigned)nb), dim3(256), 0, as_stream(s), (int)nrows,
                       (int)k, lpr, rowptr, col, (const double2 *)val_complex, (const double2 *)in, in_bstride,
                       (double2 *)out, out_bstride);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_spmm_blocks");
    return RSVDT_HIP_OK;
}
