// The Horner stage of the RK4 step (out_b = ys_b y_b + c A vin_b (+ g_b),
// the SpMM of SourceCode/TSRK4.c:18-42 on the whole N x k block) on the
// tiled operator form, with the gathered vector staged in LDS.
//
// A row gather reads k contiguous complex values (512 B for k = 32 in
// FP64), and a row of a discretised operator shares most of its columns
// with its neighbours: on the 5-field 3D flow operator a run of 32
// consecutive rows touches about 170 distinct columns for 352 nonzeros.
// Gathering per nonzero moves every such row through the load path 11
// times; here a workgroup owns a tile of TR consecutive rows, loads each of
// the tile's distinct rows of vin once (coalesced, UB loads in flight per
// lane) into LDS, then forms all TR output rows from LDS.  The tile's
// column list, its rows' nonzero slots and values are copied to LDS first
// with one coalesced pass.
//
// Lanes: KC lanes cover KC consecutive columns (one complex value each);
// 256/KC row slots, RPT rows per slot, TR = RPT * 256/KC rows per tile.
// When k > KC the tile is processed in column passes of KC columns, the
// LDS staging repeated per pass.  With tile_bpw > 1 a workgroup forms its
// tile in tile_bpw consecutive blocks, one after the other, as further
// passes over the same staged metadata.  With PF (descriptor variant 10) the next
// pass's first 6 rows per slot are loaded into registers while the current
// pass computes from LDS, and written to LDS after its closing barrier.
#include "common.h"

namespace rsvdt {

template <typename T, bool AREAL> struct tile_aval { using type = double2; };
template <typename T> struct tile_aval<T, true> { using type = double; };

__host__ __device__ constexpr size_t tile_al16(size_t b) { return (b + 15) & ~(size_t)15; }

struct tile_lds {
    size_t x, v, col, rp, li, total;
};

// LDS layout: [x: hmax*KC state values][v: emax values][col: hmax int32]
//             [rp: TR+1 int32][li: emax uint16]
__host__ __device__ inline tile_lds tile_layout(int hmax, int emax, int kc, int tr, size_t vsz, size_t asz)
{
    tile_lds l;
    l.x = 0;
    l.v = tile_al16((size_t)hmax * (size_t)kc * vsz);
    l.col = l.v + tile_al16((size_t)emax * asz);
    l.rp = l.col + tile_al16((size_t)hmax * 4);
    l.li = l.rp + tile_al16((size_t)(tr + 1) * 4);
    l.total = l.li + tile_al16((size_t)emax * 2);
    return l;
}

typedef double tev2d __attribute__((ext_vector_type(2)));
typedef float tev2f __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 tld(const double2 *p, bool nt) {
    if (nt) { tev2d v = __builtin_nontemporal_load((const tev2d *)p); return make_double2(v.x, v.y); }
    return *p;
}
__device__ __forceinline__ float2 tld(const float2 *p, bool nt) {
    if (nt) { tev2f v = __builtin_nontemporal_load((const tev2f *)p); return make_float2(v.x, v.y); }
    return *p;
}
__device__ __forceinline__ void tst(double2 *p, double2 v, bool nt) {
    if (nt) { tev2d w = {v.x, v.y}; __builtin_nontemporal_store(w, (tev2d *)p); }
    else *p = v;
}

template <typename T, bool AREAL, int KC, int RPT, int UB, bool PF>
__global__ __launch_bounds__(256) void k_horner_tile(
    int nrows, int k, int nb, int bpw, int hmax, int emax, int xcd_remap, int64_t bstride,
    const int32_t *__restrict__ rowptr, const int32_t *__restrict__ tptr, const int32_t *__restrict__ tcol,
    const uint16_t *__restrict__ lidx, const void *__restrict__ vals_,
    const typename vec2<T>::type *__restrict__ vin, const typename vec2<T>::type *yv,
    const double2 *__restrict__ ys, const typename vec2<T>::type *__restrict__ g,
    typename vec2<T>::type *out, double c, double2 *__restrict__ snap, const int *__restrict__ snap_slot,
    int64_t snap_stride)
{
    using V = typename vec2<T>::type;
    using AV = typename tile_aval<T, AREAL>::type;
    constexpr int NS = 256 / KC;            // row slots
    constexpr int TR = NS * RPT;            // rows per tile
    extern __shared__ __align__(16) char tsm[];
    const tile_lds lay = tile_layout(hmax, emax, KC, TR, sizeof(V), sizeof(AV));
    V *sx = (V *)(tsm + lay.x);
    AV *sv = (AV *)(tsm + lay.v);
    int *scol = (int *)(tsm + lay.col);
    int *srp = (int *)(tsm + lay.rp);
    uint16_t *sli = (uint16_t *)(tsm + lay.li);

    const int t = xcd_block(blockIdx.x, gridDim.x, xcd_remap);
    const int r0 = t * TR;
    const int r1 = r0 + TR < nrows ? r0 + TR : nrows;
    const int h0 = tptr[t], h = tptr[t + 1] - h0;
    const int e0 = rowptr[r0], ne = rowptr[r1] - e0;
    const int tid = threadIdx.x, lane = tid % KC, rs = tid / KC;
    // passes: bpw blocks of this workgroup x the column passes of a block,
    // block-major; the tile's metadata is staged once for all of them
    const int npc = (k + KC - 1) / KC;
    const int b0 = blockIdx.y * bpw;
    const int npass = (nb - b0 < bpw ? nb - b0 : bpw) * npc;
    const AV *vals = (const AV *)vals_;
    if (snap_slot != nullptr) {                 // snapshot slot read at run time (graph replays)
        const int sl = *snap_slot;
        snap = sl >= 0 ? snap + (int64_t)sl * snap_stride : nullptr;
    }
    // the tile's metadata, coalesced
    for (int i = tid; i < h; i += 256) scol[i] = tcol[h0 + i];
    for (int i = tid; i < ne; i += 256) {
        sv[i] = vals[e0 + i];
        sli[i] = lidx[e0 + i];
    }
    for (int i = tid; i <= r1 - r0; i += 256) srp[i] = rowptr[r0 + i] - e0;
    // when y is also the gathered vector (first stage) its rows were just
    // staged through the caches: no nontemporal hint on it then
    const bool nty = (const void *)yv != (const void *)vin;
    // PF: the first NS * UB distinct rows of the next pass are loaded into
    // registers while the current pass computes from LDS (the rest, if any,
    // one at a time after the barrier)
    V xp[PF ? UB : 1];
    if constexpr (PF) {
        __syncthreads();                        // scol visible
        const V *vb0 = vin + (int64_t)b0 * bstride;
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int i = rs + u * NS;
            xp[u].x = 0; xp[u].y = 0;
            if (i < h && lane < k) xp[u] = vb0[(int64_t)scol[i] * k + lane];
        }
    }

    for (int ps = 0; ps < npass; ++ps) {
        const int bl = b0 + ps / npc, j0 = (ps % npc) * KC;
        const int64_t boff = (int64_t)bl * bstride;
        const V *vb = vin + boff;
        double ysr = 1.0, ysi = 0.0;
        if (ys != nullptr) { ysr = ys[bl].x; ysi = ys[bl].y; }
        const int j = j0 + lane;
        const bool jok = j < k;
        // row-local operands first, so that they are in flight during the staging
        V yr[RPT], gr[RPT];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int row = r0 + rs + q * NS;
            yr[q].x = 0; yr[q].y = 0; gr[q] = yr[q];
            if (row < r1 && jok) {
                const int64_t idx = boff + (int64_t)row * k + j;
                yr[q] = tld(yv + idx, nty);
                if (g != nullptr) gr[q] = tld(g + idx, true);
            }
        }
        if constexpr (PF) {
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int i = rs + u * NS;
                if (i < h) sx[i * KC + lane] = xp[u];
            }
        } else {
            __syncthreads();    // metadata visible; the previous pass's reads of sx are done
        }
        // stage the tile's distinct rows of vin, columns j0 .. j0+KC-1 (with
        // PF only those beyond the first NS * UB, one load at a time)
        constexpr int UL = PF ? 1 : UB;
        for (int i0 = rs + (PF ? NS * UB : 0); i0 < h; i0 += NS * UL) {
            V xb[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const int i = i0 + u * NS;
                xb[u].x = 0; xb[u].y = 0;
                if (i < h && jok) xb[u] = vb[(int64_t)scol[i] * k + j];
            }
#pragma unroll
            for (int u = 0; u < UL; ++u) {
                const int i = i0 + u * NS;
                if (i < h) sx[i * KC + lane] = xb[u];
            }
        }
        __syncthreads();
        if constexpr (PF) {                     // the next pass: next columns, or the next block
            const int pn = ps + 1;
            const int jn = (pn % npc) * KC + lane;
            const bool more = pn < npass && jn < k;
            const V *vbn = more ? vin + (int64_t)(b0 + pn / npc) * bstride : vin;
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int i = rs + u * NS;
                xp[u].x = 0; xp[u].y = 0;
                if (i < h && more) xp[u] = vbn[(int64_t)scol[i] * k + jn];
            }
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int lr = rs + q * NS;
            const int row = r0 + lr;
            if (row >= r1) continue;
            double sr = 0, si = 0;
            const int eb = srp[lr], ee = srp[lr + 1];
            for (int e = eb; e < ee; ++e) {
                const V x = sx[(int)sli[e] * KC + lane];
                const double xr = (double)x.x, xi = (double)x.y;
                const AV a = sv[e];
                if constexpr (AREAL) {
                    sr += a * xr;
                    si += a * xi;
                } else {
                    sr += a.x * xr - a.y * xi;
                    si += a.x * xi + a.y * xr;
                }
            }
            if (!jok) continue;
            const int64_t idx = boff + (int64_t)row * k + j;
            double yx = (double)yr[q].x, yy = (double)yr[q].y;
            if (ys != nullptr) {
                const double tx = ysr * yx - ysi * yy;
                yy = ysr * yy + ysi * yx;
                yx = tx;
            }
            double ox = yx + c * sr, oy = yy + c * si;
            if (g != nullptr) { ox += (double)gr[q].x; oy += (double)gr[q].y; }
            V o;
            o.x = (T)ox;
            o.y = (T)oy;
            out[idx] = o;                       // gathered by the next stage: keep it cacheable
            if (snap != nullptr) tst(snap + idx, make_double2((double)o.x, (double)o.y), true);
        }
        if constexpr (PF) __syncthreads();      // this pass's reads of sx are done
    }
}

template <typename T, bool AREAL, int KC, int RPT, bool PF>
static int launch_tile(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g,
                       double c, void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    using V = typename vec2<T>::type;
    using AV = typename tile_aval<T, AREAL>::type;
    constexpr int TR = (256 / KC) * RPT;
    if (d->tile_rows != TR) {
        set_error_msg("rsvdt_hip_horner_stage: tile_rows does not match the launch shape");
        return RSVDT_HIP_ERR_ARG;
    }
    const tile_lds lay = tile_layout(d->tile_hmax, d->tile_emax, KC, TR, sizeof(V), sizeof(AV));
    if (lay.total > 160 * 1024) {
        set_error_msg("rsvdt_hip_horner_stage: a tile's columns do not fit in LDS (use fewer tile rows)");
        return RSVDT_HIP_ERR_ARG;
    }
    const int bpw = d->tile_bpw > 1 ? d->tile_bpw : 1;
    const unsigned gx = ceil_div(d->nrows, TR);
    const unsigned gy = ceil_div(d->nb, bpw);
    const int remap = (d->xcd_remap > 0 && d->xcd_remap <= 3 && gy == 1 && gx >= 8) ? d->xcd_remap : 0;
    hipLaunchKernelGGL((k_horner_tile<T, AREAL, KC, RPT, PF ? 6 : 8, PF>), dim3(gx, gy), dim3(256), lay.total, s,
                       (int)d->nrows, (int)d->k, (int)d->nb, bpw, d->tile_hmax, d->tile_emax, remap,
                       d->nrows * d->k, d->rowptr,
                       d->tptr, d->tcol, d->lidx, d->val, (const V *)vin, (const V *)y, (const double2 *)ys,
                       (const V *)g, (V *)out, c, (double2 *)snap, slot, sstride);
    RSVDT_LAUNCH_CHECK("rsvdt_hip_horner_stage(tiled)");
    return RSVDT_HIP_OK;
}

template <typename T, bool AREAL>
static int dispatch_tile(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g,
                         double c, void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    int kc = d->lanes;
    if (kc == 0) kc = d->k <= 8 ? 8 : 16;
    const int rpt = d->tile_rows * kc / 256;
#define RSVDT_TCASE(KC, RPT)                                                                                 \
    if (kc == KC && rpt == RPT)                                                                              \
        return d->variant == 10 ? launch_tile<T, AREAL, KC, RPT, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s) \
                                : launch_tile<T, AREAL, KC, RPT, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
    RSVDT_TCASE(8, 1) RSVDT_TCASE(16, 1) RSVDT_TCASE(16, 2) RSVDT_TCASE(16, 3) RSVDT_TCASE(16, 4)
    RSVDT_TCASE(32, 1) RSVDT_TCASE(32, 2) RSVDT_TCASE(32, 4)
#undef RSVDT_TCASE
    set_error_msg("rsvdt_hip_horner_stage: tiled variant needs lanes 8, 16 or 32 and tile_rows = (256/lanes) x "
                  "{1, 2, 3, 4} (lanes 8: 32 rows; 32 lanes: 8, 16 or 32 rows)");
    return RSVDT_HIP_ERR_ARG;
}

int horner_tile(const rsvdt_spmm_desc *d, const void *vin, const void *y, const double *ys, const void *g, double c,
                void *out, double *snap, const int *slot, int64_t sstride, hipStream_t s)
{
    if (d->tile_bpw < 0 || d->tile_bpw > 65535) {
        set_error_msg("rsvdt_hip_horner_stage: tile_bpw must be 0..65535");
        return RSVDT_HIP_ERR_ARG;
    }
    if (!d->tptr || !d->tcol || !d->lidx || d->tile_rows <= 0 || d->tile_hmax <= 0 || d->tile_emax < 0 ||
        d->tile_hmax > 65536) {
        set_error_msg("rsvdt_hip_horner_stage: variant 9 needs the tiled operator arrays");
        return RSVDT_HIP_ERR_ARG;
    }
    if (d->prec == 64) {
        if (d->a_real) return dispatch_tile<double, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
        return dispatch_tile<double, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
    }
    if (d->prec == 32) {
        if (d->a_real) return dispatch_tile<float, true>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
        return dispatch_tile<float, false>(d, vin, y, ys, g, c, out, snap, slot, sstride, s);
    }
    set_error_msg("rsvdt_hip_horner_stage: prec must be 64 or 32");
    return RSVDT_HIP_ERR_ARG;
}

}  // namespace rsvdt
