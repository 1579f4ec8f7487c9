// ct2_project.hip -- Joseph forward projection, one ray per lane.
//
// A 256-thread workgroup covers a 4 (u) x 64 (v) tile of one view: each
// wavefront is 64 consecutive detector rows v of one column u, so its rays
// sample neighbouring voxels (v runs along z for circular orbits, and z is
// the fastest volume axis) and its sinogram stores are one coalesced row.
//
// Reference: emil_code/view_proj3D.f:703-910 (linear), :915-1092 (nearest),
// view_proj3D_tomo.f:4-148 (z forced), ray_proj3D.f:514-804 (ray lists).
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>

#include <type_traits>

#include "ct2_internal.h"

namespace ct2 {

// Volumes addressed with 24-bit index multiplies and 32-bit byte offsets:
// every stride and index below 2^24 and the volume below 4 GiB.  The check
// is uniform (kernel arguments), so each ray picks its loop once.
template <typename VT>
__device__ __forceinline__ bool grid_small(const Grid& g) {
    // every stride and dimension a signed 24-bit operand, every element
    // within 32-bit byte offsets (the largest n * s spans the volume, in any
    // axis order)
    unsigned long long ext = 0;
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        ok &= g.s[q] < (1 << 23) && g.n[q] < (1 << 23);
        const unsigned long long e = (unsigned long long)g.n[q] * (unsigned)g.s[q];
        ext = e > ext ? e : ext;
    }
    return ok && ext * sizeof(VT) < (1ull << 32);
}

// Element offset of cell (ib, ic) of a plane (SMALL: 24-bit multiplies).
template <bool SMALL>
__device__ __forceinline__ int cell_offset(int ib, int sb, int ic, int sc) {
    return SMALL ? __mul24(ib, sb) + __mul24(ic, sc) : ib * sb + ic * sc;
}

// March one ray and return its sum; reference FP32 order (sequential sum
// over planes), so the result is bit-identical to the Fortran.
template <int INTERP, typename VT, bool SMALL>
__device__ __forceinline__ float march_exact(const AxisSel& A, const VT* __restrict__ vol,
                                             bool strict) {
#pragma clang fp contract(off)
    float total = 0.0f;
    int base = 0;   // l * sa
    for (int l = 0; l < A.na; ++l, base += A.sa) {
        Sample sm;
        if (INTERP == CT2_LINEAR) {
            if (!sample_linear_exact(A, l, sm)) continue;
            int i00 = base + cell_offset<SMALL>(sm.ib, A.sb, sm.ic, A.sc);
            float v00 = Store<VT>::ld(vol, i00);
            float v10 = Store<VT>::ld(vol, i00 + A.sb);
            float v01 = Store<VT>::ld(vol, i00 + A.sc);
            float v11 = Store<VT>::ld(vol, i00 + A.sb + A.sc);
            float vi1 = sm.wb1 * v00 + sm.wb2 * v10;
            float vi2 = sm.wb1 * v01 + sm.wb2 * v11;
            float val = sm.wc1 * vi1 + sm.wc2 * vi2;
            total = total + A.trav * val;
        } else {
            if (!sample_nearest_exact(A, l, strict, sm)) continue;
            int i00 = base + cell_offset<SMALL>(sm.ib, A.sb, sm.ic, A.sc);
            total = total + A.trav * Store<VT>::ld(vol, i00);
        }
    }
    return total;
}

#ifndef CT2_FWD_BATCH
#define CT2_FWD_BATCH 2
#endif

// Bilinear value of a cell from its corners (x_bc) in the fast arithmetic:
// for f16 storage (1-wb) v00 + wb v10 with the half operands fed to
// mixed-precision FMAs (no separate conversions), for f32 and bf16
// v00 + wb (v10 - v00); then the same along c.
template <typename VT>
__device__ __forceinline__ float lerp_cell(VT x00, VT x10, VT x01, VT x11, float wb, float wc);

// f16 pairs: each corner's own value (the corners were loaded one by one)
template <>
__device__ __forceinline__ float lerp_cell(h16x2 x00, h16x2 x10, h16x2 x01, h16x2 x11, float wb,
                                           float wc) {
    return lerp_cell<__half>(x00.lo, x10.lo, x01.lo, x11.lo, wb, wc);
}

template <typename VT>
__device__ __forceinline__ float lerp_cell(VT x00, VT x10, VT x01, VT x11, float wb, float wc) {
    float vi1, vi2;
    if constexpr (std::is_same<VT, __half>::value) {
        const float wb1 = 1.0f - wb;
        vi1 = __builtin_fmaf(wb, __half2float(x10), __builtin_fmaf(wb1, __half2float(x00), 0.0f));
        vi2 = __builtin_fmaf(wb, __half2float(x11), __builtin_fmaf(wb1, __half2float(x01), 0.0f));
    } else {
        const float v00 = static_cast<float>(x00), v10 = static_cast<float>(x10);
        const float v01 = static_cast<float>(x01), v11 = static_cast<float>(x11);
        vi1 = __builtin_fmaf(wb, v10 - v00, v00);
        vi2 = __builtin_fmaf(wb, v11 - v01, v01);
    }
    return __builtin_fmaf(wc, vi2 - vi1, vi1);
}

// The fast ray sum: the samples of each batch of CT2_SUMB planes aligned at
// multiples of CT2_SUMB (planes [8m, 8m+8)) are added in plane order in
// FP32, and each batch's sum that holds a sample enters a compensated
// (Kahan) sum.  A batch sum of at most eight samples carries rounding far
// below the ray sum's, and the compensation costs one step per batch
// instead of one per sample.  Every fast forward kernel sums this way, so
// their projections agree bit for bit.
constexpr int CT2_SUMB = 8;

// Kahan step of one batch sum, taken where `any` (the batch held a sample).
__device__ __forceinline__ void kahan_add(float part, float& sum, float& comp, bool any) {
    const float y = part - comp;
    const float ts = sum + y;
    const float cn = (ts - sum) - y;
    sum = any ? ts : sum;
    comp = any ? cn : comp;
}

// march_fast for bilinear sampling of volumes with 32-bit byte offsets: the
// planes are taken UB at a time and all their loads are issued before the
// first is used, so a wavefront keeps 4*UB loads in flight instead of 4.  A
// sample outside the volume loads element 0 and leaves the sum alone, so the
// samples, their arithmetic (per storage type, as in march_fast), their order
// and the compensated sum are exactly those of the one-plane loop.
template <typename VT, bool ZC, int UB = CT2_FWD_BATCH>
__device__ __forceinline__ float march_fast_batched(const RayFast& R, const VT* __restrict__ vol,
                                                    int l0, int l1) {
    static_assert(CT2_SUMB % UB == 0, "plane steps within a batch");
    constexpr unsigned E = (unsigned)sizeof(VT);
    float sum = 0.0f, comp = 0.0f;
    const unsigned hb = (unsigned)(R.nb - 1), hc = (unsigned)(R.nc - 1);
    const unsigned db = (unsigned)R.sb * E;
    const unsigned dc = ZC ? E : (unsigned)R.sc * E;
    for (int m0 = l0 & ~(CT2_SUMB - 1); m0 <= l1; m0 += CT2_SUMB) {
        float part = 0.0f;
        bool any = false;
#pragma unroll
        for (int h = 0; h < CT2_SUMB; h += UB) {
            const int l = m0 + h;
            float ub[UB], uc[UB];
            bool ok[UB];
            unsigned o[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const float lfu = (float)(l + u);   // integer-valued: exact
                ub[u] = __builtin_fmaf(lfu, R.dub, R.ub0);
                uc[u] = __builtin_fmaf(lfu, R.duc, R.uc0);
                const int ib = floor_i(ub[u]), ic = floor_i(uc[u]);
                ok[u] = (l + u >= l0) & (l + u <= l1) & ((unsigned)ib < hb) & ((unsigned)ic < hc);
                const int i00 = (l + u) * R.sa + (ZC ? __mul24(ib, R.sb) + ic
                                                     : cell_offset<true>(ib, R.sb, ic, R.sc));
                o[u] = ok[u] ? (unsigned)i00 * E : 0u;
            }
            // f16 pairs: along a contiguous c axis one aligned 4-byte load gives
            // a corner and its c+1 neighbour
            using XT = typename std::conditional<std::is_same<VT, h16x2>::value, __half, VT>::type;
            XT x00[UB], x10[UB], x01[UB], x11[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                if constexpr (std::is_same<VT, h16x2>::value) {
                    const h16x2 p0 = *at_byte(vol, o[u]), p1 = *at_byte(vol, o[u] + db);
                    x00[u] = p0.lo;
                    x10[u] = p1.lo;
                    if constexpr (ZC) {
                        x01[u] = p0.hi;
                        x11[u] = p1.hi;
                    } else {
                        x01[u] = at_byte(vol, o[u] + dc)->lo;
                        x11[u] = at_byte(vol, o[u] + db + dc)->lo;
                    }
                } else {
                    x00[u] = *at_byte(vol, o[u]);
                    x10[u] = *at_byte(vol, o[u] + db);
                    x01[u] = *at_byte(vol, o[u] + dc);
                    x11[u] = *at_byte(vol, o[u] + db + dc);
                }
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const float val = lerp_cell(x00[u], x10[u], x01[u], x11[u], frac_pos(ub[u]),
                                            frac_pos(uc[u]));
                part += ok[u] ? val : 0.0f;
                any |= ok[u];
            }
        }
        kahan_add(part, sum, comp, any);
    }
    return sum * R.trav;
}

// Fast arithmetic: per-ray setup in FP64, the transverse cell coordinate of
// plane l is one FMA (u(l) = u0 + l*du), FP32 bilinear lerps, compensated sum.
template <int INTERP, typename VT, bool SMALL, bool ZC>
__device__ __forceinline__ float march_fast(const RayFast& R, const VT* __restrict__ vol,
                                            bool strict, int l0, int l1) {
    float sum = 0.0f, comp = 0.0f;
    const unsigned hb = (unsigned)(INTERP == CT2_LINEAR ? R.nb - 1 : (strict ? R.nb - 1 : R.nb));
    const unsigned hc = (unsigned)(INTERP == CT2_LINEAR ? R.nc - 1 : (strict ? R.nc - 1 : R.nc));
    // [l0, l1] is wavefront-uniform: all lanes step the same plane together,
    // so one load instruction touches one plane's neighbourhood
    for (int m0 = l0 & ~(CT2_SUMB - 1); m0 <= l1; m0 += CT2_SUMB) {
        float part = 0.0f;
        bool any = false;
        for (int l = max(m0, l0); l <= min(m0 + CT2_SUMB - 1, l1); ++l) {
            const float lf = (float)l;   // integer-valued: exact
            float ub = __builtin_fmaf(lf, R.dub, R.ub0);
            float uc = __builtin_fmaf(lf, R.duc, R.uc0);
            const int ib = floor_i(ub), ic = floor_i(uc);
            if ((unsigned)ib >= hb || (unsigned)ic >= hc) continue;
            const int base = l * R.sa;
            // ZC: the c axis is the contiguous one (stride 1), so the c+1
            // neighbours sit one element further, an immediate load offset
            int i00 = base + (ZC ? (SMALL ? __mul24(ib, R.sb) : ib * R.sb) + ic
                                 : cell_offset<SMALL>(ib, R.sb, ic, R.sc));
            float val;
            if (INTERP == CT2_LINEAR) {
                // (linear sampling reaches this loop only for volumes beyond
                // 32-bit byte offsets; march_fast_batched covers the rest)
                const int sc = ZC ? 1 : R.sc;
                VT x00, x10, x01, x11;
                if (SMALL) {
                    const unsigned o = (unsigned)i00 * (unsigned)sizeof(VT);
                    const unsigned db = (unsigned)R.sb * (unsigned)sizeof(VT);
                    const unsigned dc = (unsigned)sc * (unsigned)sizeof(VT);
                    x00 = *at_byte(vol, o);
                    x10 = *at_byte(vol, o + db);
                    x01 = *at_byte(vol, o + dc);
                    x11 = *at_byte(vol, o + db + dc);
                } else {
                    x00 = vol[i00];
                    x10 = vol[(size_t)i00 + R.sb];
                    x01 = vol[(size_t)i00 + sc];
                    x11 = vol[(size_t)i00 + R.sb + sc];
                }
                val = lerp_cell(x00, x10, x01, x11, frac_pos(ub), frac_pos(uc));
            } else {
                val = Store<VT>::ld(vol, i00);
            }
            part += val;
            any = true;
        }
        kahan_add(part, sum, comp, any);
    }
    return sum * R.trav;
}

// One lane's ray (view p, detector column j, row k) of a view-indexed
// projection.  Every lane of the wavefront must call it (the fast path
// shares the plane range across the wavefront); lanes with use == false
// return 0.
template <int INTERP, int ARITH, typename VT>
__device__ __forceinline__ float project_lane(const Grid& g, const Det& dt, int p, int j, int k,
                                              bool use, bool zonly,
                                              const float* __restrict__ fv,
                                              const VT* __restrict__ vol) {
    float out = 0.0f;
    if (ARITH == CT2_ARITH_EXACT) {
        if (use) {
            float S[3], B[3];
            view_ray(fv, p, j, k, dt, S, B);
            AxisSel A;
            setup_exact(g, S, B, zonly, A);
            out = grid_small<VT>(g) ? march_exact<INTERP, VT, true>(A, vol, false)
                                    : march_exact<INTERP, VT, false>(A, vol, false);
        }
    } else {
        RayFast R;
        int l0 = 0x7fffffff, l1 = -1;
        if (use) {
            double S[3], B[3];
            view_ray_f64(fv, p, j, k, dt, S, B);
            setup_fast(g, S, B, zonly, INTERP == CT2_LINEAR, R);
            // planes where the ray can be inside the volume
            int a0 = 0, a1 = R.na - 1;
            const float hb = INTERP == CT2_LINEAR ? R.nb - 1 : R.nb;
            const float hc = INTERP == CT2_LINEAR ? R.nc - 1 : R.nc;
            clip_planes_f(R.ub0, R.dub, 0.0f, hb, 0.05f, a0, a1);
            clip_planes_f(R.uc0, R.duc, 0.0f, hc, 0.05f, a0, a1);
            if (a0 <= a1) { l0 = a0; l1 = a1; }
        }
        // union over the wavefront (every lane takes part), in scalar registers
        l0 = wave_min(l0);
        l1 = wave_max(l1);
        // the contiguous-c variant when every used lane's c axis has stride 1
        // (driving axis x or y in the default layout)
        const bool zc = __all(!use || R.sc == 1);
        // a lane whose driving axis is shorter than the union's end stops at
        // its own last plane; when no lane needs that the loop stays uniform
        const bool fits = __all(!use || R.na - 1 >= l1);
        if (use) {
            const bool small = grid_small<VT>(g);
            auto march = [&](int m1) {
                if constexpr (INTERP == CT2_LINEAR) {
                    if (small)
                        return zc ? march_fast_batched<VT, true>(R, vol, l0, m1)
                                  : march_fast_batched<VT, false>(R, vol, l0, m1);
                }
                if (zc)
                    return small ? march_fast<INTERP, VT, true, true>(R, vol, false, l0, m1)
                                 : march_fast<INTERP, VT, false, true>(R, vol, false, l0, m1);
                return small ? march_fast<INTERP, VT, true, false>(R, vol, false, l0, m1)
                             : march_fast<INTERP, VT, false, false>(R, vol, false, l0, m1);
            };
            out = fits ? march(l1) : march(min(l1, R.na - 1));
        }
    }
    return out;
}

// A workgroup covers TU detector columns x 64 rows of one view.
template <int INTERP, int ARITH, typename VT, int TU>
__global__ __launch_bounds__(64 * TU) void k_project_views(Grid g, Det dt, int view0,
                                                           int vstride, int zonly,
                                                           float* __restrict__ sino,
                                                           const int8_t* __restrict__ ind,
                                                           const float* __restrict__ fv,
                                                           const VT* __restrict__ vol) {
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * TU + (threadIdx.x >> 6);
    const int p = view0 + blockIdx.z * vstride;
    const bool active = k < dt.nv && j < dt.nu;
    const size_t r = ((size_t)p * dt.nu + j) * dt.nv + k;
    const bool use = active && (ind == nullptr || ind[r] == 1);
    const float out = project_lane<INTERP, ARITH, VT>(g, dt, p, j, k, use, zonly != 0, fv, vol);
    if (active) sino[r] = out;
}

// ---- batched, persistent forward projection ----
// Work items are detector tiles (TU columns x 64 rows) of every problem's
// views in subset `sub` (views sub + pv*nsub), numbered problem by problem:
// off[p] is the first item of problem p and off[nprob] the total.  The grid
// is the number of resident workgroups; each takes its next item from the
// plan's work queue (queue_next), so the items' uneven costs balance out.
template <int INTERP, int ARITH, int TU, int MODE>
__global__ __launch_bounds__(64 * TU) void k_batch_project(const BatchProb* __restrict__ probs,
                                                           const long long* __restrict__ off,
                                                           int nprob, int sub, int nsub,
                                                           const float* __restrict__ ones,
                                                           unsigned long long* queue) {
    __shared__ long long s_it;
    const long long total = off[nprob];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (;;) {
        const long long it = queue_next(queue, &s_it);
        if (it >= total) break;
        const int pi = batch_find(off, nprob, it);
        const BatchProb& P = probs[pi];
        const long long local = it - off[pi];
        const int ntv = (P.dt.nv + 63) / 64, ntu = (P.dt.nu + TU - 1) / TU;
        const int per_view = ntv * ntu;
        const int pv = (int)(local / per_view);
        const int rem = (int)(local - (long long)pv * per_view);
        const int p = sub + pv * nsub;
        const int j = (rem / ntv) * TU + wave;
        const int k = (rem % ntv) * 64 + lane;
        const bool active = k < P.dt.nv && j < P.dt.nu;
        const size_t r = ((size_t)p * P.dt.nu + j) * P.dt.nv + k;
        const bool used = active && (P.ind == nullptr || P.ind[r] == 1);
        const float* vol = MODE == BFWD_ROWSUM ? ones : P.vol;
        const float out = project_lane<INTERP, ARITH, float>(P.g, P.dt, p, j, k, used,
                                                             P.zonly != 0, P.fv, vol);
        if (active) {
            if (MODE == BFWD_PLAIN) P.proj[r] = out;
            else if (MODE == BFWD_ROWSUM) P.rowsum[r] = out;
            else P.resid[r] = sart_resid(P.sino[r], out, P.rowsum[r], used);
        }
    }
    queue_leave(queue);
}

template <int INTERP, int ARITH, int TU, int MODE>
static const void* batch_project_kernel() {
    return (const void*)&k_batch_project<INTERP, ARITH, TU, MODE>;
}

template <int INTERP, int ARITH, int TU>
static const void* pick_mode(int mode) {
    switch (mode) {
        case BFWD_PLAIN: return batch_project_kernel<INTERP, ARITH, TU, BFWD_PLAIN>();
        case BFWD_ROWSUM: return batch_project_kernel<INTERP, ARITH, TU, BFWD_ROWSUM>();
        case BFWD_RESID: return batch_project_kernel<INTERP, ARITH, TU, BFWD_RESID>();
    }
    return nullptr;
}

template <int INTERP, int ARITH>
static const void* pick_tu(int tu, int mode) {
    switch (tu) {
        case 2: return pick_mode<INTERP, ARITH, 2>(mode);
        case 4: return pick_mode<INTERP, ARITH, 4>(mode);
        case 8: return pick_mode<INTERP, ARITH, 8>(mode);
    }
    return nullptr;
}

static const void* batch_project_fn(int interp, int arith, int tu, int mode) {
    if (interp == CT2_LINEAR)
        return arith == CT2_ARITH_FAST ? pick_tu<CT2_LINEAR, CT2_ARITH_FAST>(tu, mode)
                                       : pick_tu<CT2_LINEAR, CT2_ARITH_EXACT>(tu, mode);
    return arith == CT2_ARITH_FAST ? pick_tu<CT2_NEAREST, CT2_ARITH_FAST>(tu, mode)
                                   : pick_tu<CT2_NEAREST, CT2_ARITH_EXACT>(tu, mode);
}

int batch_project_occupancy(int interp, int arith, int tu, int mode, int* blocks_per_cu) {
    const void* fn = batch_project_fn(interp, arith, tu, mode);
    if (!fn) { set_error("batched projection: no kernel for tile %d, mode %d", tu, mode); return CT2_EINVAL; }
    CT2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, 64 * tu, 0));
    return CT2_OK;
}

int launch_batch_project(int interp, int arith, int tu, int mode, const BatchProb* probs,
                         const long long* off, int nprob, int sub, int nsub, const float* ones,
                         unsigned long long* queue, int grid, hipStream_t st) {
    const void* fn = batch_project_fn(interp, arith, tu, mode);
    if (!fn || grid <= 0 || (grid & 7)) {
        set_error("batched projection: bad launch (tile %d, mode %d, grid %d)", tu, mode, grid);
        return CT2_EINVAL;
    }
    void* args[] = {(void*)&probs, (void*)&off, (void*)&nprob, (void*)&sub, (void*)&nsub,
                    (void*)&ones, (void*)&queue};
    CT2_HIP(hipLaunchKernel(fn, dim3(grid), dim3(64 * tu), args, 0, st));
    return CT2_OK;
}

template <int INTERP>
__global__ __launch_bounds__(256) void k_project_rays(Grid g, int nr, int strict,
                                                      float* __restrict__ sino,
                                                      const int8_t* __restrict__ ind,
                                                      const float* __restrict__ rv,
                                                      const float* __restrict__ vol) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nr) return;
    float out = 0.0f;
    if (ind == nullptr || ind[r] == 1) {
        float S[3], B[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            S[q] = rv[(size_t)r * 6 + q];
            B[q] = rv[(size_t)r * 6 + 3 + q];
        }
        AxisSel A;
        setup_exact(g, S, B, false, A);
        out = grid_small<float>(g) ? march_exact<INTERP, float, true>(A, vol, strict != 0)
                                   : march_exact<INTERP, float, false>(A, vol, strict != 0);
    }
    sino[r] = out;
}

// ---- column-separable fast forward (views whose detector rows run along z) ----
// With e_v parallel to z (view_separable), the rays of detector column j
// share their driving axis a (x or y) and the fast set-up of the other
// transverse axis b: on plane l every one of them samples b cell ib(l) with
// weight wb(l) = frac(fma(l, dub, ub0)).  The fast bilinear sample lerps
// along b first and then along z, so the b lerp of the two voxel rows
// (l, ib, :) and (l, ib+1, :) is the same for every ray of the column:
//     Rw(c) = lerp_b(v(l, ib, c), v(l, ib+1, c)),
// and a ray's sample on plane l is lerp_z(Rw(ic), Rw(ic+1), wc).  One
// wavefront takes one column.  Per batch of planes it blends the rows into
// its own LDS (coalesced row loads, the b lerp once per (column, plane, z)
// instead of once per ray; the next batch's rows are loaded while the
// current one is sampled), then each lane marches RPL rays of the column
// through those rows.  Every sample, its arithmetic and the ray sum (the
// batches of CT2_SUMB planes) are those of march_fast_batched, so the
// projections are bit-identical to k_project_views'.  A wavefront whose view is not
// separable, or whose rays unexpectedly disagree on a, ub0 or dub, runs the
// per-ray march of k_project_views instead (same rows per call, so the same
// results).
constexpr int FS_WAVES = 4;   // detector columns per workgroup, one per wavefront
constexpr int FS_P = 8;       // planes per batch (one row quad per lane and plane: nz <= 256)
static_assert(FS_P == CT2_SUMB, "the row batches are the ray sum's batches");
constexpr int FS_RPLMAX = 8;  // rays per lane at most
// LDS floats per wavefront: the batch's blended rows, then the rays' weights
constexpr int FS_WSLAB = FS_P * 256 + 64 * FS_RPLMAX;

// Four consecutive values of a volume row along z: the raw load, and the
// lerp across b with the fast sample's arithmetic for the storage type (as
// lerp_cell: (1-wb) x0 + wb x1 by mixed-precision FMAs for f16 and for the
// f16 pairs' own values, x0 + wb (x1 - x0) otherwise).
struct __align__(8) h16x4 { __half v[4]; };
struct __align__(8) bf16x4 { hip_bfloat16 v[4]; };

__device__ __forceinline__ float blend_half(__half x0, __half x1, float wb, float wb1) {
    return __builtin_fmaf(wb, __half2float(x1), __builtin_fmaf(wb1, __half2float(x0), 0.0f));
}

template <typename VT> struct RowQuad;
template <> struct RowQuad<float> {
    using raw = float4;
    __device__ __forceinline__ static float4 blend(const raw& a, const raw& b, float wb) {
        return make_float4(__builtin_fmaf(wb, b.x - a.x, a.x), __builtin_fmaf(wb, b.y - a.y, a.y),
                           __builtin_fmaf(wb, b.z - a.z, a.z), __builtin_fmaf(wb, b.w - a.w, a.w));
    }
};
template <> struct RowQuad<__half> {
    using raw = h16x4;
    __device__ __forceinline__ static float4 blend(const raw& a, const raw& b, float wb) {
        const float wb1 = 1.0f - wb;
        return make_float4(blend_half(a.v[0], b.v[0], wb, wb1), blend_half(a.v[1], b.v[1], wb, wb1),
                           blend_half(a.v[2], b.v[2], wb, wb1), blend_half(a.v[3], b.v[3], wb, wb1));
    }
};
template <> struct RowQuad<h16x2> {
    // four pairs as one 16-byte load; each pair's own value is its low half
    using raw = uint4;
    __device__ __forceinline__ static __half lo(unsigned w) {
        return __ushort_as_half((unsigned short)(w & 0xffffu));
    }
    __device__ __forceinline__ static float4 blend(const raw& a, const raw& b, float wb) {
        const float wb1 = 1.0f - wb;
        return make_float4(blend_half(lo(a.x), lo(b.x), wb, wb1), blend_half(lo(a.y), lo(b.y), wb, wb1),
                           blend_half(lo(a.z), lo(b.z), wb, wb1), blend_half(lo(a.w), lo(b.w), wb, wb1));
    }
};
template <> struct RowQuad<hip_bfloat16> {
    using raw = bf16x4;
    __device__ __forceinline__ static float4 blend(const raw& a, const raw& b, float wb) {
        float4 o;
        float* op = &o.x;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float x0 = static_cast<float>(a.v[i]), x1 = static_cast<float>(b.v[i]);
            op[i] = __builtin_fmaf(wb, x1 - x0, x0);
        }
        return o;
    }
};

// Orders one wavefront's LDS writes before its later reads (and its reads
// before later writes) without a workgroup barrier: the hardware keeps a
// wavefront's LDS operations in order, the fences keep the compiler from
// moving them across.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Grid: one workgroup per (view, chunk of 64*RPL rows, FS_WAVES columns),
// ncg column groups and nrc row chunks per view.  Requires nz % 4 == 0,
// nz <= 256, ny*nz < 2^23, a volume below 4 GiB with the default strides
// and a 16-byte aligned base; LDS
// FS_WAVES * FS_WSLAB floats.
// 16-bit rows take half the registers of the batch in flight: four
// wavefronts per SIMD then fit without spills (5.9 against 6.7 ms for the
// f16 volume at 256^3 x 360 x 512^2); 32-bit rows at that limit spill.
#ifndef CT2_FS_WPE16
#define CT2_FS_WPE16 4
#endif
#ifndef CT2_FS_WPE32
#define CT2_FS_WPE32 1
#endif
template <typename VT>
constexpr int fs_waves_per_eu() {
    return sizeof(typename RowQuad<VT>::raw) == 8 ? CT2_FS_WPE16 : CT2_FS_WPE32;
}

template <typename VT, int RPL>
__global__ __launch_bounds__(64 * FS_WAVES)
__attribute__((amdgpu_waves_per_eu(fs_waves_per_eu<VT>()))) void k_project_sep(Grid g, Det dt, int view0,
                                                               int vstride, int ncg, int nrc,
                                                               float* __restrict__ sino,
                                                               const int8_t* __restrict__ ind,
                                                               const float* __restrict__ fv,
                                                               const VT* __restrict__ vol) {
    using RQ = RowQuad<VT>;
    using raw = typename RQ::raw;
    extern __shared__ float s_rows[];
    const unsigned t = xcd_tile(blockIdx.x, gridDim.x);
    const int cg = (int)(t % (unsigned)ncg);
    const unsigned rest = t / (unsigned)ncg;
    const int rc = (int)(rest % (unsigned)nrc), pv = (int)(rest / (unsigned)nrc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = cg * FS_WAVES + wave;
    if (j >= dt.nu) return;   // the whole wavefront; no workgroup barrier follows
    const int p = view0 + pv * vstride;
    const int k0 = rc * 64 * RPL;
    const size_t col = ((size_t)p * dt.nu + j) * dt.nv;

    bool sep = view_separable(fv + (size_t)p * 18, g, dt);
    float uc0[RPL], duc[RPL];
    float* rows = s_rows + (size_t)wave * FS_WSLAB;
    float* trav = rows + FS_P * 256;   // ray weights, lane-private (no sync needed)
    float ub0 = 0.0f, dub = 0.0f;
    int a = 0;
    if (sep) {
        // setup_fast_ax's expressions, split into the column's part and each
        // row's: with e_v = (0, 0, e_vz) the bin's x and y do not depend on
        // the row (fma(0, v, x) = x), so neither do a, ub0 and dub
        const ViewFrame F = load_frame(fv, p);
        const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
        const double dd0 = fma((double)F.eu[0], u, (double)F.d[0]) - (double)F.s[0];
        const double dd1 = fma((double)F.eu[1], u, (double)F.d[1]) - (double)F.s[1];
        const double X2 = fma((double)F.eu[2], u, (double)F.d[2]);
        const double S2 = F.s[2], ev2 = F.ev[2];
        auto ddz = [&](int k) {
            return fma(ev2, fma((double)k + 0.5, (double)dt.dv, (double)dt.v0), X2) - S2;
        };
        a = fast_axis(g, dd0, dd1, ddz(k0), false);   // row k0 exists
        const double dda = a == 0 ? dd0 : dd1, ddb = a == 0 ? dd1 : dd0;
        const float da = a == 0 ? g.d[0] : g.d[1], oa = a == 0 ? g.o[0] : g.o[1];
        const float ob = a == 0 ? g.o[1] : g.o[0];
        const double Sa = a == 0 ? F.s[0] : F.s[1], Sb = a == 0 ? F.s[1] : F.s[0];
        const double ia = rcp_f64(dda);
        const double id0 = rcp_f64((double)g.d[0]), id1 = rcp_f64((double)g.d[1]);
        const double idb = a == 0 ? id1 : id0, idc = rcp_f64((double)g.d[2]);
        const double r1 = ddb * ia;
        const double ta = fma(0.5, (double)da, (double)oa) - Sa;
        ub0 = (float)fma(fma(r1, ta, Sb) - (double)ob, idb, -0.5);
        dub = (float)(r1 * ((double)da * idb));
        const double dacd = (double)da * idc;
        const float r1f = (float)r1;
        const float r1sq = 1.0f + r1f * r1f;
        const float tra = fabsf(da);
        bool agree = true;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int k = k0 + 64 * r + lane;
            const bool active = k < dt.nv;
            const bool use = active && (ind == nullptr || ind[col + k] == 1);
            const double d2 = ddz(k);
            agree &= !active || fast_axis(g, dd0, dd1, d2, false) == a;
            const double r2 = d2 * ia;
            const float r2f = (float)r2;
            // a ray that is not used marches cell 0 of every plane (in
            // bounds) with weight 0 and is stored as 0
            uc0[r] = use ? (float)fma(fma(r2, ta, S2) - (double)g.o[2], idc, -0.5) : 0.5f;
            duc[r] = use ? (float)(r2 * dacd) : 0.0f;
            trav[64 * r + lane] = use ? tra * fast_sqrt(r1sq + r2f * r2f) : 0.0f;
        }
        sep = __all(agree) && a < 2;
    }
    if (!sep) {
        // per-ray march, 64 consecutive rows per call as in k_project_views
#pragma unroll 1
        for (int r = 0; r < RPL; ++r) {
            const int k = k0 + 64 * r + lane;
            const bool active = k < dt.nv;
            const bool use = active && (ind == nullptr || ind[col + k] == 1);
            const float out =
                project_lane<CT2_LINEAR, CT2_ARITH_FAST, VT>(g, dt, p, j, k, use, false, fv, vol);
            if (active) sino[col + k] = out;
        }
        return;
    }
    const int na = a == 0 ? g.n[0] : g.n[1], nb = a == 0 ? g.n[1] : g.n[0];
    // element strides of a and b (below 2^23, and the volume below 4 GiB:
    // signed 24-bit multiplies and 32-bit byte offsets)
    const int sa = a == 0 ? g.s[0] : g.s[1], sb = a == 0 ? g.s[1] : g.s[0];
    const int nc = g.n[2];
    // planes whose b cell is inside the volume: an interval, since
    // fma(l, dub, ub0) is monotonic in l
    int lo = 0x7fffffff, hi = -1;
    for (int l = lane; l < na; l += 64) {
        const int ib = floor_i(__builtin_fmaf((float)l, dub, ub0));
        if ((unsigned)ib < (unsigned)(nb - 1)) { lo = min(lo, l); hi = max(hi, l); }
    }
    const int lb0 = wave_min(lo), lb1 = wave_max(hi);
    float sum[RPL], comp[RPL];
#pragma unroll
    for (int r = 0; r < RPL; ++r) sum[r] = comp[r] = 0.0f;
    const unsigned hc = (unsigned)(nc - 1);
    const int c4 = 4 * lane;   // this lane's z values c4 .. c4+3 of every row
    // the row quads of the batch starting at plane lb (a plane past the
    // volume, or a lane past nz, loads element 0 and is not used)
    raw x0[FS_P], x1[FS_P];
    const char* vb = reinterpret_cast<const char*>(vol);
    const unsigned esz = (unsigned)sizeof(VT);
    auto load_batch = [&](int lb) {
#pragma unroll
        for (int q = 0; q < FS_P; ++q) {
            const int l = lb + q;
            const bool ok = l >= lb0 && l <= lb1 && c4 < nc;
            const int ib = floor_i(__builtin_fmaf((float)l, dub, ub0));
            const unsigned o = ok ? (unsigned)(__mul24(l, sa) + __mul24(ib, sb) + c4) * esz : 0u;
            x0[q] = *reinterpret_cast<const raw*>(vb + o);
            x1[q] = *reinterpret_cast<const raw*>(vb + o + (ok ? (unsigned)sb * esz : 0u));
        }
    };
    // batches aligned at multiples of FS_P = CT2_SUMB planes: the batches of
    // the fast ray sum
    const int lbs = lb0 & ~(FS_P - 1);
    if (lb0 <= lb1) load_batch(lbs);
    for (int lb = lbs; lb <= lb1; lb += FS_P) {
        // planes of the batch whose b cell is inside the volume: [qa, qe]
        const int qa = max(lb0 - lb, 0), qe = min(lb1 - lb, FS_P - 1);
        float lfq[FS_P];   // the batch's plane indices as floats (exact), in scalar registers
#pragma unroll
        for (int q = 0; q < FS_P; ++q) lfq[q] = uniform_f((float)(lb + q));
        // the batch's rows, lerped across b, into LDS
#pragma unroll
        for (int q = 0; q < FS_P; ++q) {
            if (q >= qa && q <= qe && c4 < nc)
                *reinterpret_cast<float4*>(rows + q * nc + c4) =
                    RQ::blend(x0[q], x1[q], frac_pos(__builtin_fmaf(lfq[q], dub, ub0)));
        }
        wave_lds_sync();
        if (lb + FS_P <= lb1) load_batch(lb + FS_P);   // in flight while this batch is sampled
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            // every ray of this slot inside the volume on the batch's first
            // and last plane is inside on every plane between (the valid
            // planes of a ray are an interval): then no sample needs a mask
            const int ia = floor_i(__builtin_fmaf(lfq[0], duc[r], uc0[r]));
            const int ie = floor_i(__builtin_fmaf(lfq[FS_P - 1], duc[r], uc0[r]));
            float part = 0.0f;
            bool any = true;
            if (qa == 0 && qe == FS_P - 1 && __all((unsigned)ia < hc && (unsigned)ie < hc)) {
#pragma unroll
                for (int h = 0; h < FS_P; h += 4) {
                    float val[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float uc = __builtin_fmaf(lfq[h + q], duc[r], uc0[r]);
                        const int ic = floor_i(uc);
                        const float* row = rows + (h + q) * nc;
                        const float vi1 = row[ic], vi2 = row[ic + 1];
                        val[q] = __builtin_fmaf(frac_pos(uc), vi2 - vi1, vi1);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) part += val[q];
                }
            } else {
                any = false;
#pragma unroll
                for (int h = 0; h < FS_P; h += 4) {
                    float val[4];
                    bool ok[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float uc = __builtin_fmaf(lfq[h + q], duc[r], uc0[r]);
                        const int ic = floor_i(uc);
                        ok[q] = (unsigned)ic < hc && h + q >= qa && h + q <= qe;
                        const int icc = ok[q] ? ic : 0;
                        const float* row = rows + (h + q) * nc;
                        const float vi1 = row[icc], vi2 = row[icc + 1];
                        val[q] = __builtin_fmaf(frac_pos(uc), vi2 - vi1, vi1);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        part += ok[q] ? val[q] : 0.0f;
                        any |= ok[q];
                    }
                }
            }
            kahan_add(part, sum[r], comp[r], any);
        }
        wave_lds_sync();
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int k = k0 + 64 * r + lane;
        const float tw = trav[64 * r + lane];
        if (k < dt.nv) sino[col + k] = tw != 0.0f ? sum[r] * tw : 0.0f;
    }
}

// ---------------------------------------------------------------- host ----
template <int INTERP, int ARITH, typename VT, int TU>
static void launch_tile(const Grid& g, const Det& dt, const ViewSet& vs, int zonly, float* sino,
                        const int8_t* ind, const float* fv, const void* vol, hipStream_t st) {
    dim3 grid((dt.nv + 63) / 64, (dt.nu + TU - 1) / TU, vs.nviews);
    hipLaunchKernelGGL((k_project_views<INTERP, ARITH, VT, TU>), grid, dim3(64 * TU), 0, st, g,
                       dt, vs.view0, vs.vstride, zonly, sino, ind, fv, (const VT*)vol);
}

template <int INTERP, int ARITH, typename VT>
static void launch_views(const Grid& g, const Det& dt, const ViewSet& vs, int zonly, int tu,
                         float* sino, const int8_t* ind, const float* fv, const void* vol,
                         hipStream_t st) {
    switch (tu) {
        case 1: launch_tile<INTERP, ARITH, VT, 1>(g, dt, vs, zonly, sino, ind, fv, vol, st); break;
        case 2: launch_tile<INTERP, ARITH, VT, 2>(g, dt, vs, zonly, sino, ind, fv, vol, st); break;
        case 8: launch_tile<INTERP, ARITH, VT, 8>(g, dt, vs, zonly, sino, ind, fv, vol, st); break;
        default: launch_tile<INTERP, ARITH, VT, 4>(g, dt, vs, zonly, sino, ind, fv, vol, st); break;
    }
}

template <int INTERP, int ARITH>
static int dispatch_dtype(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                          int tu, float* sino, const int8_t* ind, const float* fv,
                          const void* vol, hipStream_t st) {
    int z = d->axis == CT2_AXIS_Z;
    switch (d->vol_dtype) {
        case CT2_F32: launch_views<INTERP, ARITH, float>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        case CT2_F16: launch_views<INTERP, ARITH, __half>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        case CT2_BF16: launch_views<INTERP, ARITH, hip_bfloat16>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        case CT2_F16Z2: launch_views<INTERP, ARITH, h16x2>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        // plane-major storage: the same kernels, the grid's strides say the layout
        case CT2_F32_ZPLANES: launch_views<INTERP, ARITH, float>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        case CT2_F16_ZPLANES: launch_views<INTERP, ARITH, __half>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        case CT2_BF16_ZPLANES: launch_views<INTERP, ARITH, hip_bfloat16>(g, dt, vs, z, tu, sino, ind, fv, vol, st); break;
        default: set_error("ct2_project: unknown vol_dtype %d", d->vol_dtype); return CT2_EINVAL;
    }
    return CT2_OK;
}

static int launch_tu(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs, int tu,
                     float* s, const int8_t* ind, const float* fv, const void* vol,
                     hipStream_t st) {
    int rc;
    if (d->arith == CT2_ARITH_EXACT) {
        rc = d->interp == CT2_LINEAR
                 ? dispatch_dtype<CT2_LINEAR, CT2_ARITH_EXACT>(d, g, dt, vs, tu, s, ind, fv, vol, st)
                 : dispatch_dtype<CT2_NEAREST, CT2_ARITH_EXACT>(d, g, dt, vs, tu, s, ind, fv, vol, st);
    } else if (d->arith == CT2_ARITH_FAST) {
        rc = d->interp == CT2_LINEAR
                 ? dispatch_dtype<CT2_LINEAR, CT2_ARITH_FAST>(d, g, dt, vs, tu, s, ind, fv, vol, st)
                 : dispatch_dtype<CT2_NEAREST, CT2_ARITH_FAST>(d, g, dt, vs, tu, s, ind, fv, vol, st);
    } else {
        set_error("ct2_project: unknown arith %d", d->arith);
        return CT2_EINVAL;
    }
    if (rc) return rc;
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// ---- tile autotuner ----
// The first projection of a geometry times tile heights 4, 2, 8 and 1 after
// one untimed launch (every launch writes the same output, so the last one
// leaves the right result) and caches the winner per geometry and device; the
// default of 4 is used while the stream is being captured or with
// CT2_AUTOTUNE=0.  (At 256^3 x 360 views the four lie within 3%; 16 was 5%
// slower.)
struct TileKey {
    int v[14];
    bool operator<(const TileKey& o) const { return memcmp(v, o.v, sizeof v) < 0; }
};
static std::mutex g_tile_mu;
static std::map<TileKey, int>& tile_cache() {
    static std::map<TileKey, int>* m = new std::map<TileKey, int>();
    return *m;
}

// The column-separable forward takes fast linear projections along free
// axes of volumes below 4 GiB with nz % 4 == 0, nz <= 256, ny*nz < 2^23 and
// a 16-byte aligned base
// (CT2_FWD_SEP=0 keeps k_project_views, which gives the same results).
static bool fsep_supported(const ct2_desc* d, const Grid& g, const void* vol) {
    if (d->arith != CT2_ARITH_FAST || d->interp != CT2_LINEAR || d->axis != CT2_AXIS_FREE)
        return false;
    if (is_zplanes(d->vol_dtype)) return false;   // it reads C-ordered z rows
    if (g.n[2] % 4 != 0 || g.n[2] > 256 || ((uintptr_t)vol & 15u) != 0) return false;
    const size_t esz = d->vol_dtype == CT2_F32 || d->vol_dtype == CT2_F16Z2 ? 4 : 2;
    if (g.s[0] >= (1 << 23) || g.n[0] >= (1 << 23) || g.n[1] >= (1 << 23) ||
        (size_t)g.n[0] * (size_t)g.s[0] * esz >= (1ull << 32))
        return false;
    const char* env = getenv("CT2_FWD_SEP");
    return !(env && env[0] == '0');
}

template <typename VT, int RPL>
static void launch_fsep_t(const Grid& g, const Det& dt, const ViewSet& vs, float* sino,
                          const int8_t* ind, const float* fv, const void* vol, hipStream_t st) {
    const int ncg = (dt.nu + FS_WAVES - 1) / FS_WAVES;
    const int nrc = (dt.nv + 64 * RPL - 1) / (64 * RPL);
    const size_t nblk = (size_t)ncg * nrc * vs.nviews;
    const size_t lds = (size_t)FS_WAVES * FS_WSLAB * sizeof(float);
    hipLaunchKernelGGL((k_project_sep<VT, RPL>), dim3((unsigned)nblk), dim3(64 * FS_WAVES), lds,
                       st, g, dt, vs.view0, vs.vstride, ncg, nrc, sino, ind, fv, (const VT*)vol);
}

template <typename VT>
static void launch_fsep_rpl(const Grid& g, const Det& dt, const ViewSet& vs, float* sino,
                            const int8_t* ind, const float* fv, const void* vol,
                            hipStream_t st) {
    if (dt.nv <= 64) launch_fsep_t<VT, 1>(g, dt, vs, sino, ind, fv, vol, st);
    else if (dt.nv <= 128) launch_fsep_t<VT, 2>(g, dt, vs, sino, ind, fv, vol, st);
    else if (dt.nv <= 256) launch_fsep_t<VT, 4>(g, dt, vs, sino, ind, fv, vol, st);
    else launch_fsep_t<VT, 8>(g, dt, vs, sino, ind, fv, vol, st);
}

static int launch_fsep(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                       float* sino, const int8_t* ind, const float* fv, const void* vol,
                       hipStream_t st) {
    if ((size_t)((dt.nu + FS_WAVES - 1) / FS_WAVES) * ((dt.nv + 63) / 64) * vs.nviews >=
        (1ull << 31)) {
        set_error("ct2_project: too many detector tiles");
        return CT2_EINVAL;
    }
    switch (d->vol_dtype) {
        case CT2_F32: launch_fsep_rpl<float>(g, dt, vs, sino, ind, fv, vol, st); break;
        case CT2_F16: launch_fsep_rpl<__half>(g, dt, vs, sino, ind, fv, vol, st); break;
        case CT2_BF16: launch_fsep_rpl<hip_bfloat16>(g, dt, vs, sino, ind, fv, vol, st); break;
        case CT2_F16Z2: launch_fsep_rpl<h16x2>(g, dt, vs, sino, ind, fv, vol, st); break;
        default: set_error("ct2_project: unknown vol_dtype %d", d->vol_dtype); return CT2_EINVAL;
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

static int tuned_project(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                         float* s, const int8_t* ind, const float* fv, const void* vol,
                         hipStream_t st) {
    if (fsep_supported(d, g, vol)) return launch_fsep(d, g, dt, vs, s, ind, fv, vol, st);
    int dev = 0;
    CT2_HIP(hipGetDevice(&dev));
    TileKey key;
    memset(&key, 0, sizeof key);
    int i = 0;
    for (int v : {d->nx, d->ny, d->nz, d->ns, d->nu, d->nv, vs.nviews, d->interp, d->arith,
                  d->axis, d->vol_dtype, dev})
        key.v[i++] = v;
    memcpy(&key.v[i++], &d->dx, sizeof(float));
    key.v[i++] = ind != nullptr;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        auto it = tile_cache().find(key);
        if (it != tile_cache().end()) return launch_tu(d, g, dt, vs, it->second, s, ind, fv, vol, st);
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    CT2_HIP(hipStreamIsCapturing(st, &cs));
    const char* env = getenv("CT2_AUTOTUNE");
    if (cs != hipStreamCaptureStatusNone || (env && env[0] == '0'))
        return launch_tu(d, g, dt, vs, 4, s, ind, fv, vol, st);
    // one untimed launch first, so that no candidate is timed with a cold
    // volume in the caches; then two rounds over the candidates, each scored
    // by its faster launch, so that one noisy launch does not decide the pick
    const int cand[4] = {4, 2, 8, 1};
    constexpr int NROUND = 2, NEV = 2 * 4 * NROUND;
    hipEvent_t ev[NEV];
    for (int c = 0; c < NEV; ++c) CT2_HIP(hipEventCreate(&ev[c]));
    int rc = launch_tu(d, g, dt, vs, cand[0], s, ind, fv, vol, st);
    for (int e = 0; e < 4 * NROUND && rc == CT2_OK; ++e) {
        hipEventRecord(ev[2 * e], st);
        rc = launch_tu(d, g, dt, vs, cand[e % 4], s, ind, fv, vol, st);
        hipEventRecord(ev[2 * e + 1], st);
    }
    int best = 4;
    if (rc == CT2_OK) {
        hipError_t he = hipEventSynchronize(ev[NEV - 1]);
        if (he != hipSuccess) rc = hip_fail(he, "tile autotune");
        float tbest = 1e30f;
        for (int c = 0; c < 4 && rc == CT2_OK; ++c) {
            float tc = 1e30f;
            for (int r = 0; r < NROUND; ++r) {
                float ms = 0;
                hipEventElapsedTime(&ms, ev[2 * (r * 4 + c)], ev[2 * (r * 4 + c) + 1]);
                tc = fminf(tc, ms);
            }
            if (tc < tbest) { tbest = tc; best = cand[c]; }
        }
    }
    for (int c = 0; c < NEV; ++c) hipEventDestroy(ev[c]);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tile_mu);
        tile_cache()[key] = best;
    }
    // the last candidate's launch already wrote the output
    return CT2_OK;
}

bool project_sep_route(const ct2_desc* d, const void* vol) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    return fsep_supported(d, g, vol);
}

int launch_project(const ct2_desc* d, void* sino, const int8_t* ind, const float* fv,
                   const void* vol, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    if (is_zplanes(d->vol_dtype)) {
        // element (i, j, k) at (k*nx + i)*ny + j
        g.s[0] = d->ny;
        g.s[1] = 1;
        g.s[2] = d->nx * d->ny;
    }
    ViewSet vs;
    if ((rc = view_set(d, "ct2_project", vs))) return rc;
    if (!sino || !fv || !vol) { set_error("ct2_project: null pointer"); return CT2_EINVAL; }
    if (d->arith != CT2_ARITH_EXACT && d->arith != CT2_ARITH_FAST) {
        set_error("ct2_project: unknown arith %d", d->arith);
        return CT2_EINVAL;
    }
    return tuned_project(d, g, dt, vs, (float*)sino, ind, fv, vol, st);
}

int launch_ray_project(const ct2_desc* d, float* sino, const int8_t* ind, const float* rv,
                       int nr, const float* vol, int strict, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    int rc;
    if ((rc = check_grid(g))) return rc;
    if (nr < 0) { set_error("ray projection: nr < 0"); return CT2_EINVAL; }
    if (nr == 0) return CT2_OK;
    if (!sino || !rv || !vol) { set_error("ray projection: null pointer"); return CT2_EINVAL; }
    dim3 grid((nr + 255) / 256);
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_project_rays<CT2_LINEAR>, grid, dim3(256), 0, st, g, nr, 0, sino, ind, rv, vol);
    else
        hipLaunchKernelGGL(k_project_rays<CT2_NEAREST>, grid, dim3(256), 0, st, g, nr, strict, sino, ind, rv, vol);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace ct2
