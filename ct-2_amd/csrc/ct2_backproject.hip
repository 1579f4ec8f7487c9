// ct2_backproject.hip -- exact adjoint of the Joseph projector, gathered per
// voxel brick.
//
// The reference scatters every ray's (weight, voxel) list into the volume
// (emil_code/view_proj3D.f:1228-1251, :1355-1360).  Here a workgroup owns a
// brick of voxels held in LDS.  For every view it projects the brick, grown
// to the cells whose corners touch it, onto the detector; only rays through
// that footprint can touch the brick.  Each candidate ray is set up and
// sampled with the same device functions as the forward kernel, so the
// matrix entries are the forward's own, and every sample's corner weights
// that fall inside the brick are added to an FP64 accumulator in LDS.  In
// exact arithmetic the brick is written to HBM once; in fast arithmetic with
// an f32 output two workgroups share a brick's views and each adds its
// half-sum into the cleared output with one global atomic per voxel.
//
// Linear weights per corner are (trav*wb)*wc and the contribution is
// datum*weight, the reference's order; only the summation order differs.
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>

#include "ct2_internal.h"

namespace ct2 {

struct Brick {
    int lo[3];
    int ext[3];
};

// Threads per workgroup of the brick kernels (their launches use it); a
// compile-time stride keeps the candidate walk in scalar registers.
constexpr int BRICK_THREADS = 256;

// Views a brick pass leaves to another route (fast arithmetic only): none,
// those view_separable accepts (column-separable adjoint), or those view_zsep
// accepts under the z-forced projector (plane-separable adjoint).
constexpr int BRICK_SKIP_NONE = 0, BRICK_SKIP_SEP = 1, BRICK_SKIP_ZSEP = 2;

// LDS accumulation in FP64: on gfx950 ds_add_f64 runs at ~9 cycles per
// wavefront instruction per CU while ds_add_f32 takes ~190 (measured), and
// the wider sum is also more accurate.
__device__ __forceinline__ void lds_add(double* p, float v) { atomicAdd(p, (double)v); }

// x / d for 0 <= x < 2^22 and d >= 1, with rd the FP32 reciprocal of d: the
// estimate is off by at most one either way and is corrected exactly.
__device__ __forceinline__ int div_small(int x, int d, float rd) {
    int q = (int)((float)x * rd);
    const int r = x - __mul24(q, d);
    q += (r >= d) - (r < 0);
    return q;
}

// Brick number bid of a volume tiled by bdim (nbr bricks per axis, z
// fastest, so neighbouring workgroups share z).
__device__ __forceinline__ Brick brick_of(const Grid& g, int bid, int3 bdim, int3 nbr) {
    int bz = bid % nbr.z;
    int by = (bid / nbr.z) % nbr.y;
    int bx = bid / (nbr.z * nbr.y);
    Brick br;
    br.lo[0] = bx * bdim.x; br.lo[1] = by * bdim.y; br.lo[2] = bz * bdim.z;
    br.ext[0] = min(bdim.x, g.n[0] - br.lo[0]);
    br.ext[1] = min(bdim.y, g.n[1] - br.lo[1]);
    br.ext[2] = min(bdim.z, g.n[2] - br.lo[2]);
    return br;
}

// LDS slot of local voxel (x, y, z) of a brick accumulator (halo layout).
__device__ __forceinline__ int brick_slot(const Brick& br, int x, int y, int z) {
    const int hy = br.ext[1] + 2, hz = br.ext[2] + 2;
    return (x + 1) * hy * hz + (y + 1) * hz + (z + 1);
}

// ---- pieces shared by the adjoint's accumulation paths ----
// Sample points that can touch brick br lie in the box of voxel centres
// lo-1 .. lo+ext; LDS keeps the brick with a one-voxel halo on every side:
// local voxel (x,y,z), -1 <= x <= ext, lives at org + x*ls[0] + y*ls[1] + z,
// so every corner of a sample whose cell touches the brick has a slot (halo
// slots belong to neighbouring bricks and are dropped at write-back).
struct BrickCtx {
    double box_lo[3], box_hi[3];
    int ls[3];
    int org, nhalo;
};

__device__ __forceinline__ BrickCtx brick_ctx(const Grid& g, const Brick& br) {
    BrickCtx c;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c.box_lo[q] = (double)g.o[q] + ((double)(br.lo[q] - 1) + 0.5) * g.d[q];
        c.box_hi[q] = (double)g.o[q] + ((double)(br.lo[q] + br.ext[q]) + 0.5) * g.d[q];
    }
    const int hy = br.ext[1] + 2, hz = br.ext[2] + 2;
    c.ls[0] = hy * hz; c.ls[1] = hz; c.ls[2] = 1;
    c.org = c.ls[0] + c.ls[1] + 1;
    c.nhalo = (br.ext[0] + 2) * hy * hz;
    return c;
}

// Detector bins [jlo,jhi] x [klo,khi] whose rays can touch the box for view
// frame f (full detector when the box is not entirely in front of the
// source).  Whole workgroup; call after a barrier; ends with a barrier, after
// which s_range holds the range.
__device__ void view_footprint(const Det& dt, const float* __restrict__ f, const BrickCtx& c,
                               int* s_range) {
    __shared__ double s_fp[8][2];
    __shared__ int s_ok[8];
    if (threadIdx.x < 8) {
        int k = threadIdx.x;
        double P[3] = {(k & 1) ? c.box_hi[0] : c.box_lo[0], (k & 2) ? c.box_hi[1] : c.box_lo[1],
                       (k & 4) ? c.box_hi[2] : c.box_lo[2]};
        double u = 0, v = 0;
        int ok = 0;
        project_corner(f, P, u, v, ok);
        s_fp[k][0] = u;
        s_fp[k][1] = v;
        s_ok[k] = ok;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int allok = 1;
        double umin = 1e300, umax = -1e300, vmin = 1e300, vmax = -1e300;
        for (int k = 0; k < 8; ++k) {
            allok &= s_ok[k];
            umin = fmin(umin, s_fp[k][0]); umax = fmax(umax, s_fp[k][0]);
            vmin = fmin(vmin, s_fp[k][1]); vmax = fmax(vmax, s_fp[k][1]);
        }
        int jlo = 0, jhi = dt.nu - 1, klo = 0, khi = dt.nv - 1;
        if (allok) {
            // bin j has centre u0 + (j+0.5) du: the bins whose centre projects
            // into the box's silhouette bounds, with 0.25 bin to spare for
            // the rounding of the reference arithmetic's sample positions
            double jl = (umin - dt.u0) / dt.du - 0.5, jh = (umax - dt.u0) / dt.du - 0.5;
            double kl = (vmin - dt.v0) / dt.dv - 0.5, kh = (vmax - dt.v0) / dt.dv - 0.5;
            if (dt.du < 0) { double t = jl; jl = jh; jh = t; }
            if (dt.dv < 0) { double t = kl; kl = kh; kh = t; }
            jl = fmax(jl, -2.0); jh = fmin(jh, (double)dt.nu + 1.0);
            kl = fmax(kl, -2.0); kh = fmin(kh, (double)dt.nv + 1.0);
            jlo = max(0, (int)ceil(jl - 0.25)); jhi = min(dt.nu - 1, (int)floor(jh + 0.25));
            klo = max(0, (int)ceil(kl - 0.25)); khi = min(dt.nv - 1, (int)floor(kh + 0.25));
        }
        s_range[0] = jlo; s_range[1] = jhi; s_range[2] = klo; s_range[3] = khi;
    }
    __syncthreads();
}

// Footprints of FP_BATCH views at once: thread t projects corner t%8 of view
// t/8 of the batch, the eight corners of a view reduce by lane shuffles, and
// s_range[b] receives view b's detector range as view_footprint computes it.
// Whole workgroup (256 threads); call after a barrier; ends with a barrier.
constexpr int FP_BATCH = 32;
static_assert(FP_BATCH <= 64, "one view per lane of a wavefront");
__device__ void view_footprints(const Det& dt, const float* __restrict__ fv, int view0, int pv0,
                                int nb, int vstride, const BrickCtx& c, int (*s_range)[4]) {
    const int b = threadIdx.x >> 3, k = threadIdx.x & 7;
    double u = 0, v = 0;
    int ok = 0;
    if (b < nb) {
        const int p = view0 + (pv0 + b) * vstride;
        double P[3] = {(k & 1) ? c.box_hi[0] : c.box_lo[0], (k & 2) ? c.box_hi[1] : c.box_lo[1],
                       (k & 4) ? c.box_hi[2] : c.box_lo[2]};
        project_corner(fv + (size_t)p * 18, P, u, v, ok);
    }
    double umin = u, umax = u, vmin = v, vmax = v;
    int allok = ok;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        umin = fmin(umin, __shfl_xor(umin, off, 64));
        umax = fmax(umax, __shfl_xor(umax, off, 64));
        vmin = fmin(vmin, __shfl_xor(vmin, off, 64));
        vmax = fmax(vmax, __shfl_xor(vmax, off, 64));
        allok &= __shfl_xor(allok, off, 64);
    }
    if (b < nb && k == 0) {
        int jlo = 0, jhi = dt.nu - 1, klo = 0, khi = dt.nv - 1;
        if (allok) {
            double jl = (umin - dt.u0) / dt.du - 0.5, jh = (umax - dt.u0) / dt.du - 0.5;
            double kl = (vmin - dt.v0) / dt.dv - 0.5, kh = (vmax - dt.v0) / dt.dv - 0.5;
            if (dt.du < 0) { double t = jl; jl = jh; jh = t; }
            if (dt.dv < 0) { double t = kl; kl = kh; kh = t; }
            jl = fmax(jl, -2.0); jh = fmin(jh, (double)dt.nu + 1.0);
            kl = fmax(kl, -2.0); kh = fmin(kh, (double)dt.nv + 1.0);
            // bins whose centre projects into the box's silhouette bounds,
            // with 0.05 bin to spare for rounding
            jlo = max(0, (int)ceil(jl - 0.05)); jhi = min(dt.nu - 1, (int)floor(jh + 0.05));
            klo = max(0, (int)ceil(kl - 0.05)); khi = min(dt.nv - 1, (int)floor(kh + 0.05));
        }
        s_range[b][0] = jlo; s_range[b][1] = jhi; s_range[b][2] = klo; s_range[b][3] = khi;
    }
    __syncthreads();
}

// Candidate rays of a footprint, in the order lanes take them: every
// KSTRIDE-th bin of a detector row (then the next phase), so neighbouring
// lanes' rays sit about a voxel apart and their four corners mostly hit
// distinct LDS words (duplicate addresses in one ds_add instruction
// serialise).  Measured at 256^3 x 360 views: stride 1 63 ms, 2 50.5 ms,
// 3 53.7 ms, 4 56.3 ms -- wider strides scatter the sinogram reads of a
// wavefront over more cache lines.
constexpr int KSTRIDE = 2;
// The fast path's candidate numbering for KSTRIDE 2, without dead slots: a
// row of nk bins is the mk = ceil(nk/2) even offsets, then the odd ones.
struct CandSet {
    int jlo, klo, khi, row, mk, ncand;
    float rrow;
    bool small;   // candidate numbers below 2^22: division by reciprocal
};

__device__ __forceinline__ CandSet cand_set(int r0, int r1, int r2, int r3) {
    CandSet cs;
    cs.jlo = r0;
    cs.klo = r2;
    cs.khi = r3;
    const int nj = r1 - r0 + 1, nk = r3 - r2 + 1;
    static_assert(KSTRIDE == 2, "two phases per row");
    cs.mk = (nk + 1) / 2;
    cs.row = nk;
    cs.ncand = (nj > 0 && nk > 0) ? nj * cs.row : 0;
    cs.rrow = __builtin_amdgcn_rcpf((float)max(cs.row, 1));
    cs.small = cs.ncand < (1 << 22);
    return cs;
}

// Candidate cidx of view p: detector bin (j, k), ray index r, datum; live is
// false past the end, for bins outside the row, unused rays and zero data.
struct Cand {
    int j, k;
    size_t r;
    float datum;
    bool live;
};

__device__ __forceinline__ Cand cand_at(const CandSet& cs, int cidx, int p, const Det& dt,
                                        const int8_t* __restrict__ ind,
                                        const float* __restrict__ sino) {
    Cand c;
    c.j = 0; c.k = 0; c.r = 0; c.datum = 0.0f;
    c.live = cidx < cs.ncand;
    if (c.live) {
        int jj, q;
        if (cs.small) {   // uniform
            jj = div_small(cidx, cs.row, cs.rrow);
            q = cidx - __mul24(jj, cs.row);
        } else {
            jj = cidx / cs.row;
            q = cidx - jj * cs.row;
        }
        c.j = cs.jlo + jj;
        c.k = cs.klo + (q < cs.mk ? 2 * q : 2 * (q - cs.mk) + 1);
    }
    if (c.live) {
        // ray index: a 24-bit multiply while both factors fit 24 bits and the
        // sinogram has fewer than 2^32 bins
        const int rowi = p * dt.nu + c.j;
        c.r = (rowi < (1 << 24) && dt.nv < (1 << 24) &&
               (unsigned long long)dt.ns * dt.nu * dt.nv < (1ull << 32))
                  ? (size_t)(unsigned)__umul24((unsigned)rowi, (unsigned)dt.nv) + (unsigned)c.k
                  : (size_t)rowi * dt.nv + c.k;
        c.live = ind == nullptr || ind[c.r] == 1;
    }
    if (c.live) {
        c.datum = sino[c.r];
        c.live = c.datum != 0.0f;
    }
    return c;
}

// One lane's candidate ray in fast arithmetic, relative to a brick: its
// brick-relative plane range [t0, t0+tcnt) and the cell windows
// [bl, bl+cntb) x [cl, cl+cntc) (the brick's cells intersected with the
// forward's validity window), so one unsigned compare per axis decides a
// sample; aoff folds the brick corner into the LDS address.
struct LaneRay {
    RayFast R;
    int la0, lsa, lsb, lsc, aoff, t0, tcnt, bl, cntb, cl, cntc;
};

template <int INTERP>
__device__ __forceinline__ void lane_ray_ax(const Grid& g, const double Sd[3], double dd0,
                                            double dd1, double dd2, int a, const Brick& br,
                                            const BrickCtx& bc, LaneRay& L) {
    setup_fast_ax(g, Sd, dd0, dd1, dd2, a, INTERP == CT2_LINEAR, L.R);
    const RayFast& R = L.R;
    L.la0 = sel3i(R.a, br.lo[0], br.lo[1], br.lo[2]);
    const int la1 = L.la0 + sel3i(R.a, br.ext[0], br.ext[1], br.ext[2]);
    const int lob = sel3i(R.b, br.lo[0], br.lo[1], br.lo[2]);
    const int lobext = sel3i(R.b, br.ext[0], br.ext[1], br.ext[2]);
    const int loc = sel3i(R.c, br.lo[0], br.lo[1], br.lo[2]);
    const int locext = sel3i(R.c, br.ext[0], br.ext[1], br.ext[2]);
    L.lsa = sel3i(R.a, bc.ls[0], bc.ls[1], bc.ls[2]);
    L.lsb = sel3i(R.b, bc.ls[0], bc.ls[1], bc.ls[2]);
    L.lsc = sel3i(R.c, bc.ls[0], bc.ls[1], bc.ls[2]);
    L.aoff = -(__mul24(lob, L.lsb) + __mul24(loc, L.lsc));
    // cells that touch the brick: [lob-1, lob+ext) for linear cells (corners
    // ib, ib+1), [lob, lob+ext) for nearest
    const int clo = INTERP == CT2_LINEAR ? -1 : 0;
    int l0 = L.la0, l1 = la1 - 1;
    clip_planes_f(R.ub0, R.dub, (float)(lob + clo), (float)(lob + lobext), 0.05f, l0, l1);
    clip_planes_f(R.uc0, R.duc, (float)(loc + clo), (float)(loc + locext), 0.05f, l0, l1);
    L.t0 = l0 - L.la0;
    L.tcnt = max(0, l1 - l0 + 1);
    // the forward's validity window (march_fast): ib < nb-1 (linear), < nb (nearest)
    const int hb = INTERP == CT2_LINEAR ? R.nb - 1 : R.nb;
    const int hc = INTERP == CT2_LINEAR ? R.nc - 1 : R.nc;
    L.bl = max(0, lob + clo);
    L.cntb = max(0, min(hb - 1, lob + lobext - 1) - L.bl + 1);
    L.cl = max(0, loc + clo);
    L.cntc = max(0, min(hc - 1, loc + locext - 1) - L.cl + 1);
}

// One lane's candidate ray: the fast set-up (bit-identical to the forward's
// setup_fast), clipped to the brick.
template <int INTERP>
__device__ __forceinline__ void lane_ray(const Grid& g, const Det& dt, const ViewFrame& F,
                                         const Cand& c, bool zonly, const Brick& br,
                                         const BrickCtx& bc, LaneRay& L) {
    L.R = RayFast{};
    L.la0 = L.lsa = L.lsb = L.lsc = L.aoff = 0;
    L.t0 = L.tcnt = L.bl = L.cntb = L.cl = L.cntc = 0;
    if (!c.live) return;
    double Sd[3], Bd[3];
    view_ray_frame(F, c.j, c.k, dt, Sd, Bd);
    const double dd0 = Bd[0] - Sd[0], dd1 = Bd[1] - Sd[1], dd2 = Bd[2] - Sd[2];
    lane_ray_ax<INTERP>(g, Sd, dd0, dd1, dd2, fast_axis(g, dd0, dd1, dd2, zonly), br, bc, L);
}

// Union of the wavefront's plane ranges (every lane takes part).
__device__ __forceinline__ void wave_range(const LaneRay& L, int& w0, int& w1) {
    w0 = L.tcnt > 0 ? L.t0 : 0x7fffffff;
    w1 = L.tcnt > 0 ? L.t0 + L.tcnt - 1 : -1;
    // scalar registers, so the plane loop is uniform
    w0 = wave_min(w0);
    w1 = wave_max(w1);
}

// March the lanes' rays over brick planes [w0, w1] together (same LDS plane
// per instruction; planes outside a ray's own range are masked) and hand
// every corner contribution dtr*weight to add(slot, value).  lf is the plane
// index as a float (exact), as the forward evaluates it.
template <int INTERP>
__device__ __forceinline__ void lane_march(const LaneRay& L, float dtr, double* acc, int org,
                                           int w0, int w1) {
    const RayFast& R = L.R;
    float lf = (float)(L.la0 + w0);
    // byte offsets into acc; every sample's offset is non-negative
    const int sa8 = L.lsa * 8, sb8 = L.lsb * 8, sc8 = L.lsc * 8, sbc8 = sb8 + sc8;
    int base = (org + __mul24(w0, L.lsa) + L.aoff) * 8;
    char* const accb = reinterpret_cast<char*>(acc);
    for (int t = w0; t <= w1; ++t, lf += 1.0f, base += sa8) {
        const float ub = __builtin_fmaf(lf, R.dub, R.ub0);
        const float uc = __builtin_fmaf(lf, R.duc, R.uc0);
        const int ib = floor_i(ub), ic = floor_i(uc);
        if (!(((unsigned)(t - L.t0) < (unsigned)L.tcnt) &
              ((unsigned)(ib - L.bl) < (unsigned)L.cntb) &
              ((unsigned)(ic - L.cl) < (unsigned)L.cntc)))
            continue;
        const int a00 = base + __mul24(ib, sb8) + __mul24(ic, sc8);
        __builtin_assume(a00 >= 0);
        double* const p = reinterpret_cast<double*>(accb + a00);
        if (INTERP == CT2_LINEAR) {
            const float wb = frac_pos(ub), wc = frac_pos(uc);
            const float tb0 = dtr * (1.0f - wb), tb1 = dtr * wb;
            const float w1c = 1.0f - wc;
            // cells in [lob-1, lob+ext): all four corners have a slot
            lds_add(p, tb0 * w1c);
            lds_add(reinterpret_cast<double*>(accb + (a00 + sb8)), tb1 * w1c);
            lds_add(reinterpret_cast<double*>(accb + (a00 + sc8)), tb0 * wc);
            lds_add(reinterpret_cast<double*>(accb + (a00 + sbc8)), tb1 * wc);
        } else {
            lds_add(p, dtr);
        }
    }
}

// Accumulate A^T sino over views view0 + pv*vstride (pv < nviews) for the
// voxels of brick br into the FP64 LDS accumulator acc (zeroed here).  Called
// by the whole workgroup; ends with a barrier, after which acc holds the sums.
// With skip (fast arithmetic only), the views another route takes are left
// out (BRICK_SKIP_SEP: those view_separable accepts; BRICK_SKIP_ZSEP: those
// view_zsep accepts).  Returns whether any view was taken (the same in every
// thread).
template <int INTERP, int ARITH>
__device__ bool brick_accumulate(const Grid& g, const Det& dt, int view0, int nviews,
                                 int vstride, bool zonly, const Brick& br, double* acc,
                                 const float* __restrict__ sino,
                                 const int8_t* __restrict__ ind,
                                 const float* __restrict__ fv, int skip = 0) {
    __shared__ int s_range[4];
    const BrickCtx bc = brick_ctx(g, br);
    const int* ls = bc.ls;
    const int org = bc.org;
    for (int q = threadIdx.x; q < bc.nhalo; q += BRICK_THREADS) acc[q] = 0.0;
    const double cell_lo = (INTERP == CT2_LINEAR) ? -1.0 : 0.0;

    if constexpr (ARITH == CT2_ARITH_FAST) {
        // footprints a batch of views at a time: between batches no barrier,
        // so a wavefront moves on to the next view as soon as its own
        // candidates are done (the LDS accumulation is atomic)
        __shared__ int s_ranges[FP_BATCH][4];
        const int lane = threadIdx.x & 63;
        bool any = false;
        for (int pv0 = 0; pv0 < nviews; pv0 += FP_BATCH) {
            const int nb = min(FP_BATCH, nviews - pv0);
            // views the separable kernel takes: lane b classifies view b; every
            // wavefront finds the same mask, so skipping a batch is uniform
            unsigned long long sepm = 0ull;
            if (skip) {
                const float* fb = fv + (size_t)(view0 + (pv0 + lane) * vstride) * 18;
                const bool sb = lane < nb && (skip == BRICK_SKIP_ZSEP ? view_zsep(fb, g, dt)
                                                                      : view_separable(fb, g, dt));
                sepm = __ballot(sb);
                if (__popcll(sepm) == nb) continue;
            }
            any = true;
            __syncthreads();   // the previous batch's readers of s_ranges are done
            view_footprints(dt, fv, view0, pv0, nb, vstride, bc, s_ranges);
            // lane b holds view b's footprint; each view's is taken from it
            // with v_readlane, into scalar registers
            int fr0 = 0, fr1 = -1, fr2 = 0, fr3 = -1;
            if (lane < nb) {
                fr0 = s_ranges[lane][0]; fr1 = s_ranges[lane][1];
                fr2 = s_ranges[lane][2]; fr3 = s_ranges[lane][3];
            }
            for (int b = 0; b < nb; ++b) {
                if ((sepm >> b) & 1ull) continue;
                const int p = view0 + (pv0 + b) * vstride;
                const int r0 = __builtin_amdgcn_readlane(fr0, b);
                const int r1 = __builtin_amdgcn_readlane(fr1, b);
                const int r2 = __builtin_amdgcn_readlane(fr2, b);
                const int r3 = __builtin_amdgcn_readlane(fr3, b);
                if (r1 < r0 || r3 < r2) continue;
                const CandSet cs = cand_set(r0, r1, r2, r3);
                const ViewFrame F = load_frame(fv, p);
                const int first = __builtin_amdgcn_readfirstlane((int)threadIdx.x - lane);
                for (int c0 = first; c0 < cs.ncand; c0 += BRICK_THREADS) {
                    const Cand c = cand_at(cs, c0 + lane, p, dt, ind, sino);
                    LaneRay L;
                    lane_ray<INTERP>(g, dt, F, c, zonly, br, bc, L);
                    int w0, w1;
                    wave_range(L, w0, w1);
                    if (w0 > w1) continue;   // wavefront-uniform
                    lane_march<INTERP>(L, c.datum * L.R.trav, acc, org, w0, w1);
                }
            }
        }
        __syncthreads();
        return any;
    }
    for (int pv = 0; pv < nviews; ++pv) {
        const int p = view0 + pv * vstride;
        __syncthreads();   // previous view's readers of s_range are done
        view_footprint(dt, fv + (size_t)p * 18, bc, s_range);
        const int jlo = s_range[0], jhi = s_range[1], klo = s_range[2], khi = s_range[3];
        if (jhi < jlo || khi < klo) continue;
        const int nk = khi - klo + 1;
        const int mk = (nk + KSTRIDE - 1) / KSTRIDE;
        const int row = KSTRIDE * mk;
        const int ncand = (jhi - jlo + 1) * row;
        for (int cidx = threadIdx.x; cidx < ncand; cidx += BRICK_THREADS) {
            const int j = jlo + cidx / row;
            const int q = cidx % row;
            const int k = klo + (q % mk) * KSTRIDE + q / mk;
            if (k > khi) continue;
            const size_t r = ((size_t)p * dt.nu + j) * dt.nv + k;
            if (ind != nullptr && ind[r] != 1) continue;
            const float datum = sino[r];
            if (datum == 0.0f) continue;
            float S[3], B[3];
            view_ray(fv, p, j, k, dt, S, B);
            AxisSel A;
            setup_exact(g, S, B, zonly, A);
            const int a = A.a, b = (a == 0) ? 1 : 0, c = (a == 2) ? 1 : 2;
            // select per-axis brick scalars without indexing a local array
            const int la0 = sel3i(a, br.lo[0], br.lo[1], br.lo[2]);
            const int la1 = la0 + sel3i(a, br.ext[0], br.ext[1], br.ext[2]);
            const int lob = sel3i(b, br.lo[0], br.lo[1], br.lo[2]);
            const int lobext = sel3i(b, br.ext[0], br.ext[1], br.ext[2]);
            const int loc = sel3i(c, br.lo[0], br.lo[1], br.lo[2]);
            const int locext = sel3i(c, br.ext[0], br.ext[1], br.ext[2]);
            const int lsa = sel3i(a, ls[0], ls[1], ls[2]);
            const int lsb = sel3i(b, ls[0], ls[1], ls[2]);
            const int lsc = sel3i(c, ls[0], ls[1], ls[2]);
            // FP64 estimate of the cell coordinate per plane, only to skip
            // planes whose cell cannot touch the brick; the exact FP32 test
            // below decides every sample
            int l0 = la0, l1 = la1 - 1;
            {
                const double half = (INTERP == CT2_LINEAR) ? 0.5 : 0.0;
                const double t0 = (double)A.oa + 0.5 * (double)A.da - (double)A.Sa;
                const double ub0 = ((double)A.Sb + (double)A.r1 * t0 - A.ob) / A.db - half;
                const double uc0 = ((double)A.Sc + (double)A.r2 * t0 - A.oc) / A.dc - half;
                const double dub = (double)A.r1 * A.da / A.db, duc = (double)A.r2 * A.da / A.dc;
                clip_planes(ub0, dub, lob + cell_lo, lob + lobext, 0.05, l0, l1);
                clip_planes(uc0, duc, loc + cell_lo, loc + locext, 0.05, l0, l1);
            }
            for (int l = la0; l < la1; ++l) {
                if (l < l0 || l > l1) continue;
                Sample sm;
                if (INTERP == CT2_LINEAR) {
                    if (!sample_linear_exact(A, l, sm)) continue;
                    const int qb = sm.ib - lob, qc = sm.ic - loc;
                    // cell corners qb, qb+1 and qc, qc+1 must meet the brick
                    if (qb < -1 || qb >= lobext || qc < -1 || qc >= locext) continue;
                    const int base = org + (l - la0) * lsa;
                    float w;
                    {
                        if (qb >= 0 && qc >= 0) {
                            w = A.trav * sm.wb1 * sm.wc1;
                            lds_add(&acc[base + qb * lsb + qc * lsc], datum * w);
                        }
                        if (qb + 1 < lobext && qc >= 0) {
                            w = A.trav * sm.wb2 * sm.wc1;
                            lds_add(&acc[base + (qb + 1) * lsb + qc * lsc], datum * w);
                        }
                        if (qb >= 0 && qc + 1 < locext) {
                            w = A.trav * sm.wb1 * sm.wc2;
                            lds_add(&acc[base + qb * lsb + (qc + 1) * lsc], datum * w);
                        }
                        if (qb + 1 < lobext && qc + 1 < locext) {
                            w = A.trav * sm.wb2 * sm.wc2;
                            lds_add(&acc[base + (qb + 1) * lsb + (qc + 1) * lsc], datum * w);
                        }
                    }
                } else {
                    if (!sample_nearest_exact(A, l, false, sm)) continue;
                    const int qb = sm.ib - lob, qc = sm.ic - loc;
                    if (qb < 0 || qb >= lobext || qc < 0 || qc >= locext) continue;
                    lds_add(&acc[org + (l - la0) * lsa + qb * lsb + qc * lsc], datum * A.trav);
                }
            }
        }
    }
    __syncthreads();
    return nviews > 0;
}

// LDS bytes of a brick accumulator (FP64 slots, one-voxel halo)
size_t brick_lds_bytes(int3 b) { return (size_t)(b.x + 2) * (b.y + 2) * (b.z + 2) * sizeof(double); }

// skip: leave out the views another route takes (BRICK_SKIP_SEP,
// BRICK_SKIP_ZSEP) and ADD the brick's sums to vol (after that route's
// kernels wrote it), if any view was taken; requires nsplit = 1 and an f32
// output.
template <int INTERP, int ARITH, typename OT>
__global__ __launch_bounds__(BRICK_THREADS) void k_backproject_brick(Grid g, Det dt, int view0,
                                                           int nviews, int vstride, int zonly,
                                                           int3 bdim, int3 nbr, int nsplit,
                                                           int skip, OT* __restrict__ vol,
                                                           const float* __restrict__ sino,
                                                           const int8_t* __restrict__ ind,
                                                           const float* __restrict__ fv,
                                                           const int* __restrict__ nonsep) {
    // the pass after the column-separable route: nothing to add when that
    // route found every view separable
    if (nonsep != nullptr && *nonsep == 0) return;
    extern __shared__ double acc[];
    // nsplit workgroups per brick, each over its own share of the views
    const int part = blockIdx.x % nsplit;
    const Brick br = brick_of(g, blockIdx.x / nsplit, bdim, nbr);
    const int pv_lo = part * nviews / nsplit, pv_hi = (part + 1) * nviews / nsplit;
    const bool any = brick_accumulate<INTERP, ARITH>(g, dt, view0 + pv_lo * vstride,
                                                     pv_hi - pv_lo, vstride, zonly != 0, br, acc,
                                                     sino, ind, fv, skip);
    if (skip && !any) return;   // uniform: vol already holds these voxels' sums
    // write the brick: z-rows are contiguous in both LDS and HBM
    const int nloc = br.ext[0] * br.ext[1] * br.ext[2];
    for (int q = threadIdx.x; q < nloc; q += BRICK_THREADS) {
        int lz = q % br.ext[2];
        int ly = (q / br.ext[2]) % br.ext[1];
        int lx = q / (br.ext[2] * br.ext[1]);
        size_t gi = (size_t)(br.lo[0] + lx) * g.s[0] + (size_t)(br.lo[1] + ly) * g.s[1] +
                    (size_t)(br.lo[2] + lz) * g.s[2];
        if constexpr (std::is_same<OT, float>::value) {
            if (nsplit > 1) {
                atomicAdd(vol + gi, (float)acc[brick_slot(br, lx, ly, lz)]);
                continue;
            }
            if (skip) {   // one workgroup per brick: a plain add
                vol[gi] += (float)acc[brick_slot(br, lx, ly, lz)];
                continue;
            }
        }
        Store<OT>::st(vol, gi, (float)acc[brick_slot(br, lx, ly, lz)]);
    }
}

// Zero the voxels of a volume laid out with the grid's strides (a kernel, not
// hipMemsetAsync, so that a captured graph replays it correctly).
__global__ void k_zero_strided(Grid g, float* __restrict__ vol) {
    const size_t n = (size_t)g.n[0] * g.n[1] * g.n[2];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(q % g.n[2]);
        const int y = (int)((q / g.n[2]) % g.n[1]);
        const int x = (int)(q / ((size_t)g.n[2] * g.n[1]));
        vol[(size_t)x * g.s[0] + (size_t)y * g.s[1] + (size_t)z * g.s[2]] = 0.0f;
    }
}

// ---- batched, persistent adjoint ----
// Work items are the bricks of every problem (off[p] = first brick of
// problem p, off[nprob] = total; problems with no view in the subset have
// no items), taken by the resident workgroups from the plan's work queue as
// in the batched forward.  Modes: BBP_PLAIN adj = A^T sino over all views;
// BBP_COLSUM colsum_sub = A_sub^T resid (resid holds ones); BBP_SART
// vol = sart_update(vol, lambda, A_sub^T resid, colsum_sub).
// Final value c of voxel gi of problem P in batched mode MODE.
template <int MODE>
__device__ __forceinline__ void batch_store(const BatchProb& P, size_t gi, int sub, float lambda,
                                            float c) {
    if (MODE == BBP_PLAIN) {
        P.adj[gi] = c;
    } else if (MODE == BBP_COLSUM) {
        P.colsum[(size_t)sub * P.nvox + gi] = c;
    } else {
        P.vol[gi] = sart_update(P.vol[gi], lambda, c, P.colsum[(size_t)sub * P.nvox + gi]);
    }
}

template <int INTERP, int ARITH, int MODE>
__global__ __launch_bounds__(BRICK_THREADS) void k_batch_backproject(const BatchProb* __restrict__ probs,
                                                           const long long* __restrict__ off,
                                                           int nprob, int sub, int nsub,
                                                           int3 bdim, float lambda,
                                                           const float* __restrict__ lambda_dev,
                                                           unsigned long long* queue) {
    extern __shared__ double acc[];
    __shared__ long long s_it;
    if (MODE == BBP_SART && lambda_dev != nullptr) lambda = *lambda_dev;
    const long long total = off[nprob];
    for (;;) {
        const long long it = queue_next(queue, &s_it);
        if (it >= total) break;
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= it) lo = mid; else hi = mid - 1;
        }
        const BatchProb& P = probs[lo];
        const Brick br = brick_of(P.g, (int)(it - off[lo]), bdim, P.nbr);
        const int v0 = MODE == BBP_PLAIN ? 0 : sub;
        const int vstr = MODE == BBP_PLAIN ? 1 : nsub;
        const int nv = MODE == BBP_PLAIN ? P.dt.ns : subset_views(P.dt.ns, sub, nsub);
        const float* src = MODE == BBP_PLAIN ? P.sino : P.resid;
        brick_accumulate<INTERP, ARITH>(P.g, P.dt, v0, nv, vstr, P.zonly != 0, br, acc, src,
                                        P.ind, P.fv);
        const int nloc = br.ext[0] * br.ext[1] * br.ext[2];
        for (int q = threadIdx.x; q < nloc; q += BRICK_THREADS) {
            int lz = q % br.ext[2];
            int ly = (q / br.ext[2]) % br.ext[1];
            int lx = q / (br.ext[2] * br.ext[1]);
            size_t gi = (size_t)(br.lo[0] + lx) * P.g.s[0] + (size_t)(br.lo[1] + ly) * P.g.s[1] +
                        (size_t)(br.lo[2] + lz) * P.g.s[2];
            batch_store<MODE>(P, gi, sub, lambda, (float)acc[brick_slot(br, lx, ly, lz)]);
        }
        __syncthreads();   // write-back done before the next brick clears acc
    }
    queue_leave(queue);
}

template <int INTERP, int ARITH>
static const void* bbp_mode(int mode) {
    switch (mode) {
        case BBP_PLAIN: return (const void*)&k_batch_backproject<INTERP, ARITH, BBP_PLAIN>;
        case BBP_COLSUM: return (const void*)&k_batch_backproject<INTERP, ARITH, BBP_COLSUM>;
        case BBP_SART: return (const void*)&k_batch_backproject<INTERP, ARITH, BBP_SART>;
    }
    return nullptr;
}

static const void* batch_backproject_fn(int interp, int arith, int mode) {
    if (interp == CT2_LINEAR)
        return arith == CT2_ARITH_FAST ? bbp_mode<CT2_LINEAR, CT2_ARITH_FAST>(mode)
                                       : bbp_mode<CT2_LINEAR, CT2_ARITH_EXACT>(mode);
    return arith == CT2_ARITH_FAST ? bbp_mode<CT2_NEAREST, CT2_ARITH_FAST>(mode)
                                   : bbp_mode<CT2_NEAREST, CT2_ARITH_EXACT>(mode);
}

int batch_backproject_occupancy(int interp, int arith, int mode, int3 bdim, int* blocks_per_cu) {
    const void* fn = batch_backproject_fn(interp, arith, mode);
    if (!fn) { set_error("batched backprojection: no kernel for mode %d", mode); return CT2_EINVAL; }
    CT2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, 256,
                                                         brick_lds_bytes(bdim)));
    return CT2_OK;
}

int launch_batch_backproject(int interp, int arith, int mode, const BatchProb* probs,
                             const long long* off, int nprob, int sub, int nsub, int3 bdim,
                             float lambda, const float* lambda_dev, unsigned long long* queue,
                             int grid, hipStream_t st) {
    const void* fn = batch_backproject_fn(interp, arith, mode);
    size_t lds = brick_lds_bytes(bdim);
    if (!fn || grid <= 0 || (grid & 7) || lds > 63 * 1024) {
        set_error("batched backprojection: bad launch (mode %d, grid %d, brick %dx%dx%d)", mode,
                  grid, bdim.x, bdim.y, bdim.z);
        return CT2_EINVAL;
    }
    void* args[] = {(void*)&probs, (void*)&off, (void*)&nprob, (void*)&sub, (void*)&nsub,
                    (void*)&bdim, (void*)&lambda, (void*)&lambda_dev, (void*)&queue};
    CT2_HIP(hipLaunchKernel(fn, dim3(grid), dim3(BRICK_THREADS), args, lds, st));
    return CT2_OK;
}

// Ray-list adjoint: no detector structure to bound a footprint, so rays
// scatter straight into the zero-filled volume with float atomics.
template <int INTERP>
__global__ __launch_bounds__(256) void k_backproject_rays(Grid g, int nr, int strict,
                                                          float* __restrict__ vol,
                                                          const float* __restrict__ sino,
                                                          const int8_t* __restrict__ ind,
                                                          const float* __restrict__ rv) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nr) return;
    if (ind != nullptr && ind[r] != 1) return;
    const float datum = sino[r];
    float S[3], B[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S[q] = rv[(size_t)r * 6 + q];
        B[q] = rv[(size_t)r * 6 + 3 + q];
    }
    AxisSel A;
    setup_exact(g, S, B, false, A);
    for (int l = 0; l < A.na; ++l) {
        Sample sm;
        if (INTERP == CT2_LINEAR) {
            if (!sample_linear_exact(A, l, sm)) continue;
            int i00 = l * A.sa + sm.ib * A.sb + sm.ic * A.sc;
            {
                atomicAdd(&vol[i00], datum * (A.trav * sm.wb1 * sm.wc1));
                atomicAdd(&vol[i00 + A.sb], datum * (A.trav * sm.wb2 * sm.wc1));
                atomicAdd(&vol[i00 + A.sc], datum * (A.trav * sm.wb1 * sm.wc2));
                atomicAdd(&vol[i00 + A.sb + A.sc], datum * (A.trav * sm.wb2 * sm.wc2));
            }
        } else {
            if (!sample_nearest_exact(A, l, strict != 0, sm)) continue;
            atomicAdd(&vol[l * A.sa + sm.ib * A.sb + sm.ic * A.sc], datum * A.trav);
        }
    }
}

// ---------------------------------------------------------------- host ----
static int3 default_brick(const ct2_desc* d) {
    int3 b;
    b.x = d->brick[0] > 0 ? d->brick[0] : kBrickCandidates[0][0];
    b.y = d->brick[1] > 0 ? d->brick[1] : kBrickCandidates[0][1];
    b.z = d->brick[2] > 0 ? d->brick[2] : kBrickCandidates[0][2];
    return b;
}

template <int INTERP, int ARITH, typename OT>
static int launch_brick(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                        void* vol, const float* sino, const int8_t* ind, const float* fv,
                        hipStream_t st, SepWork w) {
    int3 b = default_brick(d);
    b.x = min(b.x, g.n[0]); b.y = min(b.y, g.n[1]); b.z = min(b.z, g.n[2]);
    // FP64 brick accumulator with a one-voxel halo on every side
    const int zonly = d->axis == CT2_AXIS_Z ? 1 : 0;
    size_t lds = brick_lds_bytes(b);
    if (lds > 63 * 1024) {
        set_error("ct2_backproject: brick %dx%dx%d needs %zu B of LDS (max 64512)", b.x, b.y,
                  b.z, lds);
        return CT2_EINVAL;
    }
    int3 nb = make_int3((g.n[0] + b.x - 1) / b.x, (g.n[1] + b.y - 1) / b.y,
                        (g.n[2] + b.z - 1) / b.z);
    // Fast arithmetic with an f32 output: each brick's views are split over
    // two workgroups, which add their FP64 partial sums, rounded to f32, into
    // the zeroed output with atomics.  Exactly two f32 additions onto zero
    // commute, so the result stays deterministic.  Shorter workgroups shrink
    // the kernel's tail, where bricks drain and the CUs run part-empty
    // (256^3 x 360 views: 55.2 -> 52.9 ms).
    //
    // Fast linear arithmetic into f32 with e_v along z and no brick named by
    // the caller: the column-separable kernels store the views
    // view_separable accepts, then one workgroup per brick adds any other
    // view (stream order fixes the order of the adds).
    if constexpr (ARITH == CT2_ARITH_FAST && INTERP == CT2_LINEAR &&
                  std::is_same<OT, float>::value) {
        // z-forced views with the detector parallel to the planes: the
        // plane-separable kernel stores every voxel, then one workgroup per
        // brick adds any other view
        if (zsep_route(d, g)) {
            void* own = nullptr;
            const int* nonsep = nullptr;
            int rc = launch_backproject_zsep(g, dt, vs, (float*)vol, sino, ind, fv, st, w, &own,
                                             &nonsep);
            if (rc == CT2_OK)
                hipLaunchKernelGGL((k_backproject_brick<INTERP, ARITH, OT>),
                                   dim3(nb.x * nb.y * nb.z), dim3(BRICK_THREADS), lds, st, g, dt,
                                   vs.view0, vs.nviews, vs.vstride, zonly, b, nb, 1,
                                   BRICK_SKIP_ZSEP, (OT*)vol, sino, ind, fv, nonsep);
            const int rf = sep_work_release(own, st);
            return rc ? rc : rf;
        }
        if (sep_route(d, g)) {
            void* own = nullptr;
            int rc = sep_work_acquire(dt, vs, w, &own, st);
            if (rc) return rc;
            rc = launch_backproject_sep(g, dt, vs, (float*)vol, sino, ind, fv, st, w);
            if (rc == CT2_OK)
                hipLaunchKernelGGL((k_backproject_brick<INTERP, ARITH, OT>),
                                   dim3(nb.x * nb.y * nb.z), dim3(BRICK_THREADS), lds, st, g, dt,
                                   vs.view0, vs.nviews, vs.vstride, zonly, b, nb, 1,
                                   BRICK_SKIP_SEP, (OT*)vol,
                                   sino, ind, fv, sep_nonsep_count(w));
            const int rf = sep_work_release(own, st);
            return rc ? rc : rf;
        }
    }
    const int nsplit = (ARITH == CT2_ARITH_FAST && std::is_same<OT, float>::value) ? 2 : 1;
    if (nsplit > 1)
        hipLaunchKernelGGL(k_zero_strided, dim3(2048), dim3(256), 0, st, g, (float*)vol);
    dim3 grid(nb.x * nb.y * nb.z * nsplit);
    hipLaunchKernelGGL((k_backproject_brick<INTERP, ARITH, OT>), grid, dim3(BRICK_THREADS), lds, st, g, dt,
                       vs.view0, vs.nviews, vs.vstride, zonly, b, nb, nsplit, BRICK_SKIP_NONE,
                       (OT*)vol, sino,
                       ind, fv, (const int*)nullptr);
    return CT2_OK;
}

template <int INTERP, int ARITH>
static int dispatch_out(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                        void* vol, const float* sino, const int8_t* ind, const float* fv,
                        hipStream_t st, SepWork w) {
    switch (d->vol_dtype) {
        case CT2_F32: return launch_brick<INTERP, ARITH, float>(d, g, dt, vs, vol, sino, ind, fv, st, w);
        case CT2_F16: return launch_brick<INTERP, ARITH, __half>(d, g, dt, vs, vol, sino, ind, fv, st, w);
        case CT2_BF16: return launch_brick<INTERP, ARITH, hip_bfloat16>(d, g, dt, vs, vol, sino, ind, fv, st, w);
    }
    set_error("ct2_backproject: unknown vol_dtype %d", d->vol_dtype);
    return CT2_EINVAL;
}

static int launch_with(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                       void* vol, const float* sn, const int8_t* ind, const float* fv,
                       hipStream_t st, SepWork w = SepWork());

// ---- brick autotuner ----
// The best brick depends on the geometry (footprint per view, LDS bank
// pattern of the halo layout, occupancy).  With desc->brick all zero, the
// first call for a geometry times each candidate once on the caller's stream
// (every launch overwrites the whole output, so the final launch with the
// winner leaves the right result) and caches the winner per geometry and
// device.  No tuning while the stream is being captured into a graph, or when
// CT2_AUTOTUNE=0: the default brick is used then.
struct TuneKey {
    int v[15];
    bool operator<(const TuneKey& o) const { return memcmp(v, o.v, sizeof v) < 0; }
};

static std::mutex g_tune_mu;
static std::map<TuneKey, int3>& tune_cache() {
    static std::map<TuneKey, int3>* m = new std::map<TuneKey, int3>();
    return *m;
}

static int tuned_backproject(const ct2_desc* d, const Grid& g, const Det& dt,
                             const ViewSet& vs, void* vol, const float* sn, const int8_t* ind,
                             const float* fv, hipStream_t st) {
    int dev = 0;
    CT2_HIP(hipGetDevice(&dev));
    TuneKey key;
    memset(&key, 0, sizeof key);
    int i = 0;
    for (int v : {d->nx, d->ny, d->nz, d->ns, d->nu, d->nv, vs.nviews * 64 + min(vs.vstride, 63),
                  d->interp, d->arith, d->axis, d->vol_dtype, dev})
        key.v[i++] = v;
    memcpy(&key.v[i++], &d->dx, sizeof(float));
    memcpy(&key.v[i++], &d->du, sizeof(float));
    key.v[i++] = ind != nullptr;
    ct2_desc e = *d;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = tune_cache().find(key);
        if (it != tune_cache().end()) {
            e.brick[0] = it->second.x; e.brick[1] = it->second.y; e.brick[2] = it->second.z;
            return launch_with(&e, g, dt, vs, vol, sn, ind, fv, st);
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    CT2_HIP(hipStreamIsCapturing(st, &cs));
    const char* env = getenv("CT2_AUTOTUNE");
    if (cs != hipStreamCaptureStatusNone || (env && env[0] == '0'))
        return launch_with(&e, g, dt, vs, vol, sn, ind, fv, st);
    const auto& cand = kBrickCandidates;
    const int nc = (int)(sizeof cand / sizeof cand[0]);
    hipEvent_t ev[2 * kBrickCount];
    for (int c = 0; c < 2 * nc; ++c) CT2_HIP(hipEventCreate(&ev[c]));
    int rc = CT2_OK;
    int3 tried[kBrickCount];
    int ntried = 0;
    for (int c = 0; c < nc && rc == CT2_OK; ++c) {
        int3 b = make_int3(min(cand[c][0], g.n[0]), min(cand[c][1], g.n[1]), min(cand[c][2], g.n[2]));
        bool dup = false;
        for (int t = 0; t < ntried; ++t)
            dup |= tried[t].x == b.x && tried[t].y == b.y && tried[t].z == b.z;
        if (dup) continue;
        e.brick[0] = b.x; e.brick[1] = b.y; e.brick[2] = b.z;
        hipEventRecord(ev[2 * ntried], st);
        rc = launch_with(&e, g, dt, vs, vol, sn, ind, fv, st);
        hipEventRecord(ev[2 * ntried + 1], st);
        tried[ntried++] = b;
    }
    int3 best = tried[0];
    if (rc == CT2_OK) {
        hipError_t he = hipEventSynchronize(ev[2 * ntried - 1]);
        if (he != hipSuccess) rc = hip_fail(he, "autotune");
        float tbest = 1e30f;
        for (int t = 0; t < ntried && rc == CT2_OK; ++t) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]);
            if (ms < tbest) { tbest = ms; best = tried[t]; }
        }
    }
    for (int c = 0; c < 2 * nc; ++c) hipEventDestroy(ev[c]);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_cache()[key] = best;
    }
    e.brick[0] = best.x; e.brick[1] = best.y; e.brick[2] = best.z;
    return launch_with(&e, g, dt, vs, vol, sn, ind, fv, st);
}

int launch_backproject(const ct2_desc* d, void* vol, const void* sino, const int8_t* ind,
                       const float* fv, hipStream_t st, SepWork w) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    ViewSet vs;
    if ((rc = view_set(d, "ct2_backproject", vs))) return rc;
    if (!sino || !fv || !vol) { set_error("ct2_backproject: null pointer"); return CT2_EINVAL; }
    if (d->arith != CT2_ARITH_EXACT && d->arith != CT2_ARITH_FAST) {
        set_error("ct2_backproject: unknown arith %d", d->arith);
        return CT2_EINVAL;
    }
    // (the separable routes leave the brick kernel only the other views: no
    // brick is tuned for them)
    if (d->brick[0] <= 0 && d->brick[1] <= 0 && d->brick[2] <= 0 && !sep_route(d, g) &&
        !zsep_route(d, g))
        return tuned_backproject(d, g, dt, vs, vol, (const float*)sino, ind, fv, st);
    return launch_with(d, g, dt, vs, vol, (const float*)sino, ind, fv, st, w);
}

// One launch with the brick of *d (0 entries take the defaults).
static int launch_with(const ct2_desc* d, const Grid& g, const Det& dt, const ViewSet& vs,
                       void* vol, const float* sn, const int8_t* ind, const float* fv,
                       hipStream_t st, SepWork w) {
    int rc;
    if (d->arith == CT2_ARITH_EXACT)
        rc = d->interp == CT2_LINEAR
                 ? dispatch_out<CT2_LINEAR, CT2_ARITH_EXACT>(d, g, dt, vs, vol, sn, ind, fv, st, w)
                 : dispatch_out<CT2_NEAREST, CT2_ARITH_EXACT>(d, g, dt, vs, vol, sn, ind, fv, st, w);
    else if (d->arith == CT2_ARITH_FAST)
        rc = d->interp == CT2_LINEAR
                 ? dispatch_out<CT2_LINEAR, CT2_ARITH_FAST>(d, g, dt, vs, vol, sn, ind, fv, st, w)
                 : dispatch_out<CT2_NEAREST, CT2_ARITH_FAST>(d, g, dt, vs, vol, sn, ind, fv, st, w);
    else {
        set_error("ct2_backproject: unknown arith %d", d->arith);
        return CT2_EINVAL;
    }
    if (rc) return rc;
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int launch_ray_backproject(const ct2_desc* d, float* vol, const float* sino, const int8_t* ind,
                           const float* rv, int nr, int strict, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    int rc;
    if ((rc = check_grid(g))) return rc;
    if (nr < 0 || !vol) { set_error("ray backprojection: bad arguments"); return CT2_EINVAL; }
    int rc0 = launch_fill(vol, 0.0f, (size_t)g.n[0] * g.n[1] * g.n[2], st);
    if (rc0) return rc0;
    if (nr == 0) return CT2_OK;
    if (!sino || !rv) { set_error("ray backprojection: null pointer"); return CT2_EINVAL; }
    dim3 grid((nr + 255) / 256);
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_backproject_rays<CT2_LINEAR>, grid, dim3(256), 0, st, g, nr, 0, vol, sino, ind, rv);
    else
        hipLaunchKernelGGL(k_backproject_rays<CT2_NEAREST>, grid, dim3(256), 0, st, g, nr, strict, vol, sino, ind, rv);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace ct2
