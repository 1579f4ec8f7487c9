// ct2_internal.h -- shared device/host definitions for the ct2 HIP kernels.
//
// Geometry follows ct-2's Joseph projector (emil_code/view_proj3D.f:703-910):
// a ray from source S to detector bin B marches plane by plane along its
// driving axis (the axis of largest |B-S|/voxel, strict '>' so ties go to the
// later axis), samples the two transverse axes bilinearly at the plane
// centre, and weights each sample by d_axis*sqrt(1+slope_b^2+slope_c^2).
//
// Device layout: volume element (i,j,k) at i*sx + j*sy + k*sz (32-bit
// offsets; the default C-order [nx][ny][nz] has sz = 1), sinogram element
// (p,j,k) at (p*nu + j)*nv + k, frame vectors [ns][18] float32.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bfloat16.h>
#include <stdint.h>

#include "../../include/ct2.h"

namespace ct2 {

// ---------------------------------------------------------------- errors ----
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

#define CT2_HIP(call)                                        \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return ::ct2::hip_fail(e_, #call); \
    } while (0)

// ------------------------------------------------------------- geometry ----
struct Grid {
    int n[3];
    float d[3];
    float o[3];
    int s[3];   // element strides of x, y, z
};

struct Det {
    int ns, nu, nv;
    float du, dv, u0, v0;
};

// Per-ray scalars of the driving axis a and transverse axes b, c, chosen once
// so the inner loop never indexes a runtime axis.
struct AxisSel {
    int na, nb, nc;
    int sa, sb, sc;
    float da, db, dc;
    float oa, ob, oc;
    float Sa, Sb, Sc;
    float r1, r2, trav;
    int a;
};

__device__ __forceinline__ float sel3(int a, float x, float y, float z) {
    return a == 0 ? x : (a == 1 ? y : z);
}
__device__ __forceinline__ int sel3i(int a, int x, int y, int z) {
    return a == 0 ? x : (a == 1 ? y : z);
}

// Bin position of view ray (p,j,k): B = D + e_u*u + e_v*v with
// u = u0 + (j+0.5)du, v = v0 + (k+0.5)dv (view_proj3D.f:784-797).
__device__ __forceinline__ void view_ray(const float* __restrict__ fv, int p, int j, int k,
                                         const Det& dt, float S[3], float B[3]) {
#pragma clang fp contract(off)
    const float* f = fv + (size_t)p * 18;
    float u = dt.u0 + ((float)j + 0.5f) * dt.du;
    float v = dt.v0 + ((float)k + 0.5f) * dt.dv;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S[q] = f[q];
        B[q] = f[6 + q] + f[9 + q] * u + f[12 + q] * v;
    }
}

// Ray setup in the reference's FP32 operation order (view_proj3D.f:802-815,
// :843-847, :873-876; forced z: view_proj3D_tomo.f:111-113).
__device__ __forceinline__ void setup_exact(const Grid& g, const float S[3], const float B[3],
                                            bool zonly, AxisSel& A) {
#pragma clang fp contract(off)
    float dx = B[0] - S[0], dy = B[1] - S[1], dz = B[2] - S[2];
    int a;
    if (zonly) {
        a = 2;
    } else {
        float xad = fabsf(dx) / g.d[0];
        float yad = fabsf(dy) / g.d[1];
        float zad = fabsf(dz) / g.d[2];
        a = ((xad > yad) && (xad > zad)) ? 0 : ((yad > zad) ? 1 : 2);
    }
    int b = (a == 0) ? 1 : 0;
    int c = (a == 2) ? 1 : 2;
    float diffa = sel3(a, dx, dy, dz);
    float diffb = sel3(b, dx, dy, dz);
    float diffc = sel3(c, dx, dy, dz);
    A.a = a;
    A.na = sel3i(a, g.n[0], g.n[1], g.n[2]);
    A.nb = sel3i(b, g.n[0], g.n[1], g.n[2]);
    A.nc = sel3i(c, g.n[0], g.n[1], g.n[2]);
    A.sa = sel3i(a, g.s[0], g.s[1], g.s[2]);
    A.sb = sel3i(b, g.s[0], g.s[1], g.s[2]);
    A.sc = sel3i(c, g.s[0], g.s[1], g.s[2]);
    A.da = sel3(a, g.d[0], g.d[1], g.d[2]);
    A.db = sel3(b, g.d[0], g.d[1], g.d[2]);
    A.dc = sel3(c, g.d[0], g.d[1], g.d[2]);
    A.oa = sel3(a, g.o[0], g.o[1], g.o[2]);
    A.ob = sel3(b, g.o[0], g.o[1], g.o[2]);
    A.oc = sel3(c, g.o[0], g.o[1], g.o[2]);
    A.Sa = sel3(a, S[0], S[1], S[2]);
    A.Sb = sel3(b, S[0], S[1], S[2]);
    A.Sc = sel3(c, S[0], S[1], S[2]);
    A.r1 = diffb / diffa;
    A.r2 = diffc / diffa;
    A.trav = A.da * sqrtf(1.0f + A.r1 * A.r1 + A.r2 * A.r2);
}

// One sample on plane l, reference FP32 order.  Linear: cell
// floor((p-o)/d-0.5), valid 0 <= i < n-1, w2 = (p-(d*(i+0.5)+o))/d
// (view_proj3D.f:816-826).  Nearest: floor((p-o)/d), valid 0 <= i <= hi
// with hi = n-1, or n-2 for the ray-list bp/ART/EM (ray_proj3D.f:895).
struct Sample {
    int ib, ic;
    float wb1, wb2, wc1, wc2;
};

__device__ __forceinline__ bool sample_linear_exact(const AxisSel& A, int l, Sample& sm) {
#pragma clang fp contract(off)
    float t = A.oa + A.da * ((float)l + 0.5f);
    float pb = A.Sb + A.r1 * (t - A.Sa);
    float pc = A.Sc + A.r2 * (t - A.Sa);
    int ib = (int)floorf((pb - A.ob) / A.db - 0.5f);
    int ic = (int)floorf((pc - A.oc) / A.dc - 0.5f);
    if (!((ib >= 0) && (ib < A.nb - 1) && (ic >= 0) && (ic < A.nc - 1))) return false;
    sm.ib = ib;
    sm.ic = ic;
    sm.wb2 = (pb - (A.db * ((float)ib + 0.5f) + A.ob)) / A.db;
    sm.wb1 = 1.0f - sm.wb2;
    sm.wc2 = (pc - (A.dc * ((float)ic + 0.5f) + A.oc)) / A.dc;
    sm.wc1 = 1.0f - sm.wc2;
    return true;
}

__device__ __forceinline__ bool sample_nearest_exact(const AxisSel& A, int l, bool strict,
                                                     Sample& sm) {
#pragma clang fp contract(off)
    float t = A.oa + A.da * ((float)l + 0.5f);
    float pb = A.Sb + A.r1 * (t - A.Sa);
    float pc = A.Sc + A.r2 * (t - A.Sa);
    int ib = (int)floorf((pb - A.ob) / A.db);
    int ic = (int)floorf((pc - A.oc) / A.dc);
    int hb = strict ? A.nb - 2 : A.nb - 1;
    int hc = strict ? A.nc - 2 : A.nc - 1;
    if (!((ib >= 0) && (ib <= hb) && (ic >= 0) && (ic <= hc))) return false;
    sm.ib = ib;
    sm.ic = ic;
    return true;
}

// ------------------------------------------------------ fast arithmetic ----
// Per-ray setup in FP64 (bin position, driving axis, slopes), then the
// transverse cell coordinate of plane l is one FP32 FMA, u(l) = u0 + l*du,
// with u = (p-o)/d - 0.5 (linear) or (p-o)/d (nearest).  The forward kernel
// and the adjoint call this same function, so their matrix entries agree.
struct RayFast {
    int a, b, c;
    int na, nb, nc;
    int sa, sb, sc;
    float ub0, dub, uc0, duc;
    float trav;
};

// floor(x) converted to int in one instruction (v_cvt_flr_i32_f32), for
// |x| < 2^31; the compiler emits a floor and a conversion for (int)floorf(x).
__device__ __forceinline__ int floor_i(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// x - floor(x) in one instruction (v_fract_f32); equal to x - floorf(x) for
// x >= 0, the only samples that reach a weight.
__device__ __forceinline__ float frac_pos(float x) { return __builtin_amdgcn_fractf(x); }

// Wavefront-wide min / max of an int, returned in a scalar register, with DPP
// lane moves (no LDS traffic, so no wait on the wavefront's outstanding LDS
// operations, unlike a ds_bpermute shuffle): row_shr 1, 2, 4, 8 reduce each
// 16-lane row into its lane 15, row_bcast 15 and 31 carry the rows' results
// up to lane 63.  Every lane of the wavefront must be active.
template <bool MAX>
__device__ __forceinline__ int wave_minmax(int v) {
    constexpr int ident = MAX ? (int)0x80000000 : 0x7fffffff;
    auto op = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x111, 0xf, 0xf, false));   // row_shr:1
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x112, 0xf, 0xf, false));   // row_shr:2
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x114, 0xf, 0xf, false));   // row_shr:4
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x118, 0xf, 0xf, false));   // row_shr:8
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x142, 0xa, 0xf, false));   // row_bcast:15
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x143, 0xc, 0xf, false));   // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_min(int v) { return wave_minmax<false>(v); }
__device__ __forceinline__ int wave_max(int v) { return wave_minmax<true>(v); }

// Wavefront-wide FP32 sum by the same DPP pattern (a fixed reduction tree, so
// the result is deterministic), returned in a scalar register.  Every lane
// of the wavefront must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_add(float x) {
    const int y = __builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROW_MASK, 0xf, false);
    return x + __int_as_float(y);
}
__device__ __forceinline__ float wave_sum_f(float v) {
    v = dpp_add<0x111, 0xf>(v);   // row_shr:1
    v = dpp_add<0x112, 0xf>(v);   // row_shr:2
    v = dpp_add<0x114, 0xf>(v);   // row_shr:4
    v = dpp_add<0x118, 0xf>(v);   // row_shr:8
    v = dpp_add<0x142, 0xa>(v);   // row_bcast:15
    v = dpp_add<0x143, 0xc>(v);   // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Dynamic work queue of the batched persistent kernels: q[0] is the next
// item, q[1] counts the workgroups that have left.  Thread 0 of a workgroup
// takes an item and shares it through LDS; the last workgroup to leave puts
// the queue back to zero, so every launch (graph replays included) starts
// from item 0 without a separate reset.  A queue serves one launch at a time.
__device__ __forceinline__ long long queue_next(unsigned long long* q, long long* s_it) {
    __syncthreads();   // every thread has read the previous item
    if (threadIdx.x == 0) *s_it = (long long)atomicAdd(q, 1ull);
    __syncthreads();
    return *s_it;
}

__device__ __forceinline__ void queue_leave(unsigned long long* q) {
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(q + 1, 1ull) == (unsigned long long)gridDim.x - 1) {
            atomicExch(q, 0ull);
            atomicExch(q + 1, 0ull);
        }
    }
}

// FP64 reciprocal: the FP32 hardware reciprocal (relative error ~1e-7) refined
// by one Newton-Raphson step r <- r(2 - x r), which squares the relative error
// (~1e-14), for |x| within the FP32 normal range.
__device__ __forceinline__ double rcp_f64(double x) {
    const double r = (double)__builtin_amdgcn_rcpf((float)x);
    return r * fma(-x, r, 2.0);
}

// Square root of the fast set-up's ray weight: the hardware v_sqrt_f32
// (within 1 ulp), one instruction where sqrtf's correctly rounded sequence
// takes about twenty; every fast-arithmetic kernel weighs rays with it, so
// their matrix entries agree.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Driving axis of direction dd in the fast set-up: the largest |dd_q| / d_q,
// compared cross-multiplied (z when the policy forces it).
__device__ __forceinline__ int fast_axis(const Grid& g, double dd0, double dd1, double dd2,
                                         bool zonly) {
    if (zonly) return 2;
    const float ax = fabsf((float)dd0), ay = fabsf((float)dd1), az = fabsf((float)dd2);
    const float gx = fabsf(g.d[0]), gy = fabsf(g.d[1]), gz = fabsf(g.d[2]);
    const bool xy = ax * gy > ay * gx, xz = ax * gz > az * gx, yz = ay * gz > az * gy;
    return (xy && xz) ? 0 : (yz ? 1 : 2);
}

// Fast set-up of the ray from S along dd = B - S with driving axis a.
// FP64 where cancellation needs it (the entry coordinates ub0, uc0: source
// distances of hundreds of voxels against a sub-voxel result), FP32 elsewhere;
// no FP64 division or square root.
__device__ __forceinline__ void setup_fast_ax(const Grid& g, const double S[3], double dd0,
                                              double dd1, double dd2, int a, bool linear,
                                              RayFast& R) {
    const int b = (a == 0) ? 1 : 0;
    const int c = (a == 2) ? 1 : 2;
    const double dda = (a == 0) ? dd0 : (a == 1 ? dd1 : dd2);
    const double ddb = (b == 0) ? dd0 : dd1;
    const double ddc = (c == 1) ? dd1 : dd2;
    const double Sa = (a == 0) ? S[0] : (a == 1 ? S[1] : S[2]);
    const double Sb = (b == 0) ? S[0] : S[1];
    const double Sc = (c == 1) ? S[1] : S[2];
    const float da = sel3(a, g.d[0], g.d[1], g.d[2]);
    const float oa = sel3(a, g.o[0], g.o[1], g.o[2]);
    const float db = (b == 0) ? g.d[0] : g.d[1];
    const float ob = (b == 0) ? g.o[0] : g.o[1];
    const float dc = (c == 1) ? g.d[1] : g.d[2];
    const float oc = (c == 1) ? g.o[1] : g.o[2];
    const double ia = rcp_f64(dda);
    // voxel-size reciprocals before the axis selection (loop invariant)
    const double id0 = rcp_f64((double)g.d[0]), id1 = rcp_f64((double)g.d[1]);
    const double id2 = rcp_f64((double)g.d[2]);
    const double idb = (b == 0) ? id0 : id1, idc = (c == 1) ? id1 : id2;
    const double r1 = ddb * ia, r2 = ddc * ia;
    const double ta = fma(0.5, (double)da, (double)oa) - Sa;   // centre of plane 0 from the source
    const double half = linear ? 0.5 : 0.0;
    R.ub0 = (float)fma(fma(r1, ta, Sb) - (double)ob, idb, -half);
    R.uc0 = (float)fma(fma(r2, ta, Sc) - (double)oc, idc, -half);
    R.dub = (float)(r1 * ((double)da * idb));
    R.duc = (float)(r2 * ((double)da * idc));
    const float r1f = (float)r1, r2f = (float)r2;
    R.trav = fabsf(da) * fast_sqrt(1.0f + r1f * r1f + r2f * r2f);
    R.a = a; R.b = b; R.c = c;
    R.na = sel3i(a, g.n[0], g.n[1], g.n[2]);
    R.nb = sel3i(b, g.n[0], g.n[1], g.n[2]);
    R.nc = sel3i(c, g.n[0], g.n[1], g.n[2]);
    R.sa = sel3i(a, g.s[0], g.s[1], g.s[2]);
    R.sb = sel3i(b, g.s[0], g.s[1], g.s[2]);
    R.sc = sel3i(c, g.s[0], g.s[1], g.s[2]);
}

__device__ __forceinline__ void setup_fast(const Grid& g, const double S[3], const double B[3],
                                           bool zonly, bool linear, RayFast& R) {
    const double dd0 = B[0] - S[0], dd1 = B[1] - S[1], dd2 = B[2] - S[2];
    setup_fast_ax(g, S, dd0, dd1, dd2, fast_axis(g, dd0, dd1, dd2, zonly), linear, R);
}

// FP64 source and bin of view ray (p,j,k) for the fast setup.
__device__ __forceinline__ void view_ray_f64(const float* __restrict__ fv, int p, int j, int k,
                                             const Det& dt, double S[3], double B[3]) {
    const float* f = fv + (size_t)p * 18;
    const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
    const double v = fma((double)k + 0.5, (double)dt.dv, (double)dt.v0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S[q] = f[q];
        B[q] = fma((double)f[12 + q], v, fma((double)f[9 + q], u, (double)f[6 + q]));
    }
}

// The frame entries view_ray_f64 reads (source, detector centre, e_u, e_v),
// loaded once per view into wavefront-uniform registers; view_ray_frame
// evaluates the same expressions on them, so the rays are identical.
struct ViewFrame {
    float s[3], d[3], eu[3], ev[3];
};

__device__ __forceinline__ float uniform_f(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform_d(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ ViewFrame load_frame(const float* __restrict__ fv, int p) {
    const float* f = fv + (size_t)p * 18;
    ViewFrame F;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        F.s[q] = uniform_f(f[q]);
        F.d[q] = uniform_f(f[6 + q]);
        F.eu[q] = uniform_f(f[9 + q]);
        F.ev[q] = uniform_f(f[12 + q]);
    }
    return F;
}

__device__ __forceinline__ void view_ray_frame(const ViewFrame& F, int j, int k, const Det& dt,
                                               double S[3], double B[3]) {
    const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
    const double v = fma((double)k + 0.5, (double)dt.dv, (double)dt.v0);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        S[q] = F.s[q];
        B[q] = fma((double)F.ev[q], v, fma((double)F.eu[q], u, (double)F.d[q]));
    }
}

// Detector coordinates (u, v) of point P seen from the source of frame f:
// where the ray from S through P meets the detector plane D + u e_u + v e_v.
// ok = 0 when P is not in front of the source or the ray runs parallel to
// the plane.  FP64.
__device__ inline void project_corner(const float* __restrict__ f, const double P[3], double& u,
                                      double& v, int& ok) {
    double S[3] = {f[0], f[1], f[2]};
    double D[3] = {f[6], f[7], f[8]};
    double eu[3] = {f[9], f[10], f[11]};
    double ev[3] = {f[12], f[13], f[14]};
    double n[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2],
                   eu[0] * ev[1] - eu[1] * ev[0]};
    double dp[3] = {P[0] - S[0], P[1] - S[1], P[2] - S[2]};
    double ds[3] = {D[0] - S[0], D[1] - S[1], D[2] - S[2]};
    double den = dp[0] * n[0] + dp[1] * n[1] + dp[2] * n[2];
    double num = ds[0] * n[0] + ds[1] * n[1] + ds[2] * n[2];
    if (fabs(den) < 1e-12 * (fabs(num) + 1e-30)) { ok = 0; return; }
    double t = num / den;
    if (!(t > 0.0)) { ok = 0; return; }
    double X[3] = {S[0] + t * dp[0] - D[0], S[1] + t * dp[1] - D[1], S[2] + t * dp[2] - D[2]};
    double guu = eu[0] * eu[0] + eu[1] * eu[1] + eu[2] * eu[2];
    double gvv = ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2];
    double guv = eu[0] * ev[0] + eu[1] * ev[1] + eu[2] * ev[2];
    double bu = X[0] * eu[0] + X[1] * eu[1] + X[2] * eu[2];
    double bv = X[0] * ev[0] + X[1] * ev[1] + X[2] * ev[2];
    double det = guu * gvv - guv * guv;
    if (fabs(det) < 1e-30) { ok = 0; return; }
    u = (bu * gvv - bv * guv) / det;
    v = (bv * guu - bu * guv) / det;
    ok = 1;
}

// ---- column-separable views (the adjoint's fast route) ----
// With e_v exactly parallel to z, the rays of one detector column differ
// only in their z direction component: they share the driving axis a (x or
// y) and the per-ray set-up of the other transverse axis b, so on every
// plane of a all of them sample the same b cell with the same weight, and
// their z coordinate is linear in the row index.  view_separable says whether
// every ray of view F qualifies:
//   (1) e_v = (0, 0, +-1 or any nonzero);
//   (2) no ray is z-driven: the largest |dz|/|d_z| over the detector corners
//       stays below the smallest max(|dx|/|d_x|, |dy|/|d_y|) over the bin
//       columns by a relative margin far above the rounding of the fast
//       set-up's axis comparisons;
//   (3) for x and for y, when that axis can drive some column, the source
//       lies outside the span of its plane centres widened by one voxel, so
//       a column's plane parameters (plane - source along a) share one sign;
//   (4) a z cell holds at most 48 rows of a column on any plane.
// It depends on the frame, grid and detector only, so every kernel that
// calls it classifies a view the same way.
__device__ __forceinline__ double sep_f(double ax, double bx, double ay, double by, double gx,
                                        double gy, double u) {
    return fmax(fabs(fma(u, bx, ax)) / gx, fabs(fma(u, by, ay)) / gy);
}
__device__ __forceinline__ double sep_h(double ax, double bx, double ay, double by, double gx,
                                        double gy, double u) {
    return fabs(fma(u, bx, ax)) / gx - fabs(fma(u, by, ay)) / gy;
}

__device__ inline bool view_separable(const float* __restrict__ f, const Grid& g, const Det& dt) {
    if (f[12] != 0.0f || f[13] != 0.0f || f[14] == 0.0f) return false;
    const double S0 = f[0], S1 = f[1], S2 = f[2];
    const double ax = (double)f[6] - S0, bx = f[9];     // dx(u) = ax + u bx (e_v has no x)
    const double ay = (double)f[7] - S1, by = f[10];
    const double gx = fabs((double)g.d[0]), gy = fabs((double)g.d[1]), gz = fabs((double)g.d[2]);
    double ua = fma(0.5, (double)dt.du, (double)dt.u0);
    double ub = fma((double)dt.nu - 0.5, (double)dt.du, (double)dt.u0);
    if (ua > ub) { const double t = ua; ua = ub; ub = t; }
    double va = fma(0.5, (double)dt.dv, (double)dt.v0);
    double vb = fma((double)dt.nv - 0.5, (double)dt.dv, (double)dt.v0);
    double zmax = 0.0;
    for (int c = 0; c < 4; ++c) {
        const double u = (c & 1) ? ub : ua, v = (c & 2) ? vb : va;
        zmax = fmax(zmax, fabs((double)f[8] - S2 + u * f[11] + v * f[14]) / gz);
    }
    // candidate points of the piecewise-linear functions of u: the ends and
    // the roots of dx, dy, dx/gx - dy/gy and dx/gx + dy/gy
    double cand[6] = {ua, ub, ua, ua, ua, ua};
    if (bx != 0.0) cand[2] = -ax / bx;
    if (by != 0.0) cand[3] = -ay / by;
    const double d1 = bx / gx - by / gy, d2 = bx / gx + by / gy;
    if (d1 != 0.0) cand[4] = -(ax / gx - ay / gy) / d1;
    if (d2 != 0.0) cand[5] = -(ax / gx + ay / gy) / d2;
    double fmn = 1e300, hmax = -1e300, hmin = 1e300;
    for (int c = 0; c < 6; ++c) {
        const double u = fmin(fmax(cand[c], ua), ub);
        fmn = fmin(fmn, sep_f(ax, bx, ay, by, gx, gy, u));
        const double h = sep_h(ax, bx, ay, by, gx, gy, u);
        hmax = fmax(hmax, h);
        hmin = fmin(hmin, h);
    }
    if (!(zmax < fmn * (1.0 - 1e-5))) return false;
    const double m = 1e-6 * fmn;
    const bool xmay = hmax > -m, ymay = hmin < m;
    double dist = 1e300;
    for (int a = 0; a < 2; ++a) {
        if (!(a == 0 ? xmay : ymay)) continue;
        const double d = g.d[a], o = g.o[a];
        double lo = o + 0.5 * d, hi = o + ((double)g.n[a] - 0.5) * d;
        if (lo > hi) { const double t = lo; lo = hi; hi = t; }
        const double Sa = a == 0 ? S0 : S1;
        if (Sa >= lo - fabs(d) && Sa <= hi + fabs(d)) return false;
        dist = fmin(dist, fmin(fabs(Sa - lo), fabs(Sa - hi)));
    }
    // (4) on every plane, the z spacing of a column's rows is at least 1/48
    //     of a voxel: |e_v,z dv / d_z| times the plane's fraction of the ray,
    //     which is at least the source's distance to the nearest plane centre
    //     over the longest ray
    double maxlen = 0.0;
    for (int c = 0; c < 4; ++c) {
        const double u = (c & 1) ? ub : ua, v = (c & 2) ? vb : va;
        const double e0 = ax + u * bx, e1 = ay + u * by;
        const double e2 = (double)f[8] - S2 + u * f[11] + v * f[14];
        maxlen = fmax(maxlen, sqrt(e0 * e0 + e1 * e1 + e2 * e2));
    }
    return fabs((double)f[14] * (double)dt.dv / (double)g.d[2]) * dist >= maxlen / 48.0;
}

// ---- plane-separable views (the z-driven adjoint's fast route) ----
// With e_u exactly along x, e_v exactly along y and the projector forced to
// z, the bin of ray (j, k) lies at B = (fma(eu_x, u_j, D_x), fma(ev_y, v_k,
// D_y), D_z) (the zero products of view_ray_f64 vanish exactly), so the fast
// set-up's x coordinate of a ray (ub0, dub) depends on its column j alone,
// its y coordinate (uc0, duc) on its row k alone, and on plane l the samples
// of a view form the detector grid, scaled.  view_zsep says whether view F
// qualifies: those zero entries, D_z != S_z, and on every plane the ray
// spacing in voxels, |t_l e_u,x du / d_x| and |t_l e_v,y dv / d_y| with t_l
// the plane's fraction of the ray (of one sign over the planes), at least
// ZSEP_SPACING, so that the detector rows within 0.02 row of a voxel's
// [c-1, c+1) window span less than 3 rows (k_bp_zsep takes 3 candidates).
constexpr double ZSEP_SPACING = 0.7;

__device__ inline bool view_zsep(const float* __restrict__ f, const Grid& g, const Det& dt) {
    if (f[10] != 0.0f || f[11] != 0.0f || f[12] != 0.0f || f[14] != 0.0f || f[9] == 0.0f ||
        f[13] == 0.0f)
        return false;
    const double dd2 = (double)f[8] - (double)f[2];
    if (dd2 == 0.0) return false;
    const double ta = fma(0.5, (double)g.d[2], (double)g.o[2]) - (double)f[2];
    const double t0 = ta / dd2, t1 = fma((double)(g.n[2] - 1), (double)g.d[2], ta) / dd2;
    if (!(t0 * t1 > 0.0)) return false;
    const double tmin = fmin(fabs(t0), fabs(t1));
    return tmin * fabs((double)f[9] * (double)dt.du / (double)g.d[0]) >= ZSEP_SPACING &&
           tmin * fabs((double)f[13] * (double)dt.dv / (double)g.d[1]) >= ZSEP_SPACING;
}

// Workgroup b of a grid of nblk runs on XCD b % 8 (the dispatcher deals
// workgroups to the 8 XCDs in turn): the tile it takes, such that each XCD
// gets a contiguous range of tiles, whose data then share its L2.  A
// bijection of [0, nblk).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nblk) {
    const unsigned x = b & 7u, i = b >> 3, q = nblk >> 3, rm = nblk & 7u;
    return (x < rm ? x * (q + 1) : rm * (q + 1) + (x - rm) * q) + i;
}

// Restrict the inclusive plane range [l0,l1] to planes where the cell
// coordinate u0 + l*du may lie in [lo, hi); the interval is widened by `pad`
// cells and one plane so FP32 rounding of the exact per-sample test never
// falls outside it.
__device__ __forceinline__ void clip_planes(double u0, double du, double lo, double hi, double pad,
                                            int& l0, int& l1) {
    lo -= pad;
    hi += pad;
    if (fabs(du) < 1e-12) {
        if (!(u0 >= lo && u0 < hi)) l1 = l0 - 1;
        return;
    }
    const double idu = 1.0 / du;
    double ta = (lo - u0) * idu, tb = (hi - u0) * idu;
    double tmin = fmax(fmin(ta, tb), -1e9), tmax = fmin(fmax(ta, tb), 1e9);
    if (!(tmin <= tmax)) { l1 = l0 - 1; return; }
    l0 = max(l0, (int)floor(tmin) - 1);
    l1 = min(l1, (int)ceil(tmax) + 1);
}

// FP32 form of clip_planes for the fast arithmetic (same widening; the
// hardware reciprocal's 1-ulp error is far inside the pad).
__device__ __forceinline__ void clip_planes_f(float u0, float du, float lo, float hi, float pad,
                                              int& l0, int& l1) {
    lo -= pad;
    hi += pad;
    if (fabsf(du) < 1e-20f) {
        if (!(u0 >= lo && u0 < hi)) l1 = l0 - 1;
        return;
    }
    const float idu = __builtin_amdgcn_rcpf(du);
    const float ta = (lo - u0) * idu, tb = (hi - u0) * idu;
    const float tmin = fmaxf(fminf(ta, tb), -1e9f), tmax = fminf(fmaxf(ta, tb), 1e9f);
    if (!(tmin <= tmax)) { l1 = l0 - 1; return; }
    l0 = max(l0, (int)floorf(tmin) - 1);
    l1 = min(l1, (int)ceilf(tmax) + 1);
}

// ------------------------------------------------------- storage types ----
template <typename T> struct Store;
// Element at byte offset off of a uniform base pointer: a 32-bit offset
// lets the compiler address the load as base register + offset with no
// 64-bit pointer arithmetic per sample.
template <typename VT>
__device__ __forceinline__ const VT* at_byte(const VT* base, unsigned off) {
    return (const VT*)((const char*)base + off);
}

template <> struct Store<float> {
    __device__ __forceinline__ static float ld(const float* p, int i) { return p[i]; }
    __device__ __forceinline__ static float ldb(const float* p, unsigned off) {
        return *at_byte(p, off);
    }
    __device__ __forceinline__ static void st(float* p, size_t i, float v) { p[i] = v; }
};
template <> struct Store<__half> {
    __device__ __forceinline__ static float ld(const __half* p, int i) { return __half2float(p[i]); }
    __device__ __forceinline__ static float ldb(const __half* p, unsigned off) {
        return __half2float(*at_byte(p, off));
    }
    __device__ __forceinline__ static void st(__half* p, size_t i, float v) { p[i] = __float2half_rn(v); }
};
// Element of CT2_F16Z2 storage: the f16 value of a voxel and of its
// successor along the contiguous axis.  Reads take the voxel's own value;
// the projector's bilinear loop also uses the successor.
struct __align__(4) h16x2 {
    __half lo, hi;
};
template <> struct Store<h16x2> {
    __device__ __forceinline__ static float ld(const h16x2* p, int i) { return __half2float(p[i].lo); }
    __device__ __forceinline__ static float ldb(const h16x2* p, unsigned off) {
        return __half2float(at_byte(p, off)->lo);
    }
};
template <> struct Store<hip_bfloat16> {
    __device__ __forceinline__ static float ld(const hip_bfloat16* p, int i) {
        return static_cast<float>(p[i]);
    }
    __device__ __forceinline__ static float ldb(const hip_bfloat16* p, unsigned off) {
        return static_cast<float>(*at_byte(p, off));
    }
    __device__ __forceinline__ static void st(hip_bfloat16* p, size_t i, float v) {
        p[i] = hip_bfloat16(v);
    }
};

// Division by a runtime-invariant divisor with one multiply-high (Granlund
// and Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994): q = (t + ((n - t) >> s1)) >> s2 with t = umulhi(m, n), exact
// for every 32-bit n and 1 <= d < 2^32.
struct Div32 {
    unsigned m;
    int s1, s2;
};

inline Div32 make_div32(unsigned d) {
    int l = 0;
    while ((1ull << l) < d) ++l;   // ceil(log2 d)
    Div32 v;
    v.m = (unsigned)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    v.s1 = l < 1 ? l : 1;
    v.s2 = l - 1 > 0 ? l - 1 : 0;
    return v;
}

__device__ __forceinline__ unsigned div32(unsigned n, const Div32& v) {
    const unsigned t = __umulhi(v.m, n);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}

// ------------------------------------------------------- host utilities ----
inline Grid make_grid(int nx, int ny, int nz, float dx, float dy, float dz, float x0, float y0,
                      float z0) {
    Grid g;
    g.n[0] = nx; g.n[1] = ny; g.n[2] = nz;
    g.d[0] = dx; g.d[1] = dy; g.d[2] = dz;
    g.o[0] = x0; g.o[1] = y0; g.o[2] = z0;
    g.s[0] = ny * nz; g.s[1] = nz; g.s[2] = 1;
    return g;
}

inline Det make_det(int ns, int nu, int nv, float du, float dv, float u0, float v0) {
    Det d;
    d.ns = ns; d.nu = nu; d.nv = nv;
    d.du = du; d.dv = dv; d.u0 = u0; d.v0 = v0;
    return d;
}

int check_grid(const Grid& g);
int check_det(const Det& d);

// SART residual of one ray and update of one voxel (shared by the
// single-problem and the batched SART so both compute the same values).
__device__ __forceinline__ float sart_resid(float b, float ax, float rowsum, bool used) {
    return (used && rowsum > 0.0f) ? (b - ax) / rowsum : 0.0f;
}
__device__ __forceinline__ float sart_update(float x, float lambda, float corr, float colsum) {
    return colsum > 0.0f ? x + lambda * corr / colsum : x;
}

// ---- batched plans: one problem as the persistent kernels see it ----
struct BatchProb {
    Grid g;
    Det dt;
    int zonly;
    int3 nbr;                  // bricks per axis for the plan's brick shape
    long long nvox;
    const float* fv;           // frame vectors [ns][18]
    const int8_t* ind;         // indicator or null
    const float* sino;         // measured sinogram
    float* vol;                // estimate (read by the forward, updated by SART)
    float* proj;               // forward output (plain projection) or null
    float* adj;                // adjoint output (plain backprojection) or null
    float* rowsum;             // A1 per ray
    float* resid;              // SART residual sinogram
    float* colsum;             // per subset s: A_s^T 1 at colsum + s*nvox
};
enum { BFWD_PLAIN = 0, BFWD_ROWSUM = 1, BFWD_RESID = 2 };
enum { BBP_PLAIN = 0, BBP_COLSUM = 1, BBP_SART = 2 };
// number of views p < ns with p = sub (mod nsub)
inline __host__ __device__ int subset_views(int ns, int sub, int nsub) {
    return sub < ns ? (ns - sub + nsub - 1) / nsub : 0;
}
// The problem whose items hold item it: off[p] is the first item of problem p
// (problems without items repeat the next offset), off[nprob] the total.
__device__ __forceinline__ int batch_find(const long long* __restrict__ off, int nprob,
                                          long long it) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= it) lo = mid; else hi = mid - 1;
    }
    return lo;
}
// One problem of a batched plan on the column-separable adjoint
// (ct2_bp_sep.hip): SART's adjoint of subset s reads resid over the subset's
// views, its x-driven tiles store their sums in corr, and its y-driven tiles
// add theirs and update vol with the column sums at colsum + s*nvox; the
// column sums themselves are made the same way from a ones sinogram.
struct SepBatchProb {
    Grid g;
    Det dt;
    long long nvox;
    const float* fv;
    const int8_t* ind;
    const float* resid;
    double2* tabs;             // column tables of the problem's largest subset
    float* corr;
    float* vol;
    float* colsum;
};
enum { BSEP_SART = 0, BSEP_COLSUM = 1 };
// Tile counts of the x- and y-driven launches of a grid.
int sep_tiles(const Grid& g, int axis);
// z cells per lane of a grid's tiles (1 .. 4)
inline int sep_cpl(const Grid& g) { return (g.n[2] + 63) / 64; }
size_t sep_lds_bytes(int nv);
// Column tables of subset sub of every problem: items (problem, view of the
// subset, detector column), first item of problem p at off[p].
int launch_batch_sep_tables(const SepBatchProb* probs, const long long* off, int nprob, int sub,
                            int nsub, long long items, int* nonsep, hipStream_t st);
// The tiles of driving axis ax (0: x, stores; 1: y, adds and updates) of
// the problems whose tiles off numbers (those of z-cell class cpl), for kind
// BSEP_SART or BSEP_COLSUM; max_nv sizes the LDS tables.
int launch_batch_bp_sep(int cpl, int ax, int kind, const SepBatchProb* probs,
                        const long long* off, int nprob, int sub, int nsub, long long tiles,
                        int max_nv, float lambda, const float* lambda_dev, hipStream_t st);
// Count of views p < ns that view_separable rejects, added to *count.
int launch_count_nonsep(const Grid& g, const Det& dt, const float* fv, int* count,
                        hipStream_t st);
int launch_batch_project(int interp, int arith, int tu, int mode, const BatchProb* probs,
                         const long long* off, int nprob, int sub, int nsub, const float* ones,
                         unsigned long long* queue, int grid, hipStream_t st);
int batch_project_occupancy(int interp, int arith, int tu, int mode, int* blocks_per_cu);
// lambda_dev (BBP_SART only, may be null): relaxation read at run time
int launch_batch_backproject(int interp, int arith, int mode, const BatchProb* probs,
                             const long long* off, int nprob, int sub, int nsub, int3 bdim,
                             float lambda, const float* lambda_dev, unsigned long long* queue,
                             int grid, hipStream_t st);
int batch_backproject_occupancy(int interp, int arith, int mode, int3 bdim, int* blocks_per_cu);
size_t brick_lds_bytes(int3 bdim);
// adjoint brick shapes the autotuners time (single problems and batched plans)
// (at 256^3 x 360 views with dense data: 12x12x24 63.8 ms, 10x10x24 64.9,
// 14x14x14 66.0, 12x12x14 66.4, 16x16x12 71, 16x16x8 94).  The last four are
// flat and long across the direction a source arc moves the rays: z-driven
// tomosynthesis rays (1024x768x64 at 0.1 x 0.1 x 1 mm, +-25 deg in y-z) drift
// 4.7 voxels in y per plane, so a brick long in y keeps each ray's set-up for
// more samples (16x32x8 6.8 ms, 8x64x8 6.5 ms against 10.2 ms for the best
// of the first eight).
constexpr int kBrickCount = 12;
inline constexpr int kBrickCandidates[kBrickCount][3] = {
    {12, 12, 14}, {12, 12, 24}, {10, 10, 24}, {14, 14, 14}, {16, 12, 14}, {12, 16, 14},
    {16, 16, 12}, {16, 16, 16}, {16, 32, 8},  {32, 16, 8},  {8, 64, 8},   {64, 8, 8}};

// View subset of a descriptor: views view0 + j*vstride for j < nviews.
struct ViewSet {
    int view0, nviews, vstride;
};
int view_set(const ct2_desc* d, const char* who, ViewSet& vs);

// ------------------------------------------------------ launchers (host) ----
int launch_project(const ct2_desc* desc, void* sino, const int8_t* ind, const float* fv,
                   const void* vol, hipStream_t st);
// The forward of *desc on the volume at vol runs the column-separable kernel
// (which neither tunes nor synchronises).
bool project_sep_route(const ct2_desc* desc, const void* vol);
// Caller-owned device workspace of the column-separable adjoint (its column
// tables); a null pointer lets the launch allocate one on the stream.
struct SepWork {
    void* ptr = nullptr;
    size_t bytes = 0;
};
int launch_backproject(const ct2_desc* desc, void* vol, const void* sino, const int8_t* ind,
                       const float* fv, hipStream_t st, SepWork w = SepWork());
int launch_ray_project(const ct2_desc* desc, float* sino, const int8_t* ind, const float* rv,
                       int nr, const float* vol, int strict, hipStream_t st);
int launch_ray_backproject(const ct2_desc* desc, float* vol, const float* sino,
                           const int8_t* ind, const float* rv, int nr, int strict,
                           hipStream_t st);
int launch_art(const ct2_desc* d, const Grid* g2, float reg_fac, const float* sino,
               const int8_t* ind, const float* fv, float* vol, float* vol2, hipStream_t st);
int launch_ray_art(const ct2_desc* d, int strict, float reg_fac, const float* sino,
                   const int8_t* ind, const float* rv, int nr, float* vol, hipStream_t st);
size_t em_work_floats(const ct2_desc* d);
int launch_em(const ct2_desc* d, const float* sino, const int8_t* ind, const float* fv,
              float* vol, float* work, const float* sens, hipStream_t st);
int launch_ray_em(const ct2_desc* d, int strict, const float* sino, const int8_t* ind,
                  const float* rv, int nr, float* vol, float* work, hipStream_t st);
size_t sart_work_floats(const ct2_desc* d);
int launch_sart(const ct2_desc* d, float lambda, const float* sino, const int8_t* ind,
                const float* fv, float* vol, float* work, hipStream_t st);
// Column-separable adjoint (ct2_bp_sep.hip) and when it applies.
bool sep_supported(const ct2_desc* d, const Grid& g);
// The adjoint takes the column-separable route when sep_supported holds and
// the caller names no brick (a brick shape selects the brick kernel, whose
// weights are the fast forward's own FP32 values).
inline bool sep_route(const ct2_desc* d, const Grid& g) {
    return d->brick[0] <= 0 && d->brick[1] <= 0 && d->brick[2] <= 0 && sep_supported(d, g);
}
int launch_backproject_sep(const Grid& g, const Det& dt, const ViewSet& vs, float* vol,
                           const float* sino, const int8_t* ind, const float* fv, hipStream_t st,
                           SepWork w);
// Bytes of the separable adjoint's workspace for every view of the subset (0
// when the route does not apply): a 256-byte header (the count of views the
// route leaves to the brick kernel) and the column tables.
size_t sep_work_bytes(const ct2_desc* d);
// A workspace for the route: the caller's (checked), or one allocated on the
// stream (*own, released by sep_work_release after the launches that use
// it).
int sep_work_acquire(const Det& dt, const ViewSet& vs, SepWork& w, void** own, hipStream_t st);
int sep_work_release(void* own, hipStream_t st);
// The device count of views of the last separable launch that
// view_separable rejected (in the workspace header).
inline const int* sep_nonsep_count(const SepWork& w) { return static_cast<const int*>(w.ptr); }
// Plane-separable adjoint of the z-driven projector (ct2_bp_zsep.hip): fast
// linear arithmetic into an f32 volume in C order, axis forced to z, no
// brick named.  It stores every voxel from the views view_zsep accepts; a
// brick pass then adds the others.
// Its workspace (w, or allocated on the stream when w is empty and
// returned in *own, for sep_work_release after the brick pass) holds a
// header, the device count of views the route leaves (*nonsep, for the
// brick pass), and the records of a group of views; zsep_work_bytes is the
// size for every view.
bool zsep_route(const ct2_desc* d, const Grid& g);
size_t zsep_work_bytes(const ct2_desc* d, const Det& dt, const ViewSet& vs);
int launch_backproject_zsep(const Grid& g, const Det& dt, const ViewSet& vs, float* vol,
                            const float* sino, const int8_t* ind, const float* fv, hipStream_t st,
                            SepWork w, void** own, const int** nonsep);
// Lets the device's stream-ordered pool keep freed workspaces for the next
// call instead of unmapping them at each synchronisation (set once per
// device).
void keep_pool_memory(int dev);
int launch_pd_backproject(const ct2_desc* d, float* vol, const float* sino, const float* fv,
                          hipStream_t st);
int launch_convert(const float* src, void* dst, int dtype, size_t n, int rounding, unsigned seed,
                   hipStream_t st);
int launch_convert_zplanes(const float* src, void* dst, int dtype, int nx, int ny, int nz,
                           int rounding, unsigned seed, hipStream_t st);
inline bool is_zplanes(int dtype) {
    return dtype == CT2_F32_ZPLANES || dtype == CT2_F16_ZPLANES || dtype == CT2_BF16_ZPLANES;
}
// device fill (graph-capture safe, unlike a memset node)
int launch_fill(float* p, float v, size_t n, hipStream_t st);

}  // namespace ct2
