// ct2_stencil.hip -- weighted ray-bunch ("stencil") measurements and their
// ART: the weighted and derivative kernels whose call sites and checks the
// reference keeps (tests/cweighted_art_test.py:80-140: cweighted_project_to,
// put_weights, c_ss_weighted_art_step; tests/cderiv_art_test.py:106-140:
// get_deriv_weights, cderiv_art_step) while their GPU sources are missing
// (emil_code/Makefile.Old:28-45).
//
// A coarse bin b = (I,J,K) of the grid [NS][NU][NV], NS = ceil(ns/nss), ...,
// measures m_b = sum_q w_q (A x)[ray q of b]: ray q = (a,c,d) of the kernel
// w[ws][wu][wv] is the fine ray (I*nss + a - os, J*nsu + c - ou,
// K*nsv + d - ov) with os = ws-1-ws/2 (and likewise ou, ov), rays outside the
// fine sinogram skipped.  That is the reference test's own check,
// scipy.ndimage.convolve(proj, w[::-1,::-1,::-1], mode='constant')[::nss,
// ::nsu, ::nsv].  The bin is used when the fine indicator at its anchor
// (I*nss, J*nsu, K*nsv) is 1, as the tests set it.  Its row of the system
// matrix is r_b = sum_q w_q a_q, and
//   k_stencil_correlate  m = correlation of a fine projection with w;
//   k_stencil_weights    weights[b] = |r_b|^2 and image += sum_b r_b;
//   k_stencil_art        Kaczmarz over the used bins in (I,J,K) order:
//                        x += reg_fac (m_b - r_b.x) r_b / |r_b|^2.
// The derivative kernel is the stencil w = [-1, +1] along u with unit strides.
//
// Schedule.  As for single-ray ART (ct2_iterative.hip), one workgroup of 16
// wavefronts walks levels: in coarse view I, bin (J,K) has level
// stride*J + K with stride = RK+1, where bins more than RK apart in K have
// fine rays more than the geometric conflict radius apart and so touch
// disjoint voxels.  Conflicting bins keep their sequential order and bins of
// one level are independent; each wavefront takes one bin and walks its rays
// in kernel order, lanes over planes.  A kernel spanning several fine views
// (ws > 1) makes every bin of a coarse view conflict, and the schedule is
// sequential.
//
// |r_b|^2 needs the overlaps of the bin's rays: the wavefront scatters r_b
// into a zeroed scratch volume (rays in order), then marches the rays again,
// reading and clearing each touched voxel, so every voxel of r_b is counted
// once (in FP64) and added to image once, and the scratch is zero again.
#include <math.h>

#include "ct2_internal.h"

namespace ct2 {

int conflict_radius(const Grid& g, const Grid* g2, const Det& dt, const float* fv_host,
                    int view0, int nviews);

namespace {

constexpr int kWaves = 16;
constexpr int kThreads = 256;

struct Stencil {
    int ws, wu, wv, nss, nsu, nsv, os, ou, ov, NS, NU, NV;
};

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Orders this wavefront's global stores before its later loads, across lanes.
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bool stencil_ray(const Stencil& s, const Det& dt, int I, int J, int K,
                                            int q, int& p, int& j, int& k) {
    const int d = q % s.wv, t = q / s.wv;
    const int c = t % s.wu, a = t / s.wu;
    p = I * s.nss + a - s.os;
    j = J * s.nsu + c - s.ou;
    k = K * s.nsv + d - s.ov;
    return p >= 0 && p < dt.ns && j >= 0 && j < dt.nu && k >= 0 && k < dt.nv;
}

__global__ __launch_bounds__(kThreads) void k_stencil_correlate(Stencil s, Det dt,
                                                                const float* __restrict__ w,
                                                                const float* __restrict__ fine,
                                                                float* __restrict__ small,
                                                                long long n) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const int K = (int)(o % s.NV);
    const long long t = o / s.NV;
    const int J = (int)(t % s.NU), I = (int)(t / s.NU);
    const int R = s.ws * s.wu * s.wv;
    float acc = 0.f;
    for (int q = 0; q < R; ++q) {
        int p, j, k;
        if (!stencil_ray(s, dt, I, J, K, q, p, j, k)) continue;
        acc = acc + w[q] * fine[((long long)p * dt.nu + j) * dt.nv + k];
    }
    small[o] = acc;
}

// Calls f(voxel index, matrix entry) for every entry of one ray, lanes over
// planes; within a ray every voxel appears at most once.
template <int INTERP, class F>
__device__ __forceinline__ void for_entries(const AxisSel& X, int lane, F&& f) {
    for (int l = lane; l < X.na; l += 64) {
        Sample sm;
        if (INTERP == CT2_LINEAR) {
            if (!sample_linear_exact(X, l, sm)) continue;
            const int i00 = l * X.sa + sm.ib * X.sb + sm.ic * X.sc;
            f(i00, X.trav * sm.wb1 * sm.wc1);
            f(i00 + X.sb, X.trav * sm.wb2 * sm.wc1);
            f(i00 + X.sc, X.trav * sm.wb1 * sm.wc2);
            f(i00 + X.sb + X.sc, X.trav * sm.wb2 * sm.wc2);
        } else {
            if (!sample_nearest_exact(X, l, false, sm)) continue;
            f(l * X.sa + sm.ib * X.sb + sm.ic * X.sc, X.trav);
        }
    }
}

template <int INTERP>
__device__ __forceinline__ float ray_sum(const AxisSel& X, const float* __restrict__ V, int lane) {
    float rs = 0.f;
    for (int l = lane; l < X.na; l += 64) {
        Sample sm;
        if (INTERP == CT2_LINEAR) {
            if (!sample_linear_exact(X, l, sm)) continue;
            const int i00 = l * X.sa + sm.ib * X.sb + sm.ic * X.sc;
            float vi1 = sm.wb1 * V[i00] + sm.wb2 * V[i00 + X.sb];
            float vi2 = sm.wb1 * V[i00 + X.sc] + sm.wb2 * V[i00 + X.sb + X.sc];
            rs = rs + X.trav * (sm.wc1 * vi1 + sm.wc2 * vi2);
        } else {
            if (!sample_nearest_exact(X, l, false, sm)) continue;
            rs = rs + X.trav * V[l * X.sa + sm.ib * X.sb + sm.ic * X.sc];
        }
    }
    return wave_sum_f(rs);
}

// Bins of level L in one coarse view: K = L - stride*J, J in [jlo, jhi].
__device__ __forceinline__ void level_range(int L, int stride, const Stencil& s, int& jlo,
                                            int& jhi) {
    jlo = L - s.NV + 1 <= 0 ? 0 : (L - s.NV + 1 + stride - 1) / stride;
    jhi = min(s.NU - 1, L / stride);
}

template <int INTERP>
__global__ __launch_bounds__(64 * kWaves) void k_stencil_weights(
    Grid g, Det dt, Stencil s, int zonly, int stride, const float* __restrict__ w,
    const int8_t* __restrict__ ind, const float* __restrict__ fv, float* __restrict__ weights,
    float* __restrict__ image, float* __restrict__ scratch) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nlev = stride * (s.NU - 1) + s.NV;
    const int R = s.ws * s.wu * s.wv;
    for (int I = 0; I < s.NS; ++I) {
        for (int L = 0; L < nlev; ++L) {
            int jlo, jhi;
            level_range(L, stride, s, jlo, jhi);
            for (int J = jlo + wave; J <= jhi; J += kWaves) {
                const int K = L - stride * J;
                const long long b = ((long long)I * s.NU + J) * s.NV + K;
                const long long anchor =
                    ((long long)(I * s.nss) * dt.nu + J * s.nsu) * dt.nv + K * s.nsv;
                if (ind != nullptr && ind[anchor] != 1) {
                    if (lane == 0) weights[b] = 0.f;
                    continue;
                }
                for (int q = 0; q < R; ++q) {          // scatter r_b
                    int p, j, k;
                    if (!stencil_ray(s, dt, I, J, K, q, p, j, k)) continue;
                    float S[3], B[3];
                    view_ray(fv, p, j, k, dt, S, B);
                    AxisSel X;
                    setup_exact(g, S, B, zonly != 0, X);
                    const float wq = w[q];
                    for_entries<INTERP>(X, lane, [&](int i, float a) {
                        scratch[i] = scratch[i] + wq * a;
                    });
                    wave_fence();
                }
                double acc = 0.0;
                for (int q = 0; q < R; ++q) {          // read, count once, clear
                    int p, j, k;
                    if (!stencil_ray(s, dt, I, J, K, q, p, j, k)) continue;
                    float S[3], B[3];
                    view_ray(fv, p, j, k, dt, S, B);
                    AxisSel X;
                    setup_exact(g, S, B, zonly != 0, X);
                    for_entries<INTERP>(X, lane, [&](int i, float) {
                        const float v = scratch[i];
                        if (v != 0.f) {
                            acc += (double)v * (double)v;
                            image[i] = image[i] + v;
                            scratch[i] = 0.f;
                        }
                    });
                    wave_fence();
                }
                acc = wave_sum_d(acc);
                if (lane == 0) weights[b] = (float)acc;
            }
            __syncthreads();
        }
    }
}

template <int INTERP>
__global__ __launch_bounds__(64 * kWaves) void k_stencil_art(
    Grid g, Det dt, Stencil s, int zonly, int stride, const float* __restrict__ w, float reg_fac,
    const float* __restrict__ small, const float* __restrict__ weights,
    const int8_t* __restrict__ ind, const float* __restrict__ fv, float* __restrict__ vol) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nlev = stride * (s.NU - 1) + s.NV;
    const int R = s.ws * s.wu * s.wv;
    for (int I = 0; I < s.NS; ++I) {
        for (int L = 0; L < nlev; ++L) {
            int jlo, jhi;
            level_range(L, stride, s, jlo, jhi);
            for (int J = jlo + wave; J <= jhi; J += kWaves) {
                const int K = L - stride * J;
                const long long b = ((long long)I * s.NU + J) * s.NV + K;
                const long long anchor =
                    ((long long)(I * s.nss) * dt.nu + J * s.nsu) * dt.nv + K * s.nsv;
                if (ind != nullptr && ind[anchor] != 1) continue;
                const float norm = weights[b];
                if (!(norm > 0.f)) continue;           // empty row
                float rsum = 0.f;
                for (int q = 0; q < R; ++q) {
                    int p, j, k;
                    if (!stencil_ray(s, dt, I, J, K, q, p, j, k)) continue;
                    float S[3], B[3];
                    view_ray(fv, p, j, k, dt, S, B);
                    AxisSel X;
                    setup_exact(g, S, B, zonly != 0, X);
                    rsum = rsum + w[q] * ray_sum<INTERP>(X, vol, lane);
                }
                const float resid = reg_fac * (small[b] - rsum);
                for (int q = 0; q < R; ++q) {
                    int p, j, k;
                    if (!stencil_ray(s, dt, I, J, K, q, p, j, k)) continue;
                    float S[3], B[3];
                    view_ray(fv, p, j, k, dt, S, B);
                    AxisSel X;
                    setup_exact(g, S, B, zonly != 0, X);
                    const float wq = w[q];
                    for_entries<INTERP>(X, lane, [&](int i, float a) {
                        vol[i] = vol[i] + resid * (wq * a) / norm;
                    });
                    wave_fence();
                }
            }
            __syncthreads();
        }
    }
}

unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

struct Setup {
    Grid g;
    Det dt;
    Stencil s;
    int zonly;
};

int prepare(const char* who, const ct2_desc* d, const ct2_stencil* st, Setup& u) {
    if (!d || !st) { set_error("%s: null descriptor", who); return CT2_EINVAL; }
    u.g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    u.dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(u.g)) || (rc = check_det(u.dt))) return rc;
    if (d->vol_dtype != CT2_F32) { set_error("%s: float32 volumes only", who); return CT2_EINVAL; }
    if (d->interp != CT2_LINEAR && d->interp != CT2_NEAREST) {
        set_error("%s: bad interpolation", who); return CT2_EINVAL;
    }
    const int* v = &st->ws;
    for (int q = 0; q < 6; ++q)
        if (v[q] <= 0) { set_error("%s: stencil extents and strides must be positive", who); return CT2_EINVAL; }
    if ((long long)st->ws * st->wu * st->wv > 4096) {
        set_error("%s: stencil of more than 4096 rays", who); return CT2_EINVAL;
    }
    Stencil& s = u.s;
    s.ws = st->ws; s.wu = st->wu; s.wv = st->wv;
    s.nss = st->nss; s.nsu = st->nsu; s.nsv = st->nsv;
    s.os = s.ws - 1 - s.ws / 2; s.ou = s.wu - 1 - s.wu / 2; s.ov = s.wv - 1 - s.wv / 2;
    s.NS = (d->ns + s.nss - 1) / s.nss;
    s.NU = (d->nu + s.nsu - 1) / s.nsu;
    s.NV = (d->nv + s.nsv - 1) / s.nsv;
    u.zonly = d->axis == CT2_AXIS_Z ? 1 : 0;
    return CT2_OK;
}

// Level stride of the coarse K index (see the file comment).
int level_stride(const Setup& u, const float* fv, hipStream_t stream, int& stride) {
    const Stencil& s = u.s;
    stride = s.NV;
    if (s.ws > 1) return CT2_OK;
    float* fvh = (float*)malloc(sizeof(float) * 18 * (size_t)u.dt.ns);
    if (!fvh) { set_error("stencil: host allocation failed"); return CT2_ENOMEM; }
    hipError_t e = hipMemcpyAsync(fvh, fv, sizeof(float) * 18 * (size_t)u.dt.ns,
                                  hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) { free(fvh); return hip_fail(e, "stencil: frame copy"); }
    int r = conflict_radius(u.g, nullptr, u.dt, fvh, 0, u.dt.ns);
    free(fvh);
    if (r < 0) return CT2_OK;
    long long rk = ((long long)r + s.wv - 1) / s.nsv;
    if (rk + 1 < s.NV) stride = (int)(rk + 1);
    return CT2_OK;
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_stencil_project(const ct2_desc* desc, const ct2_stencil* st, const float* w,
                        const float* vol, const float* frame_vectors, float* fine, float* small,
                        ct2_stream_t stream) {
    Setup u;
    int rc = prepare("ct2_stencil_project", desc, st, u);
    if (rc) return rc;
    if (!w || !vol || !frame_vectors || !fine || !small) {
        set_error("ct2_stencil_project: null pointer"); return CT2_EINVAL;
    }
    ct2_desc d = *desc;
    d.arith = CT2_ARITH_EXACT;
    d.view0 = 0; d.nviews = 0; d.vstride = 1;
    rc = ct2_project(&d, fine, nullptr, frame_vectors, vol, stream);
    if (rc) return rc;
    long long n = (long long)u.s.NS * u.s.NU * u.s.NV;
    hipLaunchKernelGGL(k_stencil_correlate, dim3(blocks_for(n)), dim3(kThreads), 0,
                       (hipStream_t)stream, u.s, u.dt, w, fine, small, n);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_stencil_weights(const ct2_desc* desc, const ct2_stencil* st, const float* w,
                        const int8_t* indsino, const float* frame_vectors, float* weights,
                        float* image, float* scratch, ct2_stream_t stream) {
    Setup u;
    int rc = prepare("ct2_stencil_weights", desc, st, u);
    if (rc) return rc;
    if (!w || !frame_vectors || !weights || !image || !scratch) {
        set_error("ct2_stencil_weights: null pointer"); return CT2_EINVAL;
    }
    hipStream_t hs = (hipStream_t)stream;
    int stride;
    if ((rc = level_stride(u, frame_vectors, hs, stride))) return rc;
    if (desc->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_stencil_weights<CT2_LINEAR>, dim3(1), dim3(64 * kWaves), 0, hs, u.g,
                           u.dt, u.s, u.zonly, stride, w, indsino, frame_vectors, weights, image,
                           scratch);
    else
        hipLaunchKernelGGL(k_stencil_weights<CT2_NEAREST>, dim3(1), dim3(64 * kWaves), 0, hs, u.g,
                           u.dt, u.s, u.zonly, stride, w, indsino, frame_vectors, weights, image,
                           scratch);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_stencil_art(const ct2_desc* desc, const ct2_stencil* st, const float* w, float reg_fac,
                    const float* small, const float* weights, const int8_t* indsino,
                    const float* frame_vectors, float* vol, ct2_stream_t stream) {
    Setup u;
    int rc = prepare("ct2_stencil_art", desc, st, u);
    if (rc) return rc;
    if (!w || !small || !weights || !frame_vectors || !vol) {
        set_error("ct2_stencil_art: null pointer"); return CT2_EINVAL;
    }
    hipStream_t hs = (hipStream_t)stream;
    int stride;
    if ((rc = level_stride(u, frame_vectors, hs, stride))) return rc;
    if (desc->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_stencil_art<CT2_LINEAR>, dim3(1), dim3(64 * kWaves), 0, hs, u.g, u.dt,
                           u.s, u.zonly, stride, w, reg_fac, small, weights, indsino,
                           frame_vectors, vol);
    else
        hipLaunchKernelGGL(k_stencil_art<CT2_NEAREST>, dim3(1), dim3(64 * kWaves), 0, hs, u.g,
                           u.dt, u.s, u.zonly, stride, w, reg_fac, small, weights, indsino,
                           frame_vectors, vol);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // extern "C"
