// ct2_tv.hip -- isotropic total variation value and gradient (totalvar.f).
//
// total_var3d (totalvar.f:35-62): sum over voxels with i,j,k >= 1 of
// sqrt(eps + backward-difference^2 sum) + lambda*|x|, accumulated in FP64,
// minus (nx-1)(ny-1)(nz-1)*sqrt(eps).
// total_var_grad3d (totalvar.f:476-602): per voxel the main term
// (3x - x[-i] - x[-j] - x[-k]) / n(q) + lambda*x/(eps+|x|) plus the three
// one-ahead terms (x - x[+a]) / n(q+a), added in the reference's order, each
// norm n() with the reference's operand order; then the FP64 gradient norm
// bobo (|bobo| < 1e-5 -> 1) and optional normalisation.
//
// One thread per voxel (z fastest) gathers its 13-point stencil; FP32 math
// matches the Fortran bit for bit.  The FP64 sums use a fixed two-stage
// reduction (per-block partials, then one block), so results are
// deterministic.
#include <math.h>

#include "ct2_internal.h"

namespace ct2 {

constexpr int TV_BLOCKS = 1024;
static_assert(TV_BLOCKS + 1 <= CT2_TV_WORK_DOUBLES, "TV workspace too small");
constexpr int TV_THREADS = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    return t;   // valid in thread 0
}

struct TvDims {
    int nx, ny, nz;
    float lam, eps;
    Div32 divz, divy;   // voxel index -> (i, j, k) without 64-bit division
};

// (i, j, k) of flat index q < 2^31 of a [nx][ny][nz] volume
__device__ __forceinline__ void voxel_ijk(unsigned q, int ny, int nz, const Div32& divz,
                                          const Div32& divy, int& i, int& j, int& k) {
    const unsigned t = div32(q, divz);
    k = (int)(q - t * (unsigned)nz);
    const unsigned ii = div32(t, divy);
    j = (int)(t - ii * (unsigned)ny);
    i = (int)ii;
}

__device__ __forceinline__ float tv_grad_at(const float* __restrict__ s, const TvDims& d, int i,
                                            int j, int k) {
    const int sx = d.ny * d.nz, sy = d.nz;
    const size_t q = (size_t)i * sx + (size_t)j * sy + k;
    const float eps = d.eps;
    float g = 0.0f;
    bool set = false;
    if (i >= 1 && j >= 1 && k >= 1) {
        float w5 = s[q - 1], w11 = s[q - sy], w13 = s[q - sx], w14 = s[q];
        float term1 = (-w11 - w13 + 3.0f * w14 - w5) /
                      sqrtf((w11 - w14) * (w11 - w14) + (w13 - w14) * (w13 - w14) +
                            (-w14 + w5) * (-w14 + w5) + eps);
        float term2 = d.lam * w14 / (eps + fabsf(w14));
        g = term1 + term2;
        set = true;
    }
    if (i < d.nx - 1 && j >= 1 && k >= 1) {
        float w6 = s[q + sx - 1], w12 = s[q + sx - sy], w14 = s[q], w15 = s[q + sx];
        float t = (w14 - w15) / sqrtf((w12 - w15) * (w12 - w15) + (w14 - w15) * (w14 - w15) +
                                      (-w15 + w6) * (-w15 + w6) + eps);
        g = set ? g + t : 0.0f + t;
        set = true;
    }
    if (i >= 1 && j < d.ny - 1 && k >= 1) {
        float w8 = s[q + sy - 1], w14 = s[q], w16 = s[q - sx + sy], w17 = s[q + sy];
        float t = (w14 - w17) / sqrtf((w14 - w17) * (w14 - w17) + (w16 - w17) * (w16 - w17) +
                                      (-w17 + w8) * (-w17 + w8) + eps);
        g = set ? g + t : 0.0f + t;
        set = true;
    }
    if (i >= 1 && j >= 1 && k < d.nz - 1) {
        float w14 = s[q], w20 = s[q - sy + 1], w22 = s[q - sx + 1], w23 = s[q + 1];
        float t = (w14 - w23) / sqrtf((w14 - w23) * (w14 - w23) + (w20 - w23) * (w20 - w23) +
                                      (w22 - w23) * (w22 - w23) + eps);
        g = set ? g + t : 0.0f + t;
    }
    return g;
}

__global__ __launch_bounds__(TV_THREADS) void k_tv_grad(TvDims d, const float* __restrict__ s,
                                                        float* __restrict__ g,
                                                        double* __restrict__ partial) {
    __shared__ double sh[TV_THREADS / 64];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        voxel_ijk((unsigned)q, d.ny, d.nz, d.divz, d.divy, i, j, k);
        float v = tv_grad_at(s, d, i, j, k);
        g[q] = v;
        acc += (double)(v * v);   // Real*4 square accumulated in Real*8 (totalvar.f:580)
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Finalise the norm (one block) and publish bobo.
__global__ __launch_bounds__(TV_THREADS) void k_tv_norm(int nparts,
                                                        const double* __restrict__ partial,
                                                        float* __restrict__ bobo_out,
                                                        float* __restrict__ bobo_scratch) {
    __shared__ double sh[TV_THREADS / 64];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += partial[q];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) {
        float bobo = (float)sqrt(t);
        if (fabsf(bobo) < 0.00001f) bobo = 1.0f;
        bobo_scratch[0] = bobo;
        if (bobo_out) bobo_out[0] = bobo;
    }
}

__global__ void k_tv_normalize(size_t n, float* __restrict__ g, const float* __restrict__ bobo) {
    const float b = bobo[0];
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x)
        g[q] = g[q] / b;
}

__global__ __launch_bounds__(TV_THREADS) void k_tv_value(TvDims d, const float* __restrict__ s,
                                                         double* __restrict__ partial) {
    __shared__ double sh[TV_THREADS / 64];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    const int sx = d.ny * d.nz, sy = d.nz;
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        voxel_ijk((unsigned)q, d.ny, d.nz, d.divz, d.divy, i, j, k);
        if (i < 1 || j < 1 || k < 1) continue;
        float c = s[q];
        float ax = c - s[q - sx], ay = c - s[q - sy], az = c - s[q - 1];
        acc += (double)sqrtf(d.eps + ax * ax + ay * ay + az * az);
        acc += (double)(d.lam * fabsf(c));
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(TV_THREADS) void k_tv_value_final(int nparts,
                                                               const double* __restrict__ partial,
                                                               double shift,
                                                               double* __restrict__ out) {
    __shared__ double sh[TV_THREADS / 64];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += partial[q];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = (double)(float)(t - shift);   // Real*4 function result
}

// Distance sqrt(sum mask*(a-b)^2) in FP64 (tomo3D.py:319-324 for sinograms
// with their indicator, :946-951 for images): per-block partial sums, then
// one block adds the partials in a fixed order (deterministic).
__global__ __launch_bounds__(TV_THREADS) void k_dist_partial(size_t n, const float* __restrict__ a,
                                                             const float* __restrict__ b,
                                                             const int8_t* __restrict__ mask,
                                                             double* __restrict__ partial) {
    __shared__ double sh[TV_THREADS / 64];
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        double d = (double)a[q] - (b ? (double)b[q] : 0.0);
        if (!mask || mask[q] == 1) acc += d * d;
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(TV_THREADS) void k_dist_final(int nparts,
                                                           const double* __restrict__ partial,
                                                           double* __restrict__ out) {
    __shared__ double sh[TV_THREADS / 64];
    double acc = 0.0;
    for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += partial[q];
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) out[0] = sqrt(t);
}

// ---- p-variants (totalvar.f:4-472): x**p is powf(x, p), x**2 is x*x ----
enum { PV_LP = 0, PV_TPV = 1, PV_TPDV = 2 };
struct PvDims {
    int nx, ny, nz;
    float p, eps, dx, dy, dz;
    Div32 divz, divy;
};

// gradient at voxel (i,j,k): lp_grad3d (:150-215), total_p_var_grad3d
// (:219-343) and total_p_dvar_grad3d (:347-472), terms in the reference's order
template <int K>
__device__ __forceinline__ float pv_grad_at(const float* __restrict__ s, const PvDims& d, int i,
                                            int j, int k) {
    const int sx = d.ny * d.nz, sy = d.nz;
    const size_t q = (size_t)i * sx + (size_t)j * sy + k;
    const float eps = d.eps, p = d.p, pm2 = d.p - 2.0f;
    if (K == PV_LP) {
        const float w14 = s[q];
        return p * w14 * powf(sqrtf(w14 * w14 + eps), pm2);
    }
    const float dx = d.dx, dy = d.dy, dz = d.dz;
    float g = 0.0f;
    if (i >= 1 && j >= 1 && k >= 1) {
        float w5 = s[q - 1], w11 = s[q - sy], w13 = s[q - sx], w14 = s[q];
        float term1;
        if (K == PV_TPV) {
            term1 = (-w11 - w13 + 3.0f * w14 - w5) *
                    powf(sqrtf((w11 - w14) * (w11 - w14) + (w13 - w14) * (w13 - w14) +
                               (-w14 + w5) * (-w14 + w5) + eps), pm2);
        } else {
            float a = (w11 - w14) / dy, b = (w13 - w14) / dx, c = (-w14 + w5) / dz;
            term1 = ((w14 - w11) / (dy * dy) + (w14 - w13) / (dx * dx) + (w14 - w5) / (dz * dz)) *
                    powf(sqrtf(a * a + b * b + c * c + eps), pm2);
        }
        g = p * term1;
    }
    if (i < d.nx - 1 && j >= 1 && k >= 1) {
        float w6 = s[q + sx - 1], w12 = s[q + sx - sy], w14 = s[q], w15 = s[q + sx];
        float t;
        if (K == PV_TPV) {
            t = (w14 - w15) * powf(sqrtf((w12 - w15) * (w12 - w15) + (w14 - w15) * (w14 - w15) +
                                         (-w15 + w6) * (-w15 + w6) + eps), pm2);
        } else {
            float a = (w12 - w15) / dy, b = (w14 - w15) / dx, c = (-w15 + w6) / dz;
            t = ((w14 - w15) / (dx * dx)) * powf(sqrtf(a * a + b * b + c * c + eps), pm2);
        }
        g = g + p * t;
    }
    if (i >= 1 && j < d.ny - 1 && k >= 1) {
        float w8 = s[q + sy - 1], w14 = s[q], w16 = s[q - sx + sy], w17 = s[q + sy];
        float t;
        if (K == PV_TPV) {
            t = (w14 - w17) * powf(sqrtf((w14 - w17) * (w14 - w17) + (w16 - w17) * (w16 - w17) +
                                         (-w17 + w8) * (-w17 + w8) + eps), pm2);
        } else {
            float a = (w14 - w17) / dy, b = (w16 - w17) / dx, c = (-w17 + w8) / dz;
            t = ((w14 - w17) / (dy * dy)) * powf(sqrtf(a * a + b * b + c * c + eps), pm2);
        }
        g = g + p * t;
    }
    if (i >= 1 && j >= 1 && k < d.nz - 1) {
        float w14 = s[q], w20 = s[q - sy + 1], w22 = s[q - sx + 1], w23 = s[q + 1];
        float t;
        if (K == PV_TPV) {
            t = (w14 - w23) * powf(sqrtf((w14 - w23) * (w14 - w23) + (w20 - w23) * (w20 - w23) +
                                         (w22 - w23) * (w22 - w23) + eps), pm2);
        } else {
            float a = (w14 - w23) / dz, b = (w20 - w23) / dy, c = (w22 - w23) / dx;
            t = ((w14 - w23) / (dz * dz)) * powf(sqrtf(a * a + b * b + c * c + eps), pm2);
        }
        g = g + p * t;
    }
    return g;
}

template <int K>
__global__ __launch_bounds__(TV_THREADS) void k_pv_grad(PvDims d, const float* __restrict__ s,
                                                        float* __restrict__ g,
                                                        double* __restrict__ partial) {
    __shared__ double sh[TV_THREADS / 64];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        voxel_ijk((unsigned)q, d.ny, d.nz, d.divz, d.divy, i, j, k);
        float v = pv_grad_at<K>(s, d, i, j, k);
        g[q] = v;
        acc += (double)(v * v);
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// values: calc_lp (:65-87, every voxel), total_var3d_p (:91-117) and
// total_dvar3d_p (:121-146) over voxels with all indices >= 1
template <int K>
__global__ __launch_bounds__(TV_THREADS) void k_pv_value(PvDims d, const float* __restrict__ s,
                                                         double* __restrict__ partial) {
    __shared__ double sh[TV_THREADS / 64];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    const int sx = d.ny * d.nz, sy = d.nz;
    double acc = 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        if (K == PV_LP) {
            acc += (double)powf(sqrtf(d.eps + s[q] * s[q]), d.p);
            continue;
        }
        int i, j, k;
        voxel_ijk((unsigned)q, d.ny, d.nz, d.divz, d.divy, i, j, k);
        if (i < 1 || j < 1 || k < 1) continue;
        float c = s[q];
        float ax = c - s[q - sx], ay = c - s[q - sy], az = c - s[q - 1];
        if (K == PV_TPDV) { ax = ax / d.dx; ay = ay / d.dy; az = az / d.dz; }
        acc += (double)powf(sqrtf(d.eps + ax * ax + ay * ay + az * az), d.p);
    }
    double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// gmi (:4-31): gradient magnitude at voxels with all indices >= 1, 0 elsewhere
__global__ void k_gmi(int nx, int ny, int nz, float eps, Div32 divz, Div32 divy,
                      const float* __restrict__ s, float* __restrict__ out) {
    const size_t n = (size_t)nx * ny * nz;
    const int sx = ny * nz, sy = nz;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        int i, j, k;
        voxel_ijk((unsigned)q, ny, nz, divz, divy, i, j, k);
        float v = 0.0f;
        if (i >= 1 && j >= 1 && k >= 1) {
            float c = s[q];
            float ax = c - s[q - sx], ay = c - s[q - sy], az = c - s[q - 1];
            v = sqrtf(eps + ax * ax + ay * ay + az * az);
        }
        out[q] = v;
    }
}

static int tv_blocks(int nx, int ny, int nz) {
    size_t n = (size_t)nx * ny * nz;
    int nb = (int)((n + TV_THREADS - 1) / TV_THREADS);
    return nb > TV_BLOCKS ? TV_BLOCKS : nb;
}

template <int K>
static int pv_grad(const PvDims& d, float* g, const float* s, int normalizeq, float* bobo,
                   double* work, hipStream_t st) {
    const int nb = tv_blocks(d.nx, d.ny, d.nz);
    double* partial = work;
    float* bscr = (float*)(partial + nb);
    hipLaunchKernelGGL(k_pv_grad<K>, dim3(nb), dim3(TV_THREADS), 0, st, d, s, g, partial);
    hipLaunchKernelGGL(k_tv_norm, dim3(1), dim3(TV_THREADS), 0, st, nb, partial, bobo, bscr);
    if (normalizeq == 1)
        hipLaunchKernelGGL(k_tv_normalize, dim3(nb), dim3(TV_THREADS), 0, st,
                           (size_t)d.nx * d.ny * d.nz, g, bscr);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

template <int K>
static int pv_value(const PvDims& d, const float* s, double shift, double* out, double* work,
                    hipStream_t st) {
    const int nb = tv_blocks(d.nx, d.ny, d.nz);
    hipLaunchKernelGGL(k_pv_value<K>, dim3(nb), dim3(TV_THREADS), 0, st, d, s, work);
    hipLaunchKernelGGL(k_tv_value_final, dim3(1), dim3(TV_THREADS), 0, st, nb, work, shift, out);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

static int tv_check(int nx, int ny, int nz) {
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        set_error("TV: volume shape (%d, %d, %d) must be positive", nx, ny, nz);
        return CT2_EINVAL;
    }
    if ((long long)nx * ny * nz >= (1LL << 31)) { set_error("TV: volume too large"); return CT2_EINVAL; }
    return CT2_OK;
}

}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_tv_grad(float ambdala, float* tvgmat, const float* smat, int nx, int ny, int nz,
                int normalizeq, float eps, float* bobo, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!tvgmat || !smat || !work) { set_error("ct2_tv_grad: null pointer"); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    TvDims d{nx, ny, nz, ambdala, eps, make_div32((unsigned)nz), make_div32((unsigned)ny)};
    size_t n = (size_t)nx * ny * nz;
    int nb = (int)((n + TV_THREADS - 1) / TV_THREADS);
    if (nb > TV_BLOCKS) nb = TV_BLOCKS;
    // work: nb partial sums, then the float norm used by the normalisation
    double* partial = work;
    float* bscr = (float*)(partial + nb);
    hipLaunchKernelGGL(k_tv_grad, dim3(nb), dim3(TV_THREADS), 0, st, d, smat, tvgmat, partial);
    hipLaunchKernelGGL(k_tv_norm, dim3(1), dim3(TV_THREADS), 0, st, nb, partial, bobo, bscr);
    if (normalizeq == 1)
        hipLaunchKernelGGL(k_tv_normalize, dim3(nb), dim3(TV_THREADS), 0, st, n, tvgmat, bscr);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_tv_value(float ambdala, const float* smat, int nx, int ny, int nz, float eps,
                 double* tv_out, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smat || !tv_out || !work) { set_error("ct2_tv_value: null pointer"); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    TvDims d{nx, ny, nz, ambdala, eps, make_div32((unsigned)nz), make_div32((unsigned)ny)};
    size_t n = (size_t)nx * ny * nz;
    int nb = (int)((n + TV_THREADS - 1) / TV_THREADS);
    if (nb > TV_BLOCKS) nb = TV_BLOCKS;
    // shift = (nx-1.)*(ny-1.)*(nz-1.)*sqrt(eps), Real*4 (totalvar.f:59)
    float shift = (float)(nx - 1) * (float)(ny - 1) * (float)(nz - 1) * sqrtf(eps);
    hipLaunchKernelGGL(k_tv_value, dim3(nb), dim3(TV_THREADS), 0, st, d, smat, work);
    hipLaunchKernelGGL(k_tv_value_final, dim3(1), dim3(TV_THREADS), 0, st, nb, work,
                       (double)shift, tv_out);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_distance(const float* a, const float* b, const int8_t* mask, size_t n, double* out,
                 double* work, ct2_stream_t stream) {
    if (!a || !out || !work) { set_error("ct2_distance: null pointer"); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    size_t nbz = (n + TV_THREADS - 1) / TV_THREADS;
    int nb = nbz > (size_t)TV_BLOCKS ? TV_BLOCKS : (nbz == 0 ? 1 : (int)nbz);
    hipLaunchKernelGGL(k_dist_partial, dim3(nb), dim3(TV_THREADS), 0, st, n, a, b, mask, work);
    hipLaunchKernelGGL(k_dist_final, dim3(1), dim3(TV_THREADS), 0, st, nb, work, out);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_total_var_grad3d(float ambdala, float* tvgmat, const float* smat, int nx, int ny, int nz,
                         int normalizeq, float eps, float* bobo, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!tvgmat || !smat) { set_error("ct2_total_var_grad3d: null pointer"); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    void* work = nullptr;
    CT2_HIP(hipMallocAsync(&work, sizeof(double) * CT2_TV_WORK_DOUBLES, st));
    rc = ct2_tv_grad(ambdala, tvgmat, smat, nx, ny, nz, normalizeq, eps, bobo, (double*)work,
                     stream);
    hipError_t f = hipFreeAsync(work, st);
    if (rc) return rc;
    CT2_HIP(f);
    return CT2_OK;
}

int ct2_total_var3d(float ambdala, const float* smat, int nx, int ny, int nz, float eps,
                    double* tv_out, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smat || !tv_out) { set_error("ct2_total_var3d: null pointer"); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    void* work = nullptr;
    CT2_HIP(hipMallocAsync(&work, sizeof(double) * CT2_TV_WORK_DOUBLES, st));
    rc = ct2_tv_value(ambdala, smat, nx, ny, nz, eps, tv_out, (double*)work, stream);
    hipError_t f = hipFreeAsync(work, st);
    if (rc) return rc;
    CT2_HIP(f);
    return CT2_OK;
}

int ct2_gmi(float* smatout, const float* smat, int nx, int ny, int nz, float eps,
            ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smatout || !smat) { set_error("ct2_gmi: null pointer"); return CT2_EINVAL; }
    hipLaunchKernelGGL(k_gmi, dim3(tv_blocks(nx, ny, nz)), dim3(TV_THREADS), 0,
                       (hipStream_t)stream, nx, ny, nz, eps, make_div32((unsigned)nz),
                       make_div32((unsigned)ny), smat, smatout);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_calc_lp(float p, const float* smat, int nx, int ny, int nz, float eps, double* out,
                double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smat || !out || !work) { set_error("ct2_calc_lp: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, 1.0f, 1.0f, 1.0f, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    return pv_value<PV_LP>(d, smat, 0.0, out, work, (hipStream_t)stream);
}

int ct2_total_var3d_p(float p, const float* smat, int nx, int ny, int nz, float eps,
                      double* out, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smat || !out || !work) { set_error("ct2_total_var3d_p: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, 1.0f, 1.0f, 1.0f, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    // shift = (nx-1.)*(ny-1.)*(nz-1.)*sqrt(eps)**p, Real*4 (totalvar.f:115)
    float shift = (float)(nx - 1) * (float)(ny - 1) * (float)(nz - 1) * powf(sqrtf(eps), p);
    return pv_value<PV_TPV>(d, smat, (double)shift, out, work, (hipStream_t)stream);
}

int ct2_total_dvar3d_p(float p, const float* smat, int nx, int ny, int nz, float dx, float dy,
                       float dz, float eps, double* out, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!smat || !out || !work) { set_error("ct2_total_dvar3d_p: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, dx, dy, dz, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    return pv_value<PV_TPDV>(d, smat, 0.0, out, work, (hipStream_t)stream);
}

int ct2_lp_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                  int normalizeq, float eps, float* bobo, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!tvgmat || !smat || !work) { set_error("ct2_lp_grad3d: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, 1.0f, 1.0f, 1.0f, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    return pv_grad<PV_LP>(d, tvgmat, smat, normalizeq, bobo, work, (hipStream_t)stream);
}

int ct2_total_p_var_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                           int normalizeq, float eps, float* bobo, double* work,
                           ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!tvgmat || !smat || !work) { set_error("ct2_total_p_var_grad3d: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, 1.0f, 1.0f, 1.0f, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    return pv_grad<PV_TPV>(d, tvgmat, smat, normalizeq, bobo, work, (hipStream_t)stream);
}

int ct2_total_p_dvar_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                            float dx, float dy, float dz, int normalizeq, float eps,
                            float* bobo, double* work, ct2_stream_t stream) {
    int rc = tv_check(nx, ny, nz);
    if (rc) return rc;
    if (!tvgmat || !smat || !work) { set_error("ct2_total_p_dvar_grad3d: null pointer"); return CT2_EINVAL; }
    PvDims d{nx, ny, nz, p, eps, dx, dy, dz, make_div32((unsigned)nz),
             make_div32((unsigned)ny)};
    return pv_grad<PV_TPDV>(d, tvgmat, smat, normalizeq, bobo, work, (hipStream_t)stream);
}

}  // extern "C"
