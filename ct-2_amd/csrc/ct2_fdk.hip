// ct2_fdk.hip -- FDK for a flat detector (emil_code/FDK.py).
//
// fdk_sino_weight3D (FDK.py:10-23): out = F/sqrt((F+D)^2 + vv^2 + vu^2) * sino,
//   weight evaluated in FP64 per bin.
// fdk_sino_filter3D (FDK.py:25-103): each detector line is zero-padded to
//   npad = 2^round(log2(2n-1)), multiplied in Fourier space by
//   Filt = dp*Re(FFT(ifftshift(h))) (Kak-Slaney ramp h, optional Hann window)
//   and cropped.  With x real and Filt real this is exactly the circular
//   convolution y[m] = sum_{i<n} x[i] * hr[(m-i) mod npad], hr = Re(IFFT(Filt)).
//   hr is computed once per call on the device in FP64 (direct DFTs over a
//   cosine table), then every line is filtered by an LDS-tiled batched GEMM
//   against the implicit Toeplitz matrix T[m][i] = hr[(m-i) & (npad-1)], which
//   also reproduces the reference's wrap-around when npad < 2n-1.
// fdk_backproj_linear3D (FDK.py:130-208): voxel driven; per view
//   zoom = (F+D)/(F - x cos s - y sin s), u = (-x sin s + y cos s) zoom,
//   v = z zoom, bilinear read with offsets, += val*zoom^2, times pi/ns.
#include <math.h>

#include <mutex>

#include "ct2_internal.h"

namespace ct2 {

__global__ void k_fdk_weight(int ns, int nu, int nv, double F, double FD, double u00, double du,
                             double uoff, double v00, double dv, double voff,
                             const float* __restrict__ sino, float* __restrict__ out) {
    const size_t n = (size_t)ns * nu * nv;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
         q += (size_t)gridDim.x * blockDim.x) {
        int k = (int)(q % nv);
        int j = (int)((q / nv) % nu);
        double vu = u00 + ((double)j + uoff) * du;
        double vv = v00 + ((double)k + voff) * dv;
        double w = F / sqrt(FD * FD + vv * vv + vu * vu);
        out[q] = (float)(w * (double)sino[q]);
    }
}

// Batched C[b] = A[b] (M x K) * B[b] (K x N), 64x64 output tile, 256
// threads each owning a 4x4 block, K in chunks of 16 staged in LDS.
// DIM == 1 (filter along u): A[m][i] = hr[(m-i)&msk], B = X_p (i rows, k cols)
// DIM == 2 (filter along v): A = X_p (j rows, i cols), B[i][m] = hr[(m-i)&msk]
template <int DIM>
__global__ __launch_bounds__(256) void k_fdk_filter(int nu, int nv, int msk,
                                                    const float* __restrict__ hr,
                                                    const float* __restrict__ x,
                                                    float* __restrict__ y) {
    __shared__ float As[16][64 + 4];   // [kk][row]
    __shared__ float Bs[16][64 + 4];   // [kk][col]
    const int M = nu, N = (DIM == 1) ? nv : nv, K = (DIM == 1) ? nu : nv;
    const size_t plane = (size_t)nu * nv;
    const float* X = x + blockIdx.z * plane;
    float* Y = y + blockIdx.z * plane;
    const int row0 = blockIdx.y * 64, col0 = blockIdx.x * 64;
    const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        // stage A: 64 rows x 16 k
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            int r = e / 16, kk = e % 16;
            int gr = row0 + r, gk = k0 + kk;
            float v = 0.0f;
            if (gr < M && gk < K) {
                if (DIM == 1) v = hr[(gr - gk) & msk];
                else v = X[(size_t)gr * nv + gk];
            }
            As[kk][r] = v;
        }
        // stage B: 16 k x 64 cols
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            int kk = e / 64, c = e % 64;
            int gk = k0 + kk, gc = col0 + c;
            float v = 0.0f;
            if (gk < K && gc < N) {
                if (DIM == 1) v = X[(size_t)gk * nv + gc];
                else v = hr[(gc - gk) & msk];
            }
            Bs[kk][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) { a[q] = As[kk][tr + q]; b[q] = Bs[kk][tc + q]; }
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[r][c] = __builtin_fmaf(a[r], b[c], acc[r][c]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        int gr = row0 + tr + r;
        if (gr >= M) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            int gc = col0 + tc + c;
            if (gc < N) Y[(size_t)gr * nv + gc] = acc[r][c];
        }
    }
}

// Voxel-driven FDK backprojection; a wavefront is 64 consecutive z of one
// (x, y) column, so its detector reads run along v.
__global__ __launch_bounds__(256) void k_fdk_bp(int ns, int nu, int nv, float du, float dv,
                                                float u0, float v0, float F, float FD, float uoff,
                                                float voff, double s0, double ds, int nx, int ny,
                                                int nz, float x00, float y00, float z00, float dx,
                                                float dy, float dz, float xoff, float yoff,
                                                float zoff, const float* __restrict__ sino,
                                                float* __restrict__ recon) {
    extern __shared__ float cs[];   // cos, sin per view
    for (int p = threadIdx.x; p < ns; p += blockDim.x) {
        double s = s0 + (double)p * ds;
        cs[2 * p] = (float)cos(s);
        cs[2 * p + 1] = (float)sin(s);
    }
    __syncthreads();
    const int k = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i = blockIdx.z;
    if (k >= nz || j >= ny) return;
    const float x = x00 + ((float)i + xoff) * dx;
    const float y = y00 + ((float)j + yoff) * dy;
    const float z = z00 + ((float)k + zoff) * dz;
    float acc = 0.0f;
    for (int p = 0; p < ns; ++p) {
        const float c = cs[2 * p], s = cs[2 * p + 1];
        const float zoom = FD / (F - x * c - y * s);
        const float u = (-x * s + y * c) * zoom;
        const float v = z * zoom;
        const float fu = (u - u0) / du - uoff, fv = (v - v0) / dv - voff;
        const float lu_f = floorf(fu), lv_f = floorf(fv);
        const int lu = (int)lu_f, lv = (int)lv_f;
        if (lu < 0 || lu + 1 >= nu || lv < 0 || lv + 1 >= nv) continue;
        const float wu2 = fu - lu_f, wv2 = fv - lv_f;
        const float wu1 = 1.0f - wu2, wv1 = 1.0f - wv2;
        const float* sp = sino + ((size_t)p * nu + lu) * nv + lv;
        const float t1 = wu1 * sp[0] + wu2 * sp[nv];
        const float t2 = wu1 * sp[1] + wu2 * sp[nv + 1];
        acc += (wv1 * t1 + wv2 * t2) * zoom * zoom;
    }
    recon[((size_t)i * ny + j) * nz + k] = acc * (float)(M_PI / ns);
}

// ---- filter taps on the device ----
// hr = Re(IFFT(Filt)), Filt = dp*Re(FFT(ifftshift(h))) [* Hann window], as
// two FP64 direct DFTs over a cosine table.  npad is a power of two, so the
// angle index k*d mod npad is (k*d) & (npad-1) in 32-bit unsigned arithmetic
// (npad <= 2^16).  Workspace: cos table [npad] and spectrum [npad] doubles,
// then the float taps [npad].
size_t fdk_filter_ws_bytes(int npad) { return (size_t)npad * (2 * sizeof(double) + sizeof(float)); }

__global__ void k_fdk_cos_table(int N, double* __restrict__ ct) {
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < N; m += gridDim.x * blockDim.x)
        ct[m] = cospi(2.0 * (double)m / (double)N);
}

// Filt[k]: the Kak-Slaney kernel h(d') (h(0) = 1/(4 dp^2), h(odd) =
// -1/(pi d' dp)^2, 0 at even d' != 0) at d' = d for d < N/2 and d - N
// otherwise (ifftshift), transformed; then the window on the FFT-ordered
// frequency index: kappa of index i = (k + N/2) mod N is (i - N/2)/(dp N),
// window 0.5 + 0.5 cos(pi kappa / cutoff) for wlo <= i <= whi, 0 elsewhere
// (FDK.py:33-73).
__global__ void k_fdk_ramp_spectrum(int N, double dp, int hann, int wlo, int whi, double cutoff,
                                    const double* __restrict__ ct, double* __restrict__ filt) {
    const unsigned msk = (unsigned)N - 1u;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < N; k += gridDim.x * blockDim.x) {
        double acc = 1.0 / (4.0 * dp * dp);   // d = 0: cos(0) = 1
        for (int d = 1; d < N; d += 2) {
            const double dd = (double)(d < N / 2 ? d : d - N);
            const double t = M_PI * dd * dp;
            acc += (-1.0 / (t * t)) * ct[((unsigned)k * (unsigned)d) & msk];
        }
        double f = dp * acc;
        if (hann) {
            const int i = (k + N / 2) & (int)msk;
            const double kap = (double)(i - N / 2) / (dp * (double)N);
            f *= (i >= wlo && i <= whi) ? 0.5 + 0.5 * cos(M_PI * kap / cutoff) : 0.0;
        }
        filt[k] = f;
    }
}

__global__ void k_fdk_ramp_taps(int N, const double* __restrict__ ct,
                                const double* __restrict__ filt, float* __restrict__ hr) {
    const unsigned msk = (unsigned)N - 1u;
    for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < N; m += gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int k = 0; k < N; ++k) acc += filt[k] * ct[((unsigned)k * (unsigned)m) & msk];
        hr[m] = (float)(acc / (double)N);
    }
}

// npad = 2^round(log2(2n-1)) (FDK.py:36); 2n-1 is odd, so never half-way.
int fdk_npad(int n) {
    int e = (int)lround(log2(2.0 * n - 1.0));
    return 1 << (e < 1 ? 1 : e);
}

// Window bounds on the centred frequency grid: every index for L >= 1; for
// L < 1 from the first to the last index whose |kappa| is nearest the
// cutoff L/(2 dp) (FDK.py:55-73).  Index arithmetic only.
static void hann_bounds(int N, double dp, double L, int& wlo, int& whi) {
    wlo = 0;
    whi = N - 1;
    if (L >= 1.0) return;
    const double cutoff = 1.0 / (2.0 * dp) * L;
    double mind = 1e300;
    for (int i = 0; i < N; ++i) mind = fmin(mind, fabs(fabs((double)(i - N / 2) / (dp * N)) - cutoff));
    wlo = -1;
    for (int i = 0; i < N; ++i)
        if (fabs(fabs((double)(i - N / 2) / (dp * N)) - cutoff) == mind) {
            if (wlo < 0) wlo = i;
            whi = i;
        }
}

// Filters every line of sinomat into out with the taps built in workspace
// ws (fdk_filter_ws_bytes(npad) bytes of device memory).
int fdk_filter_run(float dp, int dim, int ns, int nu, int nv, int npad, const float* sinomat,
                   float* out, int hann, float L, void* ws, hipStream_t st) {
    double* ct = (double*)ws;
    double* filt = ct + npad;
    float* hr = (float*)(filt + npad);
    const unsigned nb = (unsigned)((npad + 255) / 256);
    int wlo = 0, whi = npad - 1;
    if (hann) hann_bounds(npad, (double)dp, (double)L, wlo, whi);
    hipLaunchKernelGGL(k_fdk_cos_table, dim3(nb), dim3(256), 0, st, npad, ct);
    hipLaunchKernelGGL(k_fdk_ramp_spectrum, dim3(nb), dim3(256), 0, st, npad, (double)dp, hann,
                       wlo, whi, 1.0 / (2.0 * (double)dp) * (double)L, (const double*)ct, filt);
    hipLaunchKernelGGL(k_fdk_ramp_taps, dim3(nb), dim3(256), 0, st, npad, (const double*)ct,
                       (const double*)filt, hr);
    dim3 grid((nv + 63) / 64, (nu + 63) / 64, ns);
    if (dim == 1)
        hipLaunchKernelGGL(k_fdk_filter<1>, grid, dim3(256), 0, st, nu, nv, npad - 1,
                           (const float*)hr, sinomat, out);
    else
        hipLaunchKernelGGL(k_fdk_filter<2>, grid, dim3(256), 0, st, nu, nv, npad - 1,
                           (const float*)hr, sinomat, out);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// Argument checks shared by the filter entry points; sets *npad.
int fdk_filter_check(const char* who, float dp, int dim, int ns, int nu, int nv,
                     const float* sinomat, const float* out, int* npad) {
    if (ns <= 0 || nu <= 0 || nv <= 0 || !sinomat || !out || (dim != 1 && dim != 2) ||
        dp == 0.0f) {
        set_error("%s: bad arguments", who);
        return CT2_EINVAL;
    }
    if (sinomat == out) { set_error("%s: in-place filtering is not supported", who); return CT2_EINVAL; }
    const int n = dim == 1 ? nu : nv;
    if (n > 32768) { set_error("%s: line of %d bins is too long", who, n); return CT2_EINVAL; }
    *npad = fdk_npad(n);
    return CT2_OK;
}

// ---- scratch of the allocation-free filter ----
// ct2_fdk_sino_filter3d_noalloc builds its taps in this file-scope device
// array instead of an allocation: every GPU the library runs on holds its own
// copy until the process ends.  Host code never passes the array's name to a
// kernel; the address of each device's copy comes from hipGetSymbolAddress,
// taken once per device with that device current.  Every call on a device
// overwrites the whole workspace it uses before reading it, so nothing needs
// clearing; calls that overlap on one device race for it.
constexpr int FDK_SCRATCH_DOUBLES =
    (int)((size_t)CT2_FDK_NOALLOC_MAX_NPAD * (2 * sizeof(double) + sizeof(float)) / sizeof(double));
static __device__ __align__(256) double fdk_scratch[FDK_SCRATCH_DOUBLES];
static_assert(sizeof(double) * FDK_SCRATCH_DOUBLES >= CT2_FDK_NOALLOC_MAX_NPAD *
                                                          (2 * sizeof(double) + sizeof(float)),
              "scratch holds the workspace of the largest pad");

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices];

// Address of the current device's scratch (looked up on its first use).
static int fdk_scratch_ptr(void** out) {
    int dev = 0;
    CT2_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("ct2_fdk_sino_filter3d_noalloc: device %d beyond the %d supported", dev,
                  kMaxDevices);
        return CT2_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        CT2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(fdk_scratch)));
        if (!p) { set_error("ct2_fdk_sino_filter3d_noalloc: no scratch address"); return CT2_EHIP; }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return CT2_OK;
}

}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_fdk_sino_weight3d(float radius, float center2detector, int ns, int nu, int nv,
                          float u00, float du, float uoff, float v00, float dv, float voff,
                          const float* sinomat, float* out, ct2_stream_t stream) {
    if (ns <= 0 || nu <= 0 || nv <= 0 || !sinomat || !out) {
        set_error("ct2_fdk_sino_weight3d: bad arguments");
        return CT2_EINVAL;
    }
    size_t n = (size_t)ns * nu * nv;
    unsigned nb = (unsigned)((n + 255) / 256);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_fdk_weight, dim3(nb), dim3(256), 0, (hipStream_t)stream, ns, nu, nv,
                       (double)radius, (double)radius + (double)center2detector, (double)u00,
                       (double)du, (double)uoff, (double)v00, (double)dv, (double)voff, sinomat,
                       out);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_fdk_sino_filter3d(float dp, int dim, int ns, int nu, int nv, const float* sinomat,
                          float* out, int hann, float L, ct2_stream_t stream) {
    int npad = 0;
    int rc = fdk_filter_check("ct2_fdk_sino_filter3d", dp, dim, ns, nu, nv, sinomat, out, &npad);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    void* ws = nullptr;
    CT2_HIP(hipMallocAsync(&ws, fdk_filter_ws_bytes(npad), st));
    rc = fdk_filter_run(dp, dim, ns, nu, nv, npad, sinomat, out, hann, L, ws, st);
    hipError_t f = hipFreeAsync(ws, st);
    if (rc) return rc;
    CT2_HIP(f);
    return CT2_OK;
}

int ct2_fdk_sino_filter3d_noalloc(float dp, int dim, int ns, int nu, int nv,
                                  const float* sinomat, float* out, int hann, float L,
                                  ct2_stream_t stream) {
    const char* who = "ct2_fdk_sino_filter3d_noalloc";
    int npad = 0;
    int rc = fdk_filter_check(who, dp, dim, ns, nu, nv, sinomat, out, &npad);
    if (rc) return rc;
    if (npad > CT2_FDK_NOALLOC_MAX_NPAD) {
        set_error("%s: lines of %d bins pad to %d samples, beyond the %d of the static "
                  "scratch (use ct2_fdk_sino_filter3d)", who, dim == 1 ? nu : nv, npad,
                  CT2_FDK_NOALLOC_MAX_NPAD);
        return CT2_EINVAL;
    }
    void* ws = nullptr;
    if ((rc = fdk_scratch_ptr(&ws))) return rc;
    return fdk_filter_run(dp, dim, ns, nu, nv, npad, sinomat, out, hann, L, ws,
                          (hipStream_t)stream);
}

int ct2_fdk_backproj_linear3d(int ns, int nu, int nv, float du, float dv, float u0, float v0,
                              float radius, float source2detector, float uoffset, float voffset,
                              double s0, double ds, int nx, int ny, int nz, float x00, float y00,
                              float z00, float dx, float dy, float dz, float xoff, float yoff,
                              float zoff, const float* sinomat, float* recon,
                              ct2_stream_t stream) {
    if (ns <= 0 || nu <= 0 || nv <= 0 || nx <= 0 || ny <= 0 || nz <= 0 || !sinomat || !recon) {
        set_error("ct2_fdk_backproj_linear3d: bad arguments");
        return CT2_EINVAL;
    }
    if (ns > 8192) { set_error("ct2_fdk_backproj_linear3d: at most 8192 views"); return CT2_EINVAL; }
    // F = radius, D = source2detector - radius, F + D = source2detector (FDK.py:151-152)
    float F = radius;
    float FD = (float)((double)radius + ((double)source2detector - (double)radius));
    dim3 grid((nz + 63) / 64, (ny + 3) / 4, nx);
    hipLaunchKernelGGL(k_fdk_bp, grid, dim3(256), sizeof(float) * 2 * ns, (hipStream_t)stream, ns,
                       nu, nv, du, dv, u0, v0, F, FD, uoffset, voffset, s0, ds, nx,
                       ny, nz, x00, y00, z00, dx, dy, dz, xoff, yoff, zoff, sinomat, recon);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // extern "C"
