// ct2_raybunch.hip -- ray bunching and ray-length scaling of sinograms
// (emil_code/tomo3DfortranAux.f:1-176, used by sinogram3D.homoscale and the
// raybunch class, tomo3D.py:395-399, :484-534).
//
// All three are HBM-bound gathers with one lane per output element:
//   k_sinoscale  in place, 8 B per bin (read + write), one sqrt;
//   k_sinoave    reads its nws*nwu*nwv window (summed in the Fortran's
//                ii, jj, kk order, so bit-identical) and writes 4 B;
//   k_sinodist   one read and one write per fine bin: the Fortran scatters
//                coarse values over their windows with the last write in loop
//                order winning; the covering coarse index that wins is, per
//                axis, the largest one whose window holds the bin, so the
//                scatter becomes a race-free gather.
#include "ct2_internal.h"

namespace ct2 {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void k_sinoscale(float* __restrict__ s,
                                                        const float* __restrict__ fv, int nu,
                                                        int nv, long long n, float du, float dv,
                                                        float u0, float v0) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const long long per = (long long)nu * nv;
    const int p = (int)(o / per);
    const int q = (int)(o - (long long)p * per);
    const int j = q / nv, k = q - j * nv;
    const float* f = fv + (size_t)p * 18;
    const float u = u0 + ((float)j + 0.5f) * du;
    const float v = v0 + ((float)k + 0.5f) * dv;
    float d2 = 0.f;
    for (int c = 0; c < 3; ++c) {
        float b = f[6 + c] + f[9 + c] * u + f[12 + c] * v;
        float e = f[c] - b;
        d2 = c == 0 ? e * e : d2 + e * e;
    }
    s[o] = s[o] / sqrtf(d2);
}

struct Bunch {
    int ns, nu, nv, nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv;
};

__global__ __launch_bounds__(kThreads) void k_sinoave(const float* __restrict__ s,
                                                      float* __restrict__ out, Bunch b,
                                                      long long n) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const long long per = (long long)b.nus * b.nvs;
    const int i = (int)(o / per);
    const int r = (int)(o - (long long)i * per);
    const int j = r / b.nvs, k = r - j * b.nvs;
    float acc = 0.f;
    for (int ii = 0; ii < b.nws; ++ii)
        for (int jj = 0; jj < b.nwu; ++jj) {
            const float* row = s + ((long long)(i * b.nds + ii) * b.nu + (j * b.ndu + jj)) * b.nv
                             + k * b.ndv;
            for (int kk = 0; kk < b.nwv; ++kk) acc = acc + row[kk];
        }
    out[o] = acc / (float)(b.nws * b.nwu * b.nwv);
}

// largest coarse index c < nc with c*nd <= a < c*nd + nw, or -1
__device__ __forceinline__ int last_cover(int a, int nd, int nw, int nc) {
    int c = a / nd;
    if (c > nc - 1) c = nc - 1;
    return (a < c * nd + nw) ? c : -1;
}

__global__ __launch_bounds__(kThreads) void k_sinodist(float* __restrict__ s,
                                                       const float* __restrict__ in, Bunch b,
                                                       long long n) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const long long per = (long long)b.nu * b.nv;
    const int a0 = (int)(o / per);
    const int r = (int)(o - (long long)a0 * per);
    const int a1 = r / b.nv, a2 = r - a1 * b.nv;
    const int i = last_cover(a0, b.nds, b.nws, b.nss);
    const int j = last_cover(a1, b.ndu, b.nwu, b.nus);
    const int k = last_cover(a2, b.ndv, b.nwv, b.nvs);
    s[o] = (i < 0 || j < 0 || k < 0) ? 0.f : in[((long long)i * b.nus + j) * b.nvs + k];
}

unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int check_bunch(const char* who, const Bunch& b) {
    const int* v = &b.ns;
    for (int q = 0; q < 12; ++q)
        if (v[q] <= 0) { set_error("%s: every size, stride and window must be positive", who); return CT2_EINVAL; }
    if ((long long)(b.nss - 1) * b.nds + b.nws > b.ns || (long long)(b.nus - 1) * b.ndu + b.nwu > b.nu ||
        (long long)(b.nvs - 1) * b.ndv + b.nwv > b.nv) {
        set_error("%s: the last window ends outside the %dx%dx%d sinogram", who, b.ns, b.nu, b.nv);
        return CT2_EINVAL;
    }
    long long nfine = (long long)b.ns * b.nu * b.nv;
    if (blocks_for(nfine) > 0x7fffffffu) { set_error("%s: sinogram too large", who); return CT2_EINVAL; }
    return CT2_OK;
}

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_sinoscale(float* sinomat, const float* frame_vectors, int ns, int nu, int nv, float du,
                  float dv, float u0, float v0, ct2_stream_t stream) {
    if (ns <= 0 || nu <= 0 || nv <= 0) { set_error("ct2_sinoscale: bad sinogram shape"); return CT2_EINVAL; }
    if (!sinomat || !frame_vectors) { set_error("ct2_sinoscale: null pointer"); return CT2_EINVAL; }
    long long n = (long long)ns * nu * nv;
    if (blocks_for(n) > 0x7fffffffu) { set_error("ct2_sinoscale: sinogram too large"); return CT2_EINVAL; }
    hipLaunchKernelGGL(k_sinoscale, dim3(blocks_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                       sinomat, frame_vectors, nu, nv, n, du, dv, u0, v0);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_sinoave(const float* sinomat, int ns, int nu, int nv, float* sinoout, int nss, int nus,
                int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv,
                ct2_stream_t stream) {
    Bunch b{ns, nu, nv, nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv};
    int rc = check_bunch("ct2_sinoave", b);
    if (rc) return rc;
    if (!sinomat || !sinoout) { set_error("ct2_sinoave: null pointer"); return CT2_EINVAL; }
    long long n = (long long)nss * nus * nvs;
    hipLaunchKernelGGL(k_sinoave, dim3(blocks_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                       sinomat, sinoout, b, n);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_sinodist(float* sinomat, int ns, int nu, int nv, const float* sinoin, int nss, int nus,
                 int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv,
                 ct2_stream_t stream) {
    Bunch b{ns, nu, nv, nss, nus, nvs, nds, ndu, ndv, nws, nwu, nwv};
    int rc = check_bunch("ct2_sinodist", b);
    if (rc) return rc;
    if (!sinomat || !sinoin) { set_error("ct2_sinodist: null pointer"); return CT2_EINVAL; }
    long long n = (long long)ns * nu * nv;
    hipLaunchKernelGGL(k_sinodist, dim3(blocks_for(n)), dim3(kThreads), 0, (hipStream_t)stream,
                       sinomat, sinoin, b, n);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // extern "C"
