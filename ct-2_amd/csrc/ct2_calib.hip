// ct2_calib.hip -- issue-rate calibration for the roofline figures.
//
// A grid of wavefronts that each run `iters` iterations of CALIB_CHAINS
// independent dependency chains of one VALU instruction kind: FP32 FMA,
// FP64 FMA, or a 32-bit integer rotate.  With enough wavefronts per SIMD every
// chain's latency is hidden, so the launch's duration measures the rate at
// which the CUs issue that instruction kind.  One iteration executes exactly
// CALIB_CHAINS VALU instructions of the kind (the loop control runs on the
// scalar unit).  The result of every chain is written out so that none is
// removed.
#include "ct2_internal.h"

namespace ct2 {

constexpr int CALIB_CHAINS = 8;
constexpr int CALIB_THREADS = 256;

template <int KIND>
__global__ __launch_bounds__(CALIB_THREADS) void k_calib_issue(int iters, float* __restrict__ out) {
    const int t = blockIdx.x * CALIB_THREADS + threadIdx.x;
    float acc = 0.0f;
    if constexpr (KIND == 0) {
        // accumulate form a += x * 0.5 (v_fmac_f32, 4-byte encoding)
        float a[CALIB_CHAINS], x[CALIB_CHAINS];
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) {
            a[j] = (float)(t + j) * 1e-9f;
            x[j] = (float)(t ^ j) * 1e-12f;
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < CALIB_CHAINS; ++j) a[j] = __builtin_fmaf(x[j], 0.5f, a[j]);
        }
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) acc += a[j];
    } else if constexpr (KIND == 1) {
        double a[CALIB_CHAINS], x[CALIB_CHAINS];
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) {
            a[j] = (double)(t + j) * 1e-9;
            x[j] = (double)(t ^ j) * 1e-12;
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < CALIB_CHAINS; ++j) a[j] = fma(x[j], 0.5, a[j]);
        }
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) acc += (float)a[j];
    } else {
        unsigned a[CALIB_CHAINS];
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) a[j] = (unsigned)(t * 8 + j);
        // a rotate of the lane's own value (v_alignbit_b32): no scalar
        // operand and no closed form for the compiler to sum
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < CALIB_CHAINS; ++j) a[j] = __builtin_amdgcn_alignbit(a[j], a[j], 31);
        }
#pragma unroll
        for (int j = 0; j < CALIB_CHAINS; ++j) acc += (float)a[j];
    }
    out[t] = acc;
}

}  // namespace ct2

using namespace ct2;

extern "C" int ct2_calib_issue(int kind, int blocks, int iters, float* out,
                               ct2_stream_t stream) {
    if (kind < 0 || kind > 2 || blocks < 1 || iters < 1 || !out) {
        set_error("ct2_calib_issue: kind %d, blocks %d, iters %d, out %p", kind, blocks, iters,
                  (void*)out);
        return CT2_EINVAL;
    }
    hipStream_t st = (hipStream_t)stream;
    switch (kind) {
        case 0: hipLaunchKernelGGL(k_calib_issue<0>, dim3(blocks), dim3(CALIB_THREADS), 0, st, iters, out); break;
        case 1: hipLaunchKernelGGL(k_calib_issue<1>, dim3(blocks), dim3(CALIB_THREADS), 0, st, iters, out); break;
        default: hipLaunchKernelGGL(k_calib_issue<2>, dim3(blocks), dim3(CALIB_THREADS), 0, st, iters, out); break;
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}
