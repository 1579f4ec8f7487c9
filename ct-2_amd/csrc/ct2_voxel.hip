// ct2_voxel.hip -- voxel-driven kernels: pd_backproj (view_proj3D.f:1598-1944)
// and the device fill used instead of memset nodes.
//
// pd_backproj projects every voxel centre onto each view's detector and
// reads the sinogram there (bilinear or nearest), summing over views with no
// distance weight; it is not the adjoint of the projector.  Reference FP32
// operation order per view (view_proj3D.f:1890-1933), sequential over views,
// so results are bit-identical to the Fortran.
#include <type_traits>

#include "ct2_internal.h"

namespace ct2 {

__global__ void k_fill(float* __restrict__ p, float v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int launch_fill(float* p, float v, size_t n, hipStream_t st) {
    if (n == 0) return CT2_OK;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, v, n);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// ---- float32 -> 16-bit volume storage ----
// Round to nearest, or stochastically: the 16-bit neighbours lo <= |x| <= hi
// are chosen with probabilities (hi-|x|)/(hi-lo), (|x|-lo)/(hi-lo), so the
// stored value is unbiased and rounding errors are independent across voxels
// (a constant region no longer carries one common bias into every ray sum).
// The uniform variate is a counter-based hash of (index, seed): deterministic.
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    // integer finaliser (xorshift-multiply rounds)
    x ^= x >> 16; x *= 0x7feb352dU;
    x ^= x >> 15; x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// The 16-bit pattern that ct2_convert_volume stores for element i of value
// x: round to nearest, or stochastically between the two neighbours with a
// counter-based hash of (i, seed).
template <typename OT>
__device__ __forceinline__ uint16_t round16(float x, size_t i, int stochastic, uint32_t seed) {
    if (!stochastic) {
        OT v;
        Store<OT>::st(&v, 0, x);
        return *reinterpret_cast<const uint16_t*>(&v);
    }
    uint16_t lo_bits;
    float lo;
    if constexpr (std::is_same<OT, __half>::value) {
        __half h = __float2half_rz(x);
        lo_bits = __half_as_ushort(h);
        lo = __half2float(h);
    } else {
        uint32_t b = __float_as_uint(x);
        lo_bits = (uint16_t)(b >> 16);           // truncation = toward zero
        lo = __uint_as_float((uint32_t)lo_bits << 16);
    }
    uint16_t out = lo_bits;
    if (lo != x && isfinite(x)) {
        const uint16_t hi_bits = (uint16_t)(lo_bits + 1);   // next magnitude up
        float hi;
        if constexpr (std::is_same<OT, __half>::value) hi = __half2float(__ushort_as_half(hi_bits));
        else hi = __uint_as_float((uint32_t)hi_bits << 16);
        const float p = (fabsf(x) - fabsf(lo)) / (fabsf(hi) - fabsf(lo));
        const float u = (float)(hash32((uint32_t)i ^ hash32(seed + (uint32_t)(i >> 32))) >> 8) *
                        (1.0f / 16777216.0f);
        if (u < p && isfinite(hi)) out = hi_bits;
    }
    return out;
}

template <typename OT>
__global__ void k_convert(const float* __restrict__ src, OT* __restrict__ dst, size_t n,
                          int stochastic, uint32_t seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        reinterpret_cast<uint16_t*>(dst)[i] = round16<OT>(src[i], i, stochastic, seed);
}

// f16 pairs: element i holds the f16 patterns of values i and i + 1 (zero
// past the end), each rounded exactly as k_convert<__half> rounds it.
__global__ void k_convert_pairs(const float* __restrict__ src, uint32_t* __restrict__ dst,
                                size_t n, int stochastic, uint32_t seed) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t lo = round16<__half>(src[i], i, stochastic, seed);
        const uint32_t hi = i + 1 < n ? round16<__half>(src[i + 1], i + 1, stochastic, seed) : 0u;
        dst[i] = lo | (hi << 16);
    }
}

// Plane-major copy of a C-ordered (nx, ny, nz) float32 volume: element
// (i, j, k) to (k*nx + i)*ny + j, as float32 or rounded to 16 bits exactly as
// k_convert rounds element (i*ny + j)*nz + k.  One 32 x 32 (k, j) tile of
// x-slice i per workgroup, transposed through LDS so that the reads (along k)
// and the writes (along j) are both contiguous.
template <typename OT>
__global__ __launch_bounds__(256) void k_convert_zplanes(const float* __restrict__ src,
                                                         OT* __restrict__ dst, int nx, int ny,
                                                         int nz, int stochastic, uint32_t seed) {
    __shared__ float t[32][33];
    const int i = blockIdx.z, k0 = blockIdx.x * 32, j0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int dy = ty; dy < 32; dy += 8) {
        const int j = j0 + dy, k = k0 + tx;
        if (j < ny && k < nz) t[dy][tx] = src[((size_t)i * ny + j) * nz + k];
    }
    __syncthreads();
    for (int dy = ty; dy < 32; dy += 8) {
        const int k = k0 + dy, j = j0 + tx;
        if (j < ny && k < nz) {
            const float x = t[tx][dy];
            const size_t q = ((size_t)k * nx + i) * ny + j;
            if constexpr (std::is_same<OT, float>::value)
                dst[q] = x;
            else
                reinterpret_cast<uint16_t*>(dst)[q] =
                    round16<OT>(x, ((size_t)i * ny + j) * nz + k, stochastic, seed);
        }
    }
}

int launch_convert_zplanes(const float* src, void* dst, int dtype, int nx, int ny, int nz,
                           int rounding, unsigned seed, hipStream_t st) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > 65535 ||
        (long long)nx * ny * nz >= (1LL << 31)) {
        set_error("ct2_convert_volume_zplanes: bad shape %d x %d x %d", nx, ny, nz);
        return CT2_EINVAL;
    }
    if (!src || !dst) { set_error("ct2_convert_volume_zplanes: null pointer"); return CT2_EINVAL; }
    if (rounding != CT2_ROUND_NEAREST && rounding != CT2_ROUND_STOCHASTIC) {
        set_error("ct2_convert_volume_zplanes: unknown rounding %d", rounding);
        return CT2_EINVAL;
    }
    const int s = rounding == CT2_ROUND_STOCHASTIC;
    const dim3 grid((unsigned)((nz + 31) / 32), (unsigned)((ny + 31) / 32), (unsigned)nx);
    switch (dtype) {
        case CT2_F32_ZPLANES:
            hipLaunchKernelGGL(k_convert_zplanes<float>, grid, dim3(256), 0, st, src, (float*)dst,
                               nx, ny, nz, s, (uint32_t)seed);
            break;
        case CT2_F16_ZPLANES:
            hipLaunchKernelGGL(k_convert_zplanes<__half>, grid, dim3(256), 0, st, src,
                               (__half*)dst, nx, ny, nz, s, (uint32_t)seed);
            break;
        case CT2_BF16_ZPLANES:
            hipLaunchKernelGGL(k_convert_zplanes<hip_bfloat16>, grid, dim3(256), 0, st, src,
                               (hip_bfloat16*)dst, nx, ny, nz, s, (uint32_t)seed);
            break;
        default:
            set_error("ct2_convert_volume_zplanes: destination must be CT2_F32_ZPLANES, "
                      "CT2_F16_ZPLANES or CT2_BF16_ZPLANES");
            return CT2_EINVAL;
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// Per-view constants: S, D, e_u, e_v, e_w and the source-to-panel distance
// sd_dist (view_proj3D.f:1699-1714), computed on the device in FP32.
struct PdView {
    float S[3], D[3], eu[3], ev[3], ew[3], sdd;
};

__device__ __forceinline__ void pd_view(const float* __restrict__ f, PdView& pv) {
#pragma clang fp contract(off)
    for (int q = 0; q < 3; ++q) {
        pv.S[q] = f[q]; pv.D[q] = f[6 + q]; pv.eu[q] = f[9 + q]; pv.ev[q] = f[12 + q];
        pv.ew[q] = f[15 + q];
    }
    float udr = (pv.D[0] - pv.S[0]) * pv.eu[0] + (pv.D[1] - pv.S[1]) * pv.eu[1] +
                (pv.D[2] - pv.S[2]) * pv.eu[2];
    float vdr = (pv.D[0] - pv.S[0]) * pv.ev[0] + (pv.D[1] - pv.S[1]) * pv.ev[1] +
                (pv.D[2] - pv.S[2]) * pv.ev[2];
    float tc[3];
    for (int q = 0; q < 3; ++q) tc[q] = pv.D[q] - pv.eu[q] * udr - pv.ev[q] * vdr;
    pv.sdd = sqrtf((pv.S[0] - tc[0]) * (pv.S[0] - tc[0]) + (pv.S[1] - tc[1]) * (pv.S[1] - tc[1]) +
                   (pv.S[2] - tc[2]) * (pv.S[2] - tc[2]));
}

template <int INTERP>
__global__ __launch_bounds__(256) void k_pd_backproject(Grid g, Det dt,
                                                        float* __restrict__ vol,
                                                        const float* __restrict__ sino,
                                                        const float* __restrict__ fv) {
    // one voxel per lane, z fastest: a wavefront is 64 consecutive z
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nvox = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (idx >= nvox) return;
    const int iz = (int)(idx % g.n[2]);
    const int iy = (int)((idx / g.n[2]) % g.n[1]);
    const int ix = (int)(idx / ((size_t)g.n[2] * g.n[1]));
    const float x = g.o[0] + ((float)ix + 0.5f) * g.d[0];
    const float y = g.o[1] + ((float)iy + 0.5f) * g.d[1];
    const float z = g.o[2] + ((float)iz + 0.5f) * g.d[2];
    float total = 0.0f;
    for (int p = 0; p < dt.ns; ++p) {
        PdView pv;
        pd_view(fv + (size_t)p * 18, pv);
        float xd = x - pv.S[0], yd = y - pv.S[1], zd = z - pv.S[2];
        float zoom = fabsf((xd * pv.ew[0] + yd * pv.ew[1] + zd * pv.ew[2]) / pv.sdd);
        float xb = xd / zoom + pv.S[0], yb = yd / zoom + pv.S[1], zb = zd / zoom + pv.S[2];
        float xdt = xb - pv.D[0], ydt = yb - pv.D[1], zdt = zb - pv.D[2];
        float u = (xdt * pv.eu[0]) + (ydt * pv.eu[1]) + (zdt * pv.eu[2]);
        float v = (xdt * pv.ev[0]) + (ydt * pv.ev[1]) + (zdt * pv.ev[2]);
        const float* sp = sino + (size_t)p * dt.nu * dt.nv;
        if (INTERP == CT2_LINEAR) {
            // iu is the 1-based left bin, valid 1 <= iu < nu
            int iu = (int)floorf((u - dt.u0) / dt.du + 0.5f);
            int iv = (int)floorf((v - dt.v0) / dt.dv + 0.5f);
            if ((iu >= 1) && (iu < dt.nu) && (iv >= 1) && (iv < dt.nv)) {
                float wu2 = (u - dt.u0) / dt.du - ((float)iu - 0.5f);
                float wu1 = 1.0f - wu2;
                float wv2 = (v - dt.v0) / dt.dv - ((float)iv - 0.5f);
                float wv1 = 1.0f - wv2;
                size_t a00 = (size_t)(iu - 1) * dt.nv + (iv - 1);
                float vi1 = wu1 * sp[a00] + wu2 * sp[a00 + dt.nv];
                float vi2 = wu1 * sp[a00 + 1] + wu2 * sp[a00 + dt.nv + 1];
                float val = wv1 * vi1 + wv2 * vi2;
                total = total + val;
            }
        } else {
            int iu = (int)floorf((u - dt.u0) / dt.du);
            int iv = (int)floorf((v - dt.v0) / dt.dv);
            if ((iu >= 0) && (iu <= dt.nu - 1) && (iv >= 0) && (iv <= dt.nv - 1))
                total = total + sp[(size_t)iu * dt.nv + iv];
        }
    }
    vol[(size_t)ix * g.s[0] + (size_t)iy * g.s[1] + (size_t)iz * g.s[2]] = total;
}

int launch_convert(const float* src, void* dst, int dtype, size_t n, int rounding, unsigned seed,
                   hipStream_t st) {
    if (n == 0) return CT2_OK;
    if (!src || !dst) { set_error("ct2_convert_volume: null pointer"); return CT2_EINVAL; }
    if (rounding != CT2_ROUND_NEAREST && rounding != CT2_ROUND_STOCHASTIC) {
        set_error("ct2_convert_volume: unknown rounding %d", rounding);
        return CT2_EINVAL;
    }
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    int s = rounding == CT2_ROUND_STOCHASTIC;
    switch (dtype) {
        case CT2_F16:
            hipLaunchKernelGGL(k_convert<__half>, dim3((unsigned)blocks), dim3(256), 0, st, src,
                               (__half*)dst, n, s, (uint32_t)seed);
            break;
        case CT2_BF16:
            hipLaunchKernelGGL(k_convert<hip_bfloat16>, dim3((unsigned)blocks), dim3(256), 0, st,
                               src, (hip_bfloat16*)dst, n, s, (uint32_t)seed);
            break;
        case CT2_F16Z2:
            hipLaunchKernelGGL(k_convert_pairs, dim3((unsigned)blocks), dim3(256), 0, st, src,
                               (uint32_t*)dst, n, s, (uint32_t)seed);
            break;
        default:
            set_error("ct2_convert_volume: destination must be CT2_F16, CT2_BF16 or CT2_F16Z2");
            return CT2_EINVAL;
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int launch_pd_backproject(const ct2_desc* d, float* vol, const float* sino, const float* fv,
                          hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    if (!vol || !sino || !fv) { set_error("pd_backproj: null pointer"); return CT2_EINVAL; }
    size_t nvox = (size_t)g.n[0] * g.n[1] * g.n[2];
    dim3 grid((unsigned)((nvox + 255) / 256));
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_pd_backproject<CT2_LINEAR>, grid, dim3(256), 0, st, g, dt, vol, sino, fv);
    else
        hipLaunchKernelGGL(k_pd_backproject<CT2_NEAREST>, grid, dim3(256), 0, st, g, dt, vol, sino, fv);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace ct2
