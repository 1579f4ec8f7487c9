// ct2_bp_zsep.hip -- the adjoint of the fast z-driven Joseph projector for
// views whose flat detector is parallel to the z planes, e_u along x and e_v
// along y: the tomosynthesis frames of emil_code/configs3D.py:615-654 under
// the z-forced projector of emil_code/view_proj3D_tomo.f:4-148.
//
// For such a view (view_zsep) the fast set-up of ray (j, k) gives its x cell
// coordinate on plane l as X_l(j) = fma(l, dub_j, ub0_j), a function of the
// detector column j alone, and its y coordinate Y_l(k) = fma(l, duc_k,
// uc0_k) of the row k alone (both the forward's own FP32 values); only the
// ray weight trav = |d_z| sqrt((1 + r1_j^2) + r2_k^2) couples them.  The
// forward samples plane l of ray (j, k) when 0 <= floor(X) <= nx-2 and 0 <=
// floor(Y) <= ny-2, with bilinear weights (1-wb, wb) x (1-wc, wc).  Its
// transpose on plane l is therefore separable: voxel (i, c, l) receives
//     sum_j hb_l(j, i) sum_k hc_l(k, c) trav(j, k) datum(j, k),
// hb_l(j, i) = 1-wb if floor(X_l(j)) = i, wb if it is i-1 (0 otherwise), and
// hc_l likewise in y.  The inner sum (a resampling of detector row data onto
// y cells) is taken per (column j, voxel column c); the outer one adds it to
// the two x rows column j reaches.
//
// Records.  k_zsep_td writes one record per view into a workspace: the
// columns' (ub0, dub) and the rows' (uc0, duc), the forward's FP32 values
// from its own FP64 set-up expressions; per plane the exact-arithmetic maps
// X_l(j) = ax + bx j and Y_l(k) = ay + by k rounded to FP32, which only
// bound searches; and td = trav datum of every ray (0 where the indicator
// leaves it out), trav formed as the forward's set-up forms it (one
// v_sqrt_f32 per ray, not one per (ray, plane)).
//
// Tiles.  A workgroup owns 8 x rows, 64 y cells and 8 z planes; wavefront w
// owns plane k0 + w, lane c one y cell, and sums its 8 voxels in FP32
// registers over every view, so each voxel is stored once, with no atomics
// (an LDS transpose at the end makes each thread's store 4 contiguous z
// values).  Per (view, plane) a wavefront
//   1. bounds, from the plane's maps, the detector columns whose sample can
//      fall in its cells (a margin of one column; fewer than 9/ZSEP_SPACING
//      + 5, so at most 17) and each lane's candidate rows: from the first row
//      within 0.02 row of the window [c-1, c+1) (a margin for the FP32
//      maps), the rows up to 2/ZSEP_SPACING + 0.04 < 3 rows further, so 3
//      candidates;
//   2. lane e evaluates column e's sample on the plane, X = fma(l, dub,
//      ub0), its cell and weight;
//   3. each lane evaluates its 3 candidate rows' Y = fma(l, duc, uc0), their
//      cells and their weights hc for its own y cell (the window moved
//      inside the detector at its edges, so the 3 rows are one load);
//   4. for each column, whose cell and weight it reads from lane e
//      (v_readlane, scalar), gathers G = sum hc td over its candidate rows
//      (coalesced: neighbouring lanes read neighbouring detector rows) into
//      the lane's own LDS slot of that column; then each x row adds (1-wb) G
//      of the columns whose cell it is and wb G of those whose cell is the
//      row below (ballot masks; the slots stand in for a register array
//      indexed by the column, which the compiler would put in scratch
//      memory).
// The matrix entries are the brick kernel's (the same FP32 cell
// coordinates, weights and trav); only the grouping of the sums differs, so
// the result is the fast forward's transpose up to FP32 rounding.
#include "ct2_internal.h"

namespace ct2 {

constexpr int ZS_ROWS = 8;                   // x rows per workgroup (all in each wavefront)
constexpr int ZS_TY = 64;                    // y cells per workgroup (one per lane)
constexpr int ZS_KZ = 8;                     // z planes per workgroup (one per wavefront)
constexpr int ZS_NC = 3;                     // candidate detector rows per y cell
constexpr int ZS_THREADS = 64 * ZS_KZ;
constexpr int ZS_MAXCOL = 20;                // detector columns per (wavefront, plane), >= 17

// A view's record in the workspace (k_zsep_td writes it), 16-byte aligned
// parts: per plane k the exact-arithmetic maps X_k(j) = ax + bx j, Y_k(q) =
// ay + by q rounded to FP32 (ax, bx, ay, by), which only bound the columns
// and rows searched; per detector column j the forward's (ub0, dub) of its x
// coordinate, per row q its (uc0, duc) of y (FP32, as setup_fast_ax rounds
// them); then td [nu][nv].
struct ZsRec {
    const float4* pm;
    const float2* cx;
    const float2* cy;
    const float* td;
};

__host__ __device__ inline size_t zsep_td_offset(int nu, int nv, int nz) {
    return ((4 * (size_t)nz + 2 * (size_t)nu + 2 * (size_t)nv) + 3) & ~(size_t)3;
}

__host__ __device__ inline size_t zsep_view_floats(int nu, int nv, int nz) {
    return (zsep_td_offset(nu, nv, nz) + (size_t)nu * nv + 3) & ~(size_t)3;
}

__device__ __forceinline__ ZsRec zs_rec(const float* base, const Det& dt, int nz) {
    ZsRec R;
    R.pm = reinterpret_cast<const float4*>(base);
    R.cx = reinterpret_cast<const float2*>(base + 4 * (size_t)nz);
    R.cy = R.cx + dt.nu;
    R.td = base + zsep_td_offset(dt.nu, dt.nv, nz);
    return R;
}

// Plane k of the tile for view record V: steps 1-4 above.
__device__ __forceinline__ void zs_plane(float (&acc)[ZS_ROWS], const Grid& g,
                                         const Det& dt, const ZsRec& V, int k, int iw, int c,
                                         float* sg) {
    const int lane = threadIdx.x & 63;
    const int nx = g.n[0], ny = g.n[1];
    const float kf = (float)k;
    // 1. the plane's maps (uniform): columns whose sample falls in cells
    // [iw-1, iw+ZS_ROWS), with a margin of one column on each side (the FP32
    // maps are within 1e-3 column of the exact ones)
    const float4 pm = V.pm[k];
    if (pm.y == 0.0f) return;   // uniform: a view the brick pass takes
    const float rbx = __builtin_amdgcn_rcpf(pm.y), rby = __builtin_amdgcn_rcpf(pm.w);
    const float e0 = ((float)(iw - 1) - pm.x) * rbx, e1 = ((float)(iw + ZS_ROWS) - pm.x) * rbx;
    const float elo = fminf(fmaxf(fminf(e0, e1) - 1.0f, -1.0f), (float)dt.nu);
    const float ehi = fminf(fmaxf(fmaxf(e0, e1) + 1.0f, -1.0f), (float)dt.nu);
    const int jlo = max(0, (int)floorf(elo)), jhi = min(dt.nu - 1, (int)ceilf(ehi));
    // fewer than 9/ZSEP_SPACING + 5 columns (17), so every one has a lane and
    // an LDS slot
    const int ncol = __builtin_amdgcn_readfirstlane(min(ZS_MAXCOL, jhi - jlo + 1));
    if (ncol <= 0) return;
    // 2. lane e: column jlo + e, its sample on this plane (the forward's)
    int ibl = -2;
    float wbl = 0.0f;
    if (lane < ncol) {
        const float2 cx = V.cx[jlo + lane];
        const float X = __builtin_fmaf(kf, cx.y, cx.x);
        const int ib = floor_i(X);
        wbl = frac_pos(X);
        // rows the wavefront owns, and the forward's validity window
        ibl = (ib >= 0 && ib <= nx - 2) ? ib - iw : -2;
    }
    // 3. this lane's candidate rows and their weights for y cell c: the rows
    // from just below the window [c-1, c+1) (a margin of 0.02 row for the
    // FP32 maps)
    int q0;
    float hc[ZS_NC];
    {
        const float f0 = ((float)(c - 1) - pm.z) * rby, f1 = ((float)(c + 1) - pm.z) * rby;
        const float flo = fminf(fmaxf(fminf(f0, f1) - 0.02f, -8.0f), (float)dt.nv + 8.0f);
        // the window moved inside the detector (nv >= ZS_NC) still holds
        // every row of [q0, q0+ZS_NC) that exists
        q0 = min(max((int)ceilf(flo), 0), dt.nv - ZS_NC);
#pragma unroll
        for (int n = 0; n < ZS_NC; ++n) {
            hc[n] = 0.0f;
            const float2 cy = V.cy[q0 + n];
            const float Y = __builtin_fmaf(kf, cy.y, cy.x);
            const int ic = floor_i(Y);
            const float wc = frac_pos(Y);
            if (c < ny && ic >= 0 && ic <= ny - 2)
                hc[n] = ic == c ? 1.0f - wc : (ic == c - 1 ? wc : 0.0f);
        }
    }
    // 4a. per column: G = sum hc td over the lane's candidate rows (loads
    // unconditional, in bounds; a row with hc = 0 adds nothing), into the
    // lane's own LDS slot of that column
    for (int e = 0; e < ncol; ++e) {
        const int r = __builtin_amdgcn_readlane(ibl, e);
        if (r < -1 || r >= ZS_ROWS) continue;   // uniform: no row of the tile takes it
        // the lane's ZS_NC consecutive rows in one load (a 4-byte aligned
        // copy, which the compiler issues as one global_load_dwordx3)
        float td[ZS_NC];
        __builtin_memcpy(td, V.td + (size_t)(jlo + e) * dt.nv + q0, sizeof td);
        float G = 0.0f;
#pragma unroll
        for (int n = 0; n < ZS_NC; ++n) {
            const float Gn = __builtin_fmaf(hc[n], td[n], G);
            G = hc[n] != 0.0f ? Gn : G;
        }
        sg[e * 64 + lane] = G;
    }
    // 4b. row q takes (1-wb) G of the columns whose cell is q and wb G of
    // those whose cell is q-1 (the lane reads back only its own slots)
#pragma unroll
    for (int q = 0; q < ZS_ROWS; ++q) {
        unsigned long long m0 = __ballot(lane < ncol && ibl == q);
        unsigned long long m1 = __ballot(lane < ncol && ibl == q - 1);
        while (m0) {
            const int e = __builtin_ctzll(m0);
            m0 &= m0 - 1;
            const float wb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wbl), e));
            acc[q] = __builtin_fmaf(1.0f - wb, sg[e * 64 + lane], acc[q]);
        }
        while (m1) {
            const int e = __builtin_ctzll(m1);
            m1 &= m1 - 1;
            const float wb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wbl), e));
            acc[q] = __builtin_fmaf(wb, sg[e * 64 + lane], acc[q]);
        }
    }
}

// work: k_zsep_td's records of the nviews views of the group (those the
// brick pass takes are marked); add: add the sums to vol instead of storing
// them.
__global__ __launch_bounds__(ZS_THREADS) void k_bp_zsep(Grid g, Det dt, int nviews, int add,
                                                        float* __restrict__ vol,
                                                        const float* __restrict__ work) {
    // per wavefront, one slot per (column, lane) for step 4; then the tile's
    // sums for the write-back
    constexpr int SG = ZS_KZ * ZS_MAXCOL * 64, SO = ZS_ROWS * ZS_TY * (ZS_KZ + 1);
    __shared__ float s_lds[SG > SO ? SG : SO];
    float (*s_out)[ZS_TY][ZS_KZ + 1] = reinterpret_cast<float (*)[ZS_TY][ZS_KZ + 1]>(s_lds);
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const int ntx = (nx + ZS_ROWS - 1) / ZS_ROWS, nty = (ny + ZS_TY - 1) / ZS_TY;
    int b = blockIdx.x;
    const int tx = b % ntx;
    b /= ntx;
    const int ty = b % nty;
    const int tz = b / nty;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int iw = tx * ZS_ROWS;            // the tile's first x row
    const int c0 = ty * ZS_TY, c = c0 + lane;   // the lane's y cell
    const int k0 = tz * ZS_KZ, k = k0 + wave;   // the wavefront's plane
    float acc[ZS_ROWS];
#pragma unroll
    for (int r = 0; r < ZS_ROWS; ++r) acc[r] = 0.0f;
    if (k < nz) {   // uniform
        const size_t rec = zsep_view_floats(dt.nu, dt.nv, nz);
        for (int pv = 0; pv < nviews; ++pv) {
            zs_plane(acc, g, dt, zs_rec(work + (size_t)pv * rec, dt, nz), k, iw, c,
                     s_lds + wave * (ZS_MAXCOL * 64));
        }
    }
    // write-back through LDS: each thread stores 4 consecutive z values of
    // one (x row, y cell)
    __syncthreads();   // every wavefront is done with its column slots
#pragma unroll
    for (int r = 0; r < ZS_ROWS; ++r) s_out[r][lane][wave] = acc[r];
    __syncthreads();
    const bool quads = (nz & 3) == 0 && ((uintptr_t)vol & 15) == 0;
    for (int q = threadIdx.x; q < ZS_ROWS * ZS_TY * 2; q += ZS_THREADS) {
        const int h = q & 1, cc = (q >> 1) % ZS_TY, r = q / (2 * ZS_TY);
        const int i = iw + r, cy = c0 + cc, kq = k0 + 4 * h;
        if (i >= nx || cy >= ny || kq >= nz) continue;
        const float* sv = &s_out[r][cc][4 * h];
        const size_t gi = ((size_t)i * ny + cy) * nz + kq;
        if (quads) {   // kq + 4 <= nz: nz and kq are multiples of 4
            float4* o = reinterpret_cast<float4*>(vol + gi);
            float4 v = make_float4(sv[0], sv[1], sv[2], sv[3]);
            if (add) {
                const float4 w = *o;
                v = make_float4(w.x + v.x, w.y + v.y, w.z + v.z, w.w + v.w);
            }
            *o = v;
        } else {
            for (int e = 0; e < 4 && kq + e < nz; ++e)
                vol[gi + e] = add ? vol[gi + e] + sv[e] : sv[e];
        }
    }
}

bool zsep_route(const ct2_desc* d, const Grid& g) {
    return d->brick[0] <= 0 && d->brick[1] <= 0 && d->brick[2] <= 0 &&
           d->arith == CT2_ARITH_FAST && d->interp == CT2_LINEAR && d->vol_dtype == CT2_F32 &&
           d->axis == CT2_AXIS_Z && d->nv >= ZS_NC && g.s[2] == 1 && g.s[1] == g.n[2] &&
           g.s[0] == g.n[1] * g.n[2];
}

// The records of views view0 + pv*vstride (pv < nviews) that view_zsep
// accepts (ZsRec): the planes' maps; the columns' and rows' set-ups; td =
// trav datum of every ray, 0 where the indicator leaves it out, with trav =
// |d_z| fast_sqrt((1 + r1f^2) + r2f^2) as setup_fast_ax forms it.  One
// workgroup per (view, detector column); the workgroups of column 0 also
// write the rows and the planes, and count the views the route leaves
// into *nonsep (the brick pass after the tiles exits at once when it is 0).
__global__ __launch_bounds__(256) void k_zsep_td(Grid g, Det dt, int view0, int nviews,
                                                 int vstride, float* __restrict__ work,
                                                 int* __restrict__ nonsep,
                                                 const float* __restrict__ sino,
                                                 const int8_t* __restrict__ ind,
                                                 const float* __restrict__ fv) {
    const int pv = blockIdx.x / dt.nu, j = blockIdx.x % dt.nu;
    if (pv >= nviews) return;
    const int p = view0 + pv * vstride;
    const float* f = fv + (size_t)p * 18;
    const int nz = g.n[2];
    float* rec = work + (size_t)pv * zsep_view_floats(dt.nu, dt.nv, nz);
    float4* pm = reinterpret_cast<float4*>(rec);
    if (!view_zsep(f, g, dt)) {   // uniform
        // the brick pass takes the view; its maps marked with bx = 0 (a view
        // the route takes has |bx| >= ZSEP_SPACING) tell the tiles to skip it
        if (j == 0) {
            for (int k = threadIdx.x; k < nz; k += 256) pm[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (threadIdx.x == 0) atomicAdd(nonsep, 1);
        }
        return;
    }
    float2* cx = reinterpret_cast<float2*>(rec + 4 * (size_t)nz);
    float2* cy = cx + dt.nu;
    // the fast set-up (setup_fast_ax with a = z, b = x, c = y): e_u, e_v have
    // no z component, so every bin has z = D_z
    const double Sx = f[0], Sy = f[1], Sz = f[2];
    const double ia = rcp_f64((double)f[8] - Sz);
    const double ta = fma(0.5, (double)g.d[2], (double)g.o[2]) - Sz;
    const double idx = rcp_f64((double)g.d[0]), idy = rcp_f64((double)g.d[1]);
    const double dzidx = (double)g.d[2] * idx, dzidy = (double)g.d[2] * idy;
    const double ox = g.o[0], oy = g.o[1];
    const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
    const double r1 = (fma((double)f[9], u, (double)f[6]) - Sx) * ia;
    if (threadIdx.x == 0)
        cx[j] = make_float2((float)fma(fma(r1, ta, Sx) - ox, idx, -0.5), (float)(r1 * dzidx));
    const float r1f = (float)r1;
    const float q1 = 1.0f + r1f * r1f;
    const float adz = fabsf(g.d[2]);
    const size_t row = ((size_t)p * dt.nu + j) * dt.nv;
    float* __restrict__ out = rec + zsep_td_offset(dt.nu, dt.nv, nz) + (size_t)j * dt.nv;
    for (int k = threadIdx.x; k < dt.nv; k += 256) {
        const double v = fma((double)k + 0.5, (double)dt.dv, (double)dt.v0);
        const double r2 = (fma((double)f[13], v, (double)f[7]) - Sy) * ia;
        const float r2f = (float)r2;
        const float trav = adz * fast_sqrt(q1 + r2f * r2f);
        const bool live = ind == nullptr || ind[row + k] == 1;
        out[k] = live ? sino[row + k] * trav : 0.0f;
        if (j == 0)
            cy[k] = make_float2((float)fma(fma(r2, ta, Sy) - oy, idy, -0.5), (float)(r2 * dzidy));
    }
    if (j == 0) {
        const double eux = f[9], evy = f[13];
        for (int k = threadIdx.x; k < nz; k += 256) {
            // X_k(j) = (S_x + r1(j) T_k - o_x)/d_x - 0.5 with r1 affine in j
            const double T = fma((double)k, (double)g.d[2], ta);
            const double sx = T * ia * idx, sy = T * ia * idy;
            const double ax = fma(Sx - ox, idx, -0.5) +
                              sx * (fma(eux, fma(0.5, (double)dt.du, (double)dt.u0), (double)f[6]) - Sx);
            const double ay = fma(Sy - oy, idy, -0.5) +
                              sy * (fma(evy, fma(0.5, (double)dt.dv, (double)dt.v0), (double)f[7]) - Sy);
            pm[k] = make_float4((float)ax, (float)(sx * eux * (double)dt.du), (float)ay,
                                (float)(sy * evy * (double)dt.dv));
        }
    }
}

// Bytes of one view's record.
static size_t zsep_view_bytes(const Det& dt, int nz) {
    return zsep_view_floats(dt.nu, dt.nv, nz) * sizeof(float);
}

// An internally allocated workspace holds the td values of at most this
// many bytes of views (the tomosynthesis scan of 25 views of 2048 x 1536
// needs 315 MB).
constexpr size_t ZSEP_WS_CAP = (size_t)4 << 30;

// The workspace: a header (the count of views the route leaves to the
// brick pass), then one record per view of a group.
constexpr size_t ZSEP_WS_HEAD = 256;

size_t zsep_work_bytes(const ct2_desc* d, const Det& dt, const ViewSet& vs) {
    return ZSEP_WS_HEAD + (size_t)vs.nviews * zsep_view_bytes(dt, d->nz);
}

__global__ void k_zsep_clear(int* nonsep) {
    if (threadIdx.x == 0) *nonsep = 0;
}

int launch_backproject_zsep(const Grid& g, const Det& dt, const ViewSet& vs, float* vol,
                            const float* sino, const int8_t* ind, const float* fv, hipStream_t st,
                            SepWork w, void** own_out, const int** nonsep) {
    *own_out = nullptr;
    *nonsep = nullptr;
    const long long ntile = (long long)((g.n[0] + ZS_ROWS - 1) / ZS_ROWS) *
                            ((g.n[1] + ZS_TY - 1) / ZS_TY) * ((g.n[2] + ZS_KZ - 1) / ZS_KZ);
    if (ntile > 0x7fffffffLL) {
        set_error("ct2_backproject: volume too large for the plane-separable route");
        return CT2_EINVAL;
    }
    const size_t per = zsep_view_bytes(dt, g.n[2]);
    void* own = nullptr;
    size_t gv;
    if (w.ptr) {
        if (((uintptr_t)w.ptr & 15u) != 0) {
            set_error("ct2_backproject: workspace must be 16-byte aligned");
            return CT2_EINVAL;
        }
        if (w.bytes < ZSEP_WS_HEAD + per) {
            set_error("ct2_backproject: workspace of %zu B holds no view's record (%zu B)",
                      w.bytes, ZSEP_WS_HEAD + per);
            return CT2_EINVAL;
        }
        gv = (w.bytes - ZSEP_WS_HEAD) / per;
    } else {
        gv = max((size_t)1, ZSEP_WS_CAP / per);
    }
    gv = min(gv, (size_t)max(vs.nviews, 1));   // so the int view counts below hold it
    if (!w.ptr) {
        int dev = 0;
        CT2_HIP(hipGetDevice(&dev));
        keep_pool_memory(dev);
        CT2_HIP(hipMallocAsync(&own, ZSEP_WS_HEAD + per * gv, st));
        *own_out = own;
    }
    char* base = static_cast<char*>(own ? own : w.ptr);
    int* count = reinterpret_cast<int*>(base);
    float* td = reinterpret_cast<float*>(base + ZSEP_WS_HEAD);
    // a kernel, not hipMemsetAsync, so that a captured graph replays it
    hipLaunchKernelGGL(k_zsep_clear, dim3(1), dim3(64), 0, st, count);
    int rc = CT2_OK;
    // groups of gv views: the first stores every voxel, later ones add, in
    // stream order (so the result does not depend on the schedule)
    for (int v = 0; v == 0 || v < vs.nviews; v += (int)gv) {
        const int nvg = min((int)gv, vs.nviews - v), v0 = vs.view0 + v * vs.vstride;
        const long long items = (long long)nvg * dt.nu;
        if (items > 0)
            hipLaunchKernelGGL(k_zsep_td, dim3((unsigned)items), dim3(256), 0, st, g, dt, v0, nvg,
                               vs.vstride, td, count, sino, ind, fv);
        hipLaunchKernelGGL(k_bp_zsep, dim3((unsigned)ntile), dim3(ZS_THREADS), 0, st, g, dt, nvg,
                           v > 0 ? 1 : 0, vol, td);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { rc = hip_fail(e, "ct2_backproject"); break; }
    }
    // the caller frees an allocated workspace (sep_work_release) after the
    // brick pass, which reads the count
    *own_out = own;
    *nonsep = count;
    return rc;
}

}  // namespace ct2
