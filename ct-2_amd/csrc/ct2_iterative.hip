// ct2_iterative.hip -- ART (sequential Kaczmarz) and MLEM.
//
// ART (emil_code/view_proj3D.f:1950-2271, :2275-2503; view_proj3D_tomo.f
// :154-347, :352-612; ray_proj3D.f:987-1449) updates the volume after every
// ray, in (view, u, v) order.  Two rays commute when they touch disjoint
// voxels, so the kernel runs an order-equivalent schedule: within a view, ray
// (j,k) gets level (r+1)*j + k, where r is a conflict radius (in bins) such
// that rays farther apart than r in j or k share no voxel.  Every pair of
// overlapping rays is then ordered by level exactly as in the sequential
// loop, and rays of one level are independent.  One workgroup of 16
// wavefronts walks the levels; each wavefront takes one ray of the level,
// its 64 lanes share the ray's planes, wave reductions give raysum and the
// squared-weight total, and the lanes scatter the update.  A workgroup
// barrier separates levels (workgroup-scope release/acquire on the volume).
// Only the raysum/total summation order differs from the reference.
//
// MLEM (view_proj3D.f:2507-3075, view_proj3D_tomo.f:615-822,
// ray_proj3D.f:1454-1651) is simultaneous: forward projection, datum/raysum
// ratio with the reference's small-value guard, two exact-adjoint
// backprojections (ratio and indicator), and a voxel update.
#include <math.h>

#include "ct2_internal.h"

namespace ct2 {

constexpr int ART_WAVES = 16;

__device__ __forceinline__ float wave_sum(float v) { return wave_sum_f(v); }

struct ArtGrid2 {
    int present;
    Grid g;
};

// Per-ray work of one wavefront: march (lanes over planes), reduce, update.
template <int INTERP>
__device__ __forceinline__ void art_ray(const Grid& g, const ArtGrid2& g2, const float S[3],
                                       const float B[3], bool zonly, bool strict, float reg_fac,
                                       float datum, float* __restrict__ vol,
                                       float* __restrict__ vol2, int lane) {
    AxisSel A, A2;
    setup_exact(g, S, B, zonly, A);
    if (g2.present) setup_exact(g2.g, S, B, true, A2);
    float raysum = 0.0f, total = 0.0f;
    for (int pass = 0; pass < (g2.present ? 2 : 1); ++pass) {
        const AxisSel& X = pass ? A2 : A;
        const float* V = pass ? vol2 : vol;
        for (int l = lane; l < X.na; l += 64) {
            Sample sm;
            if (INTERP == CT2_LINEAR) {
                if (!sample_linear_exact(X, l, sm)) continue;
                int i00 = l * X.sa + sm.ib * X.sb + sm.ic * X.sc;
                float v00 = V[i00], v10 = V[i00 + X.sb], v01 = V[i00 + X.sc],
                      v11 = V[i00 + X.sb + X.sc];
                float vi1 = sm.wb1 * v00 + sm.wb2 * v10;
                float vi2 = sm.wb1 * v01 + sm.wb2 * v11;
                float val = sm.wc1 * vi1 + sm.wc2 * vi2;
                raysum = raysum + X.trav * val;
                float w;
                w = X.trav * sm.wb1 * sm.wc1; total = total + w * w;
                w = X.trav * sm.wb2 * sm.wc1; total = total + w * w;
                w = X.trav * sm.wb1 * sm.wc2; total = total + w * w;
                w = X.trav * sm.wb2 * sm.wc2; total = total + w * w;
            } else {
                if (!sample_nearest_exact(X, l, strict, sm)) continue;
                raysum = raysum + X.trav * V[l * X.sa + sm.ib * X.sb + sm.ic * X.sc];
                total = total + X.trav * X.trav;
            }
        }
    }
    raysum = wave_sum(raysum);
    total = wave_sum(total);
    if (total == 0.0f) return;   // no sample: the reference's list is empty
    const float resid = reg_fac * (datum - raysum);
    for (int pass = 0; pass < (g2.present ? 2 : 1); ++pass) {
        const AxisSel& X = pass ? A2 : A;
        float* V = pass ? vol2 : vol;
        for (int l = lane; l < X.na; l += 64) {
            Sample sm;
            if (INTERP == CT2_LINEAR) {
                if (!sample_linear_exact(X, l, sm)) continue;
                int i00 = l * X.sa + sm.ib * X.sb + sm.ic * X.sc;
                float w;
                w = X.trav * sm.wb1 * sm.wc1; V[i00] = V[i00] + resid * w / total;
                w = X.trav * sm.wb2 * sm.wc1; V[i00 + X.sb] = V[i00 + X.sb] + resid * w / total;
                w = X.trav * sm.wb1 * sm.wc2; V[i00 + X.sc] = V[i00 + X.sc] + resid * w / total;
                w = X.trav * sm.wb2 * sm.wc2;
                V[i00 + X.sb + X.sc] = V[i00 + X.sb + X.sc] + resid * w / total;
            } else {
                if (!sample_nearest_exact(X, l, strict, sm)) continue;
                int i = l * X.sa + sm.ib * X.sb + sm.ic * X.sc;
                V[i] = V[i] + resid * X.trav / total;
            }
        }
    }
}

// View geometry: levels per view = stride*(nu-1) + nv.
template <int INTERP>
__global__ __launch_bounds__(64 * ART_WAVES) void k_art_views(
    Grid g, ArtGrid2 g2, Det dt, int view0, int nviews, int zonly, int stride, float reg_fac,
    const float* __restrict__ sino, const int8_t* __restrict__ ind,
    const float* __restrict__ fv, float* __restrict__ vol, float* __restrict__ vol2) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nlev = stride * (dt.nu - 1) + dt.nv;
    for (int pv = 0; pv < nviews; ++pv) {
        const int p = view0 + pv;
        for (int L = 0; L < nlev; ++L) {
            // rays of level L: k = L - stride*j in [0, nv)
            int jlo = L - dt.nv + 1 <= 0 ? 0 : (L - dt.nv + 1 + stride - 1) / stride;
            int jhi = min(dt.nu - 1, L / stride);
            for (int j = jlo + wave; j <= jhi; j += ART_WAVES) {
                const int k = L - stride * j;
                const size_t rr = ((size_t)p * dt.nu + j) * dt.nv + k;
                if (ind != nullptr && ind[rr] != 1) continue;
                float S[3], B[3];
                view_ray(fv, p, j, k, dt, S, B);
                art_ray<INTERP>(g, g2, S, B, zonly != 0, false, reg_fac, sino[rr], vol, vol2, lane);
            }
            __syncthreads();
        }
    }
}

// Ray list: no bin grid to order by, so rays run one after another.
template <int INTERP>
__global__ __launch_bounds__(64) void k_art_rays(Grid g, int nr, int strict, float reg_fac,
                                                 const float* __restrict__ sino,
                                                 const int8_t* __restrict__ ind,
                                                 const float* __restrict__ rv,
                                                 float* __restrict__ vol) {
    const int lane = threadIdx.x & 63;
    ArtGrid2 none;
    none.present = 0;
    for (int r = 0; r < nr; ++r) {
        if (ind != nullptr && ind[r] != 1) continue;
        float S[3], B[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            S[q] = rv[(size_t)r * 6 + q];
            B[q] = rv[(size_t)r * 6 + 3 + q];
        }
        art_ray<INTERP>(g, none, S, B, false, strict != 0, reg_fac, sino[r], vol, nullptr, lane);
        __syncthreads();   // one wavefront: orders this ray's stores before the next loads
    }
}

// ---------------------------------------------------------------- host ----
// Conflict radius: a touched voxel centre lies within sqrt(db^2+dc^2) <=
// sqrt(2)*dmax of its ray, so two rays share no voxel when their lines stay
// more than 2*sqrt(2)*dmax apart inside the volume box.  For two rays from S
// through bins B1, B2 of a flat detector at perpendicular distance h,
// sin(angle) >= |B1-B2| h / Lmax^2, and the rays are t*sin(angle) apart at
// distance t from S, so t >= tmin (nearest volume point) bounds the gap.
// Rays whose bins differ by more than r in j or k are |B1-B2| >= (r+1)*pmin
// apart.  Returns -1 when the bound does not apply (source inside the box).
int conflict_radius(const Grid& g, const Grid* g2, const Det& dt, const float* fv_host,
                    int view0, int nviews) {
    double dmax = fmax(fmax(g.d[0], g.d[1]), g.d[2]);
    double lo[3], hi[3];
    for (int q = 0; q < 3; ++q) {
        double a = g.o[q], b = g.o[q] + g.n[q] * (double)g.d[q];
        lo[q] = fmin(a, b); hi[q] = fmax(a, b);
    }
    if (g2) {
        dmax = fmax(dmax, fmax(fmax(fabs(g2->d[0]), fabs(g2->d[1])), fabs(g2->d[2])));
        for (int q = 0; q < 3; ++q) {
            double a = g2->o[q], b = g2->o[q] + g2->n[q] * (double)g2->d[q];
            lo[q] = fmin(lo[q], fmin(a, b)); hi[q] = fmax(hi[q], fmax(a, b));
        }
    }
    double pmin = fmin(fabs(dt.du), fabs(dt.dv));
    int rmax = 0;
    for (int p = view0; p < view0 + nviews; ++p) {
        const float* f = fv_host + (size_t)p * 18;
        double S[3] = {f[0], f[1], f[2]}, D[3] = {f[6], f[7], f[8]};
        double eu[3] = {f[9], f[10], f[11]}, ev[3] = {f[12], f[13], f[14]};
        // nearest point of the box to S
        double t2 = 0;
        for (int q = 0; q < 3; ++q) {
            double c = fmin(fmax(S[q], lo[q]), hi[q]);
            t2 += (c - S[q]) * (c - S[q]);
        }
        double tmin = sqrt(t2);
        if (tmin <= 1e-9) return -1;
        double n[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2],
                       eu[0] * ev[1] - eu[1] * ev[0]};
        double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (nn < 1e-12) return -1;
        double h = fabs((D[0] - S[0]) * n[0] + (D[1] - S[1]) * n[1] + (D[2] - S[2]) * n[2]) / nn;
        if (h < 1e-9) return -1;
        // |e_u|, |e_v| scale bin spacing on the panel
        double su = sqrt(eu[0] * eu[0] + eu[1] * eu[1] + eu[2] * eu[2]);
        double sv = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
        // the bin-distance bound below assumes orthogonal e_u, e_v
        if (fabs(eu[0] * ev[0] + eu[1] * ev[1] + eu[2] * ev[2]) > 1e-6 * su * sv) return -1;
        double pm = fmin(fabs(dt.du) * su, fabs(dt.dv) * sv);
        pmin = pm;
        double L2 = 0;
        for (int cu = 0; cu < 2; ++cu)
            for (int cv = 0; cv < 2; ++cv) {
                double u = dt.u0 + (cu ? dt.nu : 0) * (double)dt.du;
                double v = dt.v0 + (cv ? dt.nv : 0) * (double)dt.dv;
                double l2 = 0;
                for (int q = 0; q < 3; ++q) {
                    double b = D[q] + eu[q] * u + ev[q] * v - S[q];
                    l2 += b * b;
                }
                L2 = fmax(L2, l2);
            }
        double need = 2.0 * sqrt(2.0) * dmax * 1.25;   // 25% margin
        // need (r+1) * pmin * h * tmin / L2 > need, i.e. r + 1 > rp1
        double rp1 = need * L2 / (pmin * h * tmin);
        if (!(rp1 < 1e6)) return -1;
        int r = (int)ceil(rp1);
        if (r < 1) r = 1;
        rmax = r > rmax ? r : rmax;
    }
    return rmax;
}

int launch_art(const ct2_desc* d, const Grid* g2p, float reg_fac, const float* sino,
               const int8_t* ind, const float* fv, float* vol, float* vol2, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    if (g2p && (rc = check_grid(*g2p))) return rc;
    int view0 = d->view0, nviews = d->nviews ? d->nviews : d->ns - d->view0;
    if (view0 < 0 || nviews <= 0 || view0 + nviews > d->ns) {
        set_error("ART: bad view range"); return CT2_EINVAL;
    }
    if (!sino || !fv || !vol || (g2p && !vol2)) { set_error("ART: null pointer"); return CT2_EINVAL; }
    // the conflict radius needs the frames on the host
    float* fvh = (float*)malloc(sizeof(float) * 18 * (size_t)d->ns);
    if (!fvh) { set_error("ART: host allocation failed"); return CT2_ENOMEM; }
    hipError_t e = hipMemcpyAsync(fvh, fv, sizeof(float) * 18 * (size_t)d->ns, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { free(fvh); return hip_fail(e, "ART: frame copy"); }
    int r = conflict_radius(g, g2p, dt, fvh, view0, nviews);
    free(fvh);
    // Level = stride*j + k.  Any stride >= r+1 orders every conflicting pair
    // (|dj|,|dk| <= r) as the sequential loop does and keeps same-level rays
    // more than r apart in k.  When the bound does not apply (r < 0),
    // stride = nv gives one ray per level: the plain sequential order.
    int stride = r < 0 ? dt.nv : min(r + 1, dt.nv);
    ArtGrid2 a2;
    a2.present = g2p ? 1 : 0;
    if (g2p) a2.g = *g2p; else a2.g = g;
    int zonly = (d->axis == CT2_AXIS_Z) ? 1 : 0;
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_art_views<CT2_LINEAR>, dim3(1), dim3(64 * ART_WAVES), 0, st, g, a2, dt,
                           view0, nviews, zonly, stride, reg_fac, sino, ind, fv, vol, vol2);
    else
        hipLaunchKernelGGL(k_art_views<CT2_NEAREST>, dim3(1), dim3(64 * ART_WAVES), 0, st, g, a2, dt,
                           view0, nviews, zonly, stride, reg_fac, sino, ind, fv, vol, vol2);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int launch_ray_art(const ct2_desc* d, int strict, float reg_fac, const float* sino,
                   const int8_t* ind, const float* rv, int nr, float* vol, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    int rc;
    if ((rc = check_grid(g))) return rc;
    if (nr < 0 || !vol) { set_error("ray ART: bad arguments"); return CT2_EINVAL; }
    if (nr == 0) return CT2_OK;
    if (!sino || !rv) { set_error("ray ART: null pointer"); return CT2_EINVAL; }
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL(k_art_rays<CT2_LINEAR>, dim3(1), dim3(64), 0, st, g, nr, 0, reg_fac, sino, ind, rv, vol);
    else
        hipLaunchKernelGGL(k_art_rays<CT2_NEAREST>, dim3(1), dim3(64), 0, st, g, nr, strict, reg_fac, sino, ind, rv, vol);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// ------------------------------------------------------------------ MLEM ----
// ratio[r] = datum / max-guarded raysum (0 where the ray is unused); ones[r] = 1
__global__ void k_em_ratio(size_t n, float small, const float* __restrict__ sino,
                           const int8_t* __restrict__ ind, float* __restrict__ ratio,
                           float* __restrict__ ones) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bool use = ind == nullptr || ind[i] == 1;
    float rs = ratio[i];
    if (rs < small) rs = 1.0f;
    ratio[i] = use ? sino[i] / rs : 0.0f;
    if (ones) ones[i] = use ? 1.0f : 0.0f;
}

// x = x * W / O with O < small -> 1 (view_proj3D.f:2819-2827)
__global__ void k_em_update(size_t n, float small, float* __restrict__ vol,
                            const float* __restrict__ W, const float* __restrict__ O) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o = O[i];
    if (o < small) o = 1.0f;
    vol[i] = vol[i] * W[i] / o;
}

// Work layout (floats): [ratio: nsino][W: nvox][O: nvox][ones: nsino]
size_t em_work_floats(const ct2_desc* d) {
    size_t nsino = (size_t)d->ns * d->nu * d->nv, nvox = (size_t)d->nx * d->ny * d->nz;
    return 2 * nsino + 2 * nvox;
}

int launch_em(const ct2_desc* d, const float* sino, const int8_t* ind, const float* fv,
              float* vol, float* work, const float* sens, hipStream_t st) {
    if (!work || !vol || !sino || !fv) { set_error("MLEM: null pointer"); return CT2_EINVAL; }
    size_t nsino = (size_t)d->ns * d->nu * d->nv, nvox = (size_t)d->nx * d->ny * d->nz;
    float* ratio = work;
    float* W = work + nsino;
    float* O = W + nvox;
    float* ones = O + nvox;
    float small = d->interp == CT2_LINEAR ? 0.00001f : 0.00000001f;
    ct2_desc e = *d;
    e.vol_dtype = CT2_F32;
    e.view0 = 0;
    e.nviews = 0;
    int rc = launch_project(&e, ratio, ind, fv, vol, st);
    if (rc) return rc;
    unsigned nb = (unsigned)((nsino + 255) / 256);
    hipLaunchKernelGGL(k_em_ratio, dim3(nb), dim3(256), 0, st, nsino, small, sino, ind, ratio,
                       sens ? nullptr : ones);
    if ((rc = launch_backproject(&e, W, ratio, ind, fv, st))) return rc;
    const float* Ouse = sens;
    if (!sens) {
        if ((rc = launch_backproject(&e, O, ones, ind, fv, st))) return rc;
        Ouse = O;
    }
    hipLaunchKernelGGL(k_em_update, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, nvox,
                       small, vol, W, Ouse);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// ------------------------------------------------------------------ SART ----
// Simultaneous ART over the view subset V = {view0 + j*vstride} (Andersen &
// Kak; with vstride > 1, ordered subsets of interleaved views):
//   x_j += lambda * sum_{i in V} a_ij (b_i - a_i.x) / (sum_k a_ik)  /  sum_{i in V} a_ij
// with the row sums A1 and column sums A^T 1 of the same subset; rows or
// columns with a zero sum contribute nothing.  The reference has no SART;
// it is composed from the pinned forward and exact adjoint.
// resid[q] <- sart_resid(sino[q], resid[q], rowsum[q]) on the rays of views
// view0 + blockIdx.y*vstride (resid holds the forward projection on entry)
__global__ void k_sart_resid(int view0, int vstride, size_t per_view,
                             const float* __restrict__ sino, const int8_t* __restrict__ ind,
                             const float* __restrict__ rowsum, float* __restrict__ resid) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_view) return;
    size_t q = (size_t)(view0 + blockIdx.y * vstride) * per_view + e;
    resid[q] = sart_resid(sino[q], resid[q], rowsum[q], ind == nullptr || ind[q] == 1);
}

__global__ void k_fill_views(int view0, int vstride, size_t per_view, float v,
                             float* __restrict__ p) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_view) return;
    p[(size_t)(view0 + blockIdx.y * vstride) * per_view + e] = v;
}

__global__ void k_sart_update(size_t n, float lambda, float* __restrict__ vol,
                              const float* __restrict__ corr, const float* __restrict__ colsum) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    vol[q] = sart_update(vol[q], lambda, corr[q], colsum[q]);
}

size_t sart_work_floats(const ct2_desc* d) {
    size_t nsino = (size_t)d->ns * d->nu * d->nv, nvox = (size_t)d->nx * d->ny * d->nz;
    return 2 * nsino + 3 * nvox;
}

int launch_sart(const ct2_desc* d, float lambda, const float* sino, const int8_t* ind,
                const float* fv, float* vol, float* work, hipStream_t st) {
    if (!work || !vol || !sino || !fv) { set_error("SART: null pointer"); return CT2_EINVAL; }
    if (d->vol_dtype != CT2_F32) { set_error("SART: updates a float32 volume"); return CT2_EINVAL; }
    int rc;
    ViewSet vs;
    if ((rc = view_set(d, "SART", vs))) return rc;
    size_t nsino = (size_t)d->ns * d->nu * d->nv, nvox = (size_t)d->nx * d->ny * d->nz;
    float* resid = work;
    float* rowsum = work + nsino;
    float* corr = rowsum + nsino;
    float* colsum = corr + nvox;
    float* ones = colsum + nvox;
    size_t per_view = (size_t)d->nu * d->nv;
    dim3 vgrid((unsigned)((per_view + 255) / 256), (unsigned)vs.nviews);
    if ((rc = launch_fill(ones, 1.0f, nvox, st))) return rc;
    if ((rc = launch_project(d, rowsum, ind, fv, ones, st))) return rc;
    if ((rc = launch_project(d, resid, ind, fv, vol, st))) return rc;
    hipLaunchKernelGGL(k_sart_resid, vgrid, dim3(256), 0, st, vs.view0, vs.vstride, per_view,
                       sino, ind, rowsum, resid);
    if ((rc = launch_backproject(d, corr, resid, ind, fv, st))) return rc;
    // column sums: backprojection of ones (reuse rowsum as a ones sinogram)
    hipLaunchKernelGGL(k_fill_views, vgrid, dim3(256), 0, st, vs.view0, vs.vstride, per_view,
                       1.0f, rowsum);
    if ((rc = launch_backproject(d, colsum, rowsum, ind, fv, st))) return rc;
    hipLaunchKernelGGL(k_sart_update, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st,
                       nvox, lambda, vol, corr, colsum);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// Ray-list MLEM (ray_proj3D.f:1454-1651): strict-bounds nearest forward,
// atomics scatter for W and O.
__global__ void k_ray_em_ratio(int nr, float small, const float* __restrict__ sino,
                               const int8_t* __restrict__ ind, float* __restrict__ ratio,
                               float* __restrict__ ones) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    bool use = ind == nullptr || ind[i] == 1;
    float rs = ratio[i];
    if (rs < small) rs = 1.0f;
    ratio[i] = use ? sino[i] / rs : 0.0f;
    ones[i] = use ? 1.0f : 0.0f;
}

int launch_ray_em(const ct2_desc* d, int strict, const float* sino, const int8_t* ind,
                  const float* rv, int nr, float* vol, float* work, hipStream_t st) {
    size_t nvox = (size_t)d->nx * d->ny * d->nz;
    float* ratio = work;
    float* ones = work + nr;
    float* W = ones + nr;
    float* O = W + nvox;
    float small = d->interp == CT2_LINEAR ? 0.00001f : 0.00000001f;
    int rc = launch_ray_project(d, ratio, ind, rv, nr, vol, strict, st);
    if (rc) return rc;
    if (nr > 0)
        hipLaunchKernelGGL(k_ray_em_ratio, dim3((nr + 255) / 256), dim3(256), 0, st, nr, small,
                           sino, ind, ratio, ones);
    if ((rc = launch_ray_backproject(d, W, ratio, ind, rv, nr, strict, st))) return rc;
    if ((rc = launch_ray_backproject(d, O, ones, ind, rv, nr, strict, st))) return rc;
    hipLaunchKernelGGL(k_em_update, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, st, nvox,
                       small, vol, W, O);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace ct2
