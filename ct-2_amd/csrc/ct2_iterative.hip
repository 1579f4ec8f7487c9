// ct2_iterative.hip -- ART (sequential Kaczmarz) and MLEM.
//
// ART (emil_code/view_proj3D.f:1950-2271, :2275-2503; view_proj3D_tomo.f
// :154-347, :352-612; ray_proj3D.f:987-1449) updates the volume after every
// ray, in (view, u, v) order.  Two rays commute when they touch disjoint
// voxels, so the kernels run order-equivalent schedules.  Per ray, one
// wavefront's 64 lanes share the ray's planes, wave reductions give raysum
// and the squared-weight total, and the lanes scatter the update; only the
// raysum/total summation order differs from the reference.
// * k_art_flow (the default): a dataflow schedule over the whole GPU,
//   described above the kernel.
// * k_art_views (the fallback): within a view, ray (j,k) gets level
//   (r+1)*j + k, where r is a conflict radius (in bins) such that rays
//   farther apart than r in j or k share no voxel.  Every pair of
//   overlapping rays is then ordered by level exactly as in the sequential
//   loop, and rays of one level are independent.  One workgroup of 16
//   wavefronts walks the levels, a workgroup barrier between levels.
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

// ---- ART across the GPU: a dataflow schedule in the sequential order ----
// Persistent wavefronts take detector columns (view, j) in the reference's
// order from a counter and run each column's rays k = 0, 1, ... in order.
// Before ray (p, j, k) a wavefront waits only for the earlier rays whose
// voxels it may share:
//   * its own view: column j-1 has finished rows up to k + r (r the conflict
//     radius); the column ahead of it waited the same way, so every earlier
//     column has too;
//   * each of the ARTF_DEPTH-1 previous views q: the line of ray (p, j, k)
//     inside the voxel box, grown by the conflict distance, projected
//     through q's source onto q's detector bounds the bins of q whose rays
//     can share a voxel with it; column jmax of q has finished rows up to
//     kmax (columns of a view finish in order, so every earlier column has);
//   * views ARTF_DEPTH and more back have finished (checked once per column).
// Every pair of rays that share a voxel therefore runs in the sequential
// order, and each ray reads exactly the values the sequential loop reads.
// Progress per column is published with an agent-scope release after the
// ray's stores; waits acquire at agent scope before the ray's loads.  A wait
// gives up after ARTF_WAIT_TICKS of the constant 100 MHz clock (30 s, far
// beyond any wait a sweep needs): it sets the error word, and every other
// wait sees it and leaves, so a broken schedule ends in an error, not a hang.
// Items are handed out in order and each waits only on earlier ones, so the
// earliest unfinished item can always proceed.
constexpr int ARTF_WAVES = 4;       // wavefronts per workgroup, each a worker
constexpr int ARTF_DEPTH = 8;       // views in flight
constexpr unsigned long long ARTF_WAIT_TICKS = 3000000000ull;

struct ArtFlow {
    int* prog;        // [nviews * nu]: rows finished per column item
    int* next;        // next column item to hand out
    int* err;         // nonzero once a wait gave up
    int r;            // same-view conflict radius in bins
    unsigned long long wait_ticks;   // give-up time of one wait (100 MHz ticks)
    double delta;     // conflict distance (two rays' tubes), mm
    double lo[3], hi[3];   // box of every voxel centre of the grid(s)
};

// Bins of view frame fq whose rays may share a voxel with the line S + t(B-S):
// its part inside the box grown by delta, projected through fq's source onto
// fq's detector and grown by delta's image there.  Returns the largest column
// and row of that window (-1: none), or nu-1, nv-1 when the projection does
// not apply (a point level with the source).
__device__ void art_window(const double S[3], const double B[3], const ArtFlow& af,
                           const float* __restrict__ fq, const Det& dt, int& jmax, int& kmax) {
    jmax = -1;
    kmax = -1;
    double dir[3], t0 = -1e300, t1 = 1e300;
    for (int q = 0; q < 3; ++q) {
        dir[q] = B[q] - S[q];
        const double a = af.lo[q] - af.delta, b = af.hi[q] + af.delta;
        if (fabs(dir[q]) < 1e-300) {
            if (S[q] < a || S[q] > b) return;
            continue;
        }
        double ta = (a - S[q]) / dir[q], tb = (b - S[q]) / dir[q];
        if (ta > tb) { const double t = ta; ta = tb; tb = t; }
        t0 = fmax(t0, ta);
        t1 = fmin(t1, tb);
    }
    if (!(t0 <= t1)) return;   // the line misses the box: it touches no voxel
    const double Sq[3] = {fq[0], fq[1], fq[2]}, Dq[3] = {fq[6], fq[7], fq[8]};
    const double eu[3] = {fq[9], fq[10], fq[11]}, ev[3] = {fq[12], fq[13], fq[14]};
    const double n[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2],
                         eu[0] * ev[1] - eu[1] * ev[0]};
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double eu2 = eu[0] * eu[0] + eu[1] * eu[1] + eu[2] * eu[2];
    const double ev2 = ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2];
    const double hq = (Dq[0] - Sq[0]) * n[0] + (Dq[1] - Sq[1]) * n[1] + (Dq[2] - Sq[2]) * n[2];
    double jm = -1e300, km = -1e300;
    for (int e = 0; e < 2; ++e) {
        const double t = e ? t1 : t0;
        double X[3], w[3];
        for (int q = 0; q < 3; ++q) {
            X[q] = S[q] + t * dir[q];
            w[q] = X[q] - Sq[q];
        }
        const double den = w[0] * n[0] + w[1] * n[1] + w[2] * n[2];
        const double s = hq / den;
        if (!(den * hq > 0.0) || !(s < 1e12)) {   // behind or level with the source
            jmax = dt.nu - 1;
            kmax = dt.nv - 1;
            return;
        }
        double P[3], L2 = 0.0, dotu = 0.0, dotv = 0.0, tx2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            P[q] = Sq[q] + s * w[q];
            L2 += (P[q] - Sq[q]) * (P[q] - Sq[q]);
            tx2 += w[q] * w[q];
            dotu += (P[q] - Dq[q]) * eu[q];
            dotv += (P[q] - Dq[q]) * ev[q];
        }
        const double L = sqrt(L2), cosq = fabs(hq) / (L * nn);
        // the image of the tube radius on the detector plane, in mm
        const double m = af.delta * L / (sqrt(tx2) * cosq);
        const double u = dotu / eu2, v = dotv / ev2;
        const double mu = m / sqrt(eu2), mv = m / sqrt(ev2);
        // bins whose centre u0 + (j+0.5) du lies within [u-mu, u+mu]
        for (int sg = -1; sg <= 1; sg += 2) {
            jm = fmax(jm, (u + sg * mu - dt.u0) / dt.du - 0.5);
            km = fmax(km, (v + sg * mv - dt.v0) / dt.dv - 0.5);
        }
    }
    jmax = (int)fmin(floor(jm), (double)(dt.nu - 1));
    kmax = (int)fmin(floor(km), (double)(dt.nv - 1));
    if (jmax < 0 || kmax < 0) jmax = kmax = -1;
}

// Lane 0 polls *p until it reaches need (acquire, agent scope); false when
// the wait gave up or another wait already had.
__device__ __forceinline__ bool art_wait(const ArtFlow& af, int* p, int need, int lane) {
    int ok = 1;
    if (lane == 0) {
        unsigned n = 0;
        unsigned long long t0 = 0;
        while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if ((n++ & 255u) == 0u) {
                const unsigned long long t = wall_clock64();
                if (n == 1u) t0 = t;
                if (__hip_atomic_load(af.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                    t - t0 > af.wait_ticks) {
                    __hip_atomic_store(af.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    return __builtin_amdgcn_readfirstlane(ok) != 0;
}

// Waits of one ray, one dependency per lane: lane 0 the previous column of
// the ray's view, lane d (1 <= d < DEPTH) its window in view p-d.  Each lane
// keeps the column and progress it last read (seen_col, seen); columns of a
// view finish in order, so progress seen on a column at or after the one
// needed also satisfies the need without another read.  Returns false when
// a wait gave up (the error word is set).
__device__ __forceinline__ bool art_deps(const ArtFlow& af, int* col_base, int col, int need,
                                         int& seen_col, int& seen) {
    bool want = col >= 0 && !(seen_col >= col && seen >= need);
    unsigned n = 0;
    unsigned long long t0 = 0;
    while (__ballot(want) != 0ull) {
        if (want) {
            const int v = __hip_atomic_load(col_base + col, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            seen_col = col;
            seen = v;
            want = v < need;
        }
        if (__ballot(want) == 0ull) break;
        if ((n++ & 255u) == 0u) {
            const unsigned long long t = wall_clock64();
            if (n == 1u) t0 = t;
            const int e = __hip_atomic_load(af.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__builtin_amdgcn_readfirstlane(e) != 0 || t - t0 > af.wait_ticks) {
                if ((threadIdx.x & 63) == 0)
                    __hip_atomic_store(af.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

template <int INTERP, int DEPTH>
__global__ __launch_bounds__(64 * ARTF_WAVES) void k_art_flow(
    Grid g, ArtGrid2 g2, Det dt, int view0, int nviews, int zonly, float reg_fac, ArtFlow af,
    const float* __restrict__ sino, const int8_t* __restrict__ ind,
    const float* __restrict__ fv, float* __restrict__ vol, float* __restrict__ vol2) {
    static_assert(DEPTH <= 64, "one dependency per lane");
    const int lane = threadIdx.x & 63;
    const int nitems = nviews * dt.nu;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(af.next, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= nitems) return;
        const int pv = item / dt.nu, j = item - pv * dt.nu, p = view0 + pv;
        // every view DEPTH or more back has finished
        if (pv >= DEPTH &&
            !art_wait(af, af.prog + (size_t)(pv - DEPTH) * dt.nu + dt.nu - 1, dt.nv, lane))
            return;
        const float* f = fv + (size_t)p * 18;
        // lane 0: column j-1 of this view; lane d: view p-d (columns from its base)
        const int dv = lane < DEPTH ? lane : 0;
        int* col_base = af.prog + (size_t)(pv - dv) * dt.nu;
        int seen_col = -1, seen = 0;
        for (int k = 0; k < dt.nv; ++k) {
            const size_t rr = ((size_t)p * dt.nu + j) * dt.nv + k;
            if (ind == nullptr || ind[rr] == 1) {
                float S[3], B[3];
                view_ray(fv, p, j, k, dt, S, B);
                int col = -1, need = 0;
                if (lane == 0) {
                    if (j > 0) { col = j - 1; need = min(k + af.r + 1, dt.nv); }
                } else if (lane < DEPTH && pv - lane >= 0) {
                    const double Sd[3] = {S[0], S[1], S[2]}, Bd[3] = {B[0], B[1], B[2]};
                    int jw, kw;
                    art_window(Sd, Bd, af, f - (size_t)lane * 18, dt, jw, kw);
                    if (jw >= 0) { col = jw; need = kw + 1; }
                }
                if (!art_deps(af, col_base, col, need, seen_col, seen)) return;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                art_ray<INTERP>(g, g2, S, B, zonly != 0, false, reg_fac, sino[rr], vol, vol2,
                                lane);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            }
            if (lane == 0)
                __hip_atomic_store(af.prog + item, k + 1, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void k_art_flow_clear(int* __restrict__ w, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        w[i] = 0;
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

// Box of every voxel centre of the grid(s) and the conflict distance of
// conflict_radius: rays whose lines stay farther apart inside the box share
// no voxel.
static void art_box(const Grid& g, const Grid* g2, double lo[3], double hi[3], double& delta) {
    double dmax = fmax(fmax(g.d[0], g.d[1]), g.d[2]);
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
    delta = 2.0 * sqrt(2.0) * dmax * 1.25;
}

// Row radius of the dataflow schedule's same-view wait: rays of one view
// whose rows differ by more than rv share no voxel.  When every view's e_v
// is parallel to z, e_u horizontal, and every ray's z slope is at most 0.2
// (so z never drives), two rays of a view from the same source differ in z
// by dv*|dk| * t / L at distance t from the source (L the source-to-bin
// distance), whatever their columns.  A shared voxel needs both rays' z
// cells within one cell of it, at points at most about a cell apart along
// the rays (x- and y-driven rays sample different planes), so the rays'
// z gap there is below 2 dz plus 0.4 dz of slope: |dk| > rv with
// rv + 1 >= 2 dz L_max / (dv t_min) * 1.35 keeps them apart.  Otherwise
// rv = r, the conflict radius.
static int art_row_radius(const Grid& g, const Grid* g2, const Det& dt, const float* fvh,
                          int view0, int nviews, int r) {
    double lo[3], hi[3], delta;
    art_box(g, g2, lo, hi, delta);
    double dz = fabs((double)g.d[2]), dxy = fmax(fabs((double)g.d[0]), fabs((double)g.d[1]));
    if (g2) {
        dz = fmax(dz, fabs((double)g2->d[2]));
        dxy = fmax(dxy, fmax(fabs((double)g2->d[0]), fabs((double)g2->d[1])));
    }
    double need = 0.0;
    for (int p = view0; p < view0 + nviews; ++p) {
        const float* f = fvh + (size_t)p * 18;
        const double evn = sqrt((double)f[12] * f[12] + (double)f[13] * f[13] +
                                (double)f[14] * f[14]);
        if (!(evn > 0.0) || fabs(f[12]) > 1e-7 * evn || fabs(f[13]) > 1e-7 * evn) return r;
        double tmin2 = 0.0;
        for (int q = 0; q < 3; ++q) {
            const double c = fmin(fmax((double)f[q], lo[q]), hi[q]);
            tmin2 += (c - f[q]) * (c - f[q]);
        }
        if (!(tmin2 > 0.0)) return r;
        double L2 = 0.0, zmax = 0.0, hmin = 1e300, smax = 0.0;
        for (int cu = 0; cu < 2; ++cu)
            for (int cv = 0; cv < 2; ++cv) {
                const double u = dt.u0 + (cu ? dt.nu : 0) * (double)dt.du;
                const double v = dt.v0 + (cv ? dt.nv : 0) * (double)dt.dv;
                double b[3];
                for (int q = 0; q < 3; ++q) b[q] = f[6 + q] + f[9 + q] * u + f[12 + q] * v - f[q];
                L2 = fmax(L2, b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
                zmax = fmax(zmax, fabs(b[2]) / dz);
                // the larger of |b_x|/dx, |b_y|/dy is at least |b_xy| / (sqrt2 dxy)
                hmin = fmin(hmin, sqrt(b[0] * b[0] + b[1] * b[1]) / (sqrt(2.0) * dxy));
                smax = fmax(smax, fabs(b[2]) / sqrt(b[0] * b[0] + b[1] * b[1] + 1e-300));
            }
        // z drives no ray and its slope stays below 0.2: the corners bound
        // |b_z| from above (it is linear in v) and |b_xy| from below when e_u
        // is horizontal
        if (fabs(f[11]) > 1e-7 * evn || !(zmax < 0.9 * hmin) || !(smax <= 0.2)) return r;
        need = fmax(need, 2.0 * dz * sqrt(L2) / (fabs((double)dt.dv) * evn * sqrt(tmin2)) * 1.35);
    }
    const int rv = (int)ceil(need) - 1;
    return rv >= 1 ? min(rv, r) : 1;
}

// The dataflow schedule projects lines through each source: every source of
// the views must lie outside the box grown by the conflict distance, and
// e_u, e_v must span a plane.
static bool art_flow_geometry_ok(const Grid& g, const Grid* g2, const float* fvh, int view0,
                                 int nviews) {
    double lo[3], hi[3], delta;
    art_box(g, g2, lo, hi, delta);
    for (int p = view0; p < view0 + nviews; ++p) {
        const float* f = fvh + (size_t)p * 18;
        bool inside = true;
        for (int q = 0; q < 3; ++q)
            inside &= f[q] >= lo[q] - 2 * delta && f[q] <= hi[q] + 2 * delta;
        if (inside) return false;
        const double n0 = (double)f[10] * f[14] - (double)f[11] * f[13];
        const double n1 = (double)f[11] * f[12] - (double)f[9] * f[14];
        const double n2 = (double)f[9] * f[13] - (double)f[10] * f[12];
        if (n0 * n0 + n1 * n1 + n2 * n2 < 1e-24) return false;
    }
    return true;
}

static int launch_art_flow(const ct2_desc* d, const Grid& g, const ArtGrid2& a2, const Det& dt,
                           int view0, int nviews, int zonly, int r, float reg_fac,
                           const float* sino, const int8_t* ind, const float* fv, float* vol,
                           float* vol2, hipStream_t st) {
    ArtFlow af;
    art_box(g, a2.present ? &a2.g : nullptr, af.lo, af.hi, af.delta);
    af.r = r;   // the row radius of the same-view wait
    // CT2_ART_WAIT_TICKS shortens the give-up time (tests of the error path)
    const char* wt = getenv("CT2_ART_WAIT_TICKS");
    af.wait_ticks = wt ? strtoull(wt, nullptr, 10) : ARTF_WAIT_TICKS;
    const int nitems = nviews * dt.nu;
    int* ws = nullptr;
    CT2_HIP(hipMallocAsync((void**)&ws, sizeof(int) * ((size_t)nitems + 2), st));
    af.prog = ws;
    af.next = ws + nitems;
    af.err = ws + nitems + 1;
    hipLaunchKernelGGL(k_art_flow_clear, dim3(64), dim3(256), 0, st, ws, nitems + 2);
    // at most one resident round of workgroups (each waits only on columns
    // handed out before its own, so the order in which they start is free)
    int dev = 0, cus = 0, bpc = 0;
    const void* fn = d->interp == CT2_LINEAR ? (const void*)k_art_flow<CT2_LINEAR, ARTF_DEPTH>
                                             : (const void*)k_art_flow<CT2_NEAREST, ARTF_DEPTH>;
    CT2_HIP(hipGetDevice(&dev));
    CT2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CT2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, fn, 64 * ARTF_WAVES, 0));
    const int grid = max(1, min(max(1, bpc) * cus, (nitems + ARTF_WAVES - 1) / ARTF_WAVES));
    if (d->interp == CT2_LINEAR)
        hipLaunchKernelGGL((k_art_flow<CT2_LINEAR, ARTF_DEPTH>), dim3(grid), dim3(64 * ARTF_WAVES), 0, st, g,
                           a2, dt, view0, nviews, zonly, reg_fac, af, sino, ind, fv, vol, vol2);
    else
        hipLaunchKernelGGL((k_art_flow<CT2_NEAREST, ARTF_DEPTH>), dim3(grid), dim3(64 * ARTF_WAVES), 0, st, g,
                           a2, dt, view0, nviews, zonly, reg_fac, af, sino, ind, fv, vol, vol2);
    hipError_t e = hipGetLastError();
    int err_h = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&err_h, af.err, sizeof(int), hipMemcpyDeviceToHost, st);
    hipError_t ef = hipFreeAsync(ws, st);
    if (e == hipSuccess) e = ef;
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "ART");
    if (err_h) {
        set_error("ART: a dependency wait of the dataflow schedule gave up; the volume is "
                  "incomplete");
        return CT2_EHIP;
    }
    return CT2_OK;
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
    const bool flow_ok = r >= 0 && art_flow_geometry_ok(g, g2p, fvh, view0, nviews);
    const int rv = flow_ok ? art_row_radius(g, g2p, dt, fvh, view0, nviews, r) : r;
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
    const char* env = getenv("CT2_ART_FLOW");
    if (flow_ok && !(env && env[0] == '0'))
        return launch_art_flow(d, g, a2, dt, view0, nviews, zonly, zonly ? r : rv, reg_fac,
                               sino, ind, fv, vol, vol2, st);
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
