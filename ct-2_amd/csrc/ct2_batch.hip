// ct2_batch.hip -- batched plans: many independent cone-beam problems of
// uneven size (volume shape, view count, detector) in one persistent launch.
//
// The reference runs one problem per call (emil_code/tomo3D.py:1127-1168,
// :432-472).  A plan binds every problem's buffers once; each entry point
// then launches ONE persistent kernel that walks the detector tiles (forward)
// or voxel bricks (adjoint) of all problems, so small problems no longer
// leave most of the GPU idle and the launches are few enough to capture an
// iteration loop in a graph.  SART runs over nsub ordered subsets of
// interleaved views (subset s = views s, s+nsub, ...), with the residual
// fused into the forward and the update fused into the adjoint's brick
// write-back; row sums A1 and the per-subset column sums A_s^T 1 are computed
// once when the plan is made.  The entry points after creation allocate
// nothing and synchronise nothing, so they can be captured.
//
// A plan of ONE problem whose forward and adjoint both take the
// column-separable kernels (fast linear arithmetic, detector rows along z)
// runs SART subset by subset through them instead: the forward of the subset
// (ct2_project's route), the residual, the adjoint into a plan-owned
// correction volume with a plan-owned column-table workspace (ct2_backproject's
// route), and the update.  Row and column sums come from the same kernels, so
// each subset step equals ct2_sart_step on that subset bit for bit.  At
// 256^3 x 360 views x 512^2 (10 subsets) the batched kernels take 5.6 ms per
// subset adjoint and 1.55 ms per subset forward, the separable ones about
// 1.8 and 0.6 ms.
//
// In a plan of several problems, SART's adjoint takes the column-separable
// tiles for every problem whose views all have their rows along z (fast
// linear arithmetic, nz <= 256, at most 512 views per subset): one launch
// builds the column tables of all those problems' subset views, then per
// z-cell class (nz up to 64, 128, 192, 256) one launch runs their x-driven
// tiles, storing the sums in a plan-owned correction volume, and one their
// y-driven tiles, adding their sums and applying the update.  Each problem's
// voxel sums are then those of ct2_backproject's two commuting atomic adds,
// so these problems' SART steps equal ct2_sart_step bit for bit too.  The
// persistent brick kernel keeps the other problems (its offsets leave the
// separable ones out), the all-view backprojection, and the forward.
#include <string.h>

#include <algorithm>
#include <vector>

#include "ct2_internal.h"

using namespace ct2;

struct ct2_batch {
    int nprob = 0, nsub = 1, interp = CT2_LINEAR, arith = CT2_ARITH_EXACT, device = 0;
    int tu = 4, fgrid = 8, bgrid = 8;
    int3 bdim = make_int3(kBrickCandidates[0][0], kBrickCandidates[0][1], kBrickCandidates[0][2]);
    std::vector<BatchProb> h;
    BatchProb* d_probs = nullptr;
    long long* d_foff = nullptr;   // (nsub+1) rows of nprob+1; row nsub = every view
    long long* d_boff = nullptr;
    float* rowsum = nullptr;
    float* resid = nullptr;
    float* colsum = nullptr;
    unsigned long long* d_queue = nullptr;   // work queues: [0,1] forward, [2,3] adjoint
    bool has_proj = false, has_adj = false;
    size_t bytes = 0;
    // one-problem plans on the column-separable kernels
    bool sep = false;
    ct2_desc sd{};             // the problem's descriptor (subset fields set per call)
    float* corr = nullptr;     // A_s^T r of the current subset
    void* work = nullptr;      // column tables of the largest subset
    size_t work_bytes = 0;
    // several problems: the separable tiles for problems psep[p]
    std::vector<char> psep;
    bool sepb = false, brick_any = true;
    SepBatchProb* d_sprobs = nullptr;
    long long* d_stoff = nullptr;    // per subset: table items of each problem (nprob+1)
    long long* d_sxoff = nullptr;    // per (subset, class, axis): tiles of each problem
    std::vector<long long> stab_items, stiles;
    int max_nv[5] = {0, 0, 0, 0, 0};
    void* sep_mem = nullptr;         // tables and correction volumes of those problems
    int* d_nonsep = nullptr;
};

namespace {

void release(ct2_batch* b) {
    if (!b) return;
    hipFree(b->d_probs);
    hipFree(b->d_foff);
    hipFree(b->d_boff);
    hipFree(b->rowsum);
    hipFree(b->resid);
    hipFree(b->colsum);
    hipFree(b->d_queue);
    hipFree(b->corr);
    hipFree(b->work);
    hipFree(b->d_sprobs);
    hipFree(b->d_stoff);
    hipFree(b->d_sxoff);
    hipFree(b->sep_mem);
    hipFree(b->d_nonsep);
    delete b;
}

// Item offsets of every subset row (and the all-views row) for tile height tu.
std::vector<long long> fwd_offsets(const ct2_batch* b, int tu) {
    const int n1 = b->nprob + 1;
    std::vector<long long> off((size_t)(b->nsub + 1) * n1, 0);
    for (int row = 0; row <= b->nsub; ++row) {
        long long acc = 0;
        for (int p = 0; p < b->nprob; ++p) {
            const Det& dt = b->h[p].dt;
            off[(size_t)row * n1 + p] = acc;
            int nv = row == b->nsub ? dt.ns : subset_views(dt.ns, row, b->nsub);
            acc += (long long)nv * ((dt.nu + tu - 1) / tu) * ((dt.nv + 63) / 64);
        }
        off[(size_t)row * n1 + b->nprob] = acc;
    }
    return off;
}

std::vector<long long> bp_offsets(ct2_batch* b, int3 bd) {
    const int n1 = b->nprob + 1;
    std::vector<long long> off((size_t)(b->nsub + 1) * n1, 0);
    for (int p = 0; p < b->nprob; ++p) {
        const Grid& g = b->h[p].g;
        b->h[p].nbr = make_int3((g.n[0] + bd.x - 1) / bd.x, (g.n[1] + bd.y - 1) / bd.y,
                                (g.n[2] + bd.z - 1) / bd.z);
    }
    for (int row = 0; row <= b->nsub; ++row) {
        long long acc = 0;
        for (int p = 0; p < b->nprob; ++p) {
            const BatchProb& P = b->h[p];
            off[(size_t)row * n1 + p] = acc;
            int nv = row == b->nsub ? P.dt.ns : subset_views(P.dt.ns, row, b->nsub);
            // (the separable tiles take these problems' subsets)
            if (row < b->nsub && b->sepb && b->psep[p]) nv = 0;
            if (nv > 0) acc += (long long)P.nbr.x * P.nbr.y * P.nbr.z;
        }
        off[(size_t)row * n1 + b->nprob] = acc;
    }
    return off;
}

int upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
    CT2_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    CT2_HIP(hipStreamSynchronize(st));
    return CT2_OK;
}

int persistent_grid(int blocks_per_cu, int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cus <= 0)
        cus = 256;
    int g = std::max(1, blocks_per_cu) * cus;
    return std::max(8, g & ~7);
}

int set_fwd(ct2_batch* b, int tu, hipStream_t st) {
    std::vector<long long> off = fwd_offsets(b, tu);
    int rc = upload(b->d_foff, off.data(), off.size() * sizeof(long long), st);
    if (rc) return rc;
    int occ = 0;
    if ((rc = batch_project_occupancy(b->interp, b->arith, tu, BFWD_RESID, &occ))) return rc;
    b->tu = tu;
    b->fgrid = persistent_grid(occ, b->device);
    return CT2_OK;
}

int set_bp(ct2_batch* b, int3 bd, hipStream_t st) {
    std::vector<long long> off = bp_offsets(b, bd);
    int rc;
    if ((rc = upload(b->d_probs, b->h.data(), sizeof(BatchProb) * b->nprob, st))) return rc;
    if ((rc = upload(b->d_boff, off.data(), off.size() * sizeof(long long), st))) return rc;
    int occ = 0;
    if ((rc = batch_backproject_occupancy(b->interp, b->arith, BBP_SART, bd, &occ))) return rc;
    b->bdim = bd;
    b->bgrid = persistent_grid(occ, b->device);
    return CT2_OK;
}

const long long* frow(const ct2_batch* b, int row) { return b->d_foff + (size_t)row * (b->nprob + 1); }
const long long* brow(const ct2_batch* b, int row) { return b->d_boff + (size_t)row * (b->nprob + 1); }

// r = sart_resid(b, A_s x, rowsum, used) over the views of subset view0,
// in place on the forward's output
__global__ void k_plan_resid(int view0, int vstride, size_t per_view,
                             const float* __restrict__ sino, const int8_t* __restrict__ ind,
                             const float* __restrict__ rowsum, float* __restrict__ resid) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_view) return;
    const size_t q = (size_t)(view0 + blockIdx.y * vstride) * per_view + e;
    resid[q] = sart_resid(sino[q], resid[q], rowsum[q], ind == nullptr || ind[q] == 1);
}

// x = sart_update(x, lambda, corr, colsum); lambda_dev, when given, holds
// the relaxation
__global__ void k_plan_update(size_t n, float lambda, const float* __restrict__ lambda_dev,
                              float* __restrict__ vol, const float* __restrict__ corr,
                              const float* __restrict__ colsum) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (lambda_dev) lambda = *lambda_dev;
    vol[q] = sart_update(vol[q], lambda, corr[q], colsum[q]);
}

ct2_desc subset_desc(const ct2_batch* b, int sub) {
    ct2_desc d = b->sd;
    d.view0 = sub;
    d.nviews = 0;
    d.vstride = b->nsub;
    return d;
}

// One SART subset of a one-problem separable plan.
int sep_sart_subset(ct2_batch* b, int sub, float lambda, const float* lambda_dev,
                    hipStream_t st) {
    const BatchProb& P = b->h[0];
    const int nv = subset_views(P.dt.ns, sub, b->nsub);
    if (nv == 0) return CT2_OK;
    const ct2_desc d = subset_desc(b, sub);
    int rc = launch_project(&d, P.resid, P.ind, P.fv, P.vol, st);
    if (rc) return rc;
    const size_t per_view = (size_t)P.dt.nu * P.dt.nv;
    hipLaunchKernelGGL(k_plan_resid, dim3((unsigned)((per_view + 255) / 256), (unsigned)nv),
                       dim3(256), 0, st, sub, b->nsub, per_view, P.sino, P.ind, P.rowsum,
                       P.resid);
    SepWork w;
    w.ptr = b->work;
    w.bytes = b->work_bytes;
    if ((rc = launch_backproject(&d, b->corr, P.resid, P.ind, P.fv, st, w))) return rc;
    const size_t n = (size_t)P.nvox;
    hipLaunchKernelGGL(k_plan_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                       lambda, lambda_dev, P.vol, b->corr, P.colsum + (size_t)sub * n);
    return CT2_OK;
}

const long long* sxrow(const ct2_batch* b, int sub, int cls, int ax) {
    return b->d_sxoff + ((size_t)(sub * 4 + cls - 1) * 2 + ax) * (b->nprob + 1);
}

// Subset sub's adjoint on the separable tiles of the problems that take them
// (kind BSEP_SART: update those estimates; BSEP_COLSUM: their column sums of
// the subset, from resid holding ones).
int sepb_subset(ct2_batch* b, int kind, int sub, float lambda, const float* lambda_dev,
                hipStream_t st) {
    const size_t n1 = (size_t)b->nprob + 1;
    int rc = launch_batch_sep_tables(b->d_sprobs, b->d_stoff + (size_t)sub * n1, b->nprob, sub,
                                     b->nsub, b->stab_items[sub], b->d_nonsep, st);
    for (int cls = 1; cls <= 4 && !rc; ++cls)
        for (int ax = 0; ax < 2 && !rc; ++ax)
            rc = launch_batch_bp_sep(cls, ax, kind, b->d_sprobs, sxrow(b, sub, cls, ax), b->nprob,
                                     sub, b->nsub,
                                     b->stiles[((size_t)sub * 4 + cls - 1) * 2 + ax],
                                     b->max_nv[cls], lambda, lambda_dev, st);
    return rc;
}

// Which problems of a several-problem plan take the separable tiles, and
// their tables, correction volumes and offsets (synchronises the stream).
int sepb_setup(ct2_batch* b, const ct2_desc* descs, hipStream_t st) {
    const char* env = getenv("CT2_PLAN_SEP");
    if (env && env[0] == '0') return CT2_OK;
    const int np = b->nprob, ns1 = b->nsub;
    b->psep.assign(np, 0);
    std::vector<char> cand(np, 0);
    bool any = false;
    for (int p = 0; p < np; ++p) {
        const BatchProb& P = b->h[p];
        cand[p] = sep_route(&descs[p], P.g) && subset_views(P.dt.ns, 0, ns1) <= 512;
        any |= cand[p] != 0;
    }
    if (!any) return CT2_OK;
    // every view of a candidate must be separable
    CT2_HIP(hipMalloc(&b->d_nonsep, sizeof(int) * (np + 1)));
    int rc = CT2_OK;
    {
        std::vector<int> zero(np + 1, 0);
        CT2_HIP(hipMemcpyAsync(b->d_nonsep, zero.data(), sizeof(int) * (np + 1),
                               hipMemcpyHostToDevice, st));
        for (int p = 0; p < np && !rc; ++p)
            if (cand[p]) rc = launch_count_nonsep(b->h[p].g, b->h[p].dt, b->h[p].fv, b->d_nonsep + p, st);
        if (rc) return rc;
        std::vector<int> cnt(np + 1, 0);
        CT2_HIP(hipMemcpyAsync(cnt.data(), b->d_nonsep, sizeof(int) * (np + 1),
                               hipMemcpyDeviceToHost, st));
        CT2_HIP(hipStreamSynchronize(st));
        for (int p = 0; p < np; ++p) b->psep[p] = cand[p] && cnt[p] == 0;
    }
    // memory: tables for each problem's largest subset (subset 0), and a
    // correction volume
    size_t total = 0;
    std::vector<size_t> tab_off(np, 0), corr_off(np, 0);
    for (int p = 0; p < np; ++p) {
        if (!b->psep[p]) continue;
        const BatchProb& P = b->h[p];
        b->sepb = true;
        tab_off[p] = total;
        total += (size_t)subset_views(P.dt.ns, 0, ns1) * P.dt.nu * (P.dt.nv + 1) * 16;
    }
    if (!b->sepb) return CT2_OK;
    for (int p = 0; p < np; ++p) {
        if (!b->psep[p]) continue;
        corr_off[p] = total;
        total += ((size_t)b->h[p].nvox * sizeof(float) + 255) & ~(size_t)255;
    }
    CT2_HIP(hipMalloc(&b->sep_mem, total));
    b->bytes += total;
    b->brick_any = false;
    std::vector<SepBatchProb> sp(np);
    char* base = static_cast<char*>(b->sep_mem);
    for (int p = 0; p < np; ++p) {
        memset(&sp[p], 0, sizeof(SepBatchProb));
        const BatchProb& P = b->h[p];
        if (!b->psep[p]) { b->brick_any = true; continue; }
        sp[p].g = P.g;
        sp[p].dt = P.dt;
        sp[p].nvox = P.nvox;
        sp[p].fv = P.fv;
        sp[p].ind = P.ind;
        sp[p].resid = P.resid;
        sp[p].tabs = reinterpret_cast<double2*>(base + tab_off[p]);
        sp[p].corr = reinterpret_cast<float*>(base + corr_off[p]);
        sp[p].vol = P.vol;
        sp[p].colsum = P.colsum;
        const int cls = sep_cpl(P.g);
        b->max_nv[cls] = std::max(b->max_nv[cls], P.dt.nv);
    }
    // offsets: table items per subset; tiles per (subset, class, axis)
    const size_t n1 = (size_t)np + 1;
    std::vector<long long> toff((size_t)ns1 * n1), xoff((size_t)ns1 * 8 * n1);
    b->stab_items.assign(ns1, 0);
    b->stiles.assign((size_t)ns1 * 8, 0);
    for (int s = 0; s < ns1; ++s) {
        long long acc = 0;
        for (int p = 0; p < np; ++p) {
            toff[(size_t)s * n1 + p] = acc;
            if (b->psep[p])
                acc += (long long)subset_views(b->h[p].dt.ns, s, ns1) * b->h[p].dt.nu;
        }
        toff[(size_t)s * n1 + np] = acc;
        b->stab_items[s] = acc;
        for (int cls = 1; cls <= 4; ++cls)
            for (int ax = 0; ax < 2; ++ax) {
                const size_t row = ((size_t)s * 4 + cls - 1) * 2 + ax;
                long long t = 0;
                for (int p = 0; p < np; ++p) {
                    xoff[row * n1 + p] = t;
                    const BatchProb& P = b->h[p];
                    if (b->psep[p] && sep_cpl(P.g) == cls && subset_views(P.dt.ns, s, ns1) > 0)
                        t += sep_tiles(P.g, ax);
                }
                xoff[row * n1 + np] = t;
                b->stiles[row] = t;
            }
    }
    CT2_HIP(hipMalloc(&b->d_sprobs, sizeof(SepBatchProb) * np));
    CT2_HIP(hipMalloc(&b->d_stoff, sizeof(long long) * toff.size()));
    CT2_HIP(hipMalloc(&b->d_sxoff, sizeof(long long) * xoff.size()));
    if ((rc = upload(b->d_sprobs, sp.data(), sizeof(SepBatchProb) * np, st))) return rc;
    if ((rc = upload(b->d_stoff, toff.data(), sizeof(long long) * toff.size(), st))) return rc;
    return upload(b->d_sxoff, xoff.data(), sizeof(long long) * xoff.size(), st);
}

// Row sums over every view and the column sums of each subset, by the
// separable kernels (resid serves as the ones sinogram).
int sep_sums(ct2_batch* b, const float* ones, size_t nsino, hipStream_t st) {
    const BatchProb& P = b->h[0];
    ct2_desc all = b->sd;
    all.view0 = 0;
    all.nviews = 0;
    all.vstride = 1;
    int rc = launch_project(&all, P.rowsum, P.ind, P.fv, ones, st);
    if (!rc) rc = launch_fill(P.resid, 1.0f, nsino, st);
    SepWork w;
    w.ptr = b->work;
    w.bytes = b->work_bytes;
    for (int s = 0; s < b->nsub && !rc; ++s) {
        if (subset_views(P.dt.ns, s, b->nsub) == 0) continue;
        const ct2_desc d = subset_desc(b, s);
        rc = launch_backproject(&d, P.colsum + (size_t)s * P.nvox, P.resid, P.ind, P.fv, st, w);
    }
    return rc;
}


int fwd(ct2_batch* b, int mode, int sub, const float* ones, hipStream_t st) {
    const bool all = mode != BFWD_RESID;
    return launch_batch_project(b->interp, b->arith, b->tu, mode, b->d_probs,
                                frow(b, all ? b->nsub : sub), b->nprob, all ? 0 : sub,
                                all ? 1 : b->nsub, ones, b->d_queue, b->fgrid, st);
}

int bp(ct2_batch* b, int mode, int sub, float lambda, hipStream_t st,
       const float* lambda_dev = nullptr) {
    const bool all = mode == BBP_PLAIN;
    return launch_batch_backproject(b->interp, b->arith, mode, b->d_probs,
                                    brow(b, all ? b->nsub : sub), b->nprob, all ? 0 : sub,
                                    all ? 1 : b->nsub, b->bdim, lambda, lambda_dev,
                                    b->d_queue + 2, b->bgrid, st);
}

// time one call of fn on the stream, in ms
template <typename F>
int time_ms(hipStream_t st, F fn, float* ms) {
    hipEvent_t e0, e1;
    CT2_HIP(hipEventCreate(&e0));
    CT2_HIP(hipEventCreate(&e1));
    hipEventRecord(e0, st);
    int rc = fn();
    hipEventRecord(e1, st);
    hipError_t he = hipEventSynchronize(e1);
    if (!rc && he == hipSuccess) hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    CT2_HIP(he);
    return CT2_OK;
}

int create(ct2_batch* b, const ct2_desc* descs, const float* const* fv, const int8_t* const* ind,
           const float* const* sino, float* const* vol, float* const* proj,
           float* const* adj, int tune, hipStream_t st) {
    int rc;
    CT2_HIP(hipGetDevice(&b->device));
    b->interp = descs[0].interp;
    b->arith = descs[0].arith;
    size_t nsino_tot = 0, nvox_tot = 0, nvox_max = 0;
    b->h.resize(b->nprob);
    b->has_proj = proj != nullptr;
    b->has_adj = adj != nullptr;
    for (int p = 0; p < b->nprob; ++p) {
        const ct2_desc* d = &descs[p];
        BatchProb& P = b->h[p];
        memset(&P, 0, sizeof P);
        P.g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
        P.dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
        if ((rc = check_grid(P.g)) || (rc = check_det(P.dt))) return rc;
        if (d->interp != b->interp || d->arith != b->arith) {
            set_error("ct2_batch_create: problem %d differs in interpolation or arithmetic", p);
            return CT2_EINVAL;
        }
        if (d->vol_dtype != CT2_F32) {
            set_error("ct2_batch_create: problem %d: volumes are float32", p);
            return CT2_EINVAL;
        }
        if (!fv || !fv[p] || !sino || !sino[p] || !vol || !vol[p]) {
            set_error("ct2_batch_create: problem %d: null frame vectors, sinogram or volume", p);
            return CT2_EINVAL;
        }
        P.zonly = d->axis == CT2_AXIS_Z;
        P.fv = fv[p];
        P.ind = ind ? ind[p] : nullptr;
        P.sino = sino[p];
        P.vol = vol[p];
        P.proj = proj ? proj[p] : nullptr;
        P.adj = adj ? adj[p] : nullptr;
        b->has_proj &= P.proj != nullptr;
        b->has_adj &= P.adj != nullptr;
        P.nvox = (long long)d->nx * d->ny * d->nz;
        nsino_tot += (size_t)d->ns * d->nu * d->nv;
        nvox_tot += (size_t)P.nvox;
        nvox_max = std::max(nvox_max, (size_t)P.nvox);
    }
    const size_t n1 = (size_t)(b->nsub + 1) * (b->nprob + 1);
    CT2_HIP(hipMalloc(&b->d_probs, sizeof(BatchProb) * b->nprob));
    CT2_HIP(hipMalloc(&b->d_foff, sizeof(long long) * n1));
    CT2_HIP(hipMalloc(&b->d_boff, sizeof(long long) * n1));
    CT2_HIP(hipMalloc(&b->rowsum, sizeof(float) * nsino_tot));
    CT2_HIP(hipMalloc(&b->resid, sizeof(float) * nsino_tot));
    CT2_HIP(hipMalloc(&b->colsum, sizeof(float) * nvox_tot * b->nsub));
    CT2_HIP(hipMalloc(&b->d_queue, 4 * sizeof(unsigned long long)));
    {
        const unsigned long long zero[4] = {0, 0, 0, 0};
        if ((rc = upload(b->d_queue, zero, sizeof zero, st))) return rc;
    }
    b->bytes = sizeof(float) * (2 * nsino_tot + nvox_tot * b->nsub);
    if (b->nprob == 1) {
        const char* env = getenv("CT2_PLAN_SEP");
        b->sd = descs[0];
        const ct2_desc d0 = subset_desc(b, 0);
        b->sep = !(env && env[0] == '0') && project_sep_route(&d0, vol[0]) &&
                 sep_route(&d0, b->h[0].g);
        if (b->sep) {
            b->work_bytes = sep_work_bytes(&d0);   // subset 0 has the most views
            CT2_HIP(hipMalloc(&b->corr, sizeof(float) * nvox_tot));
            CT2_HIP(hipMalloc(&b->work, b->work_bytes));
            b->bytes += sizeof(float) * nvox_tot + b->work_bytes;
        }
    }
    size_t so = 0, vo = 0;
    for (int p = 0; p < b->nprob; ++p) {
        BatchProb& P = b->h[p];
        P.rowsum = b->rowsum + so;
        P.resid = b->resid + so;
        P.colsum = b->colsum + vo * b->nsub;
        so += (size_t)P.dt.ns * P.dt.nu * P.dt.nv;
        vo += (size_t)P.nvox;
    }
    if (!b->sep && (rc = sepb_setup(b, descs, st))) return rc;
    float* ones = nullptr;
    CT2_HIP(hipMalloc(&ones, sizeof(float) * nvox_max));
    rc = launch_fill(ones, 1.0f, nvox_max, st);
    if (!rc) rc = set_bp(b, b->bdim, st);
    if (!rc) rc = set_fwd(b, 4, st);
    const char* env = getenv("CT2_AUTOTUNE");
    // (a one-problem separable plan's SART uses neither the tile nor the
    // brick; with every problem on the separable tiles, SART uses no brick)
    if (!rc && tune && !(env && env[0] == '0') && !(b->sep && !b->has_proj && !b->has_adj)) {
        // tile height of the forward: time the all-views row-sum projection
        float best = 1e30f;
        int best_tu = 4;
        for (int tu : {2, 4, 8}) {
            if ((rc = set_fwd(b, tu, st))) break;
            float ms = 0;
            if ((rc = time_ms(st, [&] { return fwd(b, BFWD_ROWSUM, 0, ones, st); }, &ms))) break;
            if (ms < best) { best = ms; best_tu = tu; }
        }
        if (!rc) rc = set_fwd(b, best_tu, st);
        // brick of the adjoint: time the column sums of subset 0
        if (!rc) rc = launch_fill(b->resid, 1.0f, nsino_tot, st);
        best = 1e30f;
        int3 best_b = b->bdim;
        for (const auto& c : kBrickCandidates) {
            if (rc || (!b->brick_any && !b->has_adj)) break;
            int3 bd = make_int3(c[0], c[1], c[2]);
            if ((rc = set_bp(b, bd, st))) break;
            float ms = 0;
            if ((rc = time_ms(st, [&] { return bp(b, BBP_COLSUM, 0, 0.0f, st); }, &ms))) break;
            if (ms < best) { best = ms; best_b = bd; }
        }
        if (!rc) rc = set_bp(b, best_b, st);
    }
    // row sums over every view, then the column sums of each subset
    if (b->sep) {
        if (!rc) rc = sep_sums(b, ones, nsino_tot, st);
    } else {
        if (!rc) rc = fwd(b, BFWD_ROWSUM, 0, ones, st);
        if (!rc) rc = launch_fill(b->resid, 1.0f, nsino_tot, st);
        for (int s = 0; s < b->nsub && !rc; ++s) {
            if (b->brick_any) rc = bp(b, BBP_COLSUM, s, 0.0f, st);
            if (!rc && b->sepb) rc = sepb_subset(b, BSEP_COLSUM, s, 0.0f, nullptr, st);
        }
    }
    hipError_t he = hipStreamSynchronize(st);
    hipFree(ones);
    if (rc) return rc;
    CT2_HIP(he);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace

extern "C" {

int ct2_batch_create(ct2_batch** plan, int nprob, const ct2_desc* descs,
                     const float* const* frame_vectors, const int8_t* const* indicators,
                     const float* const* sinos, float* const* vols, float* const* proj_out,
                     float* const* adj_out, int nsub, int tune, ct2_stream_t stream) {
    if (!plan) { set_error("ct2_batch_create: null plan pointer"); return CT2_EINVAL; }
    *plan = nullptr;
    if (nprob <= 0 || !descs) { set_error("ct2_batch_create: no problems"); return CT2_EINVAL; }
    if (nsub < 1) { set_error("ct2_batch_create: nsub must be >= 1"); return CT2_EINVAL; }
    ct2_batch* b = new ct2_batch;
    b->nprob = nprob;
    b->nsub = nsub;
    int rc = create(b, descs, frame_vectors, indicators, sinos, vols, proj_out, adj_out, tune,
                    (hipStream_t)stream);
    if (rc) {
        release(b);
        return rc;
    }
    *plan = b;
    return CT2_OK;
}

int ct2_batch_destroy(ct2_batch* plan) {
    release(plan);
    return CT2_OK;
}

int ct2_batch_info(const ct2_batch* plan, int* info, size_t* bytes) {
    if (!plan) { set_error("ct2_batch_info: null plan"); return CT2_EINVAL; }
    if (info) {
        info[0] = plan->tu;
        info[1] = plan->fgrid;
        info[2] = plan->bdim.x;
        info[3] = plan->bdim.y;
        info[4] = plan->bdim.z;
        info[5] = plan->bgrid;
        info[6] = plan->nsub;
        info[7] = plan->nprob;
    }
    if (bytes) *bytes = plan->bytes;
    return CT2_OK;
}

int ct2_batch_project(ct2_batch* plan, ct2_stream_t stream) {
    if (!plan) { set_error("ct2_batch_project: null plan"); return CT2_EINVAL; }
    if (!plan->has_proj) { set_error("ct2_batch_project: plan has no projection outputs"); return CT2_EINVAL; }
    int rc = fwd(plan, BFWD_PLAIN, 0, nullptr, (hipStream_t)stream);
    if (rc) return rc;
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_batch_backproject(ct2_batch* plan, ct2_stream_t stream) {
    if (!plan) { set_error("ct2_batch_backproject: null plan"); return CT2_EINVAL; }
    if (!plan->has_adj) { set_error("ct2_batch_backproject: plan has no adjoint outputs"); return CT2_EINVAL; }
    int rc = bp(plan, BBP_PLAIN, 0, 0.0f, (hipStream_t)stream);
    if (rc) return rc;
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

static int batch_sart(const char* who, ct2_batch* plan, float lambda, const float* lambda_dev,
                      int niter, ct2_stream_t stream) {
    if (!plan) { set_error("%s: null plan", who); return CT2_EINVAL; }
    if (niter < 0) { set_error("%s: niter < 0", who); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    for (int it = 0; it < niter; ++it)
        for (int s = 0; s < plan->nsub; ++s) {
            int rc;
            if (plan->sep) {
                rc = sep_sart_subset(plan, s, lambda, lambda_dev, st);
            } else {
                rc = fwd(plan, BFWD_RESID, s, nullptr, st);
                if (!rc && plan->brick_any) rc = bp(plan, BBP_SART, s, lambda, st, lambda_dev);
                if (!rc && plan->sepb)
                    rc = sepb_subset(plan, BSEP_SART, s, lambda, lambda_dev, st);
            }
            if (rc) return rc;
        }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int ct2_batch_sart(ct2_batch* plan, float lambda, int niter, ct2_stream_t stream) {
    return batch_sart("ct2_batch_sart", plan, lambda, nullptr, niter, stream);
}

int ct2_batch_sart_dev(ct2_batch* plan, const float* lambda, int niter, ct2_stream_t stream) {
    if (!lambda) { set_error("ct2_batch_sart_dev: null relaxation"); return CT2_EINVAL; }
    return batch_sart("ct2_batch_sart_dev", plan, 0.0f, lambda, niter, stream);
}

}  // extern "C"
