/*
 * ct2.h -- C ABI of libct2.so, the MI355X (gfx950) implementation of ct-2's
 * reconstruction hot path.
 *
 * The reference reaches its Fortran kernels through f2py extension modules
 * (emil_code/tomo3D.py:10-14, dispatch tables tomo3D.py:55-103).  Each
 * entry point below replaces one f2py-wrapped subroutine and keeps its
 * argument list and order (outputs first, as in the Fortran), with three
 * differences:
 *   - every array is a DEVICE pointer (hipMalloc'd or a torch tensor's
 *     data_ptr), C-ordered over the reference's logical indices:
 *       volume   smat[nx][ny][nz]   element (i,j,k) at (i*ny+j)*nz+k
 *       sinogram sinomat[ns][nu][nv]
 *       indicator indsino[ns][nu][nv] as int8 (a ray is used iff == 1);
 *                NULL means every ray is used
 *       frame_vectors[ns][18], ray_vectors[nr][6] float32;
 *   - outputs that f2py would return are written in place (zero-filled by
 *     the library where f2py zero-fills an intent(out) array);
 *   - a trailing `stream` (a hipStream_t, NULL = default stream); every call
 *     is asynchronous on that stream and allocates nothing unless stated.
 * Every function returns CT2_OK (0) or a negative status; ct2_last_error()
 * then describes the failure (thread-local).
 */
#ifndef CT2_H
#define CT2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* ct2_stream_t;

enum { CT2_OK = 0, CT2_EINVAL = -1, CT2_EHIP = -2, CT2_ENOMEM = -3 };
enum { CT2_LINEAR = 0, CT2_NEAREST = 1 };
enum { CT2_AXIS_FREE = 0, CT2_AXIS_Z = 1 };
enum { CT2_F32 = 0, CT2_F16 = 1, CT2_BF16 = 2, CT2_F16Z2 = 3 };
/* Plane-major volume storage (forward projection only): element (i,j,k) at
 * (k*nx + i)*ny + j, i.e. the volume as nz planes of C-ordered (x, y); made
 * by ct2_convert_volume_zplanes. */
enum { CT2_F32_ZPLANES = 4, CT2_F16_ZPLANES = 5, CT2_BF16_ZPLANES = 6 };
/* CT2_ARITH_EXACT: the reference's FP32 operation order, correctly rounded
 * divisions, no FMA contraction -- bit-identical to the reference Fortran
 * for the forward projector.  CT2_ARITH_FAST: per-ray setup in FP64, FMA
 * plane stepping in FP32, compensated (Kahan) ray sums. */
enum { CT2_ARITH_EXACT = 0, CT2_ARITH_FAST = 1 };
enum { CT2_ROUND_NEAREST = 0, CT2_ROUND_STOCHASTIC = 1 };

/* Operator descriptor of the extended API. */
typedef struct ct2_desc {
    int ns, nu, nv;                    /* views, detector columns (u), rows (v) */
    float du, dv, u0, v0;              /* bin pitch and left edge of the detector */
    int nx, ny, nz;                    /* volume shape */
    float dx, dy, dz, x0, y0, z0;      /* voxel size and lower volume corner */
    int interp;                        /* CT2_LINEAR | CT2_NEAREST */
    int axis;                          /* CT2_AXIS_FREE | CT2_AXIS_Z */
    int vol_dtype;                     /* storage type of the volume operand */
    int arith;                         /* CT2_ARITH_EXACT | CT2_ARITH_FAST */
    int view0, nviews;                 /* view subset: views view0 + j*vstride, j < nviews; */
                                       /* nviews 0 = every such view up to ns-1          */
    int brick[3];                      /* adjoint voxel brick (0 = autotuned/default); */
                                       /* naming one also keeps the fast linear f32    */
                                       /* adjoint on the brick kernel (exact transpose */
                                       /* of the fast forward's FP32 weights)          */
    int vstride;                       /* view step of the subset (0 or 1 = contiguous) */
} ct2_desc;

/* ---------------------------------------------------------------- misc -- */
int ct2_version(void);
const char* ct2_last_error(void);
/* Fills *desc with defaults for the given geometry. */
int ct2_desc_init(ct2_desc* desc, int ns, int nu, int nv, float du, float dv, float u0,
                  float v0, int nx, int ny, int nz, float dx, float dy, float dz, float x0,
                  float y0, float z0);

/* ----------------------------------------------------- extended API ---- */
/* Joseph forward projection of vol (storage desc->vol_dtype) into the
 * float32 sinogram; rays with indicator != 1 are set to 0. */
int ct2_project(const ct2_desc* desc, float* sino, const int8_t* ind, const float* fv,
                const void* vol, ct2_stream_t stream);
/* Exact adjoint of ct2_project (same matrix entries), written over vol
 * (storage desc->vol_dtype).  Two routes build tables in device memory: the
 * column-separable route (fast linear arithmetic into float32, views whose
 * detector rows run along z) per-(view, column) prefix sums, and the
 * plane-separable route (the same arithmetic with desc->axis CT2_AXIS_Z,
 * views whose detector is parallel to the z planes with e_u along x and e_v
 * along y) per-view set-ups and trav*datum of every ray.  This entry
 * allocates them on the stream (hipMallocAsync, at most 4 GiB, freed on the
 * stream; the device's default stream-ordered pool is set to keep up to
 * 4 GiB of freed memory for the next call).  ct2_backproject_ws takes them
 * from the caller instead. */
int ct2_backproject(const ct2_desc* desc, void* vol, const float* sino, const int8_t* ind,
                    const float* fv, ct2_stream_t stream);
/* ct2_backproject with a caller-owned device workspace (16-byte aligned) for
 * the separable routes.  Column-separable: a 256-byte header and the column
 * tables, work_bytes >= 256 + one view's tables (nu * (nv+1) * 16 bytes).
 * Plane-separable: a 256-byte header and one record per view of
 * 4 * (r4(4 nz + 2 nu + 2 nv) + nu nv) bytes rounded up to 16 (r4: up to a
 * multiple of 4), work_bytes >= 256 + one record.  ct2_backproject_work_bytes(desc) bytes hold every view of the
 * subset, and a smaller workspace processes the views in groups.  The call
 * allocates nothing; it must not overlap another call that uses the same
 * workspace.  work == NULL behaves as
 * ct2_backproject.  No counterpart in the reference (view_proj3D.f:1096-1370
 * has no workspace); for callers that keep a workspace across calls (graph
 * capture, iteration loops). */
int ct2_backproject_ws(const ct2_desc* desc, void* vol, const float* sino, const int8_t* ind,
                       const float* fv, void* work, size_t work_bytes, ct2_stream_t stream);
/* Bytes of ct2_backproject_ws's workspace for every view of desc's subset; 0
 * when neither separable route applies (then no workspace is used). */
size_t ct2_backproject_work_bytes(const ct2_desc* desc);
/* float32 volume -> 16-bit storage (dtype CT2_F16 or CT2_BF16) for the
 * mixed-precision forward projection.  CT2_ROUND_STOCHASTIC picks the upper
 * or lower 16-bit neighbour with probability proportional to proximity,
 * driven by a counter-based hash of (index, seed), so the stored volume is
 * unbiased and deterministic.  CT2_F16Z2 stores f16 pairs: element i (4
 * bytes) holds the f16 values of elements i and i+1 (zero past the end),
 * rounded as CT2_F16 rounds them, so that the fast forward projector reads a
 * voxel and its neighbour along the contiguous axis with one aligned load;
 * ct2_project takes it with desc->vol_dtype = CT2_F16Z2 and returns the
 * same projections as from the CT2_F16 volume.  No counterpart in the
 * reference (all REAL*4). */
int ct2_convert_volume(const float* src, void* dst, int dtype, size_t n, int rounding,
                       unsigned seed, ct2_stream_t stream);
/* float32 (nx, ny, nz) volume -> plane-major storage (dtype CT2_F32_ZPLANES,
 * CT2_F16_ZPLANES or CT2_BF16_ZPLANES: element (i,j,k) at (k*nx + i)*ny + j),
 * 16-bit values rounded exactly as ct2_convert_volume rounds the same
 * element.  ct2_project reads it with desc->vol_dtype set to that dtype and
 * returns the same projections as from the C-ordered volume of the same
 * values.  For z-driven projection (tomosynthesis), where a wavefront's rays
 * sample neighbouring (x, y) of one z plane, the plane-major copy turns their
 * gathers into contiguous reads.  No counterpart in the reference. */
int ct2_convert_volume_zplanes(const float* src, void* dst, int dtype, int nx, int ny, int nz,
                               int rounding, unsigned seed, ct2_stream_t stream);
/* SART update over the view subset of desc (views view0 + j*vstride; no
 * counterpart in the reference, which has only sequential ART):
 * x += lambda * A^T((b-Ax)/A1) / A^T1 restricted to those views, with
 * work = 2*ns*nu*nv + 3*nx*ny*nz floats of device scratch supplied by the
 * caller. */
int ct2_sart_step(const ct2_desc* desc, float lambda, const float* sino, const int8_t* ind,
                  const float* fv, float* vol, float* work, ct2_stream_t stream);

/* ------------------------------------------------------ batched plans -- */
/* Many independent problems of uneven size in one persistent launch.  No
 * reference counterpart (the reference runs one problem per call); the
 * per-problem results equal those of ct2_project / ct2_backproject /
 * ct2_sart_step.  A plan binds, per problem p: descriptor descs[p] (all
 * problems share interp and arith; vol_dtype float32), device frame vectors,
 * indicator (the array or its entries may be NULL), measured sinogram
 * sinos[p], volume vols[p] (read by the forward, updated by SART), and
 * optionally outputs proj_out[p] (forward of vols[p]) and adj_out[p]
 * (adjoint of sinos[p]); the arrays of pointers are read at creation only.
 * SART uses nsub ordered subsets of interleaved views (subset s = views s,
 * s+nsub, ...; the same update as ct2_sart_step with view0 = s, vstride =
 * nsub).  Creation allocates the row sums, residuals and nsub column-sum
 * volumes, and with tune != 0 (unless CT2_AUTOTUNE=0) times forward tile
 * heights {2,4,8} and adjoint bricks to pick the fastest; it synchronises
 * the stream.  The other entry points only launch kernels on the stream
 * (graph-capturable).  Calls on one plan must not overlap: the plan's
 * kernels hand out work items from a counter in the plan, which each launch
 * leaves at zero for the next one, so a plan is used on one stream at a
 * time (and a launch that fails part-way leaves the plan unusable).
 * A plan of one problem whose forward and adjoint both take the
 * column-separable kernels (fast linear arithmetic, free axis, rows along
 * z, nz % 4 == 0, nz <= 256, 16-byte aligned volume; CT2_PLAN_SEP=0 turns
 * this off) runs SART through them instead, with a correction volume and
 * the column tables of one subset allocated at creation: each subset step
 * then equals ct2_sart_step bit for bit, and still allocates nothing.  In a
 * plan of several problems, SART's adjoint runs on batched separable tiles
 * for every problem whose views all have their rows along z (fast linear
 * arithmetic, free axis, nz <= 256, at most 512 views per subset; creation
 * counts the views on the device): those problems' steps equal
 * ct2_sart_step bit for bit too, and creation allocates their column tables
 * (largest subset) and a correction volume each.  CT2_PLAN_SEP=0 keeps the
 * brick kernel for all. */
typedef struct ct2_batch ct2_batch;
int ct2_batch_create(ct2_batch** plan, int nprob, const ct2_desc* descs,
                     const float* const* frame_vectors, const int8_t* const* indicators,
                     const float* const* sinos, float* const* vols, float* const* proj_out,
                     float* const* adj_out, int nsub, int tune, ct2_stream_t stream);
int ct2_batch_destroy(ct2_batch* plan);
/* info[8] = {forward tile height, forward grid, brick x, y, z, adjoint grid,
 * nsub, nprob}; *bytes = device memory held by the plan. */
int ct2_batch_info(const ct2_batch* plan, int* info, size_t* bytes);
int ct2_batch_project(ct2_batch* plan, ct2_stream_t stream);
int ct2_batch_backproject(ct2_batch* plan, ct2_stream_t stream);
/* niter SART iterations (each one pass over the nsub subsets) on every
 * problem's volume in place. */
int ct2_batch_sart(ct2_batch* plan, float lambda, int niter, ct2_stream_t stream);
/* ct2_batch_sart with the relaxation read from device memory (one float)
 * when the kernels run, so a captured graph of the iteration can change it
 * between replays (the ASD-POCS driver's reg_fac schedule,
 * examples/jawPhantomCCB.py:325).  No counterpart in the reference. */
int ct2_batch_sart_dev(ct2_batch* plan, const float* lambda, int niter, ct2_stream_t stream);

/* ------------------------------------------ view_proj3D.f replacements -- */
/* view_proj3D.f:703 sinoproj_linear */
int ct2_sinoproj_linear(float* sinomat, const int8_t* indsino, const float* frame_vectors,
                        int ns, int nu, int nv, float du, float dv, float u0, float v0,
                        const float* smat, float dx, float dy, float dz, float x0, float y0,
                        float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D.f:915 sinoproj_nearest */
int ct2_sinoproj_nearest(float* sinomat, const int8_t* indsino, const float* frame_vectors,
                         int ns, int nu, int nv, float du, float dv, float u0, float v0,
                         const float* smat, float dx, float dy, float dz, float x0, float y0,
                         float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D.f:1096 sinobackproject_linear */
int ct2_sinobackproject_linear(const float* sinomat, const int8_t* indsino,
                               const float* frame_vectors, int ns, int nu, int nv, float du,
                               float dv, float u0, float v0, float* smat, float dx, float dy,
                               float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                               ct2_stream_t stream);
/* view_proj3D.f:1374 sinobackproject_nearest */
int ct2_sinobackproject_nearest(const float* sinomat, const int8_t* indsino,
                                const float* frame_vectors, int ns, int nu, int nv, float du,
                                float dv, float u0, float v0, float* smat, float dx, float dy,
                                float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                                ct2_stream_t stream);
/* view_proj3D.f:1769 pd_backproj_linear (voxel driven, not an adjoint) */
int ct2_pd_backproj_linear(const float* sinomat, const int8_t* indsino,
                           const float* frame_vectors, int ns, int nu, int nv, float du,
                           float dv, float u0, float v0, float* smat, float dx, float dy,
                           float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                           ct2_stream_t stream);
/* view_proj3D.f:1598 pd_backproj_nearest */
int ct2_pd_backproj_nearest(const float* sinomat, const int8_t* indsino,
                            const float* frame_vectors, int ns, int nu, int nv, float du,
                            float dv, float u0, float v0, float* smat, float dx, float dy,
                            float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                            ct2_stream_t stream);
/* view_proj3D.f:1950 artstep_linear (sequential Kaczmarz; smat in/out) */
int ct2_artstep_linear(float reg_fac, const float* sinomat, const int8_t* indsino,
                       const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                       float u0, float v0, float* smat, float dx, float dy, float dz, float x0,
                       float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D.f:2275 artstep_nearest */
int ct2_artstep_nearest(float reg_fac, const float* sinomat, const int8_t* indsino,
                        const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                        float u0, float v0, float* smat, float dx, float dy, float dz, float x0,
                        float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D.f:2507 emstep_linear (MLEM; smat in/out).  Allocates two
 * volume-sized scratch arrays stream-ordered (hipMallocAsync). */
int ct2_emstep_linear(const float* sinomat, const int8_t* indsino, const float* frame_vectors,
                      int ns, int nu, int nv, float du, float dv, float u0, float v0,
                      float* smat, float dx, float dy, float dz, float x0, float y0, float z0,
                      int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D.f:2834 emstep_nearest */
int ct2_emstep_nearest(const float* sinomat, const int8_t* indsino, const float* frame_vectors,
                       int ns, int nu, int nv, float du, float dv, float u0, float v0,
                       float* smat, float dx, float dy, float dz, float x0, float y0, float z0,
                       int nx, int ny, int nz, ct2_stream_t stream);
/* MLEM with caller-supplied scratch work[2*nx*ny*nz] and an optional cached
 * sensitivity image sens = A^T 1 (NULL: computed into the scratch). */
int ct2_em_step(const ct2_desc* desc, const float* sino, const int8_t* ind, const float* fv,
                float* vol, float* work, const float* sens, ct2_stream_t stream);
/* ART over a view-geometry with descriptor options (interp, axis). */
int ct2_art_step(const ct2_desc* desc, float reg_fac, const float* sino, const int8_t* ind,
                 const float* fv, float* vol, ct2_stream_t stream);

/* ------------------------------------ view_proj3D_tomo.f replacements -- */
/* view_proj3D_tomo.f:4 sinoprojz (driving axis forced to z) */
int ct2_sinoprojz(float* sinomat, const int8_t* indsino, const float* frame_vectors, int ns,
                  int nu, int nv, float du, float dv, float u0, float v0, const float* smat,
                  float dx, float dy, float dz, float x0, float y0, float z0, int nx, int ny,
                  int nz, ct2_stream_t stream);
/* view_proj3D_tomo.f:154 artstepz */
int ct2_artstepz(float reg_fac, const float* sinomat, const int8_t* indsino,
                 const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                 float u0, float v0, float* smat, float dx, float dy, float dz, float x0,
                 float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* view_proj3D_tomo.f:352 artstepz2vol (main + auxiliary grid, both in/out) */
int ct2_artstepz2vol(float reg_fac, const float* sinomat, const int8_t* indsino,
                     const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                     float u0, float v0, float* smat, float dx, float dy, float dz, float x0,
                     float y0, float z0, int nx, int ny, int nz, float* smata, float dxa,
                     float dya, float dza, float x0a, float y0a, float z0a, int nxa, int nya,
                     int nza, ct2_stream_t stream);
/* view_proj3D_tomo.f:615 emstep_linearz */
int ct2_emstep_linearz(const float* sinomat, const int8_t* indsino, const float* frame_vectors,
                       int ns, int nu, int nv, float du, float dv, float u0, float v0,
                       float* smat, float dx, float dy, float dz, float x0, float y0, float z0,
                       int nx, int ny, int nz, ct2_stream_t stream);

/* ------------------------------------------- ray_proj3D.f replacements -- */
/* ray_proj3D.f:514 sinoproj_linear (ray list: ray_vectors[nr][6] = source, bin) */
int ct2_ray_sinoproj_linear(float* sinomat, const int8_t* indsino, const float* ray_vectors,
                            int nr, const float* smat, float dx, float dy, float dz, float x0,
                            float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* ray_proj3D.f:677 sinoproj_nearest */
int ct2_ray_sinoproj_nearest(float* sinomat, const int8_t* indsino, const float* ray_vectors,
                             int nr, const float* smat, float dx, float dy, float dz, float x0,
                             float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);
/* ray_proj3D.f:809 sinobackproject_nearest (bounds 0 <= i < n-1, ray_proj3D.f:895) */
int ct2_ray_sinobackproject_nearest(const float* sinomat, const int8_t* indsino,
                                    const float* ray_vectors, int nr, float* smat, float dx,
                                    float dy, float dz, float x0, float y0, float z0, int nx,
                                    int ny, int nz, ct2_stream_t stream);
/* ray_proj3D.f:987 artstep_linear */
int ct2_ray_artstep_linear(float reg_fac, const float* sinomat, const int8_t* indsino,
                           const float* ray_vectors, int nr, float* smat, float dx, float dy,
                           float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                           ct2_stream_t stream);
/* ray_proj3D.f:1272 artstep_nearest (bounds < n-1) */
int ct2_ray_artstep_nearest(float reg_fac, const float* sinomat, const int8_t* indsino,
                            const float* ray_vectors, int nr, float* smat, float dx, float dy,
                            float dz, float x0, float y0, float z0, int nx, int ny, int nz,
                            ct2_stream_t stream);
/* ray_proj3D.f:1454 emstep_nearest (bounds < n-1); allocates like ct2_emstep_linear */
int ct2_ray_emstep_nearest(const float* sinomat, const int8_t* indsino, const float* ray_vectors,
                           int nr, float* smat, float dx, float dy, float dz, float x0,
                           float y0, float z0, int nx, int ny, int nz, ct2_stream_t stream);

/* ---------------------------------------------- totalvar.f replacements -- */
/* totalvar.f:35 total_var3d; eps is the reference's 1e-9 unless changed.
 * The value (float64) is written to the device pointer tv_out. */
int ct2_total_var3d(float ambdala, const float* smat, int nx, int ny, int nz, float eps,
                    double* tv_out, ct2_stream_t stream);
/* totalvar.f:476 total_var_grad3d; the gradient norm bobo (float32, with the
 * reference's |bobo| < 1e-5 -> 1 rule) is written to the device pointer bobo. */
int ct2_total_var_grad3d(float ambdala, float* tvgmat, const float* smat, int nx, int ny,
                         int nz, int normalizeq, float eps, float* bobo, ct2_stream_t stream);
/* Same two operations with a caller-owned device workspace of
 * CT2_TV_WORK_DOUBLES doubles instead of a stream-ordered allocation, so
 * that iteration loops can be captured into a graph.  No reference
 * counterpart: the reference allocates inside the f2py wrapper. */
#define CT2_TV_WORK_DOUBLES 1040
int ct2_tv_value(float ambdala, const float* smat, int nx, int ny, int nz, float eps,
                 double* tv_out, double* work, ct2_stream_t stream);
int ct2_tv_grad(float ambdala, float* tvgmat, const float* smat, int nx, int ny, int nz,
                int normalizeq, float eps, float* bobo, double* work, ct2_stream_t stream);
/* *out = sqrt(sum over q of (a[q]-b[q])^2), summed in FP64 in a fixed order,
 * over the q with mask[q] == 1 (mask NULL = all; b NULL = 0).  This is the
 * distance of sinogram3D.dist_to (tomo3D.py:319-324, indicator-masked) and
 * image3D.dist_to (tomo3D.py:946-951) on device arrays; work holds
 * CT2_TV_WORK_DOUBLES doubles. */
int ct2_distance(const float* a, const float* b, const int8_t* mask, size_t n, double* out,
                 double* work, ct2_stream_t stream);

/* The p-variants of totalvar.f.  Values are written as float64 to the device
 * pointer out (the reference's Real*4 function result, widened); gradients
 * write tvgmat and the float32 norm bobo; work holds CT2_TV_WORK_DOUBLES
 * doubles.  eps is the reference's 1e-9 unless changed. */
/* totalvar.f:4 gmi: smatout = sqrt(eps + |backward differences|^2) at
 * voxels with every index >= 1, 0 elsewhere */
int ct2_gmi(float* smatout, const float* smat, int nx, int ny, int nz, float eps,
            ct2_stream_t stream);
/* totalvar.f:65 calc_lp */
int ct2_calc_lp(float p, const float* smat, int nx, int ny, int nz, float eps, double* out,
                double* work, ct2_stream_t stream);
/* totalvar.f:91 total_var3d_p */
int ct2_total_var3d_p(float p, const float* smat, int nx, int ny, int nz, float eps,
                      double* out, double* work, ct2_stream_t stream);
/* totalvar.f:121 total_dvar3d_p */
int ct2_total_dvar3d_p(float p, const float* smat, int nx, int ny, int nz, float dx, float dy,
                       float dz, float eps, double* out, double* work, ct2_stream_t stream);
/* totalvar.f:150 lp_grad3d */
int ct2_lp_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                  int normalizeq, float eps, float* bobo, double* work, ct2_stream_t stream);
/* totalvar.f:219 total_p_var_grad3d */
int ct2_total_p_var_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                           int normalizeq, float eps, float* bobo, double* work,
                           ct2_stream_t stream);
/* totalvar.f:347 total_p_dvar_grad3d */
int ct2_total_p_dvar_grad3d(float p, float* tvgmat, const float* smat, int nx, int ny, int nz,
                            float dx, float dy, float dz, int normalizeq, float eps,
                            float* bobo, double* work, ct2_stream_t stream);

/* --------------------------------------------------- FDK.py replacements -- */
/* FDK.py:10 fdk_sino_weight3D: out[p,j,k] = F/sqrt((F+D)^2 + vu[j,k]^2 + vv[j,k]^2)
 * * sinomat[p,j,k] with vu[j,k] = u00+(j+uoff)*du, vv[j,k] = v00+(k+voff)*dv. */
int ct2_fdk_sino_weight3d(float radius, float center2detector, int ns, int nu, int nv,
                          float u00, float du, float uoff, float v00, float dv, float voff,
                          const float* sinomat, float* out, ct2_stream_t stream);
/* FDK.py:25 fdk_sino_filter3D, dim = 1 (filter along u) or 2 (along v):
 * zero-pad to npad = 2^round(log2(2n-1)), multiply the FFT by the
 * Kak-Slaney ramp (optionally Hann-windowed), crop.  Implemented as the
 * equivalent real circulant matrix applied to every detector line; its taps
 * are computed on the device (FP64 DFTs).  Allocates a workspace of
 * 20*npad bytes stream-ordered (hipMallocAsync / hipFreeAsync on the
 * stream); lines up to 32768 bins. */
int ct2_fdk_sino_filter3d(float dp, int dim, int ns, int nu, int nv, const float* sinomat,
                          float* out, int hann, float L, ct2_stream_t stream);
/* ct2_fdk_sino_filter3d without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  No
 * counterpart in the reference.  Same arguments and result; the taps are
 * built in a static device array that the library holds on every GPU it runs
 * on, so lines may pad to at most CT2_FDK_NOALLOC_MAX_NPAD samples (npad =
 * 2^round(log2(2n-1)): n <= 5792 bins); longer lines are rejected with
 * CT2_EINVAL.  Calls that overlap on one GPU share that array and race. */
#define CT2_FDK_NOALLOC_MAX_NPAD 8192
int ct2_fdk_sino_filter3d_noalloc(float dp, int dim, int ns, int nu, int nv,
                                  const float* sinomat, float* out, int hann, float L,
                                  ct2_stream_t stream);
/* FDK.py:130 fdk_backproj_linear3D over the voxel grid
 * x = x00+(i+xoff)dx, y = y00+(j+yoff)dy, z = z00+(k+zoff)dz, angles
 * vs[p] = s0 + p*ds; recon[i][j][k] (C order over x, y, z). */
int ct2_fdk_backproj_linear3d(int ns, int nu, int nv, float du, float dv, float u0, float v0,
                              float radius, float source2detector, float uoffset,
                              float voffset, double s0, double ds, int nx, int ny, int nz,
                              float x00, float y00, float z00, float dx, float dy, float dz,
                              float xoff, float yoff, float zoff, const float* sinomat,
                              float* recon, ct2_stream_t stream);

/* ----------------------------------------------------- analytic phantoms -- */
/* Exact line integrals through quadric objects and their voxel embeddings:
 * the data generators of emil_code/view_proj3D.f:11-603 (gen_shape_proj,
 * ellipsoid_proj), ray_proj3D.f:175-512 (the same over a ray list) and
 * tomo3DfortranAux.f:212-582 (ellipsoid_embed, gen_obj_embed), which
 * tomo3D.py:1507-2894 calls once per component and view.  One call here
 * evaluates a whole phantom (every component, every view) and, per bin or
 * voxel, adds the components in table order, the order of the reference's
 * Python loop (tomo3D.py:1562-1567).  The shape and surface tables are HOST
 * arrays; the library stages them in a stream-ordered device allocation. */
enum { CT2_SHAPE_ELLIPSOID = 0, CT2_SHAPE_GENERAL = 1 };
enum { CT2_SURF_PLANE = 0, CT2_SURF_ELLIPSOID = 1, CT2_SURF_CYLINDER = 2, CT2_SURF_CONE = 3 };
/* Kind flag for the voxel-value calls (ct2_phantom_embed, ct2_phantom_points)
 * only: the surface frame is given by angles, alpha, beta, gamma in xu[0..2]
 * (yu, zu unused), and a point is taken into it in the operation order of
 * framer (tomo3DfortranAux.f:281-308), which the reference's point tests
 * gen_obj_voxval (tomo3DfortranAux.f:311-416) and image3D.add_shape
 * (tomo3D.py:1073-1085, through gen_object.voxvalf, tomo3D.py:2122-2220) use. */
enum { CT2_SURF_ANGLES = 16 };

/* One bounding surface of a general object (tomo3D.py:1704-1832): the
 * surface frame's x, y and z unit vectors in space-fixed coordinates (the
 * reference's xu*, yu*, zu*) and its centre.  Plane: the half-space z >= 0 of
 * the frame.  Ellipsoid: semi-axes ax, ay, az.  Cylinder: elliptic, along
 * the frame z, semi-axes ax, ay.  Cone: the z >= 0 nappe of
 * (x/ax)^2 + (y/ay)^2 = z^2. */
typedef struct ct2_surface {
    int   kind;
    float ax, ay, az;
    float xu[3], yu[3], zu[3];
    float c[3];
} ct2_surface;

/* One additive component.
 * CT2_SHAPE_ELLIPSOID (tomo3D.py:2704-2894): semi-axes ax, ay, az, centre
 *   x0, y0, z0, rotated by beta about z and then by alpha about the new y.
 * CT2_SHAPE_GENERAL (tomo3D.py:1835-2701): the intersection of
 *   surfaces[first .. first+nsurf); they are visited planes, ellipsoids,
 *   cylinders, cones, each kind in table order, as gen_shape_proj does.
 * att scales the chord length (projection) or is the value inside (embedding). */
typedef struct ct2_shape {
    int   kind;
    float att;
    int   first, nsurf;
    float ax, ay, az;
    float x0, y0, z0;
    float alpha, beta;
} ct2_shape;

/* sinomat[ns][nu][nv] over the views of frame_vectors [ns][18]; rays with
 * indsino != 1 contribute 0 (indsino NULL: every ray is used).  accumulate = 0: the sum of the components (one
 * component gives the f2py routine's output, view_proj3D.f:175-455 or
 * :483-603); accumulate = 1: added to sinomat, which is
 * phantom3D.project_to / gen_object.project_to (tomo3D.py:1562, :2386). */
int ct2_phantom_project(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                        int nsurface, float* sinomat, const int8_t* indsino,
                        const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                        float u0, float v0, int accumulate, ct2_stream_t stream);
/* The same over a ray list: sinomat[nr], ray_vectors [nr][6]
 * (ray_proj3D.f:175-512; lino_project_to, tomo3D.py:1570, :2552, :2873). */
int ct2_phantom_project_rays(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                             int nsurface, float* sinomat, const int8_t* indsino,
                             const float* ray_vectors, int nr, int accumulate,
                             ct2_stream_t stream);
/* smat[nx][ny][nz] at voxel centres x0 + (i+0.5)dx, ...: an ellipsoid gives
 * att strictly inside and 0 outside (tomo3DfortranAux.f:212-267), a general
 * object att times the product of its surfaces' indicator steps
 * (tomo3DfortranAux.f:420-582); accumulate as above (embed_in,
 * tomo3D.py:1555, :1967, :2732). */
int ct2_phantom_embed(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                      int nsurface, float* smat, int nx, int ny, int nz, float dx, float dy,
                      float dz, float x0, float y0, float z0, int accumulate,
                      ct2_stream_t stream);

/* out[q] = the phantom's voxel value at the point xyz[q][0..2] (device
 * arrays, n points; accumulate = 1 adds it to out[q]): the scalar point tests
 * ellipse_voxval (tomo3DfortranAux.f:179-210) and gen_obj_voxval
 * (tomo3DfortranAux.f:311-416), which tomo3D.py:2122-2220 and :2725-2730
 * call once per point, evaluated for n points in one launch.  Boundaries as
 * the reference: an ellipsoid component is strict, a surface's step closed. */
int ct2_phantom_points(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                       int nsurface, const float* xyz, float* out, size_t n, int accumulate,
                       ct2_stream_t stream);

/* -------------------------------------------------------- ray bunching -- */
/* tomo3DfortranAux.f:1-67 sinoscale (sinogram3D.homoscale, tomo3D.py:395):
 * sinomat[p,j,k] /= |source_p - bin centre (j,k)|, in place. */
int ct2_sinoscale(float* sinomat, const float* frame_vectors, int ns, int nu, int nv, float du,
                  float dv, float u0, float v0, ct2_stream_t stream);
/* tomo3DfortranAux.f:71-122 sinoave (raybunch.get_data, tomo3D.py:504):
 * sinoout[i,j,k] = mean of sinomat over [i*nds, i*nds+nws) x [j*ndu,
 * j*ndu+nwu) x [k*ndv, k*ndv+nwv); every window must lie inside sinomat. */
int ct2_sinoave(const float* sinomat, int ns, int nu, int nv, float* sinoout, int nss, int nus,
                int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv,
                ct2_stream_t stream);
/* tomo3DfortranAux.f:125-176 sinodist (raybunch.put_data, tomo3D.py:520):
 * every window of sinomat takes its coarse value from sinoin (the last
 * window in loop order wins where windows overlap); other bins are 0. */
int ct2_sinodist(float* sinomat, int ns, int nu, int nv, const float* sinoin, int nss, int nus,
                 int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv,
                 ct2_stream_t stream);

/* ------------------------------------------ weighted and derivative ART -- */
/* The weighted ray-bunch and derivative measurements whose call sites and
 * checks the reference keeps (tests/cweighted_art_test.py:80-140,
 * tests/cderiv_art_test.py:106-140) while their GPU sources are missing.
 * Coarse bin (I,J,K) of [NS][NU][NV], NS = ceil(ns/nss), NU = ceil(nu/nsu),
 * NV = ceil(nv/nsv), measures m = sum_q w_q (A x)[fine ray q], where kernel
 * entry q = (a,c,d) of w[ws][wu][wv] (device memory) is the fine ray
 * (I*nss + a - os, J*nsu + c - ou, K*nsv + d - ov), os = ws-1-ws/2 (likewise
 * ou, ov), rays outside the sinogram skipped: scipy.ndimage.convolve(proj,
 * w[::-1,::-1,::-1], mode='constant')[::nss, ::nsu, ::nsv], the tests' own
 * check.  A bin is used when indsino at its anchor (I*nss, J*nsu, K*nsv) is 1
 * (NULL: every bin).  The derivative data are w = [-1, +1] along u with unit
 * strides.  desc: float32 volume, EXACT arithmetic, linear or nearest, free or
 * z axis; the view range fields are ignored. */
typedef struct ct2_stencil {
    int ws, wu, wv;      /* kernel extents */
    int nss, nsu, nsv;   /* coarse strides */
} ct2_stencil;
/* small[NS][NU][NV] = the measurements of vol (cweighted_project_to);
 * fine[ns][nu][nv] is work space for the fine projection. */
int ct2_stencil_project(const ct2_desc* desc, const ct2_stencil* st, const float* w,
                        const float* vol, const float* frame_vectors, float* fine, float* small,
                        ct2_stream_t stream);
/* weights[b] = |r_b|^2 of every used bin's row r_b = sum_q w_q a_q (0 for
 * unused bins) and image += sum of the used rows (put_weights,
 * get_deriv_weights).  scratch: nx*ny*nz floats, zero on entry, zero again
 * on return. */
int ct2_stencil_weights(const ct2_desc* desc, const ct2_stencil* st, const float* w,
                        const int8_t* indsino, const float* frame_vectors, float* weights,
                        float* image, float* scratch, ct2_stream_t stream);
/* One Kaczmarz sweep over the used bins with weights > 0, in (I,J,K) order:
 * vol += reg_fac (small[b] - r_b.vol) r_b / weights[b]
 * (c_ss_weighted_art_step, cderiv_art_step). */
int ct2_stencil_art(const ct2_desc* desc, const ct2_stencil* st, const float* w, float reg_fac,
                    const float* small, const float* weights, const int8_t* indsino,
                    const float* frame_vectors, float* vol, ct2_stream_t stream);

/* Issue-rate calibration for the roofline figures (no counterpart in the
 * reference): `blocks` workgroups of 256 threads, each thread running `iters`
 * iterations of 8 independent chains of one VALU instruction kind (kind 0:
 * FP32 FMA, 1: FP64 FMA, 2: 32-bit integer rotate), exactly 8 such instructions
 * per iteration; out[blocks*256] receives the chains' sums.  The caller
 * times the launch: blocks*4*iters*8 wavefront instructions. */
int ct2_calib_issue(int kind, int blocks, int iters, float* out, ct2_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* CT2_H */
