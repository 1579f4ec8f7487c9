// ct2_abi.hip -- extern "C" entry points (include/ct2.h): argument checks,
// thread-local error text, and the f2py-named wrappers over the launchers.
#include <stdarg.h>
#include <stdio.h>

#include "ct2_internal.h"

namespace ct2 {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char* what) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? CT2_ENOMEM : CT2_EHIP;
}

int check_grid(const Grid& g) {
    for (int q = 0; q < 3; ++q) {
        if (g.n[q] <= 0) { set_error("volume dimension %d is %d (must be > 0)", q, g.n[q]); return CT2_EINVAL; }
        if (!(g.d[q] != 0.0f)) { set_error("voxel size %d is zero", q); return CT2_EINVAL; }
    }
    if ((long long)g.n[0] * g.n[1] * g.n[2] >= (1LL << 31)) {
        set_error("volume of %d x %d x %d exceeds 2^31 voxels", g.n[0], g.n[1], g.n[2]);
        return CT2_EINVAL;
    }
    return CT2_OK;
}

int check_det(const Det& d) {
    if (d.ns <= 0 || d.nu <= 0 || d.nv <= 0) {
        set_error("sinogram shape (%d, %d, %d) must be positive", d.ns, d.nu, d.nv);
        return CT2_EINVAL;
    }
    if (d.du == 0.0f || d.dv == 0.0f) { set_error("detector pitch is zero"); return CT2_EINVAL; }
    if ((long long)d.ns * d.nu * d.nv >= (1LL << 40)) { set_error("sinogram too large"); return CT2_EINVAL; }
    return CT2_OK;
}

int view_set(const ct2_desc* d, const char* who, ViewSet& vs) {
    vs.vstride = d->vstride > 0 ? d->vstride : 1;
    vs.view0 = d->view0;
    vs.nviews = d->nviews ? d->nviews : (d->ns - d->view0 + vs.vstride - 1) / vs.vstride;
    if (d->vstride < 0 || vs.view0 < 0 || vs.nviews <= 0 ||
        (long long)vs.view0 + (long long)(vs.nviews - 1) * vs.vstride >= d->ns) {
        set_error("%s: bad view subset (view0 %d, nviews %d, vstride %d) of %d views", who,
                  d->view0, d->nviews, d->vstride, d->ns);
        return CT2_EINVAL;
    }
    return CT2_OK;
}

static void fill_desc(ct2_desc* d, int ns, int nu, int nv, float du, float dv, float u0,
                      float v0, int nx, int ny, int nz, float dx, float dy, float dz, float x0,
                      float y0, float z0) {
    *d = ct2_desc{};
    d->ns = ns; d->nu = nu; d->nv = nv;
    d->du = du; d->dv = dv; d->u0 = u0; d->v0 = v0;
    d->nx = nx; d->ny = ny; d->nz = nz;
    d->dx = dx; d->dy = dy; d->dz = dz;
    d->x0 = x0; d->y0 = y0; d->z0 = z0;
    d->interp = CT2_LINEAR;
    d->axis = CT2_AXIS_FREE;
    d->vol_dtype = CT2_F32;
    d->arith = CT2_ARITH_EXACT;
    d->vstride = 1;
}

}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_version(void) { return 100; }

const char* ct2_last_error(void) { return g_err; }

int ct2_desc_init(ct2_desc* d, int ns, int nu, int nv, float du, float dv, float u0, float v0,
                  int nx, int ny, int nz, float dx, float dy, float dz, float x0, float y0,
                  float z0) {
    if (!d) { set_error("ct2_desc_init: null descriptor"); return CT2_EINVAL; }
    fill_desc(d, ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0);
    return CT2_OK;
}

int ct2_project(const ct2_desc* d, float* sino, const int8_t* ind, const float* fv,
                const void* vol, ct2_stream_t stream) {
    if (!d) { set_error("ct2_project: null descriptor"); return CT2_EINVAL; }
    return launch_project(d, sino, ind, fv, vol, (hipStream_t)stream);
}

int ct2_backproject(const ct2_desc* d, void* vol, const float* sino, const int8_t* ind,
                    const float* fv, ct2_stream_t stream) {
    if (!d) { set_error("ct2_backproject: null descriptor"); return CT2_EINVAL; }
    return launch_backproject(d, vol, sino, ind, fv, (hipStream_t)stream);
}

int ct2_backproject_ws(const ct2_desc* d, void* vol, const float* sino, const int8_t* ind,
                       const float* fv, void* work, size_t work_bytes, ct2_stream_t stream) {
    if (!d) { set_error("ct2_backproject_ws: null descriptor"); return CT2_EINVAL; }
    SepWork w;
    w.ptr = work;
    w.bytes = work ? work_bytes : 0;
    return launch_backproject(d, vol, sino, ind, fv, (hipStream_t)stream, w);
}

size_t ct2_backproject_work_bytes(const ct2_desc* d) {
    if (!d) { set_error("ct2_backproject_work_bytes: null descriptor"); return 0; }
    return sep_work_bytes(d);
}

#define VIEW_ARGS                                                                              \
    int ns, int nu, int nv, float du, float dv, float u0, float v0
#define GRID_ARGS                                                                              \
    float dx, float dy, float dz, float x0, float y0, float z0, int nx, int ny, int nz
#define DESC(d)                                                                                \
    ct2_desc d;                                                                                \
    fill_desc(&d, ns, nu, nv, du, dv, u0, v0, nx, ny, nz, dx, dy, dz, x0, y0, z0)

int ct2_sinoproj_linear(float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                        const float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    return launch_project(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_sinoproj_nearest(float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                         const float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.interp = CT2_NEAREST;
    return launch_project(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_sinoprojz(float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                  const float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.axis = CT2_AXIS_Z;
    return launch_project(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_sinobackproject_linear(const float* sinomat, const int8_t* indsino, const float* fv,
                               VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    return launch_backproject(&d, smat, sinomat, indsino, fv, (hipStream_t)stream);
}

int ct2_sinobackproject_nearest(const float* sinomat, const int8_t* indsino, const float* fv,
                                VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.interp = CT2_NEAREST;
    return launch_backproject(&d, smat, sinomat, indsino, fv, (hipStream_t)stream);
}

#define RAY_DESC(d)                                                                            \
    ct2_desc d;                                                                                \
    fill_desc(&d, 1, 1, 1, 1.f, 1.f, 0.f, 0.f, nx, ny, nz, dx, dy, dz, x0, y0, z0)

int ct2_ray_sinoproj_linear(float* sinomat, const int8_t* indsino, const float* rv, int nr,
                            const float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    return launch_ray_project(&d, sinomat, indsino, rv, nr, smat, 0, (hipStream_t)stream);
}

int ct2_ray_sinoproj_nearest(float* sinomat, const int8_t* indsino, const float* rv, int nr,
                             const float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    d.interp = CT2_NEAREST;
    return launch_ray_project(&d, sinomat, indsino, rv, nr, smat, 0, (hipStream_t)stream);
}

int ct2_ray_sinobackproject_nearest(const float* sinomat, const int8_t* indsino, const float* rv,
                                    int nr, float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    d.interp = CT2_NEAREST;
    // ray_proj3D.f:895 bounds the nearest adjoint with .lt. n-1
    return launch_ray_backproject(&d, smat, sinomat, indsino, rv, nr, 1, (hipStream_t)stream);
}

int ct2_pd_backproj_linear(const float* sinomat, const int8_t* indsino, const float* fv,
                           VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    (void)indsino;   // the reference's indicator test is commented out (view_proj3D.f:1926-1928)
    DESC(d);
    return launch_pd_backproject(&d, smat, sinomat, fv, (hipStream_t)stream);
}

int ct2_pd_backproj_nearest(const float* sinomat, const int8_t* indsino, const float* fv,
                            VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    (void)indsino;   // commented out in the reference too (view_proj3D.f:1752-1754)
    DESC(d);
    d.interp = CT2_NEAREST;
    return launch_pd_backproject(&d, smat, sinomat, fv, (hipStream_t)stream);
}

// ---- ART ----
int ct2_art_step(const ct2_desc* d, float reg_fac, const float* sino, const int8_t* ind,
                 const float* fv, float* vol, ct2_stream_t stream) {
    if (!d) { set_error("ct2_art_step: null descriptor"); return CT2_EINVAL; }
    if (d->vol_dtype != CT2_F32) { set_error("ct2_art_step: ART updates a float32 volume"); return CT2_EINVAL; }
    return launch_art(d, nullptr, reg_fac, sino, ind, fv, vol, nullptr, (hipStream_t)stream);
}

int ct2_artstep_linear(float reg_fac, const float* sinomat, const int8_t* indsino,
                       const float* fv, VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    return launch_art(&d, nullptr, reg_fac, sinomat, indsino, fv, smat, nullptr, (hipStream_t)stream);
}

int ct2_artstep_nearest(float reg_fac, const float* sinomat, const int8_t* indsino,
                        const float* fv, VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.interp = CT2_NEAREST;
    return launch_art(&d, nullptr, reg_fac, sinomat, indsino, fv, smat, nullptr, (hipStream_t)stream);
}

int ct2_artstepz(float reg_fac, const float* sinomat, const int8_t* indsino, const float* fv,
                 VIEW_ARGS, float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.axis = CT2_AXIS_Z;
    return launch_art(&d, nullptr, reg_fac, sinomat, indsino, fv, smat, nullptr, (hipStream_t)stream);
}

int ct2_artstepz2vol(float reg_fac, const float* sinomat, const int8_t* indsino, const float* fv,
                     VIEW_ARGS, float* smat, GRID_ARGS, float* smata, float dxa, float dya,
                     float dza, float x0a, float y0a, float z0a, int nxa, int nya, int nza,
                     ct2_stream_t stream) {
    DESC(d);
    d.axis = CT2_AXIS_Z;
    Grid ga = make_grid(nxa, nya, nza, dxa, dya, dza, x0a, y0a, z0a);
    return launch_art(&d, &ga, reg_fac, sinomat, indsino, fv, smat, smata, (hipStream_t)stream);
}

int ct2_ray_artstep_linear(float reg_fac, const float* sinomat, const int8_t* indsino,
                           const float* rv, int nr, float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    return launch_ray_art(&d, 0, reg_fac, sinomat, indsino, rv, nr, smat, (hipStream_t)stream);
}

int ct2_ray_artstep_nearest(float reg_fac, const float* sinomat, const int8_t* indsino,
                            const float* rv, int nr, float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    d.interp = CT2_NEAREST;
    // ray_proj3D.f:1351 bounds the nearest ART with .lt. n-1
    return launch_ray_art(&d, 1, reg_fac, sinomat, indsino, rv, nr, smat, (hipStream_t)stream);
}

int ct2_convert_volume(const float* src, void* dst, int dtype, size_t n, int rounding,
                       unsigned seed, ct2_stream_t stream) {
    return launch_convert(src, dst, dtype, n, rounding, seed, (hipStream_t)stream);
}

int ct2_convert_volume_zplanes(const float* src, void* dst, int dtype, int nx, int ny, int nz,
                               int rounding, unsigned seed, ct2_stream_t stream) {
    return launch_convert_zplanes(src, dst, dtype, nx, ny, nz, rounding, seed,
                                  (hipStream_t)stream);
}

// ---- SART ----
int ct2_sart_step(const ct2_desc* d, float lambda, const float* sino, const int8_t* ind,
                  const float* fv, float* vol, float* work, ct2_stream_t stream) {
    if (!d) { set_error("ct2_sart_step: null descriptor"); return CT2_EINVAL; }
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    return launch_sart(d, lambda, sino, ind, fv, vol, work, (hipStream_t)stream);
}

// ---- MLEM ----
int ct2_em_step(const ct2_desc* d, const float* sino, const int8_t* ind, const float* fv,
                float* vol, float* work, const float* sens, ct2_stream_t stream) {
    if (!d) { set_error("ct2_em_step: null descriptor"); return CT2_EINVAL; }
    if (d->vol_dtype != CT2_F32) { set_error("ct2_em_step: MLEM updates a float32 volume"); return CT2_EINVAL; }
    return launch_em(d, sino, ind, fv, vol, work, sens, (hipStream_t)stream);
}

// The f2py-named MLEM steps own their scratch, like the reference's
// automatic arrays (view_proj3D.f:2519): stream-ordered allocation.
static int em_owned(const ct2_desc* d, const float* sino, const int8_t* ind, const float* fv,
                    float* vol, hipStream_t st) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    int rc;
    if ((rc = check_grid(g)) || (rc = check_det(dt))) return rc;
    void* work = nullptr;
    CT2_HIP(hipMallocAsync(&work, sizeof(float) * em_work_floats(d), st));
    rc = launch_em(d, sino, ind, fv, vol, (float*)work, nullptr, st);
    hipError_t e = hipFreeAsync(work, st);
    if (rc) return rc;
    CT2_HIP(e);
    return CT2_OK;
}

int ct2_emstep_linear(const float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                      float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    return em_owned(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_emstep_nearest(const float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                       float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.interp = CT2_NEAREST;
    return em_owned(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_emstep_linearz(const float* sinomat, const int8_t* indsino, const float* fv, VIEW_ARGS,
                       float* smat, GRID_ARGS, ct2_stream_t stream) {
    DESC(d);
    d.axis = CT2_AXIS_Z;
    return em_owned(&d, sinomat, indsino, fv, smat, (hipStream_t)stream);
}

int ct2_ray_emstep_nearest(const float* sinomat, const int8_t* indsino, const float* rv, int nr,
                           float* smat, GRID_ARGS, ct2_stream_t stream) {
    RAY_DESC(d);
    d.interp = CT2_NEAREST;
    int rc;
    Grid g = make_grid(nx, ny, nz, dx, dy, dz, x0, y0, z0);
    if ((rc = check_grid(g))) return rc;
    if (nr < 0) { set_error("ray MLEM: nr < 0"); return CT2_EINVAL; }
    size_t nvox = (size_t)nx * ny * nz;
    void* work = nullptr;
    hipStream_t st = (hipStream_t)stream;
    CT2_HIP(hipMallocAsync(&work, sizeof(float) * (2 * (size_t)nr + 2 * nvox), st));
    // ray_proj3D.f:1543 bounds the nearest MLEM with .lt. n-1
    rc = launch_ray_em(&d, 1, sinomat, indsino, rv, nr, smat, (float*)work, st);
    hipError_t e = hipFreeAsync(work, st);
    if (rc) return rc;
    CT2_HIP(e);
    return CT2_OK;
}

}  // extern "C"
