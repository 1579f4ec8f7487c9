// ct2_phantom.hip -- analytic phantoms on the GPU: exact line integrals
// through quadric objects (emil_code/view_proj3D.f:11-603,
// emil_code/ray_proj3D.f:175-512) and their voxel embeddings
// (emil_code/tomo3DfortranAux.f:212-582).
//
// The reference calls one Fortran routine per (component, view) from Python
// and adds the results into the sinogram (tomo3D.py:1562-1567, :2531-2550).
// Here one launch evaluates every component for every bin: a lane owns one
// bin (or voxel), walks the component table (wave-uniform, so the records
// come in through scalar loads) and adds the components in table order,
// which is the order of the reference's Python loop, so the sums are
// bit-identical.  Each component's value follows the Fortran's FP32
// expressions operation by operation (compiled with -ffp-contract=off);
// cos/sin of the ellipsoid angles are taken on the host with the C
// library's cosf/sinf, which the compiled Fortran also calls.
//
// Bound: FP32 VALU (tens of flops per bin and surface, one division and one
// square root per root), not HBM: a bin costs 4 B written, 1 B of indicator
// and 4 B read when accumulating.
#include "ct2_internal.h"

#include <math.h>

#include <algorithm>
#include <vector>

namespace ct2 {
namespace {

constexpr float kLarge = 100000000.0f;   // the reference's stand-in for infinity
constexpr int kThreads = 256;

struct QSurf {       // 64 B
    int kind;
    float ax, ay, az;
    float u[12];     // frame rows xu, yu, zu, then the centre
};

struct QShape {      // 64 B
    int kind;
    float att;
    int first, nsurf;
    float ax, ay, az, x0, y0, z0;
    float ca, sa, cb, sb;
    int pad[2];
};

static_assert(sizeof(QSurf) == 64 && sizeof(QShape) == 64, "table records are 64 B");

// root_plane (view_proj3D.f:11-39): the frame's half-space z >= 0
__device__ __forceinline__ void root_plane(float* t, float zs, float zb) {
    if (fabsf(zs - zb) < 0.00000001f) {
        if (zs >= 0.f) { t[0] = -kLarge; t[1] = kLarge; }
        else           { t[0] = 1.0f; t[1] = -1.0f; }
    } else {
        float root1 = zs / (zs - zb);
        if (zs < zb) { t[0] = root1; t[1] = kLarge; }
        else         { t[0] = -kLarge; t[1] = root1; }
    }
}

// root_ellipsoid (view_proj3D.f:42-72)
__device__ __forceinline__ void root_ellipsoid(float* t, float xs, float ys, float zs, float xb,
                                               float yb, float zb, float ax, float ay, float az) {
    float ax2 = ax * ax, ay2 = ay * ay, az2 = az * az;
    float dx = xb - xs, dy = yb - ys, dz = zb - zs;
    float c1 = xs * yb - xb * ys, c2 = xs * zb - xb * zs, c3 = ys * zb - yb * zs;
    float den = ax2 * az2 * (dy * dy) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz));
    float s = -(az2 * (c1 * c1)) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz) - c2 * c2)
            + ax2 * (az2 * (dy * dy) - c3 * c3);
    float dist2 = (4.f * ax2 * ay2 * az2 * s) / (den * den);
    float avet = (ax2 * az2 * ys * (ys - yb) + ay2 * (az2 * xs * (xs - xb) + ax2 * zs * (zs - zb)))
               / den;
    if (dist2 < 0.f) { t[0] = 1.0f; t[1] = -1.0f; }
    else {
        float h = 0.5f * sqrtf(dist2);
        t[0] = avet - h;
        t[1] = avet + h;
    }
}

// root_cylinder (view_proj3D.f:75-97): elliptic cylinder along the frame z
__device__ __forceinline__ void root_cylinder(float* t, float xs, float ys, float xb, float yb,
                                              float ax, float ay) {
    float ax2 = ax * ax, ay2 = ay * ay;
    float dx = xb - xs, dy = yb - ys;
    float c1 = xs * yb - xb * ys;
    float den = ay2 * (dx * dx) + ax2 * (dy * dy);
    float dist2 = (4.f * ax2 * ay2 * (ay2 * (dx * dx) + ax2 * (dy * dy) - c1 * c1)) / (den * den);
    float avet = (ay2 * xs * (xs - xb) + ax2 * ys * (ys - yb)) / den;
    if (dist2 < 0.f) { t[0] = 1.0f; t[1] = -1.0f; }
    else {
        float h = 0.5f * sqrtf(dist2);
        t[0] = avet - h;
        t[1] = avet + h;
    }
}

// root_cone (view_proj3D.f:100-171): the z >= 0 nappe of
// (x/ax)^2 + (y/ay)^2 = z^2.  The two one-root cases are sequential tests as
// in the Fortran; t keeps its contents when neither applies (NaN roots only).
__device__ __forceinline__ void root_cone(float* t, float xs, float ys, float zs, float xb,
                                          float yb, float zb, float ax, float ay) {
    float ax2 = ax * ax, ay2 = ay * ay;
    float dy = yb - ys, dz = zb - zs;
    float c3 = ys * zb - yb * zs;
    float s = xs * xs * (-(yb * yb) + ay2 * (zb * zb)) + ax2 * (c3 * c3)
            + 2.f * xb * xs * (yb * ys - ay2 * zb * zs) + xb * xb * (-(ys * ys) + ay2 * (zs * zs));
    float den = ax2 * (dy * dy) + ay2 * (xb * xb - 2.f * xb * xs + xs * xs - ax2 * (dz * dz));
    float dist2 = (4.f * ax2 * ay2 * s) / (den * den);
    float avet = (ax2 * ys * (ys - yb) + ay2 * (-(xb * xs) + xs * xs + ax2 * (zb - zs) * zs)) / den;
    if (dist2 < 0.f) { t[0] = 1.0f; t[1] = -1.0f; return; }
    float sd2 = sqrtf(dist2) * 0.5f;
    float tr1 = avet - sd2, tr2 = avet + sd2;
    float zr1 = zs * (1.f - tr1) + zb * tr1;
    float zr2 = zs * (1.f - tr2) + zb * tr2;
    if ((zr1 <= 0.f) && (zr2 <= 0.f)) { t[0] = 1.0f; t[1] = -1.0f; return; }
    if ((zr1 > 0.f) && (zr2 > 0.f)) { t[0] = tr1; t[1] = tr2; return; }
    if (zr1 >= 0.f) {
        if (zs < zb) { t[0] = tr1; t[1] = kLarge; } else { t[0] = -kLarge; t[1] = tr1; }
    }
    if (zr2 >= 0.f) {
        if (zs < zb) { t[0] = tr2; t[1] = kLarge; } else { t[0] = -kLarge; t[1] = tr2; }
    }
}

__device__ __forceinline__ void to_frame(const float* u, float x, float y, float z, float* o) {
    float xt = x - u[9], yt = y - u[10], zt = z - u[11];
    o[0] = u[0] * xt + u[1] * yt + u[2] * zt;
    o[1] = u[3] * xt + u[4] * yt + u[5] * zt;
    o[2] = u[6] * xt + u[7] * yt + u[8] * zt;
}

// A point in a surface's frame for the voxel value.  Unit-vector frames as
// to_frame; angle frames (CT2_SURF_ANGLES) in the operation order of framer
// (tomo3DfortranAux.f:281-308), u = cos/sin of alpha, beta, gamma, then the
// centre in u[9..11].
__device__ __forceinline__ void surf_frame(const QSurf& sf, float x, float y, float z, float* o) {
    if (!(sf.kind & CT2_SURF_ANGLES)) { to_frame(sf.u, x, y, z, o); return; }
    const float* u = sf.u;
    const float xp = (x - u[9]) * u[0] - (u[10] - y) * u[1];
    const float yp = (y - u[10]) * u[0] + (u[9] - x) * u[1];
    const float zp = z - u[11];
    const float xpp = xp * u[2] - zp * u[3], ypp = yp, zpp = xp * u[3] + zp * u[2];
    o[0] = xpp * u[4] + ypp * u[5];
    o[1] = -(xpp * u[5]) + ypp * u[4];
    o[2] = zpp;
}

// tlen of gen_shape_proj (view_proj3D.f:306-455).  MAX/MIN as the Fortran
// compiler lowers them: compare, then select the second operand when the
// compare fails.
__device__ float gen_tlen(const QSurf* __restrict__ sf, int nsurf, const float* S, const float* B,
                          float* troot) {
    float f1 = -kLarge, f2 = kLarge;
    for (int q = 0; q < nsurf; ++q) {
        const QSurf s = sf[q];
        float a[3], b[3];
        to_frame(s.u, S[0], S[1], S[2], a);
        to_frame(s.u, B[0], B[1], B[2], b);
        switch (s.kind) {
        case CT2_SURF_PLANE: root_plane(troot, a[2], b[2]); break;
        case CT2_SURF_ELLIPSOID:
            root_ellipsoid(troot, a[0], a[1], a[2], b[0], b[1], b[2], s.ax, s.ay, s.az); break;
        case CT2_SURF_CYLINDER: root_cylinder(troot, a[0], a[1], b[0], b[1], s.ax, s.ay); break;
        default: root_cone(troot, a[0], a[1], a[2], b[0], b[1], b[2], s.ax, s.ay); break;
        }
        f1 = (f1 > troot[0]) ? f1 : troot[0];
        f2 = (f2 < troot[1]) ? f2 : troot[1];
        if (f1 >= f2) return 0.f;
    }
    return f2 - f1;
}

// ellipsoid_proj body (view_proj3D.f:557-597)
__device__ float ell_value(const QShape& e, const float* S, const float* B) {
    float xspp = (S[0] - e.x0) * e.cb - (e.y0 - S[1]) * e.sb;
    float yspp = (S[1] - e.y0) * e.cb + (e.x0 - S[0]) * e.sb;
    float zspp = S[2] - e.z0;
    float xsp = xspp * e.ca + zspp * e.sa, ysp = yspp, zsp = -(xspp * e.sa) + zspp * e.ca;
    float xbpp = (B[0] - e.x0) * e.cb - (e.y0 - B[1]) * e.sb;
    float ybpp = (B[1] - e.y0) * e.cb + (e.x0 - B[0]) * e.sb;
    float zbpp = B[2] - e.z0;
    float xbp = xbpp * e.ca + zbpp * e.sa, ybp = ybpp, zbp = -(xbpp * e.sa) + zbpp * e.ca;
    float ax2 = e.ax * e.ax, ay2 = e.ay * e.ay, az2 = e.az * e.az;
    float dx = xbp - xsp, dy = ybp - ysp, dz = zbp - zsp;
    float c1 = xsp * ybp - xbp * ysp, c2 = xsp * zbp - xbp * zsp, c3 = ysp * zbp - ybp * zsp;
    float s = -(az2 * (c1 * c1)) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz) - c2 * c2)
            + ax2 * (az2 * (dy * dy) - c3 * c3);
    float den = ax2 * az2 * (dy * dy) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz));
    float pl2 = (4.f * ax2 * ay2 * az2 * (dx * dx + dy * dy + dz * dz) * s) / (den * den);
    if (pl2 < 0.f) pl2 = 0.f;
    return sqrtf(pl2) * e.att;
}

// Line integrals.  RAYS=false: element o of a [ns][nu][nv] sinogram, ray
// from the view's source to bin (j,k) (view_proj3D.f:273-292).  RAYS=true:
// element o of a ray list [nr] with ray_vectors [nr][6] (ray_proj3D.f:244-256).
template <bool RAYS>
__global__ __launch_bounds__(kThreads) void k_phantom_lines(
    const QShape* __restrict__ shp, int nshape, const QSurf* __restrict__ srf,
    float* __restrict__ out, const int8_t* __restrict__ ind, const float* __restrict__ geom,
    long long nbins, int nu, int nv, float du, float dv, float u0, float v0, int accumulate) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= nbins) return;
    float S[3], B[3];
    if (RAYS) {
        const float* r = geom + o * 6;
        S[0] = r[0]; S[1] = r[1]; S[2] = r[2];
        B[0] = r[3]; B[1] = r[4]; B[2] = r[5];
    } else {
        const long long per = (long long)nu * nv;
        const int p = (int)(o / per);
        const int q = (int)(o - (long long)p * per);
        const int j = q / nv, k = q - j * nv;
        const float* f = geom + (size_t)p * 18;
        float u = u0 + ((float)j + 0.5f) * du;
        float v = v0 + ((float)k + 0.5f) * dv;
        for (int c = 0; c < 3; ++c) {
            S[c] = f[c];
            B[c] = f[6 + c] + f[9 + c] * u + f[12 + c] * v;
        }
    }
    const bool use = !ind || ind[o] == 1;
    const float sbdist = sqrtf((S[0] - B[0]) * (S[0] - B[0]) + (S[1] - B[1]) * (S[1] - B[1])
                               + (S[2] - B[2]) * (S[2] - B[2]));
    float acc = accumulate ? out[o] : 0.f;
    bool first = !accumulate;
    float troot[2] = {0.f, 0.f};
    for (int s = 0; s < nshape; ++s) {
        const QShape sh = shp[s];
        float c = 0.f;
        if (use) {
            if (sh.kind == CT2_SHAPE_ELLIPSOID)
                c = ell_value(sh, S, B);
            else
                c = sbdist * gen_tlen(srf + sh.first, sh.nsurf, S, B, troot) * sh.att;
        }
        acc = first ? c : acc + c;
        first = false;
    }
    out[o] = acc;
}

// The phantom's value at (x, y, z): the components' voxel values
// (ellipsoid_embed, gen_obj_embed, ellipse_voxval, gen_obj_voxval of
// tomo3DfortranAux.f:179-582) added in table order onto acc, or, when first
// is set, starting from the first component's value.
__device__ float point_value(const QShape* __restrict__ shp, int nshape,
                             const QSurf* __restrict__ srf, float x, float y, float z, float acc,
                             bool first) {
    for (int s = 0; s < nshape; ++s) {
        const QShape sh = shp[s];
        float c;
        if (sh.kind == CT2_SHAPE_ELLIPSOID) {
            float rx = x - sh.x0, ry = y - sh.y0, rz = z - sh.z0;
            float r1x = sh.cb * rx + sh.sb * ry, r1y = -(sh.sb * rx) + sh.cb * ry, r1z = rz;
            float r2x = sh.ca * r1x + sh.sa * r1z, r2y = r1y, r2z = -(sh.sa * r1x) + sh.ca * r1z;
            float qx = r2x / sh.ax, qy = r2y / sh.ay, qz = r2z / sh.az;
            float lhs = qx * qx + qy * qy + qz * qz;
            c = (lhs < 1.f) ? sh.att : 0.f;
        } else {
            float si = 1.0f;
            for (int q = 0; q < sh.nsurf; ++q) {
                const QSurf sf = srf[sh.first + q];
                float f[3];
                surf_frame(sf, x, y, z, f);
                float e, a, b;
                switch (sf.kind & 3) {
                case CT2_SURF_PLANE: si = si * (f[2] >= 0.f ? 1.f : 0.f); break;
                case CT2_SURF_ELLIPSOID:
                    a = f[0] / sf.ax; b = f[1] / sf.ay;
                    e = a * a + b * b + (f[2] / sf.az) * (f[2] / sf.az);
                    si = si * ((1.0f - e) >= 0.f ? 1.f : 0.f);
                    break;
                case CT2_SURF_CYLINDER:
                    a = f[0] / sf.ax; b = f[1] / sf.ay;
                    e = a * a + b * b;
                    si = si * ((1.0f - e) >= 0.f ? 1.f : 0.f);
                    break;
                default:
                    a = f[0] / sf.ax; b = f[1] / sf.ay;
                    e = a * a + b * b;
                    si = si * ((f[2] * f[2] - e) >= 0.f ? 1.f : 0.f) * (f[2] >= 0.f ? 1.f : 0.f);
                    break;
                }
            }
            c = sh.att * si;
        }
        acc = first ? c : acc + c;
        first = false;
    }
    return acc;
}

// Voxel embedding of [nx][ny][nz] at voxel centres (tomo3DfortranAux.f:
// 226-266 for the ellipsoid, :508-580 for general objects).
__global__ __launch_bounds__(kThreads) void k_phantom_embed(
    const QShape* __restrict__ shp, int nshape, const QSurf* __restrict__ srf,
    float* __restrict__ out, int nx, int ny, int nz, float dx, float dy, float dz, float x0i,
    float y0i, float z0i, int accumulate) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    const long long n = (long long)nx * ny * nz;
    if (o >= n) return;
    const long long r = o / nz;
    const int k = (int)(o - r * nz);
    const int i = (int)(r / ny);
    const int j = (int)(r - (long long)i * ny);
    const float x = x0i + ((float)i + 0.5f) * dx;
    const float y = y0i + ((float)j + 0.5f) * dy;
    const float z = z0i + ((float)k + 0.5f) * dz;
    out[o] = point_value(shp, nshape, srf, x, y, z, accumulate ? out[o] : 0.f, !accumulate);
}

// Voxel values at arbitrary points xyz[n][3] (the scalar point tests of
// tomo3DfortranAux.f:179-210 and :311-416, one lane per point).
__global__ __launch_bounds__(kThreads) void k_phantom_points(
    const QShape* __restrict__ shp, int nshape, const QSurf* __restrict__ srf,
    const float* __restrict__ xyz, float* __restrict__ out, long long n, int accumulate) {
    const long long o = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (o >= n) return;
    const float x = xyz[3 * o], y = xyz[3 * o + 1], z = xyz[3 * o + 2];
    out[o] = point_value(shp, nshape, srf, x, y, z, accumulate ? out[o] : 0.f, !accumulate);
}

// Validates the host tables and copies them, packed, into one stream-ordered
// device allocation: QShape[nshape] then the surfaces of every general shape,
// stably ordered planes, ellipsoids, cylinders, cones (the order in which
// gen_shape_proj visits them).
struct Staged {
    void* dev = nullptr;
    int nshape = 0;
    const QShape* shapes() const { return (const QShape*)dev; }
    const QSurf* surfs() const { return (const QSurf*)((const char*)dev + sizeof(QShape) * nshape); }
};

int stage(const char* who, const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
          int nsurface, hipStream_t st, Staged& out, bool allow_angles = false) {
    if (nshape < 0 || nsurface < 0) { set_error("%s: negative table size", who); return CT2_EINVAL; }
    if (nshape > 0 && !shapes) { set_error("%s: null shape table", who); return CT2_EINVAL; }
    std::vector<QShape> qs((size_t)nshape);
    std::vector<QSurf> qf;
    for (int s = 0; s < nshape; ++s) {
        const ct2_shape& in = shapes[s];
        QShape& q = qs[s];
        q = QShape{};
        q.kind = in.kind;
        q.att = in.att;
        if (in.kind == CT2_SHAPE_ELLIPSOID) {
            q.ax = in.ax; q.ay = in.ay; q.az = in.az;
            q.x0 = in.x0; q.y0 = in.y0; q.z0 = in.z0;
            q.ca = cosf(in.alpha); q.sa = sinf(in.alpha);
            q.cb = cosf(in.beta);  q.sb = sinf(in.beta);
        } else if (in.kind == CT2_SHAPE_GENERAL) {
            if (in.nsurf < 0 || in.first < 0 || (long long)in.first + in.nsurf > nsurface) {
                set_error("%s: shape %d: surfaces [%d, %d) outside the table of %d", who, s,
                          in.first, in.first + in.nsurf, nsurface);
                return CT2_EINVAL;
            }
            if (in.nsurf > 0 && !surfaces) { set_error("%s: null surface table", who); return CT2_EINVAL; }
            q.first = (int)qf.size();
            q.nsurf = in.nsurf;
            for (int kind = CT2_SURF_PLANE; kind <= CT2_SURF_CONE; ++kind)
                for (int m = 0; m < in.nsurf; ++m) {
                    const ct2_surface& f = surfaces[in.first + m];
                    const bool angles = (f.kind & CT2_SURF_ANGLES) != 0;
                    const int base = f.kind & ~CT2_SURF_ANGLES;
                    if (base < CT2_SURF_PLANE || base > CT2_SURF_CONE || (angles && !allow_angles)) {
                        set_error("%s: surface %d: unknown kind %d%s", who, in.first + m, f.kind,
                                  angles ? " (angle frames are for voxel values only)" : "");
                        return CT2_EINVAL;
                    }
                    if (base != kind) continue;
                    QSurf r{};
                    r.kind = f.kind;
                    r.ax = f.ax; r.ay = f.ay; r.az = f.az;
                    if (angles) {
                        // the angles in xu; their cosines and sines by the C
                        // library, as the compiled Fortran's framer takes them
                        for (int c = 0; c < 3; ++c) {
                            r.u[2 * c] = cosf(f.xu[c]);
                            r.u[2 * c + 1] = sinf(f.xu[c]);
                            r.u[9 + c] = f.c[c];
                        }
                    } else {
                        for (int c = 0; c < 3; ++c) {
                            r.u[c] = f.xu[c];
                            r.u[3 + c] = f.yu[c];
                            r.u[6 + c] = f.zu[c];
                            r.u[9 + c] = f.c[c];
                        }
                    }
                    qf.push_back(r);
                }
        } else {
            set_error("%s: shape %d: unknown kind %d", who, s, in.kind);
            return CT2_EINVAL;
        }
    }
    out.nshape = nshape;
    size_t bytes = sizeof(QShape) * qs.size() + sizeof(QSurf) * qf.size();
    if (bytes == 0) return CT2_OK;
    std::vector<char> host(bytes);
    if (!qs.empty()) std::copy((const char*)qs.data(), (const char*)(qs.data() + qs.size()), host.data());
    if (!qf.empty())
        std::copy((const char*)qf.data(), (const char*)(qf.data() + qf.size()),
                  host.data() + sizeof(QShape) * qs.size());
    CT2_HIP(hipMallocAsync(&out.dev, bytes, st));
    // pageable source: the call returns once the bytes are staged, so the
    // vector may go out of scope afterwards
    hipError_t e = hipMemcpyAsync(out.dev, host.data(), bytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        (void)hipFreeAsync(out.dev, st);
        out.dev = nullptr;
        return hip_fail(e, "hipMemcpyAsync(phantom table)");
    }
    return CT2_OK;
}

int release(Staged& s, hipStream_t st) {
    if (s.dev) CT2_HIP(hipFreeAsync(s.dev, st));
    s.dev = nullptr;
    return CT2_OK;
}

unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace
}  // namespace ct2

using namespace ct2;

extern "C" {

int ct2_phantom_project(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                        int nsurface, float* sinomat, const int8_t* indsino,
                        const float* frame_vectors, int ns, int nu, int nv, float du, float dv,
                        float u0, float v0, int accumulate, ct2_stream_t stream) {
    const char* who = "ct2_phantom_project";
    if (ns <= 0 || nu <= 0 || nv <= 0) { set_error("%s: bad sinogram shape", who); return CT2_EINVAL; }
    if (!sinomat || !frame_vectors) { set_error("%s: null pointer", who); return CT2_EINVAL; }
    long long nbins = (long long)ns * nu * nv;
    if (blocks_for(nbins) > 0x7fffffffu) { set_error("%s: sinogram too large", who); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    Staged t;
    int rc = stage(who, shapes, nshape, surfaces, nsurface, st, t);
    if (rc) return rc;
    hipLaunchKernelGGL(k_phantom_lines<false>, dim3(blocks_for(nbins)), dim3(kThreads), 0, st,
                       t.shapes(), nshape, t.surfs(), sinomat, indsino, frame_vectors, nbins, nu,
                       nv, du, dv, u0, v0, accumulate);
    hipError_t e = hipGetLastError();
    rc = release(t, st);
    CT2_HIP(e);
    return rc;
}

int ct2_phantom_project_rays(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                             int nsurface, float* sinomat, const int8_t* indsino,
                             const float* ray_vectors, int nr, int accumulate,
                             ct2_stream_t stream) {
    const char* who = "ct2_phantom_project_rays";
    if (nr <= 0) { set_error("%s: nr must be positive", who); return CT2_EINVAL; }
    if (!sinomat || !ray_vectors) { set_error("%s: null pointer", who); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    Staged t;
    int rc = stage(who, shapes, nshape, surfaces, nsurface, st, t);
    if (rc) return rc;
    hipLaunchKernelGGL(k_phantom_lines<true>, dim3(blocks_for(nr)), dim3(kThreads), 0, st,
                       t.shapes(), nshape, t.surfs(), sinomat, indsino, ray_vectors,
                       (long long)nr, 0, 0, 0.f, 0.f, 0.f, 0.f, accumulate);
    hipError_t e = hipGetLastError();
    rc = release(t, st);
    CT2_HIP(e);
    return rc;
}

int ct2_phantom_embed(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                      int nsurface, float* smat, int nx, int ny, int nz, float dx, float dy,
                      float dz, float x0, float y0, float z0, int accumulate,
                      ct2_stream_t stream) {
    const char* who = "ct2_phantom_embed";
    if (nx <= 0 || ny <= 0 || nz <= 0) { set_error("%s: bad volume shape", who); return CT2_EINVAL; }
    if (!smat) { set_error("%s: null pointer", who); return CT2_EINVAL; }
    long long n = (long long)nx * ny * nz;
    if (blocks_for(n) > 0x7fffffffu) { set_error("%s: volume too large", who); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    Staged t;
    int rc = stage(who, shapes, nshape, surfaces, nsurface, st, t, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_phantom_embed, dim3(blocks_for(n)), dim3(kThreads), 0, st, t.shapes(),
                       nshape, t.surfs(), smat, nx, ny, nz, dx, dy, dz, x0, y0, z0, accumulate);
    hipError_t e = hipGetLastError();
    rc = release(t, st);
    CT2_HIP(e);
    return rc;
}

int ct2_phantom_points(const ct2_shape* shapes, int nshape, const ct2_surface* surfaces,
                       int nsurface, const float* xyz, float* out, size_t n, int accumulate,
                       ct2_stream_t stream) {
    const char* who = "ct2_phantom_points";
    if (n == 0) return CT2_OK;
    if (!xyz || !out) { set_error("%s: null pointer", who); return CT2_EINVAL; }
    if (n > (size_t)0x7fffffff * kThreads) { set_error("%s: too many points", who); return CT2_EINVAL; }
    hipStream_t st = (hipStream_t)stream;
    Staged t;
    int rc = stage(who, shapes, nshape, surfaces, nsurface, st, t, true);
    if (rc) return rc;
    hipLaunchKernelGGL(k_phantom_points, dim3(blocks_for((long long)n)), dim3(kThreads), 0, st,
                       t.shapes(), nshape, t.surfs(), xyz, out, (long long)n, accumulate);
    hipError_t e = hipGetLastError();
    rc = release(t, st);
    CT2_HIP(e);
    return rc;
}

}  // extern "C"
