/*
 * ora_body.h -- body of the CPU oracle, included twice by ct2_oracle.c:
 * once with REAL=float (Fortran-faithful FP32: the reference's own loop order
 * and operation order, no FMA contraction) and once with REAL=double (FP64).
 *
 * TEST INFRASTRUCTURE ONLY. This restatement is the checker for the HIP
 * kernels; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it. The product path never calls it.
 *
 * Array conventions (logical indices equal the reference's 1-based Fortran
 * indices minus one; f2py hands the Fortran an F-ordered copy of the same
 * logical array):
 *   vol  [nx][ny][nz]   C order, element (i,j,k) at (i*ny+j)*nz+k
 *   sino [ns][nu][nv]   C order, element (p,j,k) at (p*nu+j)*nv+k
 *   ind  [ns][nu][nv]   int8, a ray is used iff ind==1
 *   fv   [ns][18]       frame vectors S, tangent, D, e_u, e_v, e_w
 *   rv   [nr][6]        ray list: source (3), bin (3)
 */

#define R(c) ((REAL)(c))

typedef struct {
    REAL s[3];      /* source */
    REAL b[3];      /* detector bin */
} FN(ray_t);

typedef struct {
    REAL d[3], o[3];
    int n[3];
} FN(grid_t);

/* Per-ray Joseph setup: driving axis choice and slopes.
 * view_proj3D.f:802-815 (x), :843-847 (y), :873-876 (z);
 * z-forced variant view_proj3D_tomo.f:111-113.  Axis x wins only when
 * strictly larger than both others; y when strictly larger than z; else z. */
typedef struct {
    int a, b, c;    /* driving axis and the two transverse axes */
    REAL r1, r2;    /* slopes d(b)/d(a), d(c)/d(a) */
    REAL trav;      /* d_a * sqrt(1 + r1^2 + r2^2) */
} FN(setup_t);

static void FN(ray_setup)(const FN(ray_t) *ry, const FN(grid_t) *g, int zonly,
                          FN(setup_t) *st)
{
    REAL diff[3];
    diff[0] = ry->b[0] - ry->s[0];
    diff[1] = ry->b[1] - ry->s[1];
    diff[2] = ry->b[2] - ry->s[2];
    int a;
    if (zonly) {
        a = 2;
    } else {
        REAL xad = ABS(diff[0]) / g->d[0];
        REAL yad = ABS(diff[1]) / g->d[1];
        REAL zad = ABS(diff[2]) / g->d[2];
        if ((xad > yad) && (xad > zad)) a = 0;
        else if (yad > zad) a = 1;
        else a = 2;
    }
    int b = (a == 0) ? 1 : 0;
    int c = (a == 2) ? 1 : 2;
    st->a = a; st->b = b; st->c = c;
    st->r1 = diff[b] / diff[a];
    st->r2 = diff[c] / diff[a];
    st->trav = g->d[a] * SQRT(R(1.0) + st->r1 * st->r1 + st->r2 * st->r2);
}

/* Sample of a ray on plane l of its driving axis.
 * Linear (view_proj3D.f:816-826): cell index floor((p-o)/d - 0.5), valid
 * 0 <= i < n-1, weights w2 = (p - (d*(i+0.5)+o))/d, w1 = 1-w2.
 * Nearest (view_proj3D.f:1028-1035): index floor((p-o)/d), valid
 * 0 <= i <= n-1 (or < n-1 for the ray-list bp/ART/EM, ray_proj3D.f:895). */
typedef struct {
    int ib, ic;
    REAL wb1, wb2, wc1, wc2;
} FN(sample_t);

static int FN(sample_linear)(const FN(ray_t) *ry, const FN(grid_t) *g,
                             const FN(setup_t) *st, int l, FN(sample_t) *sm)
{
    int a = st->a, b = st->b, c = st->c;
    REAL t = g->o[a] + g->d[a] * ((REAL)l + R(0.5));
    REAL pb = ry->s[b] + st->r1 * (t - ry->s[a]);
    REAL pc = ry->s[c] + st->r2 * (t - ry->s[a]);
    int ib = (int)FLOOR((pb - g->o[b]) / g->d[b] - R(0.5));
    int ic = (int)FLOOR((pc - g->o[c]) / g->d[c] - R(0.5));
    if (!((ib >= 0) && (ib < g->n[b] - 1) && (ic >= 0) && (ic < g->n[c] - 1)))
        return 0;
    sm->ib = ib; sm->ic = ic;
    sm->wb2 = (pb - (g->d[b] * ((REAL)ib + R(0.5)) + g->o[b])) / g->d[b];
    sm->wb1 = R(1.0) - sm->wb2;
    sm->wc2 = (pc - (g->d[c] * ((REAL)ic + R(0.5)) + g->o[c])) / g->d[c];
    sm->wc1 = R(1.0) - sm->wc2;
    return 1;
}

static int FN(sample_nearest)(const FN(ray_t) *ry, const FN(grid_t) *g,
                              const FN(setup_t) *st, int l, int strict,
                              FN(sample_t) *sm)
{
    int a = st->a, b = st->b, c = st->c;
    REAL t = g->o[a] + g->d[a] * ((REAL)l + R(0.5));
    REAL pb = ry->s[b] + st->r1 * (t - ry->s[a]);
    REAL pc = ry->s[c] + st->r2 * (t - ry->s[a]);
    int ib = (int)FLOOR((pb - g->o[b]) / g->d[b]);
    int ic = (int)FLOOR((pc - g->o[c]) / g->d[c]);
    int hb = strict ? g->n[b] - 2 : g->n[b] - 1;
    int hc = strict ? g->n[c] - 2 : g->n[c] - 1;
    if (!((ib >= 0) && (ib <= hb) && (ic >= 0) && (ic <= hc)))
        return 0;
    sm->ib = ib; sm->ic = ic;
    return 1;
}

static inline long FN(vidx)(const FN(grid_t) *g, int a, int b, int c,
                            int l, int ib, int ic)
{
    int q[3];
    q[a] = l; q[b] = ib; q[c] = ic;
    return ((long)q[0] * g->n[1] + q[1]) * g->n[2] + q[2];
}

/* Bin position of view ray (p,j,k): view_proj3D.f:784-797. */
static void FN(view_ray)(const float *fv, int p, int j, int k,
                         REAL du, REAL dv, REAL u0, REAL v0, FN(ray_t) *ry)
{
    const float *f = fv + (long)p * 18;
    REAL u = u0 + ((REAL)j + R(0.5)) * du;
    REAL v = v0 + ((REAL)k + R(0.5)) * dv;
    for (int q = 0; q < 3; ++q) {
        ry->s[q] = (REAL)f[q];
        ry->b[q] = (REAL)f[6 + q] + (REAL)f[9 + q] * u + (REAL)f[12 + q] * v;
    }
}

static void FN(list_ray)(const float *rv, int r, FN(ray_t) *ry)
{
    for (int q = 0; q < 3; ++q) {
        ry->s[q] = (REAL)rv[(long)r * 6 + q];
        ry->b[q] = (REAL)rv[(long)r * 6 + 3 + q];
    }
}

/* Weighted corner list of one ray, in the reference's order: per plane the
 * corners (b,c) = (0,0),(1,0),(0,1),(1,1) with weight trav*wb*wc
 * (view_proj3D.f:1228-1251); nearest: one entry of weight trav. */
typedef struct {
    long n, cap;
    long *idx;
    REAL *w;
} FN(list_t);

static void FN(list_push)(FN(list_t) *L, long idx, REAL w)
{
    if (L->n == L->cap) {
        L->cap = L->cap ? 2 * L->cap : 1024;
        L->idx = (long *)realloc(L->idx, L->cap * sizeof(long));
        L->w = (REAL *)realloc(L->w, L->cap * sizeof(REAL));
    }
    L->idx[L->n] = idx;
    L->w[L->n] = w;
    L->n++;
}

/* March one ray. Returns the ray sum of vol (if vol != NULL), fills the
 * corner list (if L != NULL) and the ART normaliser sum of squared weights
 * (*tot; linear: sum of raypix^2 per corner, view_proj3D.f:2086-2115;
 * nearest: sum of trav^2, view_proj3D.f:2407). */
static REAL FN(march)(const FN(ray_t) *ry, const FN(grid_t) *g, int zonly,
                      int interp, int strict, const REAL *vol,
                      FN(list_t) *L, REAL *tot)
{
    FN(setup_t) st;
    FN(ray_setup)(ry, g, zonly, &st);
    REAL total = R(0.0), sq = R(0.0);
    int a = st.a, b = st.b, c = st.c;
    for (int l = 0; l < g->n[a]; ++l) {
        FN(sample_t) sm;
        if (interp == ORA_LINEAR) {
            if (!FN(sample_linear)(ry, g, &st, l, &sm)) continue;
            long i00 = FN(vidx)(g, a, b, c, l, sm.ib, sm.ic);
            long i10 = FN(vidx)(g, a, b, c, l, sm.ib + 1, sm.ic);
            long i01 = FN(vidx)(g, a, b, c, l, sm.ib, sm.ic + 1);
            long i11 = FN(vidx)(g, a, b, c, l, sm.ib + 1, sm.ic + 1);
            if (vol) {
                REAL vi1 = sm.wb1 * vol[i00] + sm.wb2 * vol[i10];
                REAL vi2 = sm.wb1 * vol[i01] + sm.wb2 * vol[i11];
                REAL val = sm.wc1 * vi1 + sm.wc2 * vi2;
                total = total + st.trav * val;
            }
            if (L) {
                REAL w;
                w = st.trav * sm.wb1 * sm.wc1; FN(list_push)(L, i00, w); sq = sq + w * w;
                w = st.trav * sm.wb2 * sm.wc1; FN(list_push)(L, i10, w); sq = sq + w * w;
                w = st.trav * sm.wb1 * sm.wc2; FN(list_push)(L, i01, w); sq = sq + w * w;
                w = st.trav * sm.wb2 * sm.wc2; FN(list_push)(L, i11, w); sq = sq + w * w;
            }
        } else {
            if (!FN(sample_nearest)(ry, g, &st, l, strict, &sm)) continue;
            long i00 = FN(vidx)(g, a, b, c, l, sm.ib, sm.ic);
            if (vol) total = total + st.trav * vol[i00];
            if (L) { FN(list_push)(L, i00, st.trav); sq = sq + st.trav * st.trav; }
        }
    }
    if (tot) *tot = sq;
    return total;
}

static void FN(grid_init)(FN(grid_t) *g, REAL dx, REAL dy, REAL dz,
                          REAL x0, REAL y0, REAL z0, int nx, int ny, int nz)
{
    g->d[0] = dx; g->d[1] = dy; g->d[2] = dz;
    g->o[0] = x0; g->o[1] = y0; g->o[2] = z0;
    g->n[0] = nx; g->n[1] = ny; g->n[2] = nz;
}

/* ---- views: forward (sinoproj_linear view_proj3D.f:703-910,
 *      sinoproj_nearest :915-1092, sinoprojz view_proj3D_tomo.f:4-148) ---- */
EXPORT void FN(ora_sinoproj)(int interp, int zonly, REAL *sino, const int8_t *ind,
                             const float *fv, int ns, int nu, int nv,
                             REAL du, REAL dv, REAL u0, REAL v0, const REAL *vol,
                             REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                             int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    long nray = (long)ns * nu * nv;
#pragma omp parallel for schedule(dynamic, 64)
    for (long r = 0; r < nray; ++r) {
        int p = (int)(r / ((long)nu * nv));
        int j = (int)((r / nv) % nu);
        int k = (int)(r % nv);
        if (ind && ind[r] != 1) { sino[r] = R(0.0); continue; }
        FN(ray_t) ry;
        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
        sino[r] = FN(march)(&ry, &g, zonly, interp, 0, vol, NULL, NULL);
    }
}

/* ---- views: exact-adjoint backprojection (sinobackproject_linear
 *      view_proj3D.f:1096-1370, _nearest :1374-1595). Output zero-filled
 *      first (f2py intent(out) zero-fill for the linear one, explicit loop
 *      view_proj3D.f:1411-1417 for the nearest one). Scatter order = ray
 *      order (p, j, k), corners in list order. zonly gives the transpose of
 *      sinoprojz (no reference counterpart). ---- */
EXPORT void FN(ora_sinobackproject)(int interp, int zonly, REAL *vol, const REAL *sino,
                                    const int8_t *ind, const float *fv,
                                    int ns, int nu, int nv, REAL du, REAL dv,
                                    REAL u0, REAL v0,
                                    REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                    int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    memset(vol, 0, sizeof(REAL) * (size_t)nx * ny * nz);
    FN(list_t) L = {0, 0, NULL, NULL};
    long nray = (long)ns * nu * nv;
    for (long r = 0; r < nray; ++r) {
        if (ind && ind[r] != 1) continue;
        int p = (int)(r / ((long)nu * nv));
        int j = (int)((r / nv) % nu);
        int k = (int)(r % nv);
        FN(ray_t) ry;
        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
        L.n = 0;
        FN(march)(&ry, &g, zonly, interp, 0, NULL, &L, NULL);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e)
            vol[L.idx[e]] = vol[L.idx[e]] + dataval * L.w[e];
    }
    free(L.idx); free(L.w);
}

/* ---- views: voxel-driven backprojection, not an adjoint
 *      (pd_backproj_linear view_proj3D.f:1769-1944,
 *       pd_backproj_nearest :1598-1766) ---- */
EXPORT void FN(ora_pd_backproj)(int interp, REAL *vol, const REAL *sino,
                                const float *fv, int ns, int nu, int nv,
                                REAL du, REAL dv, REAL u0, REAL v0,
                                REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                int nx, int ny, int nz)
{
    REAL *pv = (REAL *)malloc(sizeof(REAL) * 16 * (size_t)(ns > 0 ? ns : 1));
    /* per-view: S(3) D(3) eu(3) ev(3) ew(3) sd_dist (view_proj3D.f:1699-1714) */
    for (int p = 0; p < ns; ++p) {
        const float *f = fv + (long)p * 18;
        REAL S[3] = {f[0], f[1], f[2]}, D[3] = {f[6], f[7], f[8]};
        REAL eu[3] = {f[9], f[10], f[11]}, ev[3] = {f[12], f[13], f[14]};
        REAL ew[3] = {f[15], f[16], f[17]};
        REAL udr = (D[0] - S[0]) * eu[0] + (D[1] - S[1]) * eu[1] + (D[2] - S[2]) * eu[2];
        REAL vdr = (D[0] - S[0]) * ev[0] + (D[1] - S[1]) * ev[1] + (D[2] - S[2]) * ev[2];
        REAL tc[3];
        for (int q = 0; q < 3; ++q) tc[q] = D[q] - eu[q] * udr - ev[q] * vdr;
        REAL sdd = SQRT((S[0] - tc[0]) * (S[0] - tc[0]) + (S[1] - tc[1]) * (S[1] - tc[1])
                        + (S[2] - tc[2]) * (S[2] - tc[2]));
        REAL *o = pv + 16 * p;
        for (int q = 0; q < 3; ++q) {
            o[q] = S[q]; o[3 + q] = D[q]; o[6 + q] = eu[q]; o[9 + q] = ev[q]; o[12 + q] = ew[q];
        }
        o[15] = sdd;
    }
#pragma omp parallel for collapse(2) schedule(static)
    for (int ix = 0; ix < nx; ++ix) {
        for (int iy = 0; iy < ny; ++iy) {
            REAL x = x0 + ((REAL)ix + R(0.5)) * dx;
            REAL y = y0 + ((REAL)iy + R(0.5)) * dy;
            for (int iz = 0; iz < nz; ++iz) {
                REAL z = z0 + ((REAL)iz + R(0.5)) * dz;
                REAL total = R(0.0);
                for (int p = 0; p < ns; ++p) {
                    const REAL *o = pv + 16 * p;
                    REAL xd = x - o[0], yd = y - o[1], zd = z - o[2];
                    REAL zoom = ABS((xd * o[12] + yd * o[13] + zd * o[14]) / o[15]);
                    REAL xb = xd / zoom + o[0], yb = yd / zoom + o[1], zb = zd / zoom + o[2];
                    REAL xdt = xb - o[3], ydt = yb - o[4], zdt = zb - o[5];
                    REAL u = (xdt * o[6]) + (ydt * o[7]) + (zdt * o[8]);
                    REAL v = (xdt * o[9]) + (ydt * o[10]) + (zdt * o[11]);
                    const REAL *sp = sino + (long)p * nu * nv;
                    if (interp == ORA_LINEAR) {
                        /* iu is the 1-based left bin, valid 1 <= iu < nu */
                        int iu = (int)FLOOR((u - u0) / du + R(0.5));
                        int iv = (int)FLOOR((v - v0) / dv + R(0.5));
                        if ((iu >= 1) && (iu < nu) && (iv >= 1) && (iv < nv)) {
                            REAL wu2 = (u - u0) / du - ((REAL)iu - R(0.5));
                            REAL wu1 = R(1.0) - wu2;
                            REAL wv2 = (v - v0) / dv - ((REAL)iv - R(0.5));
                            REAL wv1 = R(1.0) - wv2;
                            long a00 = (long)(iu - 1) * nv + (iv - 1);
                            REAL vi1 = wu1 * sp[a00] + wu2 * sp[a00 + nv];
                            REAL vi2 = wu1 * sp[a00 + 1] + wu2 * sp[a00 + nv + 1];
                            REAL val = wv1 * vi1 + wv2 * vi2;
                            total = total + val;
                        }
                    } else {
                        int iu = (int)FLOOR((u - u0) / du);
                        int iv = (int)FLOOR((v - v0) / dv);
                        if ((iu >= 0) && (iu <= nu - 1) && (iv >= 0) && (iv <= nv - 1))
                            total = total + sp[(long)iu * nv + iv];
                    }
                }
                vol[((long)ix * ny + iy) * nz + iz] = total;
            }
        }
    }
    free(pv);
}

/* ---- views: sequential Kaczmarz ART (artstep_linear view_proj3D.f:1950-2271,
 *      artstep_nearest :2275-2503, artstepz view_proj3D_tomo.f:154-347).
 *      x += reg_fac*(b - a.x)*w/total, ray after ray in (p,j,k) order. ---- */
EXPORT void FN(ora_artstep)(int interp, int zonly, REAL reg_fac, const REAL *sino,
                            const int8_t *ind, const float *fv, int ns, int nu, int nv,
                            REAL du, REAL dv, REAL u0, REAL v0, REAL *vol,
                            REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                            int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    FN(list_t) L = {0, 0, NULL, NULL};
    long nray = (long)ns * nu * nv;
    for (long r = 0; r < nray; ++r) {
        if (ind && ind[r] != 1) continue;
        int p = (int)(r / ((long)nu * nv));
        int j = (int)((r / nv) % nu);
        int k = (int)(r % nv);
        FN(ray_t) ry;
        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
        L.n = 0;
        REAL total;
        REAL raysum = FN(march)(&ry, &g, zonly, interp, 0, vol, &L, &total);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e)
            vol[L.idx[e]] = vol[L.idx[e]] + reg_fac * (dataval - raysum) * L.w[e] / total;
    }
    free(L.idx); free(L.w);
}

/* ---- views: two-volume z-driven ART (artstepz2vol view_proj3D_tomo.f:352-612):
 *      one raysum / total over the main then the auxiliary grid, then both
 *      grids updated with the same residual. ---- */
EXPORT void FN(ora_artstepz2vol)(REAL reg_fac, const REAL *sino, const int8_t *ind,
                                 const float *fv, int ns, int nu, int nv,
                                 REAL du, REAL dv, REAL u0, REAL v0,
                                 REAL *vol, REAL dx, REAL dy, REAL dz,
                                 REAL x0, REAL y0, REAL z0, int nx, int ny, int nz,
                                 REAL *vola, REAL dxa, REAL dya, REAL dza,
                                 REAL x0a, REAL y0a, REAL z0a, int nxa, int nya, int nza)
{
    FN(grid_t) g, ga;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    FN(grid_init)(&ga, dxa, dya, dza, x0a, y0a, z0a, nxa, nya, nza);
    FN(list_t) L = {0, 0, NULL, NULL}, La = {0, 0, NULL, NULL};
    long nray = (long)ns * nu * nv;
    for (long r = 0; r < nray; ++r) {
        if (ind && ind[r] != 1) continue;
        int p = (int)(r / ((long)nu * nv));
        int j = (int)((r / nv) % nu);
        int k = (int)(r % nv);
        FN(ray_t) ry;
        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
        /* The reference accumulates raysum and total across both grids in
         * one running sum: main grid first, auxiliary second. */
        FN(setup_t) st, sta;
        FN(ray_setup)(&ry, &g, 1, &st);
        FN(ray_setup)(&ry, &ga, 1, &sta);
        REAL raysum = R(0.0), total = R(0.0);
        L.n = 0; La.n = 0;
        for (int pass = 0; pass < 2; ++pass) {
            const FN(grid_t) *gg = pass ? &ga : &g;
            const FN(setup_t) *ss = pass ? &sta : &st;
            const REAL *vv = pass ? vola : vol;
            FN(list_t) *LL = pass ? &La : &L;
            for (int l = 0; l < gg->n[2]; ++l) {
                FN(sample_t) sm;
                if (!FN(sample_linear)(&ry, gg, ss, l, &sm)) continue;
                long i00 = FN(vidx)(gg, 2, 0, 1, l, sm.ib, sm.ic);
                long i10 = FN(vidx)(gg, 2, 0, 1, l, sm.ib + 1, sm.ic);
                long i01 = FN(vidx)(gg, 2, 0, 1, l, sm.ib, sm.ic + 1);
                long i11 = FN(vidx)(gg, 2, 0, 1, l, sm.ib + 1, sm.ic + 1);
                REAL vi1 = sm.wb1 * vv[i00] + sm.wb2 * vv[i10];
                REAL vi2 = sm.wb1 * vv[i01] + sm.wb2 * vv[i11];
                REAL val = sm.wc1 * vi1 + sm.wc2 * vi2;
                raysum = raysum + ss->trav * val;
                REAL w;
                w = ss->trav * sm.wb1 * sm.wc1; FN(list_push)(LL, i00, w); total = total + w * w;
                w = ss->trav * sm.wb2 * sm.wc1; FN(list_push)(LL, i10, w); total = total + w * w;
                w = ss->trav * sm.wb1 * sm.wc2; FN(list_push)(LL, i01, w); total = total + w * w;
                w = ss->trav * sm.wb2 * sm.wc2; FN(list_push)(LL, i11, w); total = total + w * w;
            }
        }
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e)
            vol[L.idx[e]] = vol[L.idx[e]] + reg_fac * (dataval - raysum) * L.w[e] / total;
        for (long e = 0; e < La.n; ++e)
            vola[La.idx[e]] = vola[La.idx[e]] + reg_fac * (dataval - raysum) * La.w[e] / total;
    }
    free(L.idx); free(L.w); free(La.idx); free(La.w);
}

/* ---- views: MLEM (emstep_linear view_proj3D.f:2507-2831,
 *      emstep_nearest :2834-3075, emstep_linearz view_proj3D_tomo.f:615-822).
 *      raysum < small -> 1; W += (b/raysum)*w; O += w; x = x*W/O with
 *      O < small -> 1.  small = 1e-5 (linear), 1e-8 (nearest). ---- */
EXPORT void FN(ora_emstep)(int interp, int zonly, const REAL *sino, const int8_t *ind,
                           const float *fv, int ns, int nu, int nv,
                           REAL du, REAL dv, REAL u0, REAL v0, REAL *vol,
                           REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                           int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    size_t nvox = (size_t)nx * ny * nz;
    REAL *ws = (REAL *)calloc(nvox, sizeof(REAL));
    REAL *one = (REAL *)calloc(nvox, sizeof(REAL));
    REAL small = (interp == ORA_LINEAR) ? (REAL)0.00001f : (REAL)0.00000001f;
    FN(list_t) L = {0, 0, NULL, NULL};
    long nray = (long)ns * nu * nv;
    for (long r = 0; r < nray; ++r) {
        if (ind && ind[r] != 1) continue;
        int p = (int)(r / ((long)nu * nv));
        int j = (int)((r / nv) % nu);
        int k = (int)(r % nv);
        FN(ray_t) ry;
        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
        L.n = 0;
        REAL raysum = FN(march)(&ry, &g, zonly, interp, 0, vol, &L, NULL);
        if (raysum < small) raysum = R(1.0);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e) {
            ws[L.idx[e]] = ws[L.idx[e]] + (dataval / raysum) * L.w[e];
            one[L.idx[e]] = one[L.idx[e]] + L.w[e];
        }
    }
    for (size_t q = 0; q < nvox; ++q) {
        if (one[q] < small) one[q] = R(1.0);
        vol[q] = vol[q] * ws[q] / one[q];
    }
    free(ws); free(one); free(L.idx); free(L.w);
}

/* ---- ray list ("linogram"): ray_proj3D.f.  Forward linear :514-671,
 *      forward nearest :677-804 (bounds <= n-1); backprojection nearest
 *      :809-980, ART linear :987-1267, ART nearest :1272-1449, EM nearest
 *      :1454-1651 -- the nearest ones there use bounds < n-1. ---- */
EXPORT void FN(ora_ray_sinoproj)(int interp, REAL *sino, const int8_t *ind,
                                 const float *rv, int nr, const REAL *vol,
                                 REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                 int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
#pragma omp parallel for schedule(dynamic, 64)
    for (int r = 0; r < nr; ++r) {
        if (ind && ind[r] != 1) { sino[r] = R(0.0); continue; }
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        sino[r] = FN(march)(&ry, &g, 0, interp, 0, vol, NULL, NULL);
    }
}

EXPORT void FN(ora_ray_sinobackproject)(int interp, int strict, REAL *vol, const REAL *sino,
                                        const int8_t *ind, const float *rv, int nr,
                                        REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                        int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    memset(vol, 0, sizeof(REAL) * (size_t)nx * ny * nz);
    FN(list_t) L = {0, 0, NULL, NULL};
    for (int r = 0; r < nr; ++r) {
        if (ind && ind[r] != 1) continue;
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        L.n = 0;
        FN(march)(&ry, &g, 0, interp, strict, NULL, &L, NULL);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e)
            vol[L.idx[e]] = vol[L.idx[e]] + dataval * L.w[e];
    }
    free(L.idx); free(L.w);
}

EXPORT void FN(ora_ray_artstep)(int interp, int strict, REAL reg_fac, const REAL *sino,
                                const int8_t *ind, const float *rv, int nr, REAL *vol,
                                REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    FN(list_t) L = {0, 0, NULL, NULL};
    for (int r = 0; r < nr; ++r) {
        if (ind && ind[r] != 1) continue;
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        L.n = 0;
        REAL total;
        REAL raysum = FN(march)(&ry, &g, 0, interp, strict, vol, &L, &total);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e)
            vol[L.idx[e]] = vol[L.idx[e]] + reg_fac * (dataval - raysum) * L.w[e] / total;
    }
    free(L.idx); free(L.w);
}

EXPORT void FN(ora_ray_emstep)(int interp, int strict, const REAL *sino, const int8_t *ind,
                               const float *rv, int nr, REAL *vol,
                               REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                               int nx, int ny, int nz)
{
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    size_t nvox = (size_t)nx * ny * nz;
    REAL *ws = (REAL *)calloc(nvox, sizeof(REAL));
    REAL *one = (REAL *)calloc(nvox, sizeof(REAL));
    REAL small = (interp == ORA_LINEAR) ? (REAL)0.00001f : (REAL)0.00000001f;
    FN(list_t) L = {0, 0, NULL, NULL};
    for (int r = 0; r < nr; ++r) {
        if (ind && ind[r] != 1) continue;
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        L.n = 0;
        REAL raysum = FN(march)(&ry, &g, 0, interp, strict, vol, &L, NULL);
        if (raysum < small) raysum = R(1.0);
        REAL dataval = sino[r];
        for (long e = 0; e < L.n; ++e) {
            ws[L.idx[e]] = ws[L.idx[e]] + (dataval / raysum) * L.w[e];
            one[L.idx[e]] = one[L.idx[e]] + L.w[e];
        }
    }
    for (size_t q = 0; q < nvox; ++q) {
        if (one[q] < small) one[q] = R(1.0);
        vol[q] = vol[q] * ws[q] / one[q];
    }
    free(ws); free(one); free(L.idx); free(L.w);
}

/* ---- TV (totalvar.f).  total_var3d :35-62: Real*8 sum over i,j,k >= 1 of
 *      sqrt(eps + dx^2 + dy^2 + dz^2) + lambda*|x|, minus the sqrt(eps)
 *      shift.  total_var_grad3d :476-602: main term + three one-ahead
 *      terms, Real*8 norm, |norm|<1e-5 -> 1, optional normalisation.
 *      eps is a parameter (the reference hard-codes 1e-9). ---- */
EXPORT double FN(ora_total_var3d)(REAL lam, const REAL *s, int nx, int ny, int nz, REAL eps)
{
    double tot = 0.0;
#define V(i, j, k) s[((long)(i) * ny + (j)) * nz + (k)]
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL c = V(i, j, k);
                REAL ax = c - V(i - 1, j, k), ay = c - V(i, j - 1, k), az = c - V(i, j, k - 1);
                /* Fortran: (tot + sqrt(..)) + lambda*|x|, Real*4 terms promoted */
                tot = tot + (double)SQRT(eps + ax * ax + ay * ay + az * az);
                tot = tot + (double)(lam * ABS(c));
            }
    REAL shift = (REAL)(nx - 1) * (REAL)(ny - 1) * (REAL)(nz - 1) * SQRT(eps);
    return (double)(REAL)(tot - (double)shift);   /* Real*4 function result */
}

EXPORT double FN(ora_total_var_grad3d)(REAL lam, REAL *g, const REAL *s, int nx, int ny,
                                       int nz, int normalizeq, REAL eps)
{
    size_t nvox = (size_t)nx * ny * nz;
#define G(i, j, k) g[((long)(i) * ny + (j)) * nz + (k)]
    memset(g, 0, sizeof(REAL) * nvox);
    /* main TV term and l1 term (:505-521) */
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w5 = V(i, j, k - 1), w11 = V(i, j - 1, k), w13 = V(i - 1, j, k), w14 = V(i, j, k);
                REAL term1 = (-w11 - w13 + R(3) * w14 - w5)
                           / SQRT((w11 - w14) * (w11 - w14) + (w13 - w14) * (w13 - w14)
                                  + (-w14 + w5) * (-w14 + w5) + eps);
                REAL term2 = lam * w14 / (eps + ABS(w14));
                G(i, j, k) = term1 + term2;
            }
    /* one-ahead in x (:524-539) */
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w6 = V(i + 1, j, k - 1), w12 = V(i + 1, j - 1, k), w14 = V(i, j, k), w15 = V(i + 1, j, k);
                REAL t = (w14 - w15) / SQRT((w12 - w15) * (w12 - w15) + (w14 - w15) * (w14 - w15)
                                            + (-w15 + w6) * (-w15 + w6) + eps);
                G(i, j, k) = G(i, j, k) + t;
            }
    /* one-ahead in y (:542-556) */
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w8 = V(i, j + 1, k - 1), w14 = V(i, j, k), w16 = V(i - 1, j + 1, k), w17 = V(i, j + 1, k);
                REAL t = (w14 - w17) / SQRT((w14 - w17) * (w14 - w17) + (w16 - w17) * (w16 - w17)
                                            + (-w17 + w8) * (-w17 + w8) + eps);
                G(i, j, k) = G(i, j, k) + t;
            }
    /* one-ahead in z (:559-574) */
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 0; k < nz - 1; ++k) {
                REAL w14 = V(i, j, k), w20 = V(i, j - 1, k + 1), w22 = V(i - 1, j, k + 1), w23 = V(i, j, k + 1);
                REAL t = (w14 - w23) / SQRT((w14 - w23) * (w14 - w23) + (w20 - w23) * (w20 - w23)
                                            + (w22 - w23) * (w22 - w23) + eps);
                G(i, j, k) = G(i, j, k) + t;
            }
#undef V
#undef G
    /* Real*8 norm (:576-586); the reference stores it in a Real*4 */
    double bb = 0.0;
    for (size_t q = 0; q < nvox; ++q) bb = bb + (double)(g[q] * g[q]);   /* Real*4 square */
    REAL bobo = (REAL)sqrt(bb);
    if (ABS(bobo) < R(0.00001)) bobo = R(1.0);
    if (normalizeq == 1)
        for (size_t q = 0; q < nvox; ++q) g[q] = g[q] / bobo;
    return (double)bobo;
}

/* ---- the p-variants of totalvar.f.  REAL x**p is pow(x, p); x**2 is x*x.
 *      Values accumulate Real*4 terms in a Real*8 sum and return Real*4. ---- */
#define V(i, j, k) s[((long)(i) * ny + (j)) * nz + (k)]
#define G(i, j, k) g[((long)(i) * ny + (j)) * nz + (k)]

/* gmi (totalvar.f:4-31): gradient magnitude at voxels with all indices >= 1
 * (1-based >= 2); zero elsewhere (f2py zero-fills the intent(out) array) */
EXPORT void FN(ora_gmi)(REAL *g, const REAL *s, int nx, int ny, int nz, REAL eps)
{
    memset(g, 0, sizeof(REAL) * (size_t)nx * ny * nz);
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL c = V(i, j, k);
                REAL ax = c - V(i - 1, j, k), ay = c - V(i, j - 1, k), az = c - V(i, j, k - 1);
                G(i, j, k) = SQRT(eps + ax * ax + ay * ay + az * az);
            }
}

/* calc_lp (totalvar.f:65-87): sum of sqrt(eps + x^2)**p over every voxel */
EXPORT double FN(ora_calc_lp)(REAL p, const REAL *s, int nx, int ny, int nz, REAL eps)
{
    double tot = 0.0;
    size_t n = (size_t)nx * ny * nz;
    for (size_t q = 0; q < n; ++q) tot = tot + (double)POW(SQRT(eps + s[q] * s[q]), p);
    return (double)(REAL)tot;
}

/* total_var3d_p (totalvar.f:91-117): sum of sqrt(eps + |grad-|^2)**p minus
 * (nx-1)(ny-1)(nz-1) sqrt(eps)**p */
EXPORT double FN(ora_total_var3d_p)(REAL p, const REAL *s, int nx, int ny, int nz, REAL eps)
{
    double tot = 0.0;
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL c = V(i, j, k);
                REAL ax = c - V(i - 1, j, k), ay = c - V(i, j - 1, k), az = c - V(i, j, k - 1);
                tot = tot + (double)POW(SQRT(eps + ax * ax + ay * ay + az * az), p);
            }
    REAL shift = (REAL)(nx - 1) * (REAL)(ny - 1) * (REAL)(nz - 1) * POW(SQRT(eps), p);
    return (double)(REAL)(tot - (double)shift);
}

/* total_dvar3d_p (totalvar.f:121-146): differences divided by the voxel
 * size; no shift */
EXPORT double FN(ora_total_dvar3d_p)(REAL p, const REAL *s, int nx, int ny, int nz, REAL dx,
                                     REAL dy, REAL dz, REAL eps)
{
    double tot = 0.0;
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL c = V(i, j, k);
                REAL ax = (c - V(i - 1, j, k)) / dx, ay = (c - V(i, j - 1, k)) / dy;
                REAL az = (c - V(i, j, k - 1)) / dz;
                tot = tot + (double)POW(SQRT(eps + ax * ax + ay * ay + az * az), p);
            }
    return (double)(REAL)tot;
}

/* gradient norm bobo (Real*8 sum of Real*4 squares, |bobo| < 1e-5 -> 1) and
 * optional normalisation, as in every gradient routine of totalvar.f */
static REAL FN(tv_norm)(REAL *g, size_t nvox, int normalizeq)
{
    double bb = 0.0;
    for (size_t q = 0; q < nvox; ++q) bb = bb + (double)(g[q] * g[q]);
    REAL bobo = (REAL)sqrt(bb);
    if (ABS(bobo) < R(0.00001)) bobo = R(1.0);
    if (normalizeq == 1)
        for (size_t q = 0; q < nvox; ++q) g[q] = g[q] / bobo;
    return bobo;
}

/* lp_grad3d (totalvar.f:150-215): p*x*sqrt(x^2+eps)**(p-2) */
EXPORT double FN(ora_lp_grad3d)(REAL p, REAL *g, const REAL *s, int nx, int ny, int nz,
                                int normalizeq, REAL eps)
{
    size_t n = (size_t)nx * ny * nz;
    for (size_t q = 0; q < n; ++q) {
        REAL w14 = s[q];
        g[q] = p * w14 * POW(SQRT(w14 * w14 + eps), p - R(2));
    }
    return (double)FN(tv_norm)(g, n, normalizeq);
}

/* total_p_var_grad3d (totalvar.f:219-343): main term and the three
 * one-ahead terms, each times sqrt(..)**(p-2), in the reference's order */
EXPORT double FN(ora_total_p_var_grad3d)(REAL p, REAL *g, const REAL *s, int nx, int ny, int nz,
                                         int normalizeq, REAL eps)
{
    size_t n = (size_t)nx * ny * nz;
    REAL pm2 = p - R(2);
    memset(g, 0, sizeof(REAL) * n);
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w5 = V(i, j, k - 1), w11 = V(i, j - 1, k), w13 = V(i - 1, j, k), w14 = V(i, j, k);
                REAL term1 = (-w11 - w13 + R(3) * w14 - w5)
                           * POW(SQRT((w11 - w14) * (w11 - w14) + (w13 - w14) * (w13 - w14)
                                      + (-w14 + w5) * (-w14 + w5) + eps), pm2);
                G(i, j, k) = p * term1;
            }
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w6 = V(i + 1, j, k - 1), w12 = V(i + 1, j - 1, k), w14 = V(i, j, k), w15 = V(i + 1, j, k);
                REAL t = (w14 - w15) * POW(SQRT((w12 - w15) * (w12 - w15) + (w14 - w15) * (w14 - w15)
                                                + (-w15 + w6) * (-w15 + w6) + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w8 = V(i, j + 1, k - 1), w14 = V(i, j, k), w16 = V(i - 1, j + 1, k), w17 = V(i, j + 1, k);
                REAL t = (w14 - w17) * POW(SQRT((w14 - w17) * (w14 - w17) + (w16 - w17) * (w16 - w17)
                                                + (-w17 + w8) * (-w17 + w8) + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 0; k < nz - 1; ++k) {
                REAL w14 = V(i, j, k), w20 = V(i, j - 1, k + 1), w22 = V(i - 1, j, k + 1), w23 = V(i, j, k + 1);
                REAL t = (w14 - w23) * POW(SQRT((w14 - w23) * (w14 - w23) + (w20 - w23) * (w20 - w23)
                                                + (w22 - w23) * (w22 - w23) + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    return (double)FN(tv_norm)(g, n, normalizeq);
}

/* total_p_dvar_grad3d (totalvar.f:347-472): differences over voxel sizes */
EXPORT double FN(ora_total_p_dvar_grad3d)(REAL p, REAL *g, const REAL *s, int nx, int ny, int nz,
                                          REAL dx, REAL dy, REAL dz, int normalizeq, REAL eps)
{
    size_t n = (size_t)nx * ny * nz;
    REAL pm2 = p - R(2);
    memset(g, 0, sizeof(REAL) * n);
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w5 = V(i, j, k - 1), w11 = V(i, j - 1, k), w13 = V(i - 1, j, k), w14 = V(i, j, k);
                REAL a = (w11 - w14) / dy, b = (w13 - w14) / dx, c = (-w14 + w5) / dz;
                REAL term1 = ((w14 - w11) / (dy * dy) + (w14 - w13) / (dx * dx) + (w14 - w5) / (dz * dz))
                           * POW(SQRT(a * a + b * b + c * c + eps), pm2);
                G(i, j, k) = p * term1;
            }
    for (int i = 0; i < nx - 1; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w6 = V(i + 1, j, k - 1), w12 = V(i + 1, j - 1, k), w14 = V(i, j, k), w15 = V(i + 1, j, k);
                REAL a = (w12 - w15) / dy, b = (w14 - w15) / dx, c = (-w15 + w6) / dz;
                REAL t = ((w14 - w15) / (dx * dx)) * POW(SQRT(a * a + b * b + c * c + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    for (int i = 1; i < nx; ++i)
        for (int j = 0; j < ny - 1; ++j)
            for (int k = 1; k < nz; ++k) {
                REAL w8 = V(i, j + 1, k - 1), w14 = V(i, j, k), w16 = V(i - 1, j + 1, k), w17 = V(i, j + 1, k);
                REAL a = (w14 - w17) / dy, b = (w16 - w17) / dx, c = (-w17 + w8) / dz;
                REAL t = ((w14 - w17) / (dy * dy)) * POW(SQRT(a * a + b * b + c * c + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    for (int i = 1; i < nx; ++i)
        for (int j = 1; j < ny; ++j)
            for (int k = 0; k < nz - 1; ++k) {
                REAL w14 = V(i, j, k), w20 = V(i, j - 1, k + 1), w22 = V(i - 1, j, k + 1), w23 = V(i, j, k + 1);
                REAL a = (w14 - w23) / dz, b = (w20 - w23) / dy, c = (w22 - w23) / dx;
                REAL t = ((w14 - w23) / (dz * dz)) * POW(SQRT(a * a + b * b + c * c + eps), pm2);
                G(i, j, k) = G(i, j, k) + p * t;
            }
    return (double)FN(tv_norm)(g, n, normalizeq);
}
#undef V
#undef G

/* ---- analytic phantoms: line integrals through quadric objects
 *      (view_proj3D.f:11-603; ray_proj3D.f:11-512 holds identical root_*
 *      routines) and their voxel embeddings (tomo3DfortranAux.f:212-582).
 *      Every expression keeps the Fortran's left-to-right association;
 *      x**2 is x*x. ---- */
#define ALARGE R(100000000.0)

/* root_plane (view_proj3D.f:11-39): the half-space zs<=0 side of the plane
 * z=0 in the surface frame, as an interval of the ray parameter t */
static void FN(root_plane)(REAL *t, REAL zs, REAL zb)
{
    if (ABS(zs - zb) < R(0.00000001)) {
        if (zs >= R(0)) { t[0] = -ALARGE; t[1] = ALARGE; }
        else            { t[0] = R(1.0);  t[1] = R(-1.0); }
    } else {
        REAL root1 = zs / (zs - zb);
        if (zs < zb) { t[0] = root1; t[1] = ALARGE; }
        else         { t[0] = -ALARGE; t[1] = root1; }
    }
}

/* root_ellipsoid (view_proj3D.f:42-72) */
static void FN(root_ellipsoid)(REAL *t, REAL xs, REAL ys, REAL zs, REAL xb, REAL yb, REAL zb,
                               REAL ax, REAL ay, REAL az)
{
    REAL ax2 = ax * ax, ay2 = ay * ay, az2 = az * az;
    REAL dx = xb - xs, dy = yb - ys, dz = zb - zs;
    REAL c1 = xs * yb - xb * ys, c2 = xs * zb - xb * zs, c3 = ys * zb - yb * zs;
    REAL den = ax2 * az2 * (dy * dy) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz));
    REAL s = -(az2 * (c1 * c1)) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz) - c2 * c2)
           + ax2 * (az2 * (dy * dy) - c3 * c3);
    REAL dist2 = (R(4) * ax2 * ay2 * az2 * s) / (den * den);
    REAL avet = (ax2 * az2 * ys * (ys - yb) + ay2 * (az2 * xs * (xs - xb) + ax2 * zs * (zs - zb)))
              / den;
    if (dist2 < R(0)) { t[0] = R(1.0); t[1] = R(-1.0); }
    else {
        t[0] = avet - R(0.5) * SQRT(dist2);
        t[1] = avet + R(0.5) * SQRT(dist2);
    }
}

/* root_cylinder (view_proj3D.f:75-97): elliptic cylinder along the frame z */
static void FN(root_cylinder)(REAL *t, REAL xs, REAL ys, REAL xb, REAL yb, REAL ax, REAL ay)
{
    REAL ax2 = ax * ax, ay2 = ay * ay;
    REAL dx = xb - xs, dy = yb - ys;
    REAL c1 = xs * yb - xb * ys;
    REAL den = ay2 * (dx * dx) + ax2 * (dy * dy);
    REAL dist2 = (R(4) * ax2 * ay2 * (ay2 * (dx * dx) + ax2 * (dy * dy) - c1 * c1)) / (den * den);
    REAL avet = (ay2 * xs * (xs - xb) + ax2 * ys * (ys - yb)) / den;
    if (dist2 < R(0)) { t[0] = R(1.0); t[1] = R(-1.0); }
    else {
        t[0] = avet - R(0.5) * SQRT(dist2);
        t[1] = avet + R(0.5) * SQRT(dist2);
    }
}

/* root_cone (view_proj3D.f:100-171): the z>0 nappe of the elliptic cone
 * (x/ax)^2 + (y/ay)^2 = z^2.  The "one root on the cone" cases run as two
 * sequential tests, as in the Fortran; t keeps its previous contents when
 * neither applies. */
static void FN(root_cone)(REAL *t, REAL xs, REAL ys, REAL zs, REAL xb, REAL yb, REAL zb,
                          REAL ax, REAL ay)
{
    REAL ax2 = ax * ax, ay2 = ay * ay;
    REAL dy = yb - ys, dz = zb - zs;
    REAL c3 = ys * zb - yb * zs;
    REAL s = xs * xs * (-(yb * yb) + ay2 * (zb * zb)) + ax2 * (c3 * c3)
           + R(2) * xb * xs * (yb * ys - ay2 * zb * zs) + xb * xb * (-(ys * ys) + ay2 * (zs * zs));
    REAL den = ax2 * (dy * dy) + ay2 * (xb * xb - R(2) * xb * xs + xs * xs - ax2 * (dz * dz));
    REAL dist2 = (R(4) * ax2 * ay2 * s) / (den * den);
    REAL avet = (ax2 * ys * (ys - yb) + ay2 * (-(xb * xs) + xs * xs + ax2 * (zb - zs) * zs)) / den;
    REAL sd2 = SQRT(dist2) * R(0.5);
    if (dist2 < R(0)) { t[0] = R(1.0); t[1] = R(-1.0); return; }
    REAL tr1 = avet - sd2, tr2 = avet + sd2;
    REAL zr1 = zs * (R(1.) - tr1) + zb * tr1;
    REAL zr2 = zs * (R(1.) - tr2) + zb * tr2;
    if ((zr1 <= R(0)) && (zr2 <= R(0))) { t[0] = R(1.0); t[1] = R(-1.0); return; }
    if ((zr1 > R(0)) && (zr2 > R(0))) { t[0] = tr1; t[1] = tr2; return; }
    if (zr1 >= R(0)) {
        if (zs < zb) { t[0] = tr1; t[1] = ALARGE; } else { t[0] = -ALARGE; t[1] = tr1; }
    }
    if (zr2 >= R(0)) {
        if (zs < zb) { t[0] = tr2; t[1] = ALARGE; } else { t[0] = -ALARGE; t[1] = tr2; }
    }
}

/* Fortran MAX/MIN of two reals as the compiler lowers them (compare, then
 * select the second operand when the compare fails) */
static inline REAL FN(fmax2)(REAL a, REAL b) { return (a > b) ? a : b; }
static inline REAL FN(fmin2)(REAL a, REAL b) { return (a < b) ? a : b; }

/* One object of a general quadric phantom: sf[nsurf][12] holds, per
 * surface, the frame rows xu1..xu3, yu1..yu3, zu1..zu3 and the centre
 * x0,y0,z0, ordered planes, ellipsoids, cylinders, cones; the per-kind
 * counts and semi-axis arrays carry the reference's leading dummy entry
 * (its loops start at 2). */
typedef struct {
    const REAL *sf;
    int nplane, nell, ncyl, ncone;
    const REAL *axe, *aye, *aze, *axc, *ayc, *axk, *ayk;
} FN(genobj_t);

static void FN(to_frame)(const REAL *u, REAL x, REAL y, REAL z, REAL *o)
{
    REAL xt = x - u[9], yt = y - u[10], zt = z - u[11];
    o[0] = u[0] * xt + u[1] * yt + u[2] * zt;
    o[1] = u[3] * xt + u[4] * yt + u[5] * zt;
    o[2] = u[6] * xt + u[7] * yt + u[8] * zt;
}

/* tlen of gen_shape_proj (view_proj3D.f:306-455): the parameter length of
 * the chord inside every surface, 0 as soon as the interval empties */
static REAL FN(gen_tlen)(const FN(genobj_t) *ob, const REAL S[3], const REAL B[3], REAL *troot)
{
    REAL f1 = -ALARGE, f2 = ALARGE;
    int counts[4] = {ob->nplane, ob->nell, ob->ncyl, ob->ncone};
    int is = 0;
    for (int kind = 0; kind < 4; ++kind)
        for (int q = 1; q < counts[kind]; ++q, ++is) {
            const REAL *u = ob->sf + 12 * is;
            REAL s[3], b[3];
            FN(to_frame)(u, S[0], S[1], S[2], s);
            FN(to_frame)(u, B[0], B[1], B[2], b);
            switch (kind) {
            case 0: FN(root_plane)(troot, s[2], b[2]); break;
            case 1: FN(root_ellipsoid)(troot, s[0], s[1], s[2], b[0], b[1], b[2],
                                       ob->axe[q], ob->aye[q], ob->aze[q]); break;
            case 2: FN(root_cylinder)(troot, s[0], s[1], b[0], b[1], ob->axc[q], ob->ayc[q]); break;
            default: FN(root_cone)(troot, s[0], s[1], s[2], b[0], b[1], b[2],
                                   ob->axk[q], ob->ayk[q]); break;
            }
            f1 = FN(fmax2)(f1, troot[0]);
            f2 = FN(fmin2)(f2, troot[1]);
            if (f1 >= f2) return R(0.);
        }
    return f2 - f1;
}

static REAL FN(gen_value)(const FN(genobj_t) *ob, const REAL S[3], const REAL B[3], REAL att,
                          REAL *troot)
{
    REAL sbdist = SQRT((S[0] - B[0]) * (S[0] - B[0]) + (S[1] - B[1]) * (S[1] - B[1])
                       + (S[2] - B[2]) * (S[2] - B[2]));
    REAL tlen = FN(gen_tlen)(ob, S, B, troot);
    return sbdist * tlen * att;
}

/* ellipsoid_proj body (view_proj3D.f:557-597): chord length through the
 * ellipsoid rotated by beta about z, then alpha about the new y */
static REAL FN(ell_value)(const REAL S[3], const REAL B[3], REAL ax, REAL ay, REAL az,
                          REAL x0, REAL y0, REAL z0, REAL alpha, REAL beta, REAL att)
{
    REAL ca = COS(alpha), sa = SIN(alpha), cb = COS(beta), sb = SIN(beta);
    REAL xspp = (S[0] - x0) * cb - (y0 - S[1]) * sb;
    REAL yspp = (S[1] - y0) * cb + (x0 - S[0]) * sb;
    REAL zspp = S[2] - z0;
    REAL xsp = xspp * ca + zspp * sa, ysp = yspp, zsp = -(xspp * sa) + zspp * ca;
    REAL xbpp = (B[0] - x0) * cb - (y0 - B[1]) * sb;
    REAL ybpp = (B[1] - y0) * cb + (x0 - B[0]) * sb;
    REAL zbpp = B[2] - z0;
    REAL xbp = xbpp * ca + zbpp * sa, ybp = ybpp, zbp = -(xbpp * sa) + zbpp * ca;
    REAL ax2 = ax * ax, ay2 = ay * ay, az2 = az * az;
    REAL dx = xbp - xsp, dy = ybp - ysp, dz = zbp - zsp;
    REAL c1 = xsp * ybp - xbp * ysp, c2 = xsp * zbp - xbp * zsp, c3 = ysp * zbp - ybp * zsp;
    REAL s = -(az2 * (c1 * c1)) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz) - c2 * c2)
           + ax2 * (az2 * (dy * dy) - c3 * c3);
    REAL den = ax2 * az2 * (dy * dy) + ay2 * (az2 * (dx * dx) + ax2 * (dz * dz));
    REAL pl2 = (R(4) * ax2 * ay2 * az2 * (dx * dx + dy * dy + dz * dz) * s) / (den * den);
    if (pl2 < R(0.)) pl2 = R(0.);
    return SQRT(pl2) * att;
}

/* gen_shape_proj over ns views (view_proj3D.f:175-455, called once per view
 * from tomo3D.py:2531-2550); unused bins are zero as f2py leaves them */
EXPORT void FN(ora_gen_shape_proj)(REAL *sino, const int8_t *ind, const float *fv, int ns, int nu,
                                   int nv, REAL du, REAL dv, REAL u0, REAL v0, REAL att,
                                   const REAL *sf, int nplane, int nell, const REAL *axe,
                                   const REAL *aye, const REAL *aze, int ncyl, const REAL *axc,
                                   const REAL *ayc, int ncone, const REAL *axk, const REAL *ayk)
{
    FN(genobj_t) ob = {sf, nplane, nell, ncyl, ncone, axe, aye, aze, axc, ayc, axk, ayk};
    REAL troot[2] = {R(0), R(0)};
    for (int p = 0; p < ns; ++p)
        for (int j = 0; j < nu; ++j)
            for (int k = 0; k < nv; ++k) {
                long q = ((long)p * nu + j) * nv + k;
                sino[q] = R(0);
                if (ind[q] != 1) continue;
                FN(ray_t) ry;
                FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                sino[q] = FN(gen_value)(&ob, ry.s, ry.b, att, troot);
            }
}

/* ellipsoid_proj over ns views (view_proj3D.f:483-603) */
EXPORT void FN(ora_ellipsoid_proj)(REAL *sino, const int8_t *ind, const float *fv, int ns, int nu,
                                   int nv, REAL du, REAL dv, REAL u0, REAL v0, REAL ax, REAL ay,
                                   REAL az, REAL x0, REAL y0, REAL z0, REAL alpha, REAL beta,
                                   REAL att)
{
    for (int p = 0; p < ns; ++p)
        for (int j = 0; j < nu; ++j)
            for (int k = 0; k < nv; ++k) {
                long q = ((long)p * nu + j) * nv + k;
                sino[q] = R(0);
                if (ind[q] != 1) continue;
                FN(ray_t) ry;
                FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                sino[q] = FN(ell_value)(ry.s, ry.b, ax, ay, az, x0, y0, z0, alpha, beta, att);
            }
}

/* ray-list forms (ray_proj3D.f:175-431, :435-512) */
EXPORT void FN(ora_ray_gen_shape_proj)(REAL *sino, const int8_t *ind, const float *rv, int nr,
                                       REAL att, const REAL *sf, int nplane, int nell,
                                       const REAL *axe, const REAL *aye, const REAL *aze,
                                       int ncyl, const REAL *axc, const REAL *ayc, int ncone,
                                       const REAL *axk, const REAL *ayk)
{
    FN(genobj_t) ob = {sf, nplane, nell, ncyl, ncone, axe, aye, aze, axc, ayc, axk, ayk};
    REAL troot[2] = {R(0), R(0)};
    for (int r = 0; r < nr; ++r) {
        sino[r] = R(0);
        if (ind[r] != 1) continue;
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        sino[r] = FN(gen_value)(&ob, ry.s, ry.b, att, troot);
    }
}

EXPORT void FN(ora_ray_ellipsoid_proj)(REAL *sino, const int8_t *ind, const float *rv, int nr,
                                       REAL ax, REAL ay, REAL az, REAL x0, REAL y0, REAL z0,
                                       REAL alpha, REAL beta, REAL att)
{
    for (int r = 0; r < nr; ++r) {
        sino[r] = R(0);
        if (ind[r] != 1) continue;
        FN(ray_t) ry;
        FN(list_ray)(rv, r, &ry);
        sino[r] = FN(ell_value)(ry.s, ry.b, ax, ay, az, x0, y0, z0, alpha, beta, att);
    }
}

/* ellipsoid_embed (tomo3DfortranAux.f:212-267): att where the voxel centre
 * is strictly inside, zero elsewhere */
EXPORT void FN(ora_ellipsoid_embed)(REAL *s, int nx, int ny, int nz, REAL dx, REAL dy, REAL dz,
                                    REAL x0i, REAL y0i, REAL z0i, REAL att, REAL alpha, REAL beta,
                                    REAL x0, REAL y0, REAL z0, REAL ax, REAL ay, REAL az)
{
    REAL ca = COS(alpha), sa = SIN(alpha), cb = COS(beta), sb = SIN(beta);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                REAL x = x0i + ((REAL)i + R(0.5)) * dx;
                REAL y = y0i + ((REAL)j + R(0.5)) * dy;
                REAL z = z0i + ((REAL)k + R(0.5)) * dz;
                REAL rx = x - x0, ry = y - y0, rz = z - z0;
                REAL r1x = cb * rx + sb * ry, r1y = -(sb * rx) + cb * ry, r1z = rz;
                REAL r2x = ca * r1x + sa * r1z, r2y = r1y, r2z = -(sa * r1x) + ca * r1z;
                REAL qx = r2x / ax, qy = r2y / ay, qz = r2z / az;
                REAL lhs = qx * qx + qy * qy + qz * qz;
                s[((long)i * ny + j) * nz + k] = (lhs < R(1.)) ? att : R(0);
            }
}

/* gen_obj_embed (tomo3DfortranAux.f:420-582): att times the product of
 * the surfaces' step functions at the voxel centre */
EXPORT void FN(ora_gen_obj_embed)(REAL *s, int nx, int ny, int nz, REAL dx, REAL dy, REAL dz,
                                  REAL x0i, REAL y0i, REAL z0i, REAL att, const REAL *sf,
                                  int nplane, int nell, const REAL *axe, const REAL *aye,
                                  const REAL *aze, int ncyl, const REAL *axc, const REAL *ayc,
                                  int ncone, const REAL *axk, const REAL *ayk)
{
    int counts[4] = {nplane, nell, ncyl, ncone};
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                REAL x = x0i + ((REAL)i + R(0.5)) * dx;
                REAL y = y0i + ((REAL)j + R(0.5)) * dy;
                REAL z = z0i + ((REAL)k + R(0.5)) * dz;
                REAL si = R(1.0);
                int is = 0;
                for (int kind = 0; kind < 4; ++kind)
                    for (int q = 1; q < counts[kind]; ++q, ++is) {
                        REAL f[3];
                        FN(to_frame)(sf + 12 * is, x, y, z, f);
                        REAL e;
                        switch (kind) {
                        case 0: si = si * (f[2] >= R(0) ? R(1) : R(0)); break;
                        case 1:
                            e = (f[0] / axe[q]) * (f[0] / axe[q]) + (f[1] / aye[q]) * (f[1] / aye[q])
                              + (f[2] / aze[q]) * (f[2] / aze[q]);
                            si = si * ((R(1.0) - e) >= R(0) ? R(1) : R(0));
                            break;
                        case 2:
                            e = (f[0] / axc[q]) * (f[0] / axc[q]) + (f[1] / ayc[q]) * (f[1] / ayc[q]);
                            si = si * ((R(1.0) - e) >= R(0) ? R(1) : R(0));
                            break;
                        default:
                            e = (f[0] / axk[q]) * (f[0] / axk[q]) + (f[1] / ayk[q]) * (f[1] / ayk[q]);
                            si = si * ((f[2] * f[2] - e) >= R(0) ? R(1) : R(0))
                                    * (f[2] >= R(0) ? R(1) : R(0));
                            break;
                        }
                    }
                s[((long)i * ny + j) * nz + k] = att * si;
            }
}
/* ellipse_voxval (tomo3DfortranAux.f:179-210) at n points xyz[n][3]: att
 * strictly inside the ellipsoid, zero elsewhere */
EXPORT void FN(ora_ellipse_voxval)(REAL *out, const REAL *xyz, long n, REAL att, REAL alpha,
                                   REAL beta, REAL x0, REAL y0, REAL z0, REAL ax, REAL ay,
                                   REAL az)
{
    REAL ca = COS(alpha), sa = SIN(alpha), cb = COS(beta), sb = SIN(beta);
    for (long q = 0; q < n; ++q) {
        REAL rx = xyz[3 * q] - x0, ry = xyz[3 * q + 1] - y0, rz = xyz[3 * q + 2] - z0;
        REAL r1x = cb * rx + sb * ry, r1y = -(sb * rx) + cb * ry, r1z = rz;
        REAL r2x = ca * r1x + sa * r1z, r2y = r1y, r2z = -(sa * r1x) + ca * r1z;
        REAL lhs = (r2x / ax) * (r2x / ax) + (r2y / ay) * (r2y / ay) + (r2z / az) * (r2z / az);
        out[q] = (lhs < R(1.)) ? att : R(0);
    }
}

/* framer (tomo3DfortranAux.f:281-308): (x, y, z) in the frame turned by
 * alpha about z, then beta, then gam, about the centre a[3..5]; a = (alpha,
 * beta, gam, x0, y0, z0) */
static void FN(framer)(REAL x, REAL y, REAL z, const REAL *a, REAL *o)
{
    REAL ca = COS(a[0]), sa = SIN(a[0]), cb = COS(a[1]), sb = SIN(a[1]);
    REAL cg = COS(a[2]), sg = SIN(a[2]);
    REAL xp = (x - a[3]) * ca - (a[4] - y) * sa;
    REAL yp = (y - a[4]) * ca + (a[3] - x) * sa;
    REAL zp = z - a[5];
    REAL xpp = xp * cb - zp * sb, ypp = yp, zpp = xp * sb + zp * cb;
    o[0] = xpp * cg + ypp * sg;
    o[1] = -(xpp * sg) + ypp * cg;
    o[2] = zpp;
}

/* gen_obj_voxval (tomo3DfortranAux.f:311-416) at n points: att times the
 * product of the surfaces' steps, the surfaces' frames by framer from the
 * rows sf[nsurf][6] = (alpha, beta, gam, x0, y0, z0); counts and semi-axis
 * arrays carry the reference's leading dummy entry */
EXPORT void FN(ora_gen_obj_voxval)(REAL *out, const REAL *xyz, long n, REAL att, const REAL *sf,
                                   int nplane, int nell, const REAL *axe, const REAL *aye,
                                   const REAL *aze, int ncyl, const REAL *axc, const REAL *ayc,
                                   int ncone, const REAL *axk, const REAL *ayk)
{
    int counts[4] = {nplane, nell, ncyl, ncone};
    for (long p = 0; p < n; ++p) {
        REAL si = R(1.0);
        int is = 0;
        for (int kind = 0; kind < 4; ++kind)
            for (int q = 1; q < counts[kind]; ++q, ++is) {
                REAL f[3], e;
                FN(framer)(xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2], sf + 6 * is, f);
                if (kind == 0) {
                    si = si * (f[2] >= R(0) ? R(1) : R(0));
                } else if (kind == 1) {
                    e = (f[0] / axe[q]) * (f[0] / axe[q]) + (f[1] / aye[q]) * (f[1] / aye[q])
                      + (f[2] / aze[q]) * (f[2] / aze[q]);
                    si = si * ((R(1.0) - e) >= R(0) ? R(1) : R(0));
                } else if (kind == 2) {
                    e = (f[0] / axc[q]) * (f[0] / axc[q]) + (f[1] / ayc[q]) * (f[1] / ayc[q]);
                    si = si * ((R(1.0) - e) >= R(0) ? R(1) : R(0));
                } else {
                    e = (f[0] / axk[q]) * (f[0] / axk[q]) + (f[1] / ayk[q]) * (f[1] / ayk[q]);
                    si = si * ((f[2] * f[2] - e) >= R(0) ? R(1) : R(0))
                            * (f[2] >= R(0) ? R(1) : R(0));
                }
            }
        out[p] = att * si;
    }
}

#undef ALARGE

/* ---- ray bunching (tomo3DfortranAux.f:1-176; tomo3D.py:395-399, :484-534) ---- */

/* sinoscale (tomo3DfortranAux.f:1-67): each bin divided by the distance
 * from the source to the bin centre */
EXPORT void FN(ora_sinoscale)(REAL *sino, const float *fv, int ns, int nu, int nv, REAL du,
                              REAL dv, REAL u0, REAL v0)
{
    for (int p = 0; p < ns; ++p)
        for (int j = 0; j < nu; ++j)
            for (int k = 0; k < nv; ++k) {
                FN(ray_t) ry;
                FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                REAL d = SQRT((ry.s[0] - ry.b[0]) * (ry.s[0] - ry.b[0])
                              + (ry.s[1] - ry.b[1]) * (ry.s[1] - ry.b[1])
                              + (ry.s[2] - ry.b[2]) * (ry.s[2] - ry.b[2]));
                long q = ((long)p * nu + j) * nv + k;
                sino[q] = sino[q] / d;
            }
}

/* sinoave (tomo3DfortranAux.f:71-122): out[i,j,k] = mean of the window
 * [i*nds, i*nds+nws) x [j*ndu, +nwu) x [k*ndv, +nwv), summed in the
 * Fortran's ii, jj, kk order */
EXPORT void FN(ora_sinoave)(const REAL *s, int ns, int nu, int nv, REAL *out, int nss, int nus,
                            int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv)
{
    (void)ns;
    for (int i = 0; i < nss; ++i)
        for (int j = 0; j < nus; ++j)
            for (int k = 0; k < nvs; ++k) {
                REAL acc = R(0);
                for (int ii = 0; ii < nws; ++ii)
                    for (int jj = 0; jj < nwu; ++jj)
                        for (int kk = 0; kk < nwv; ++kk)
                            acc = acc + s[((long)(i * nds + ii) * nu + (j * ndu + jj)) * nv
                                          + (k * ndv + kk)];
                out[((long)i * nus + j) * nvs + k] = acc / (REAL)(nws * nwu * nwv);
            }
}

/* sinodist (tomo3DfortranAux.f:125-176): every window of the fine
 * sinogram gets its coarse value; where windows overlap the last write in
 * the Fortran's loop order wins; uncovered bins are zero */
EXPORT void FN(ora_sinodist)(REAL *s, int ns, int nu, int nv, const REAL *in, int nss, int nus,
                             int nvs, int nds, int ndu, int ndv, int nws, int nwu, int nwv)
{
    memset(s, 0, sizeof(REAL) * (size_t)ns * nu * nv);
    for (int i = 0; i < nss; ++i)
        for (int j = 0; j < nus; ++j)
            for (int k = 0; k < nvs; ++k)
                for (int ii = 0; ii < nws; ++ii)
                    for (int jj = 0; jj < nwu; ++jj)
                        for (int kk = 0; kk < nwv; ++kk)
                            s[((long)(i * nds + ii) * nu + (j * ndu + jj)) * nv + (k * ndv + kk)]
                                = in[((long)i * nus + j) * nvs + k];
}

/* ---- weighted ray-bunch ("stencil") measurements and their ART: the
 *      semantics of the reference's cweighted_project_to / put_weights /
 *      c_ss_weighted_art_step and get_deriv_weights / cderiv_art_step
 *      (tests/cweighted_art_test.py:80-140, tests/cderiv_art_test.py:106-140;
 *      the GPU sources are missing).  Kernel entry q = (a,c,d) of
 *      w[ws][wu][wv] belongs to fine ray (I*nss+a-os, J*nsu+c-ou, K*nsv+d-ov),
 *      os = ws-1-ws/2; rays outside the sinogram are skipped. ---- */
typedef struct {
    const REAL *w;
    int ws, wu, wv, nss, nsu, nsv, NS, NU, NV;
} FN(stencil_t);

static void FN(stencil_init)(FN(stencil_t) *s, const REAL *w, int ws, int wu, int wv, int nss,
                             int nsu, int nsv, int ns, int nu, int nv)
{
    s->w = w; s->ws = ws; s->wu = wu; s->wv = wv;
    s->nss = nss; s->nsu = nsu; s->nsv = nsv;
    s->NS = (ns + nss - 1) / nss; s->NU = (nu + nsu - 1) / nsu; s->NV = (nv + nsv - 1) / nsv;
}

static int FN(stencil_ray)(const FN(stencil_t) *s, int ns, int nu, int nv, int I, int J, int K,
                           int q, int *p, int *j, int *k)
{
    int d = q % s->wv, t = q / s->wv, c = t % s->wu, a = t / s->wu;
    *p = I * s->nss + a - (s->ws - 1 - s->ws / 2);
    *j = J * s->nsu + c - (s->wu - 1 - s->wu / 2);
    *k = K * s->nsv + d - (s->wv - 1 - s->wv / 2);
    return *p >= 0 && *p < ns && *j >= 0 && *j < nu && *k >= 0 && *k < nv;
}

/* small[I,J,K] = sum_q w_q fine[ray q], in kernel order */
EXPORT void FN(ora_stencil_correlate)(const REAL *fine, int ns, int nu, int nv, const REAL *w,
                                      int ws, int wu, int wv, int nss, int nsu, int nsv,
                                      REAL *small)
{
    FN(stencil_t) s;
    FN(stencil_init)(&s, w, ws, wu, wv, nss, nsu, nsv, ns, nu, nv);
    int R = ws * wu * wv;
    for (int I = 0; I < s.NS; ++I)
        for (int J = 0; J < s.NU; ++J)
            for (int K = 0; K < s.NV; ++K) {
                REAL acc = R(0);
                for (int q = 0; q < R; ++q) {
                    int p, j, k;
                    if (!FN(stencil_ray)(&s, ns, nu, nv, I, J, K, q, &p, &j, &k)) continue;
                    acc = acc + w[q] * fine[((long)p * nu + j) * nv + k];
                }
                small[((long)I * s.NU + J) * s.NV + K] = acc;
            }
}

/* weights[b] = |r_b|^2 (FP64 sum, each voxel once), image += r_b, for every
 * bin whose anchor indicator is 1 (others: weight 0) */
EXPORT void FN(ora_stencil_weights)(int interp, int zonly, const REAL *w, int ws, int wu, int wv,
                                    int nss, int nsu, int nsv, const int8_t *ind, const float *fv,
                                    int ns, int nu, int nv, REAL du, REAL dv, REAL u0, REAL v0,
                                    REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0,
                                    int nx, int ny, int nz, REAL *weights, REAL *image)
{
    FN(stencil_t) s;
    FN(stencil_init)(&s, w, ws, wu, wv, nss, nsu, nsv, ns, nu, nv);
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    REAL *scratch = (REAL *)calloc((size_t)nx * ny * nz, sizeof(REAL));
    FN(list_t) L = {0, 0, NULL, NULL};
    int R = ws * wu * wv;
    for (int I = 0; I < s.NS; ++I)
        for (int J = 0; J < s.NU; ++J)
            for (int K = 0; K < s.NV; ++K) {
                long b = ((long)I * s.NU + J) * s.NV + K;
                long anchor = ((long)(I * nss) * nu + J * nsu) * nv + K * nsv;
                if (ind && ind[anchor] != 1) { weights[b] = R(0); continue; }
                for (int pass = 0; pass < 2; ++pass) {
                    double acc = 0.0;
                    for (int q = 0; q < R; ++q) {
                        int p, j, k;
                        if (!FN(stencil_ray)(&s, ns, nu, nv, I, J, K, q, &p, &j, &k)) continue;
                        FN(ray_t) ry;
                        FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                        L.n = 0;
                        FN(march)(&ry, &g, zonly, interp, 0, NULL, &L, NULL);
                        for (long e = 0; e < L.n; ++e) {
                            long i = L.idx[e];
                            if (pass == 0) {
                                scratch[i] = scratch[i] + w[q] * L.w[e];
                            } else if (scratch[i] != R(0)) {
                                REAL v = scratch[i];
                                acc += (double)v * (double)v;
                                image[i] = image[i] + v;
                                scratch[i] = R(0);
                            }
                        }
                    }
                    if (pass == 1) weights[b] = (REAL)acc;
                }
            }
    free(L.idx); free(L.w); free(scratch);
}

/* Kaczmarz over the used bins with weights > 0 in (I,J,K) order:
 * vol += reg_fac (small[b] - sum_q w_q (a_q . vol)) (w_q a_q) / weights[b] */
EXPORT void FN(ora_stencil_art)(int interp, int zonly, REAL reg_fac, const REAL *w, int ws, int wu,
                                int wv, int nss, int nsu, int nsv, const REAL *small,
                                const REAL *weights, const int8_t *ind, const float *fv, int ns,
                                int nu, int nv, REAL du, REAL dv, REAL u0, REAL v0, REAL *vol,
                                REAL dx, REAL dy, REAL dz, REAL x0, REAL y0, REAL z0, int nx,
                                int ny, int nz)
{
    FN(stencil_t) s;
    FN(stencil_init)(&s, w, ws, wu, wv, nss, nsu, nsv, ns, nu, nv);
    FN(grid_t) g;
    FN(grid_init)(&g, dx, dy, dz, x0, y0, z0, nx, ny, nz);
    FN(list_t) L = {0, 0, NULL, NULL};
    int R = ws * wu * wv;
    for (int I = 0; I < s.NS; ++I)
        for (int J = 0; J < s.NU; ++J)
            for (int K = 0; K < s.NV; ++K) {
                long b = ((long)I * s.NU + J) * s.NV + K;
                long anchor = ((long)(I * nss) * nu + J * nsu) * nv + K * nsv;
                if (ind && ind[anchor] != 1) continue;
                REAL norm = weights[b];
                if (!(norm > R(0))) continue;
                REAL rsum = R(0);
                for (int q = 0; q < R; ++q) {
                    int p, j, k;
                    if (!FN(stencil_ray)(&s, ns, nu, nv, I, J, K, q, &p, &j, &k)) continue;
                    FN(ray_t) ry;
                    FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                    rsum = rsum + w[q] * FN(march)(&ry, &g, zonly, interp, 0, vol, NULL, NULL);
                }
                REAL resid = reg_fac * (small[b] - rsum);
                for (int q = 0; q < R; ++q) {
                    int p, j, k;
                    if (!FN(stencil_ray)(&s, ns, nu, nv, I, J, K, q, &p, &j, &k)) continue;
                    FN(ray_t) ry;
                    FN(view_ray)(fv, p, j, k, du, dv, u0, v0, &ry);
                    L.n = 0;
                    FN(march)(&ry, &g, zonly, interp, 0, NULL, &L, NULL);
                    for (long e = 0; e < L.n; ++e)
                        vol[L.idx[e]] = vol[L.idx[e]] + resid * (w[q] * L.w[e]) / norm;
                }
            }
    free(L.idx); free(L.w);
}
