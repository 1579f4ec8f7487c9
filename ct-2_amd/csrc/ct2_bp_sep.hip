// ct2_bp_sep.hip -- the adjoint of the fast Joseph projector, column by
// column, for views whose detector rows run along z (view_separable).
//
// The reference scatters each ray's samples (emil_code/view_proj3D.f:
// 1228-1251, :1355-1360).  With e_v parallel to z, the rays of one detector
// column share their driving axis a (x or y) and the fast set-up of the other
// transverse axis b, so on plane l of a every ray of the column samples b cell
// ib(l) with weights 1-wb, wb (the forward's own fma(l, dub, ub0)), and only
// the z cell coordinate differs from row to row, uc(k, l) = fma(l, duc_k,
// uc0_k).  The adjoint of plane l then factorises: the column's data
// resampled onto z with the forward's linear-interpolation weights,
//     G(c) = sum_k td_k max(0, 1 - |uc(k,l) - c|),   td_k = trav_k datum_k,
// lands on the two voxel rows (l, ib, :) and (l, ib+1, :) with weights 1-wb
// and wb.
//
// The resampling uses prefix sums.  In exact arithmetic uc(k, l) = A(l) +
// B(l) k is linear in the row (the ray's z slope is linear in the detector
// v coordinate and the plane's distance is shared), with B(l) of one sign
// over the planes; rows are stored in ascending z.  The rows of cell c (c <=
// uc < c+1) are then the stored rows [r_c, r_c+1), r_c = floor(t_c) + 1 for
// the crossing t_c = (c - A)/B, evaluated in FP32 (ceil(t_c) but at integers,
// where the choice is open anyway).  Linear interpolation is
// continuous across a cell boundary, so a row that lies within the FP32
// error of t_c from it and is taken in the neighbouring cell gets weights
// that extend that cell's interpolation by as much (about 1e-4 of a cell at
// most), the order of the forward's own rounding of uc.  The volume's two z
// faces, c = 0 and c = nc-1, are cut-offs instead: the forward samples no
// row outside [0, nc-1).  Their rows r_0 and r_nc-1, the first whose forward
// uc (its FP32 value) reaches the face, are found exactly with the table:
// the forward's own set-up of the row nearest the crossing, whose uc lies
// within its rounding (below 1e-3 of a cell) of the linear value, while a
// row is at least 1/48 of a cell from the next (view_separable), settles
// whether that row or the next is the first.  Every crossing is clamped to
// [r_0, r_nc-1] (the true rows increase with c, so this only moves an
// estimate towards them), and the faces' own estimates are pushed past the
// ends, so they take r_0 and r_nc-1.  With FP64
// prefix sums PS(r) = sum_{k<r} td_k and PK(r) = sum_{k<r}
// td_k k, cell c receives S = PS(r_c+1) - PS(r_c) and
//     H = sum td_k (uc_k - c) = (A - c) S + B (PK(r_c+1) - PK(r_c)),
// its lower voxel S - H and its upper voxel H.  The weights use the exact
// linear uc where the forward uses its FP32 rounding (about 1e-5 of a cell
// apart, and of random sign), so the kernel is the transpose of the fast
// forward up to rounding.
//
// One launch takes the x-driven columns of every separable view, a second the
// y-driven ones.  A workgroup owns SEP_TA consecutive planes of the driving
// axis a, SEP_TB cells of the other axis b, and every z; wavefront w owns
// plane a0 + w.  Lane l owns the CPL z cells [l*CPL,
// (l+1)*CPL) and voxels; the upper part of its last cell moves to the next
// lane with one DPP lane shift.  Sums accumulate in FP32 registers per
// (plane, b cell, z cell), so no LDS or global atomics are needed, and each
// voxel is written by one workgroup: the x launch stores and the y launch
// adds, in stream order, so the result is deterministic.
//
// Each view's columns are taken SEP_CHUNK at a time, one per wavefront: the
// wavefront sets the column up (FP64, the forward's expressions) and copies
// its table of prefix sums to LDS.  Eight wavefronts of one plane each, at
// 128 VGPRs, let two workgroups (16 wavefronts) share a CU; four wavefronts
// of two planes each need 168 VGPRs, which leaves 12 (16.8 against 16.2 ms
// at 256^3 x 360 views x 512^2).
#include <mutex>

#include "ct2_internal.h"

namespace ct2 {

constexpr int SEP_PPW = 1;                      // planes per wavefront
constexpr int SEP_WAVES = 8;
constexpr int SEP_WPE = 4;                      // wavefronts per SIMD (128 VGPRs)
constexpr int SEP_TA = SEP_PPW * SEP_WAVES;     // tile planes along a
constexpr int SEP_TB = 8;                       // tile cells along b
constexpr int SEP_THREADS = 64 * SEP_WAVES;
constexpr int SEP_CHUNK = SEP_WAVES;            // detector columns per LDS table set
constexpr int SEP_MAXV = 512;                   // views per launch (one LDS word each)
constexpr int SEP_RMAX = 16;                    // table rows per lane in k_sep_tables

// Per (column, tile plane): the linear model uc = A + B kk over stored rows
// kk (FP64, B > 0); one past the crossing of z, t(z) + 1 = z rB + nA1
// (FP32); the face rows r_0 and r_nc-1 (as floats); the weight of each b row
// of the tile from the forward's b cell ib and weight wb: 1-wb at ib, wb at
// ib+1, 0 elsewhere.
struct SepPlane {
    double A, B;
    float rB, nA1, rbot, rtop;
    float wr[SEP_TB];
};

// A column's table: the prefix pairs (PS, PK) of entries 0 .. nv, 16 bytes
// each, in the workspace (one per view and column) and in LDS (one per
// wavefront).  (Lanes whose entries lie a multiple of 8 apart share LDS
// banks; spreading them with spare slots costs more address arithmetic than
// the conflicts cost.)
__host__ __device__ inline size_t sep_col_bytes(int nv) {
    return (size_t)16 * (nv + 1);
}
size_t sep_lds_bytes(int nv) { return SEP_CHUNK * sep_col_bytes(nv); }

// Whether a column's rows are stored reversed (uc decreasing in the row on
// the tile's planes; the sign of the plane distance T is the same on every
// plane of a separable view): one expression, so that k_sep_tables and the
// tiles agree.
__device__ __forceinline__ bool sep_rev(double R1, int P0, double da, double ta, double idc) {
    return R1 * fma((double)P0, da, ta) * idc < 0.0;
}

// x from the lane CTRL names (DPP), 0 where that lane is out of range or
// the row is masked off by RM.
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Inclusive prefix sum over the 64 lanes (every lane active).
__device__ __forceinline__ double wave_scan_f64(double x) {
    x += dpp_f64<0x111, 0xf>(x);   // row_shr:1
    x += dpp_f64<0x112, 0xf>(x);   // row_shr:2
    x += dpp_f64<0x114, 0xf>(x);   // row_shr:4
    x += dpp_f64<0x118, 0xf>(x);   // row_shr:8
    x += dpp_f64<0x142, 0xa>(x);   // row_bcast:15 into rows 1 and 3
    x += dpp_f64<0x143, 0xc>(x);   // row_bcast:31 into rows 2 and 3
    return x;
}

// Detector columns whose rays can pass through the box x in [xa, xb],
// y in [ya, yb] (every z): the box corners projected through the source
// (u does not depend on z when e_v is parallel to z), one bin to spare,
// clamped to the detector; jlo > jhi when the box's silhouette misses it.
__device__ void sep_column_range(const float* __restrict__ f, const Det& dt, double xa,
                                 double xb, double ya, double yb, int& jlo, int& jhi) {
    double umin = 1e300, umax = -1e300;
    bool ok = true;
    for (int c = 0; c < 4; ++c) {
        const double P[3] = {(c & 1) ? xb : xa, (c & 2) ? yb : ya, 0.0};
        double u = 0, v = 0;
        int okc = 0;
        project_corner(f, P, u, v, okc);
        ok &= okc != 0;
        umin = fmin(umin, u);
        umax = fmax(umax, u);
    }
    jlo = 0;
    jhi = dt.nu - 1;
    if (!ok) return;
    double jl = (umin - dt.u0) / dt.du - 0.5, jh = (umax - dt.u0) / dt.du - 0.5;
    if (dt.du < 0) { const double t = jl; jl = jh; jh = t; }
    jl = fmin(fmax(jl, -2.0), (double)dt.nu + 1.0);
    jh = fmin(fmax(jh, -2.0), (double)dt.nu + 1.0);
    jlo = max(0, (int)ceil(jl) - 1);
    jhi = min(dt.nu - 1, (int)floor(jh) + 1);
}

// Write-back of a tile's sums: store, add, or add atomically (two atomic
// adds onto a zeroed voxel commute, so that result is deterministic too), or
// (SART) add the stored sums of the other axis's tiles, corr, and update
// vol with them: vol = sart_update(vol, lambda, corr + sum, colsum).
enum { SEP_STORE = 0, SEP_ADD = 1, SEP_ATOMIC = 2, SEP_SART = 3 };

// One tile of driving axis AX: tile number `tile` of that axis's tiling.
// tabs: the column tables of views pv = 0 .. nviews-1 (k_sep_tables);
// s_tab: the dynamic LDS tables; s_pl, s_mask, s_view: the kernel's.
template <int AX, int CPL>
__device__ __forceinline__ void bp_sep_tile(int tile, const Grid& g, const Det& dt, int view0,
                                            int nviews, int vstride, int mode,
                                            float* __restrict__ vol,
                                            const double2* __restrict__ tabs,
                                            const float* __restrict__ fv, double* s_tab,
                                            SepPlane (*s_pl)[SEP_TA], unsigned* s_mask,
                                            unsigned* s_view, const float* __restrict__ corr,
                                            const float* __restrict__ colsum, float lambda) {
    constexpr int A = AX, Bx = 1 - AX;   // driving axis, other transverse axis
    const int nv = dt.nv;
    const size_t colb = sep_col_bytes(nv);
    const int nta = (g.n[A] + SEP_TA - 1) / SEP_TA;
    const int P0 = (tile % nta) * SEP_TA, Q0 = (tile / nta) * SEP_TB;
    const int nP = min(SEP_TA, g.n[A] - P0), nQ = min(SEP_TB, g.n[Bx] - Q0);
    const int i0 = A == 0 ? P0 : Q0, j0 = A == 0 ? Q0 : P0;
    const int ni = A == 0 ? nP : nQ, nj = A == 0 ? nQ : nP;
    const int nb = g.n[Bx], nc = g.n[2];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    // the fast set-up's grid constants (setup_fast_ax), FP64 (uniform: kept
    // in scalar registers)
    const double da = uniform_d((double)g.d[A]), oa = uniform_d((double)g.o[A]);
    const double ob = uniform_d((double)g.o[Bx]), oc = uniform_d((double)g.o[2]);
    const double idb = uniform_d(rcp_f64((double)g.d[Bx]));
    const double idc = uniform_d(rcp_f64((double)g.d[2]));
    const double dacd = uniform_d(da * idc);   // the forward's (double)da * idc
    const double dvd = uniform_d((double)dt.dv), v0d = uniform_d((double)dt.v0);
    const float gxa = fabsf(g.d[0]), gya = fabsf(g.d[1]);
    // this lane's z cells c0 .. c0+CPL-1 (cell c is the forward's corner
    // pair (c, c+1), valid for 0 <= c <= nc-2)
    const int c0 = lane * CPL;
    const float c0f = (float)c0;
    // the lane's boundaries c0 .. c0+CPL as searched: the faces pushed past
    // the ends of the clamp [r_0, r_nc-1]
    float zs[CPL + 1];
#pragma unroll
    for (int bq = 0; bq <= CPL; ++bq)
        zs[bq] = c0 + bq == 0 ? -1e30f : (c0 + bq >= nc - 1 ? 1e30f : c0f + (float)bq);

    // the tile's cell box: sample points whose cells touch the tile
    const double xa = (double)g.o[0] + ((double)(i0 - 1) + 0.5) * g.d[0];
    const double xb = (double)g.o[0] + ((double)(i0 + ni) + 0.5) * g.d[0];
    const double ya = (double)g.o[1] + ((double)(j0 - 1) + 0.5) * g.d[1];
    const double yb = (double)g.o[1] + ((double)(j0 + nj) + 0.5) * g.d[1];
    // prologue, before the accumulators live: which views are separable and
    // which detector columns can reach the tile in each
    for (int pv = threadIdx.x; pv < nviews; pv += SEP_THREADS) {
        const float* f = fv + (size_t)(view0 + pv * vstride) * 18;
        unsigned w = ~0u;
        if (view_separable(f, g, dt)) {
            int jl, jh;
            sep_column_range(f, dt, fmin(xa, xb), fmax(xa, xb), fmin(ya, yb), fmax(ya, yb), jl,
                             jh);
            // 0 <= jl, jh < nu < 2^16; an empty range skips the view here
            if (jl <= jh) w = (unsigned)jl | ((unsigned)jh << 16);
        }
        s_view[pv] = w;
    }
    __syncthreads();

    float acc[SEP_PPW][SEP_TB][CPL];
#pragma unroll
    for (int pl = 0; pl < SEP_PPW; ++pl)
#pragma unroll
        for (int q = 0; q < SEP_TB; ++q)
#pragma unroll
            for (int c = 0; c < CPL; ++c) acc[pl][q][c] = 0.0f;

    for (int pv = 0; pv < nviews; ++pv) {
        const int p = view0 + pv * vstride;
        const float* f = fv + (size_t)p * 18;
        const unsigned vw = (unsigned)__builtin_amdgcn_readfirstlane((int)s_view[pv]);
        if (vw == ~0u) continue;   // not separable (the brick kernel takes it), or no column reaches the tile
        const int jlo = (int)(vw & 0xffffu), jhi = (int)(vw >> 16);
        const double Sa = uniform_d(f[A]), Sb = uniform_d(f[Bx]);
        const double Sc = uniform_d(f[2]), evz = uniform_d(f[14]);
        const double ta = uniform_d(fma(0.5, da, oa) - Sa);   // plane 0 centre from the source
        const double ScOc = uniform_d(Sc - oc);
        for (int cb = jlo; cb <= jhi; cb += SEP_CHUNK) {
            __syncthreads();   // the previous chunk's samples are done with the tables
            // ---- wavefront w: column cb + w, its set-up and its table ----
            {
                const int j = cb + wave;
                double2* pre = reinterpret_cast<double2*>(reinterpret_cast<char*>(s_tab) +
                                                          wave * colb);
                unsigned mask = 0u;
                if (j <= jhi && j < dt.nu) {
                    const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
                    // x and y of the bin do not depend on the row (e_v has no x, y)
                    const double dd0 = fma((double)f[9], u, (double)f[6]) - (double)f[0];
                    const double dd1 = fma((double)f[10], u, (double)f[7]) - (double)f[1];
                    // the fast set-up's axis choice (fast_axis; z never drives here)
                    const bool xy = fabsf((float)dd0) * gya > fabsf((float)dd1) * gxa;
                    if ((xy ? 0 : 1) == A) {
                        // the column's table, built once by k_sep_tables,
                        // copied into LDS while its plane records are made
                        const double2* src = tabs + ((size_t)pv * dt.nu + j) * (nv + 1);
                        for (int e0 = 0; e0 <= nv; e0 += 64)
                            if (e0 + lane <= nv)
                                __builtin_amdgcn_global_load_lds(
                                    (const __attribute__((address_space(1))) void*)(src + e0 + lane),
                                    (__attribute__((address_space(3))) void*)(pre + e0), 16, 0, 0);
                        const double dda = A == 0 ? dd0 : dd1, ddb = A == 0 ? dd1 : dd0;
                        const double ia = rcp_f64(dda);
                        const double r1 = ddb * ia;
                        const float ub0 = (float)fma(fma(r1, ta, Sb) - ob, idb, -0.5);
                        const float dub = (float)(r1 * (da * idb));
                        const double X2 = fma((double)f[11], u, (double)f[8]);
                        const double v0c = fma(0.5, dvd, v0d);
                        // z slope of row k: R0 + R1 k (real arithmetic)
                        const double R0 = (fma(evz, v0c, X2) - Sc) * ia;
                        const double R1 = evz * dvd * ia;
                        // rows in ascending z: B(l) has one sign over the planes
                        const bool rev = sep_rev(R1, P0, da, ta, idc);
                        // lane q < nP: tile plane P0 + q
                        const int lo = max(Q0 - 1, 0), hi = min(Q0 + nQ - 1, nb - 2);
                        bool ok = false;
                        if (lane < nP) {
                            const int l = P0 + lane;
                            const float ubl = __builtin_fmaf((float)l, dub, ub0);
                            SepPlane sp;
                            const int ibl = floor_i(ubl);
                            const float w1 = frac_pos(ubl), w0 = 1.0f - w1;
                            ok = ibl >= lo && ibl <= hi;
#pragma unroll
                            for (int bi = 0; bi < SEP_TB; ++bi) {
                                const int dj = bi - (ibl - Q0);
                                sp.wr[bi] = dj == 0 ? w0 : (dj == 1 ? w1 : 0.0f);
                            }
                            // uc of row k on plane l: A(l) + B(l) k, in z cells
                            const double T = fma((double)l, da, ta);
                            sp.A = fma(fma(R0, T, ScOc), idc, -0.5);
                            sp.B = R1 * T * idc;
                            if (rev) {   // stored row kk is row nv-1-kk
                                sp.A = fma(sp.B, (double)(nv - 1), sp.A);
                                sp.B = -sp.B;
                            }
                            const double rBd = rcp_f64(sp.B);
                            sp.rB = (float)rBd;
                            sp.nA1 = (float)fma(-sp.A, rBd, 1.0);
                            // the face rows: m nearest the crossing (clamped to
                            // the detector), then m or m+1 by the forward's uc
                            // of row m (setup_fast_ax)
#pragma unroll
                            for (int e = 0; e < 2; ++e) {
                                const double zc = e ? (double)(nc - 1) : 0.0;
                                const int m = (int)fmin(fmax(rint((zc - sp.A) * rBd), 0.0),
                                                        (double)(nv - 1));
                                const int k = rev ? nv - 1 - m : m;
                                const double v = fma((double)k + 0.5, dvd, v0d);
                                const double r2 = (fma(evz, v, X2) - Sc) * ia;
                                const float uc0 = (float)fma(fma(r2, ta, Sc) - oc, idc, -0.5);
                                const float duc = (float)(r2 * dacd);
                                const float rf = (float)(m + (__builtin_fmaf((float)l, duc, uc0) <
                                                              (float)zc ? 1 : 0));
                                if (e) sp.rtop = rf; else sp.rbot = rf;
                            }
                            if (ok) s_pl[wave][lane] = sp;
                        }
                        if (!ok && lane < SEP_TA) {
                            // a plane that misses the tile (or lies past the
                            // volume): every boundary clamps to row 0 (PS =
                            // PK = 0) and the weights are 0, so sampling it
                            // adds nothing
                            SepPlane sp;
                            sp.A = sp.B = 0.0;
                            sp.rB = sp.nA1 = sp.rbot = sp.rtop = 0.0f;
#pragma unroll
                            for (int bi = 0; bi < SEP_TB; ++bi) sp.wr[bi] = 0.0f;
                            s_pl[wave][lane] = sp;
                        }
                        mask = (unsigned)__ballot(ok);
                    }
                }
                if (lane == 0) s_mask[wave] = mask;
            }
            __builtin_amdgcn_s_waitcnt(0);   // this wavefront's table copy has landed
            __syncthreads();
            // ---- samples on this wavefront's planes, column by column ----
#pragma unroll
            for (int c = 0; c < SEP_CHUNK; ++c) {
                const unsigned mask = (unsigned)__builtin_amdgcn_readfirstlane((int)s_mask[c]);
                const double2* pre = reinterpret_cast<const double2*>(
                    reinterpret_cast<const char*>(s_tab) + c * colb);
                // the wavefront's planes together, when either samples the
                // tile (the record of a plane that misses it adds nothing)
                if (((mask >> (wave * SEP_PPW)) & ((1u << SEP_PPW) - 1u)) == 0u) continue;
#pragma unroll
                for (int pl = 0; pl < SEP_PPW; ++pl) {
                    const int q = wave * SEP_PPW + pl;
                    const SepPlane& sp = s_pl[c][q];
                    const double Ad = sp.A, Bd = sp.B;
                    const float rB = uniform_f(sp.rB), nA1 = uniform_f(sp.nA1);
                    const float rbot = uniform_f(sp.rbot), rtop = uniform_f(sp.rtop);
                    const float lf = (float)(P0 + q);
                    // cell boundaries z = c0 .. c0+CPL (at most nc-1, so the
                    // cells from nc-1 up, which the forward never samples, are
                    // empty): the first stored row whose uc reaches z, and the
                    // prefix sums there
                    int r[CPL + 1];
#pragma unroll
                    for (int bq = 0; bq <= CPL; ++bq)
                        r[bq] = (int)__builtin_amdgcn_fmed3f(__builtin_fmaf(zs[bq], rB, nA1), rbot,
                                                            rtop);
                    double PS[CPL + 1], PK[CPL + 1];
#pragma unroll
                    for (int bq = 0; bq <= CPL; ++bq) {
                        const double2 P = pre[r[bq]];
                        PS[bq] = P.x;
                        PK[bq] = P.y;
                    }
                    double lo[CPL], hi[CPL];
#pragma unroll
                    for (int cc = 0; cc < CPL; ++cc) {
                        const double S = PS[cc + 1] - PS[cc];
                        const double H = fma(Bd, PK[cc + 1] - PK[cc], (Ad - (double)(c0 + cc)) * S);
                        lo[cc] = S - H;
                        hi[cc] = H;
                    }
                    // voxel c0+cc: lower part of cell c0+cc, upper part of
                    // the cell below (the previous lane's last cell for cc = 0)
                    const double carry = dpp_f64<0x138, 0xf>(hi[CPL - 1]);   // wave_shr:1
                    // b rows of the tile: ib takes 1-wb, ib+1 takes wb, the
                    // others 0, as the record holds them (the same on every
                    // lane, so every row is one FMA per z cell)
                    float wr[SEP_TB];
#pragma unroll
                    for (int bi = 0; bi < SEP_TB; ++bi) wr[bi] = sp.wr[bi];
#pragma unroll
                    for (int cc = 0; cc < CPL; ++cc) {
                        const float G = (float)(lo[cc] + (cc == 0 ? carry : hi[cc - 1]));
#pragma unroll
                        for (int bi = 0; bi < SEP_TB; ++bi)
                            acc[pl][bi][cc] = __builtin_fmaf(wr[bi], G, acc[pl][bi][cc]);
                    }
                }
            }
        }
    }
    // ---- write-back: this wavefront's planes, every b cell, CPL z per lane ----
#pragma unroll
    for (int pl = 0; pl < SEP_PPW; ++pl) {
        const int q = wave * SEP_PPW + pl;
        if (q >= nP) continue;
        const int l = P0 + q;
#pragma unroll
        for (int bi = 0; bi < SEP_TB; ++bi) {
            if (bi >= nQ) continue;
            const int i = A == 0 ? l : Q0 + bi;
            const int j = A == 0 ? Q0 + bi : l;
#pragma unroll
            for (int cc = 0; cc < CPL; ++cc) {
                const int k = c0 + cc;
                if (k >= nc) continue;
                const size_t gi = (size_t)i * g.s[0] + (size_t)j * g.s[1] + (size_t)k * g.s[2];
                const float v = acc[pl][bi][cc];
                if (mode == SEP_ATOMIC)
                    atomicAdd(vol + gi, v);
                else if (mode == SEP_SART)
                    vol[gi] = sart_update(vol[gi], lambda, corr[gi] + v, colsum[gi]);
                else
                    vol[gi] = mode == SEP_ADD ? vol[gi] + v : v;
            }
        }
    }
}

// axes bit 0: the first ntx workgroups take x-driven tiles; bit 1: the
// (other) workgroups take y-driven tiles.
template <int CPL>
__global__ __launch_bounds__(SEP_THREADS) __attribute__((amdgpu_waves_per_eu(SEP_WPE))) void k_bp_sep(
    Grid g, Det dt, int view0, int nviews, int vstride, int axes, int ntx, int mode,
    float* __restrict__ vol, const double2* __restrict__ tabs, const float* __restrict__ fv) {
    extern __shared__ double s_tab[];       // SEP_CHUNK column tables
    __shared__ SepPlane s_pl[SEP_CHUNK][SEP_TA];
    __shared__ unsigned s_mask[SEP_CHUNK];  // bit q: tile plane q samples the tile
    __shared__ unsigned s_view[SEP_MAXV];   // per view: jlo | jhi << 16, or ~0u
    const int b = (int)blockIdx.x;
    if ((axes & 1) && b < ntx)
        bp_sep_tile<0, CPL>(b, g, dt, view0, nviews, vstride, mode, vol, tabs, fv, s_tab, s_pl,
                            s_mask, s_view, nullptr, nullptr, 0.0f);
    else
        bp_sep_tile<1, CPL>((axes & 1) ? b - ntx : b, g, dt, view0, nviews, vstride, mode, vol,
                            tabs, fv, s_tab, s_pl, s_mask, s_view, nullptr, nullptr, 0.0f);
}

// Batched plans: the tiles of driving axis ax of several problems (first
// tile of problem p at off[p]), subset sub of each (views sub + pv*nsub).
// Kind BSEP_SART: x-driven tiles store their sums in corr, y-driven tiles
// add theirs and update the estimate; BSEP_COLSUM: x-driven tiles store the
// column sums of the subset, y-driven tiles add theirs.  With the x launch
// first in stream order, every voxel ends with corr + sum = the same FP32
// sum as the single-problem adjoint's two commuting atomic adds.
template <int CPL>
__global__ __launch_bounds__(SEP_THREADS) __attribute__((amdgpu_waves_per_eu(SEP_WPE))) void
k_batch_bp_sep(const SepBatchProb* __restrict__ probs, const long long* __restrict__ off,
               int nprob, int sub, int nsub, int ax, int kind, float lambda,
               const float* __restrict__ lambda_dev) {
    extern __shared__ double s_tab[];
    __shared__ SepPlane s_pl[SEP_CHUNK][SEP_TA];
    __shared__ unsigned s_mask[SEP_CHUNK];
    __shared__ unsigned s_view[SEP_MAXV];
    const long long it = blockIdx.x;
    const int pi = batch_find(off, nprob, it);
    const SepBatchProb& P = probs[pi];
    const int tile = (int)(it - off[pi]);
    const int nviews = subset_views(P.dt.ns, sub, nsub);
    float* cs = P.colsum + (size_t)sub * P.nvox;
    if (kind == BSEP_SART && lambda_dev != nullptr) lambda = *lambda_dev;
    if (ax == 0)
        bp_sep_tile<0, CPL>(tile, P.g, P.dt, sub, nviews, nsub, (int)SEP_STORE,
                            kind == BSEP_SART ? P.corr : cs, P.tabs, P.fv, s_tab, s_pl, s_mask,
                            s_view, nullptr, nullptr, 0.0f);
    else
        bp_sep_tile<1, CPL>(tile, P.g, P.dt, sub, nviews, nsub,
                            kind == BSEP_SART ? (int)SEP_SART : (int)SEP_ADD,
                            kind == BSEP_SART ? P.vol : cs, P.tabs, P.fv, s_tab, s_pl, s_mask,
                            s_view, P.corr, cs, lambda);
}

// The column tables of views view0 + pv*vstride (pv < nviews) of a separable
// view set: one wavefront per (view, detector column), its driving axis a
// from the fast set-up's choice, table entry kk+1 = the prefix sums over the
// stored rows 0..kk (rows in ascending z) of td = trav * datum (0 for a ray
// the indicator drops) and of td * kk, in FP64; entry 0 = (0, 0).  Lane l
// sums the rows [l*R, l*R + R) and a wavefront scan adds the lanes below.
// Views that are not separable are skipped (their tables are not read).
// One wavefront: the table of view view0 + pv*vstride, column j.
__device__ __forceinline__ void sep_table_column(const Grid& g, const Det& dt, int view0,
                                                 int vstride, int pv, int j,
                                                 double2* __restrict__ tabs,
                                                 int* __restrict__ nonsep,
                                                 const float* __restrict__ sino,
                                                 const int8_t* __restrict__ ind,
                                                 const float* __restrict__ fv) {
    const int lane = threadIdx.x & 63;
    const int p = view0 + pv * vstride;
    const float* f = fv + (size_t)p * 18;
    if (!view_separable(f, g, dt)) {
        // one count per view, for the brick kernel's pass
        if (j == 0 && lane == 0) atomicAdd(nonsep, 1);
        return;
    }
    const int nv = dt.nv;
    const double u = fma((double)j + 0.5, (double)dt.du, (double)dt.u0);
    // x and y of the bin do not depend on the row (e_v has no x, y)
    const double dd0 = fma((double)f[9], u, (double)f[6]) - (double)f[0];
    const double dd1 = fma((double)f[10], u, (double)f[7]) - (double)f[1];
    // the fast set-up's axis choice (fast_axis; z never drives here)
    const bool xy = fabsf((float)dd0) * fabsf(g.d[1]) > fabsf((float)dd1) * fabsf(g.d[0]);
    const int a = xy ? 0 : 1;
    const double dda = a == 0 ? dd0 : dd1, ddb = a == 0 ? dd1 : dd0;
    const double da = (double)g.d[a], oa = (double)g.o[a];
    const double Sa = f[a], Sc = f[2], evz = f[14];
    const double ta = fma(0.5, da, oa) - Sa;
    const double idc = rcp_f64((double)g.d[2]);
    const double dvd = (double)dt.dv, v0d = (double)dt.v0;
    const double ia = rcp_f64(dda);
    const float r1f = (float)(ddb * ia);
    const float r1sq = 1.0f + r1f * r1f;
    const double X2 = fma((double)f[11], u, (double)f[8]);
    const double v0c = fma(0.5, dvd, v0d);
    // z slope of row k: R0 + R1 k (real arithmetic)
    const double R0 = (fma(evz, v0c, X2) - Sc) * ia;
    const double R1 = evz * dvd * ia;
    const float R0f = (float)R0, R1f = (float)R1;
    const bool rev = sep_rev(R1, 0, da, ta, idc);
    const float trav_a = fabsf(g.d[a]);
    const int R = ((nv + 63) >> 6) | 1;   // rows per lane (nv <= 959: R <= 15)
    const int kb = lane * R;
    const size_t rbase = ((size_t)p * dt.nu + j) * nv;
    double2* tab = tabs + ((size_t)pv * dt.nu + j) * (nv + 1);
    double ps[SEP_RMAX], pk[SEP_RMAX];
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < SEP_RMAX; ++i) {
        const int kk = kb + i;
        if (i < R && kk < nv) {
            const int k = rev ? nv - 1 - kk : kk;
            float val = 0.0f;
            if (ind == nullptr || ind[rbase + k] == 1) {
                // the forward's trav (setup_fast_ax) with the ray's z slope
                // from its affine form in the row (within 3e-8 of the
                // forward's trav, below fast_sqrt's own rounding), times the
                // datum
                const float r2f = __builtin_fmaf((float)k, R1f, R0f);
                val = trav_a * fast_sqrt(r1sq + r2f * r2f) * sino[rbase + k];
            }
            const double td = (double)val;
            s0 += td;
            s1 = fma(td, (double)kk, s1);
        }
        ps[i] = s0;
        pk[i] = s1;
    }
    // this lane's rows follow the sums of the lanes below
    const double e0 = wave_scan_f64(s0) - s0;
    const double e1 = wave_scan_f64(s1) - s1;
#pragma unroll
    for (int i = 0; i < SEP_RMAX; ++i) {
        const int kk = kb + i;
        if (i < R && kk < nv) tab[kk + 1] = make_double2(ps[i] + e0, pk[i] + e1);
    }
    if (lane == 0) tab[0] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_sep_tables(Grid g, Det dt, int view0, int nviews,
                                                    int vstride, double2* __restrict__ tabs,
                                                    int* __restrict__ nonsep,
                                                    const float* __restrict__ sino,
                                                    const int8_t* __restrict__ ind,
                                                    const float* __restrict__ fv) {
    const long long item = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int pv = (int)(item / dt.nu), j = (int)(item % dt.nu);
    if (pv >= nviews) return;   // the whole wavefront
    sep_table_column(g, dt, view0, vstride, pv, j, tabs, nonsep, sino, ind, fv);
}

// Batched plans: one wavefront per (problem, view of subset sub, column),
// the first item of problem p at off[p].
__global__ __launch_bounds__(256) void k_batch_sep_tables(const SepBatchProb* __restrict__ probs,
                                                          const long long* __restrict__ off,
                                                          int nprob, int sub, int nsub,
                                                          int* __restrict__ nonsep) {
    const long long item = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= off[nprob]) return;   // the whole wavefront
    const int pi = batch_find(off, nprob, item);
    const SepBatchProb& P = probs[pi];
    const long long local = item - off[pi];
    const int pv = (int)(local / P.dt.nu), j = (int)(local % P.dt.nu);
    sep_table_column(P.g, P.dt, sub, nsub, pv, j, P.tabs, nonsep, P.resid, P.ind, P.fv);
}

// One thread per view: the count of views view_separable rejects.
__global__ void k_count_nonsep(Grid g, Det dt, const float* __restrict__ fv,
                               int* __restrict__ count) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < dt.ns && !view_separable(fv + (size_t)p * 18, g, dt)) atomicAdd(count, 1);
}

// Clears vol (n floats) and the count of views left to the brick kernel.
__global__ void k_sep_zero(float* __restrict__ vol, size_t n, int* __restrict__ nonsep) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *nonsep = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        vol[i] = 0.0f;
}

// The column-separable adjoint of the fast forward over the view subset:
// every voxel of vol receives the sums of its x-driven tile and its y-driven
// tile.  Views that fail view_separable contribute nothing
// here; the caller adds them with the brick kernel.  Requires linear
// sampling, f32 output, the default C-order strides, nz <= 256, nu < 65536
// and column tables within 120 KiB of LDS (eight columns, nv <= 959).
bool sep_supported(const ct2_desc* d, const Grid& g) {
    return d->arith == CT2_ARITH_FAST && d->interp == CT2_LINEAR && d->vol_dtype == CT2_F32 &&
           d->axis == CT2_AXIS_FREE && g.n[2] <= 256 && d->nv >= 1 &&
           sep_lds_bytes(d->nv) <= 120 * 1024 && d->nu < 65536 && g.s[2] == 1 &&
           g.s[1] == g.n[2] && g.s[0] == g.n[1] * g.n[2];
}

// The column tables of one view: nu tables of nv + 1 entries.
static size_t sep_view_bytes(const Det& dt) { return (size_t)dt.nu * sep_col_bytes(dt.nv); }

// The workspace: a header (the device count of views the route leaves to the
// brick kernel), then the tables of a group of views.
constexpr size_t SEP_WS_HEAD = 256;

// Views whose tables an internally allocated workspace holds at most (the
// 256^3 x 360 x 512^2 scan needs 1.5 GB for all of them).
constexpr size_t SEP_WS_CAP = (size_t)4 << 30;

size_t sep_work_bytes(const ct2_desc* d) {
    Grid g = make_grid(d->nx, d->ny, d->nz, d->dx, d->dy, d->dz, d->x0, d->y0, d->z0);
    Det dt = make_det(d->ns, d->nu, d->nv, d->du, d->dv, d->u0, d->v0);
    ViewSet vs;
    if (check_grid(g) || check_det(dt) || view_set(d, "ct2_backproject_work_bytes", vs)) return 0;
    if (zsep_route(d, g)) return zsep_work_bytes(d, dt, vs);
    if (!sep_route(d, g)) return 0;
    return SEP_WS_HEAD + (size_t)vs.nviews * sep_view_bytes(dt);
}

// gv views per group: each group's tables are built, then its x- and
// y-driven tiles run.  One group: both axes in one launch, adding atomically
// into the zeroed volume (two commuting adds per voxel).  More groups: per
// group an x launch that stores (first group) or adds, then a y launch that
// adds, so stream order fixes the order of the additions.
template <int CPL>
static int launch_sep_cpl(const Grid& g, const Det& dt, const ViewSet& vs, float* vol,
                          const float* sino, const int8_t* ind, const float* fv, int* nonsep,
                          double2* tabs, int gv, hipStream_t st) {
    const size_t lds = sep_lds_bytes(dt.nv);
    const int ntx = ((g.n[0] + SEP_TA - 1) / SEP_TA) * ((g.n[1] + SEP_TB - 1) / SEP_TB);
    const int nty = ((g.n[0] + SEP_TB - 1) / SEP_TB) * ((g.n[1] + SEP_TA - 1) / SEP_TA);
    const bool one = gv >= vs.nviews;
    const size_t n = one ? (size_t)g.n[0] * g.n[1] * g.n[2] : 0;
    hipLaunchKernelGGL(k_sep_zero, dim3(one ? 2048 : 1), dim3(256), 0, st, vol, n, nonsep);
    for (int v = 0; v < vs.nviews; v += gv) {
        const int nvg = min(gv, vs.nviews - v), v0 = vs.view0 + v * vs.vstride;
        const long long items = (long long)nvg * dt.nu;   // one wavefront per (view, column)
        hipLaunchKernelGGL(k_sep_tables, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, st, g,
                           dt, v0, nvg, vs.vstride, tabs, nonsep, sino, ind, fv);
        if (one) {
            hipLaunchKernelGGL((k_bp_sep<CPL>), dim3(ntx + nty), dim3(SEP_THREADS), lds, st, g,
                               dt, v0, nvg, vs.vstride, 3, ntx, (int)SEP_ATOMIC, vol, tabs, fv);
        } else {
            hipLaunchKernelGGL((k_bp_sep<CPL>), dim3(ntx), dim3(SEP_THREADS), lds, st, g, dt, v0,
                               nvg, vs.vstride, 1, ntx, v > 0 ? (int)SEP_ADD : (int)SEP_STORE,
                               vol, tabs, fv);
            hipLaunchKernelGGL((k_bp_sep<CPL>), dim3(nty), dim3(SEP_THREADS), lds, st, g, dt, v0,
                               nvg, vs.vstride, 2, 0, (int)SEP_ADD, vol, tabs, fv);
        }
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

// The stream-ordered pool keeps up to SEP_WS_CAP of freed memory for the
// next allocation instead of returning it at each synchronisation (its
// default is 0), so that repeated adjoints without a caller workspace do not
// map and unmap their tables every call.  Set once per device.
void keep_pool_memory(int dev) {
    static std::mutex mu;
    static bool done[64] = {};
    std::lock_guard<std::mutex> lk(mu);
    if (dev < 0 || dev >= 64 || done[dev]) return;
    done[dev] = true;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return;
    uint64_t cur = 0, want = (uint64_t)SEP_WS_CAP;
    if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &cur) == hipSuccess &&
        cur < want)
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &want);
}

int sep_work_acquire(const Det& dt, const ViewSet& vs, SepWork& w, void** own, hipStream_t st) {
    *own = nullptr;
    const size_t per = sep_view_bytes(dt);
    if (w.ptr) {
        if (((uintptr_t)w.ptr & 15u) != 0) {
            set_error("ct2_backproject: workspace must be 16-byte aligned");
            return CT2_EINVAL;
        }
        if (w.bytes < SEP_WS_HEAD + per) {
            set_error("ct2_backproject: workspace of %zu B holds no view's tables (%zu B)",
                      w.bytes, SEP_WS_HEAD + per);
            return CT2_EINVAL;
        }
        return CT2_OK;
    }
    const size_t gv = min((size_t)min(vs.nviews, SEP_MAXV), max((size_t)1, SEP_WS_CAP / per));
    int dev = 0;
    CT2_HIP(hipGetDevice(&dev));
    keep_pool_memory(dev);
    CT2_HIP(hipMallocAsync(own, SEP_WS_HEAD + per * gv, st));
    w.ptr = *own;
    w.bytes = SEP_WS_HEAD + per * gv;
    return CT2_OK;
}

int sep_work_release(void* own, hipStream_t st) {
    if (own) CT2_HIP(hipFreeAsync(own, st));
    return CT2_OK;
}

// w: a workspace from sep_work_acquire.
int launch_backproject_sep(const Grid& g, const Det& dt, const ViewSet& vs, float* vol,
                           const float* sino, const int8_t* ind, const float* fv, hipStream_t st,
                           SepWork w) {
    const int cpl = (g.n[2] + 63) / 64;
    if (cpl < 1 || cpl > 4) {
        set_error("separable adjoint: nz = %d beyond 256", g.n[2]);
        return CT2_EINVAL;
    }
    const size_t per = sep_view_bytes(dt);
    if (!w.ptr || w.bytes < SEP_WS_HEAD + per) {
        set_error("separable adjoint: no workspace");
        return CT2_EINVAL;
    }
    const int gv = (int)min((size_t)min(vs.nviews, SEP_MAXV), (w.bytes - SEP_WS_HEAD) / per);
    int* nonsep = static_cast<int*>(w.ptr);
    double2* tabs = reinterpret_cast<double2*>(static_cast<char*>(w.ptr) + SEP_WS_HEAD);
    switch (cpl) {
        case 1: return launch_sep_cpl<1>(g, dt, vs, vol, sino, ind, fv, nonsep, tabs, gv, st);
        case 2: return launch_sep_cpl<2>(g, dt, vs, vol, sino, ind, fv, nonsep, tabs, gv, st);
        case 3: return launch_sep_cpl<3>(g, dt, vs, vol, sino, ind, fv, nonsep, tabs, gv, st);
        case 4: return launch_sep_cpl<4>(g, dt, vs, vol, sino, ind, fv, nonsep, tabs, gv, st);
    }
    return CT2_EINVAL;
}

int sep_tiles(const Grid& g, int axis) {
    return axis == 0 ? ((g.n[0] + SEP_TA - 1) / SEP_TA) * ((g.n[1] + SEP_TB - 1) / SEP_TB)
                     : ((g.n[0] + SEP_TB - 1) / SEP_TB) * ((g.n[1] + SEP_TA - 1) / SEP_TA);
}

int launch_batch_sep_tables(const SepBatchProb* probs, const long long* off, int nprob, int sub,
                            int nsub, long long items, int* nonsep, hipStream_t st) {
    if (items <= 0) return CT2_OK;
    if ((items + 3) / 4 >= (1ll << 31)) {
        set_error("batched separable adjoint: too many table columns");
        return CT2_EINVAL;
    }
    hipLaunchKernelGGL(k_batch_sep_tables, dim3((unsigned)((items + 3) / 4)), dim3(256), 0, st,
                       probs, off, nprob, sub, nsub, nonsep);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

template <int CPL>
static void launch_bbs(long long tiles, size_t lds, const SepBatchProb* probs,
                       const long long* off, int nprob, int sub, int nsub, int ax, int kind,
                       float lambda, const float* lambda_dev, hipStream_t st) {
    hipLaunchKernelGGL((k_batch_bp_sep<CPL>), dim3((unsigned)tiles), dim3(SEP_THREADS), lds, st,
                       probs, off, nprob, sub, nsub, ax, kind, lambda, lambda_dev);
}

int launch_batch_bp_sep(int cpl, int ax, int kind, const SepBatchProb* probs,
                        const long long* off, int nprob, int sub, int nsub, long long tiles,
                        int max_nv, float lambda, const float* lambda_dev, hipStream_t st) {
    if (tiles <= 0) return CT2_OK;
    if (tiles >= (1ll << 31)) {
        set_error("batched separable adjoint: too many tiles");
        return CT2_EINVAL;
    }
    const size_t lds = sep_lds_bytes(max_nv);
    switch (cpl) {
        case 1: launch_bbs<1>(tiles, lds, probs, off, nprob, sub, nsub, ax, kind, lambda, lambda_dev, st); break;
        case 2: launch_bbs<2>(tiles, lds, probs, off, nprob, sub, nsub, ax, kind, lambda, lambda_dev, st); break;
        case 3: launch_bbs<3>(tiles, lds, probs, off, nprob, sub, nsub, ax, kind, lambda, lambda_dev, st); break;
        case 4: launch_bbs<4>(tiles, lds, probs, off, nprob, sub, nsub, ax, kind, lambda, lambda_dev, st); break;
        default: set_error("batched separable adjoint: %d z cells per lane", cpl); return CT2_EINVAL;
    }
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

int launch_count_nonsep(const Grid& g, const Det& dt, const float* fv, int* count,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_count_nonsep, dim3((unsigned)((dt.ns + 255) / 256)), dim3(256), 0, st, g,
                       dt, fv, count);
    CT2_HIP(hipGetLastError());
    return CT2_OK;
}

}  // namespace ct2
