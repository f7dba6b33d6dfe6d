/*
 * fwn_oracle.c -- CPU oracle for the fast winding number path.  TEST INFRASTRUCTURE ONLY:
 * the engine never links or calls this file; tests and the CPU-baseline leg of the
 * benchmark use it as the checker.
 *
 * What it restates, and from where:
 *  - The reference wrapper converts positions to float and indices to int before building
 *    (modules/winding/src/FastWindingNumber.cpp:39-52) and answers
 *    solid_angle = UT_SolidAngle<float,float>::computeSolidAngle(q) and
 *    is_inside = solid_angle / (4 pi) > 0.5 with pi a double (FastWindingNumber.cpp:60-76,
 *    modules/core/include/lagrange/internal/constants.h:16).  All functions here take the
 *    same float/int32 inputs.
 *  - The engine behind it (third-party jdumas/WindingNumber @a48b8f5, absent from the
 *    reference checkout; recipe cmake/recipes/external/winding_number.cmake:21-26) implements
 *    Barill et al. 2018, "Fast Winding Numbers for Soups and Clouds": a 4-ary BVH over
 *    triangles (box-area heuristic), per-child order-2 dipole expansions, far field when
 *    |q - p|^2 > beta^2 r^2 (beta = 2 by default) and the exact triangle solid angle
 *    otherwise.  restated_build()/tree_eval() restate that algorithm; the tree heuristic and
 *    the exact float operation order of the dependency cannot be checked here (its source is
 *    absent): parity of those two items is unpinned and results are pinned through the FP64
 *    brute force below.
 *
 * Functions:
 *  oracle_tri_f64        exact triangle solid angle, FP64 (van Oosterom & Strackee 1983)
 *  oracle_bruteforce     sum of oracle_tri_f64 over all facets, OpenMP over points
 *  oracle_moments        FP64 order-2 moments of a set of triangles (test of tree node data)
 *  restated_build/_free  4-ary box-area BVH + per-slot moments in the engine's slot format
 *  tree_eval             FP32 traversal over a slot tree (restated engine, or a tree exported
 *                        by the GPU engine), optional FP64 near-field, per-point work counters
 *  lbvh_sorted_order     30-bit Morton codes of triangle centroids + stable sort: the triangle
 *                        order a Morton LBVH must produce (bit-exact check of the GPU sort)
 *  oracle_closest        brute-force closest point on a mesh: minimum over all facets of the
 *                        point-triangle squared distance of
 *                        modules/core/include/lagrange/utils/point_triangle_squared_distance.h
 *                        (Eberly's region method in double; degenerate triangles by their edges,
 *                        modules/core/include/lagrange/utils/point_segment_squared_distance.h)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/fwn.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#define API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------------------ */
/* exact triangle solid angle, FP64                                                       */
/* ------------------------------------------------------------------------------------ */

/* Omega_t = 2 atan2(a.(b x c), |a||b||c| + (a.b)|c| + (b.c)|a| + (c.a)|b|), a,b,c = v_i - q.
 * A vertex coincident with q, or a zero triple product, counts as on-surface: 0. */
static inline double tri_f64(const double* a, const double* b, const double* c)
{
    double la = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double lb = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
    double lc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (la == 0.0 || lb == 0.0 || lc == 0.0) return 0.0;
    double cx = b[1] * c[2] - b[2] * c[1];
    double cy = b[2] * c[0] - b[0] * c[2];
    double cz = b[0] * c[1] - b[1] * c[0];
    double num = a[0] * cx + a[1] * cy + a[2] * cz;
    if (num == 0.0) return 0.0;
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double bc = b[0] * c[0] + b[1] * c[1] + b[2] * c[2];
    double ca = c[0] * a[0] + c[1] * a[1] + c[2] * a[2];
    double den = la * lb * lc + ab * lc + bc * la + ca * lb;
    return 2.0 * atan2(num, den);
}

API double oracle_tri_f64(const double* v0, const double* v1, const double* v2, const double* q)
{
    double a[3] = {v0[0] - q[0], v0[1] - q[1], v0[2] - q[2]};
    double b[3] = {v1[0] - q[0], v1[1] - q[1], v1[2] - q[2]};
    double c[3] = {v2[0] - q[0], v2[1] - q[1], v2[2] - q[2]};
    return tri_f64(a, b, c);
}

API void oracle_bruteforce(const float* V, int64_t nV, const int32_t* F, int64_t nF, const float* Q,
                           int64_t nQ, double* out, int nthreads)
{
    (void)nV;
    double* T = (double*)malloc(sizeof(double) * 9 * (size_t)(nF > 0 ? nF : 1));
    for (int64_t t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) T[9 * t + 3 * k + d] = (double)V[3 * (int64_t)F[3 * t + k] + d];
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int64_t i = 0; i < nQ; ++i) {
        double q[3] = {Q[3 * i], Q[3 * i + 1], Q[3 * i + 2]};
        double s = 0.0;
        for (int64_t t = 0; t < nF; ++t) {
            const double* p = T + 9 * t;
            double a[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
            double b[3] = {p[3] - q[0], p[4] - q[1], p[5] - q[2]};
            double c[3] = {p[6] - q[0], p[7] - q[1], p[8] - q[2]};
            s += tri_f64(a, b, c);
        }
        out[i] = s;
    }
    free(T);
}

/* ------------------------------------------------------------------------------------ */
/* FP64 moments                                                                          */
/* ------------------------------------------------------------------------------------ */
/* Moment record (33 doubles):
 *   0 A | 1-3 p | 4-6 N | 7-15 M1[j*3+i] = sum (c-p)_j N_i (full) |
 *   16-25 S (sym): xxx yyy zzz xxy xxz yyx yyz zzx zzy xyz | 26 r2 | 27-29 lo | 30-32 hi   */
enum { MA = 0, MP = 1, MN = 4, M1 = 7, MS = 16, MR2 = 26, MLO = 27, MHI = 30, MSIZE = 33 };

static const int SYM_IDX[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                   {1, 1, 0}, {1, 1, 2}, {2, 2, 0}, {2, 2, 1}, {0, 1, 2}};

/* squared distance from p to the farthest corner of the box [lo, hi]: the radius bound
 * every slot carries (an upper bound on the distance from p to any point of the slot) */
static double box_r2(const double* p, const double* lo, const double* hi)
{
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) {
        double a = fabs(p[d] - lo[d]), b = fabs(hi[d] - p[d]);
        double e = a > b ? a : b;
        r2 += e * e;
    }
    return r2;
}

/* triangle about its own centroid: M1 = 0, X_jk,i = (1/12) sum_v e_j e_k N_i */
static void tri_moments(const double* v0, const double* v1, const double* v2, double* m)
{
    memset(m, 0, sizeof(double) * MSIZE);
    double e1[3], e2[3], c[3];
    for (int d = 0; d < 3; ++d) {
        e1[d] = v1[d] - v0[d];
        e2[d] = v2[d] - v0[d];
        c[d] = (v0[d] + v1[d] + v2[d]) / 3.0;
    }
    double N[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                   0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    m[MA] = sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    for (int d = 0; d < 3; ++d) {
        m[MP + d] = c[d];
        m[MN + d] = N[d];
    }
    const double* vs[3] = {v0, v1, v2};
    double C[3][3] = {{0}};
    for (int k = 0; k < 3; ++k) {
        double e[3] = {vs[k][0] - c[0], vs[k][1] - c[1], vs[k][2] - c[2]};
        for (int j = 0; j < 3; ++j)
            for (int kk = 0; kk < 3; ++kk) C[j][kk] += e[j] * e[kk] / 12.0;
    }
    for (int s = 0; s < 10; ++s) {
        int a = SYM_IDX[s][0], b = SYM_IDX[s][1], cc = SYM_IDX[s][2];
        m[MS + s] = (C[a][b] * N[cc] + C[b][cc] * N[a] + C[cc][a] * N[b]) / 3.0;
    }
    for (int d = 0; d < 3; ++d) {
        double lo = v0[d], hi = v0[d];
        if (v1[d] < lo) lo = v1[d];
        if (v2[d] < lo) lo = v2[d];
        if (v1[d] > hi) hi = v1[d];
        if (v2[d] > hi) hi = v2[d];
        m[MLO + d] = lo;
        m[MHI + d] = hi;
    }
    m[MR2] = box_r2(m + MP, m + MLO, m + MHI);
}

/* merge child records (each about its own p) into a parent record about the area centroid.
 * r2 = squared distance from p to the farthest corner of the merged box. */
static void merge_moments(const double* const* ch, int n, double* m)
{
    memset(m, 0, sizeof(double) * MSIZE);
    double A = 0.0, P[3] = {0, 0, 0}, Pm[3] = {0, 0, 0};
    for (int c = 0; c < n; ++c) {
        A += ch[c][MA];
        for (int d = 0; d < 3; ++d) {
            P[d] += ch[c][MA] * ch[c][MP + d];
            Pm[d] += ch[c][MP + d];
        }
    }
    for (int d = 0; d < 3; ++d) P[d] = (A > 0.0) ? P[d] / A : Pm[d] / n;
    m[MA] = A;
    for (int d = 0; d < 3; ++d) {
        m[MP + d] = P[d];
        m[MLO + d] = INFINITY;
        m[MHI + d] = -INFINITY;
    }
    for (int c = 0; c < n; ++c) {
        const double* x = ch[c];
        double dl[3] = {x[MP] - P[0], x[MP + 1] - P[1], x[MP + 2] - P[2]};
        for (int i = 0; i < 3; ++i) m[MN + i] += x[MN + i];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) m[M1 + 3 * j + i] += x[M1 + 3 * j + i] + dl[j] * x[MN + i];
        for (int s = 0; s < 10; ++s) {
            int a = SYM_IDX[s][0], b = SYM_IDX[s][1], cc = SYM_IDX[s][2];
            /* Y_jk,i = d_j M1_ki + d_k M1_ji + d_j d_k N_i ; Sym = (Y_ab,c + Y_bc,a + Y_ca,b)/3 */
#define YV(j, k, i) (dl[j] * x[M1 + 3 * (k) + (i)] + dl[k] * x[M1 + 3 * (j) + (i)] + dl[j] * dl[k] * x[MN + (i)])
            m[MS + s] += x[MS + s] + (YV(a, b, cc) + YV(b, cc, a) + YV(cc, a, b)) / 3.0;
#undef YV
        }
        for (int d = 0; d < 3; ++d) {
            if (x[MLO + d] < m[MLO + d]) m[MLO + d] = x[MLO + d];
            if (x[MHI + d] > m[MHI + d]) m[MHI + d] = x[MHI + d];
        }
    }
    m[MR2] = box_r2(P, m + MLO, m + MHI);
}

/* FP64 moments of an arbitrary triangle set computed directly (not by merging):
 * the test reference for every node of a tree.  r2 = farthest-corner bound of the set's box. */
API void oracle_moments(const float* V, const int32_t* F, const int64_t* tris, int64_t n, double* out)
{
    double A = 0.0, P[3] = {0, 0, 0};
    double* tm = (double*)malloc(sizeof(double) * MSIZE * (size_t)(n > 0 ? n : 1));
    for (int64_t t = 0; t < n; ++t) {
        const int32_t* f = F + 3 * tris[t];
        double v[3][3];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) v[k][d] = V[3 * (int64_t)f[k] + d];
        tri_moments(v[0], v[1], v[2], tm + MSIZE * t);
        A += tm[MSIZE * t + MA];
        for (int d = 0; d < 3; ++d) P[d] += tm[MSIZE * t + MA] * tm[MSIZE * t + MP + d];
    }
    if (A > 0.0)
        for (int d = 0; d < 3; ++d) P[d] /= A;
    else {
        for (int d = 0; d < 3; ++d) P[d] = 0.0;
        for (int64_t t = 0; t < n; ++t)
            for (int d = 0; d < 3; ++d) P[d] += tm[MSIZE * t + MP + d] / (double)n;
    }
    memset(out, 0, sizeof(double) * MSIZE);
    out[MA] = A;
    for (int d = 0; d < 3; ++d) {
        out[MP + d] = P[d];
        out[MLO + d] = INFINITY;
        out[MHI + d] = -INFINITY;
    }
    for (int64_t t = 0; t < n; ++t) {
        const double* x = tm + MSIZE * t;
        const int32_t* f = F + 3 * tris[t];
        double y[3] = {x[MP] - P[0], x[MP + 1] - P[1], x[MP + 2] - P[2]};
        for (int i = 0; i < 3; ++i) out[MN + i] += x[MN + i];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) out[M1 + 3 * j + i] += y[j] * x[MN + i];
        /* int (x-P)_j (x-P)_k n_i = N_i [ y_j y_k + C_jk / A ], symmetrised */
        for (int s = 0; s < 10; ++s) {
            int a = SYM_IDX[s][0], b = SYM_IDX[s][1], c = SYM_IDX[s][2];
            out[MS + s] += x[MS + s] + (y[a] * y[b] * x[MN + c] + y[b] * y[c] * x[MN + a] + y[c] * y[a] * x[MN + b]) / 3.0;
        }
        for (int k = 0; k < 3; ++k) {
            for (int d = 0; d < 3; ++d) {
                double vd = V[3 * (int64_t)f[k] + d];
                if (vd < out[MLO + d]) out[MLO + d] = vd;
                if (vd > out[MHI + d]) out[MHI + d] = vd;
            }
        }
    }
    out[MR2] = box_r2(P, out + MLO, out + MHI);
    free(tm);
}

/* record -> slot words 0..22 (the engine's format, fwn.h); the threshold word is written by
 * set_thresholds from *r2out (the squared radius bound, rounded up to float) */
static void slot_from_moments(const double* m, uint32_t* w, float* r2out)
{
    float f[FWN_W_LINK];
    f[FWN_W_PX] = (float)m[MP];
    f[FWN_W_PY] = (float)m[MP + 1];
    f[FWN_W_PZ] = (float)m[MP + 2];
    f[FWN_W_THR] = 0.0f;
    /* round r2 up so the float radius stays an upper bound */
    *r2out = nextafterf((float)m[MR2], INFINITY);
    f[FWN_W_NX] = (float)m[MN];
    f[FWN_W_NY] = (float)m[MN + 1];
    f[FWN_W_NZ] = (float)m[MN + 2];
    /* tr joins the diagonal of the quadratic (tr = tr |d|^2 / |d|^2) and C the cubic monomials
     * d_i d_j d_j ((C.d) = (C.d) |d|^2 / |d|^2), so the far field has no constant terms */
    const double* M = m + M1;
    const double tr = M[0] + M[4] + M[8];
    float* A = f + FWN_W_A;
    A[0] = (float)(-3.0 * M[0] + tr);
    A[1] = (float)(-3.0 * M[4] + tr);
    A[2] = (float)(-3.0 * M[8] + tr);
    A[3] = (float)(-3.0 * (M[1] + M[3]));
    A[4] = (float)(-3.0 * (M[5] + M[7]));
    A[5] = (float)(-3.0 * (M[6] + M[2]));
    const double* S = m + MS;
    static const double sc[10] = {7.5, 7.5, 7.5, 22.5, 22.5, 22.5, 22.5, 22.5, 22.5, 45.0};
    const double cx = -4.5 * (S[0] + S[5] + S[7]);
    const double cy = -4.5 * (S[3] + S[1] + S[8]);
    const double cz = -4.5 * (S[4] + S[6] + S[2]);
    /* monomials: xxx yyy zzz xxy xxz yyx yyz zzx zzy xyz */
    const double fold[10] = {cx, cy, cz, cy, cz, cx, cz, cx, cy, 0.0};
    for (int s = 0; s < 10; ++s) f[FWN_W_B + s] = (float)(sc[s] * S[s] + fold[s]);
    memcpy(w, f, sizeof(f));
}

static inline uint32_t leaf_link(int64_t first, int64_t count)
{
    return FWN_LINK_LEAF | ((uint32_t)first << 3) | (uint32_t)(count - 1);
}

/* Far-field thresholds and leaf near-field data of a slot tree for accuracy scale beta and FP64
 * distance tau, from each slot's squared radius bound r2 (float): thr = (beta^2) r2, and for a
 * leaf (first triangle t) tris[t] words 10-11 = (sqrt(r2) + tau)^2 and tau^2 |N|^2, the same
 * float operations as the engine's build. */
API void set_thresholds(uint32_t* slots, int64_t nslots, uint32_t* tris, const float* r2, float beta, float tau)
{
    const float beta2 = beta * beta;
    for (int64_t k = 0; k < nslots; ++k) {
        uint32_t* w = slots + k * FWN_SLOT_WORDS;
        const uint32_t link = w[FWN_W_LINK];
        if (link == FWN_LINK_EMPTY) continue;
        float f[FWN_W_LINK];
        memcpy(f, w, sizeof(f));
        const float thr = beta2 * r2[k];
        memcpy(w + FWN_W_THR, &thr, 4);
        if (link & FWN_LINK_LEAF) {
            const int64_t first = (int64_t)((link & ~FWN_LINK_LEAF) >> 3);
            const float rr = sqrtf(r2[k]) + tau;
            const float near2 = rr * rr;
            const float nn = fmaf(f[FWN_W_NX], f[FWN_W_NX], fmaf(f[FWN_W_NY], f[FWN_W_NY], f[FWN_W_NZ] * f[FWN_W_NZ]));
            const float plane = (tau * tau) * nn;
            memcpy(tris + first * FWN_TRI_WORDS + FWN_T_NEAR2, &near2, 4);
            memcpy(tris + first * FWN_TRI_WORDS + FWN_T_PLANE, &plane, 4);
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* restated engine: 4-ary BVH, box-area split heuristic, single-triangle items             */
/* ------------------------------------------------------------------------------------ */
typedef struct restated_tree {
    int64_t nnodes, cap;
    int32_t width;   /* slots per node: 4 restates the reference; 2 and 8 are for study */
    uint32_t* slots; /* nnodes * width * FWN_SLOT_WORDS words */
    float* r2;       /* nnodes * width squared radius bounds (thresholds: set_thresholds) */
    uint32_t* tris;  /* nF * FWN_TRI_WORDS words, in tree order */
    int64_t nF;
    int32_t root_ns;
    int32_t depth;
} restated_tree;

typedef struct {
    const float* V;
    const int32_t* F;
    float* cen; /* 3 per triangle */
    float* blo; /* box per triangle */
    float* bhi;
    int32_t* items;
    restated_tree* T;
    int64_t tri_out;
} rbuild;

static double box_area(const float* lo, const float* hi)
{
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (dx < 0) return 0.0;
    return dx * dy + dy * dz + dz * dx;
}

/* binned box-area split of items[lo,hi) (n >= 2); returns split position in (lo,hi) */
static int64_t split2(rbuild* B, int64_t lo, int64_t hi)
{
    enum { NB = 16 };
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i < hi; ++i) {
        const float* c = B->cen + 3 * (int64_t)B->items[i];
        for (int d = 0; d < 3; ++d) {
            if (c[d] < clo[d]) clo[d] = c[d];
            if (c[d] > chi[d]) chi[d] = c[d];
        }
    }
    int ax = 0;
    for (int d = 1; d < 3; ++d)
        if (chi[d] - clo[d] > chi[ax] - clo[ax]) ax = d;
    float ext = chi[ax] - clo[ax];
    if (!(ext > 0.0f)) return lo + (hi - lo) / 2;
    int64_t cnt[NB] = {0};
    float blo[NB][3], bhi[NB][3];
    for (int b = 0; b < NB; ++b)
        for (int d = 0; d < 3; ++d) {
            blo[b][d] = INFINITY;
            bhi[b][d] = -INFINITY;
        }
    float sc = (float)NB / ext;
#define BIN(it) ({ int _b = (int)((B->cen[3 * (int64_t)(it) + ax] - clo[ax]) * sc); _b < 0 ? 0 : (_b >= NB ? NB - 1 : _b); })
    for (int64_t i = lo; i < hi; ++i) {
        int32_t it = B->items[i];
        int b = BIN(it);
        cnt[b]++;
        for (int d = 0; d < 3; ++d) {
            if (B->blo[3 * (int64_t)it + d] < blo[b][d]) blo[b][d] = B->blo[3 * (int64_t)it + d];
            if (B->bhi[3 * (int64_t)it + d] > bhi[b][d]) bhi[b][d] = B->bhi[3 * (int64_t)it + d];
        }
    }
    /* sweep */
    double rightA[NB];
    int64_t rightN[NB];
    float rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t rn = 0;
    for (int b = NB - 1; b >= 1; --b) {
        for (int d = 0; d < 3; ++d) {
            if (blo[b][d] < rl[d]) rl[d] = blo[b][d];
            if (bhi[b][d] > rh[d]) rh[d] = bhi[b][d];
        }
        rn += cnt[b];
        rightA[b] = box_area(rl, rh);
        rightN[b] = rn;
    }
    float ll[3] = {INFINITY, INFINITY, INFINITY}, lh[3] = {-INFINITY, -INFINITY, -INFINITY};
    int64_t ln = 0;
    double best = INFINITY;
    int bestb = -1;
    for (int b = 0; b < NB - 1; ++b) {
        for (int d = 0; d < 3; ++d) {
            if (blo[b][d] < ll[d]) ll[d] = blo[b][d];
            if (bhi[b][d] > lh[d]) lh[d] = bhi[b][d];
        }
        ln += cnt[b];
        if (ln == 0 || rightN[b + 1] == 0) continue;
        double cost = box_area(ll, lh) * (double)ln + rightA[b + 1] * (double)rightN[b + 1];
        if (cost < best) {
            best = cost;
            bestb = b;
        }
    }
    if (bestb < 0) return lo + (hi - lo) / 2;
    /* stable partition: bins <= bestb first */
    int64_t n = hi - lo, m = lo;
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int64_t k = 0;
    for (int64_t i = lo; i < hi; ++i)
        if (BIN(B->items[i]) <= bestb) B->items[m++] = B->items[i];
        else tmp[k++] = B->items[i];
    memcpy(B->items + m, tmp, sizeof(int32_t) * (size_t)k);
    free(tmp);
#undef BIN
    return m;
}

static int64_t rnode_alloc(restated_tree* T)
{
    if (T->nnodes == T->cap) {
        T->cap = T->cap ? 2 * T->cap : 1024;
        T->slots = (uint32_t*)realloc(T->slots, sizeof(uint32_t) * T->width * FWN_SLOT_WORDS * (size_t)T->cap);
        T->r2 = (float*)realloc(T->r2, sizeof(float) * T->width * (size_t)T->cap);
    }
    int64_t id = T->nnodes++;
    uint32_t* w = T->slots + T->width * FWN_SLOT_WORDS * id;
    memset(w, 0, sizeof(uint32_t) * T->width * FWN_SLOT_WORDS);
    for (int k = 0; k < T->width; ++k) {
        w[k * FWN_SLOT_WORDS + FWN_W_LINK] = FWN_LINK_EMPTY;
        T->r2[T->width * id + k] = 0.0f;
    }
    return id;
}

/* builds the node for items[lo,hi) (n >= 2), fills its slots, returns moment record */
static int rbuild_node(rbuild* B, int64_t lo, int64_t hi, int64_t node, double* mout, int depth)
{
    /* split the range into up to `width` parts: log2(width) rounds of binary box-area splits,
     * each round splitting every part that holds two or more triangles (width 4: two rounds,
     * the reference's 4-ary layout) */
    const int width = B->T->width;
    int64_t bounds[9];
    int ng = 1;
    bounds[0] = lo;
    bounds[1] = hi;
    for (int round = 1; round < width; round <<= 1) {
        int64_t nb[9];
        int nn = 0;
        nb[0] = lo;
        for (int g = 0; g < ng; ++g) {
            int64_t a = bounds[g], b = bounds[g + 1];
            if (b - a >= 2) nb[++nn] = split2(B, a, b);
            nb[++nn] = b;
        }
        ng = nn;
        memcpy(bounds, nb, sizeof(int64_t) * (nn + 1));
    }
    double rec[8][MSIZE];
    const double* chp[8];
    int maxd = depth;
    for (int s = 0; s < ng; ++s) {
        int64_t a = bounds[s], b = bounds[s + 1];
        uint32_t* w = B->T->slots + ((int64_t)width * node + s) * FWN_SLOT_WORDS;
        int32_t child = -1, child_ns = 0;
        int64_t first = B->tri_out;
        if (b - a == 1) {
            int32_t t = B->items[a];
            const int32_t* f = B->F + 3 * (int64_t)t;
            double v[3][3];
            uint32_t* tw = B->T->tris + FWN_TRI_WORDS * B->tri_out;
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) {
                    v[k][d] = B->V[3 * (int64_t)f[k] + d];
                    memcpy(tw + 3 * k + d, &B->V[3 * (int64_t)f[k] + d], 4);
                }
            tw[FWN_T_FACET] = (uint32_t)t;
            tw[FWN_T_NEAR2] = tw[FWN_T_PLANE] = 0;
            B->tri_out++;
            tri_moments(v[0], v[1], v[2], rec[s]);
        } else {
            int64_t cn = rnode_alloc(B->T);
            child = (int32_t)cn;
            int d2 = rbuild_node(B, a, b, cn, rec[s], depth + 1);
            if (d2 > maxd) maxd = d2;
            /* the child's slot count */
            w = B->T->slots + ((int64_t)width * node + s) * FWN_SLOT_WORDS; /* slots may have moved */
            uint32_t* cw = B->T->slots + (int64_t)width * cn * FWN_SLOT_WORDS;
            child_ns = 0;
            for (int k = 0; k < width; ++k)
                if (cw[k * FWN_SLOT_WORDS + FWN_W_LINK] != FWN_LINK_EMPTY) child_ns++;
        }
        slot_from_moments(rec[s], w, B->T->r2 + (int64_t)width * node + s);
        w[FWN_W_LINK] = child >= 0 ? (((uint32_t)child << 4) | (uint32_t)child_ns) : leaf_link(first, b - a);
        chp[s] = rec[s];
    }
    merge_moments(chp, ng, mout);
    return maxd;
}

API restated_tree* restated_build_w(const float* V, int64_t nV, const int32_t* F, int64_t nF, int32_t width)
{
    (void)nV;
    restated_tree* T = (restated_tree*)calloc(1, sizeof(restated_tree));
    T->width = (width == 2 || width == 8) ? width : 4;
    T->nF = nF;
    T->tris = (uint32_t*)calloc((size_t)(nF > 0 ? nF : 1) * FWN_TRI_WORDS, sizeof(uint32_t));
    rbuild B;
    B.V = V;
    B.F = F;
    B.T = T;
    B.tri_out = 0;
    B.cen = (float*)malloc(sizeof(float) * 3 * (size_t)(nF > 0 ? nF : 1));
    B.blo = (float*)malloc(sizeof(float) * 3 * (size_t)(nF > 0 ? nF : 1));
    B.bhi = (float*)malloc(sizeof(float) * 3 * (size_t)(nF > 0 ? nF : 1));
    B.items = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nF > 0 ? nF : 1));
    for (int64_t t = 0; t < nF; ++t) {
        B.items[t] = (int32_t)t;
        for (int d = 0; d < 3; ++d) {
            float a = V[3 * (int64_t)F[3 * t] + d], b = V[3 * (int64_t)F[3 * t + 1] + d],
                  c = V[3 * (int64_t)F[3 * t + 2] + d];
            B.cen[3 * t + d] = ((a + b) + c) * (1.0f / 3.0f);
            B.blo[3 * t + d] = fminf(a, fminf(b, c));
            B.bhi[3 * t + d] = fmaxf(a, fmaxf(b, c));
        }
    }
    int64_t root = rnode_alloc(T);
    double rec[MSIZE];
    if (nF >= 2) {
        T->depth = rbuild_node(&B, 0, nF, root, rec, 1);
    } else if (nF == 1) {
        /* single item in one slot */
        const int32_t* f = F;
        double v[3][3];
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) {
                v[k][d] = V[3 * (int64_t)f[k] + d];
                memcpy(T->tris + 3 * k + d, &V[3 * (int64_t)f[k] + d], 4);
            }
        tri_moments(v[0], v[1], v[2], rec);
        slot_from_moments(rec, T->slots, T->r2);
        T->slots[FWN_W_LINK] = leaf_link(0, 1);
        T->tris[FWN_T_FACET] = 0;
        T->depth = 1;
    }
    T->root_ns = 0;
    for (int k = 0; k < T->width; ++k)
        if (T->slots[k * FWN_SLOT_WORDS + FWN_W_LINK] != FWN_LINK_EMPTY) T->root_ns++;
    free(B.cen);
    free(B.blo);
    free(B.bhi);
    free(B.items);
    return T;
}

API restated_tree* restated_build(const float* V, int64_t nV, const int32_t* F, int64_t nF)
{
    return restated_build_w(V, nV, F, nF, 4);
}

API void restated_free(restated_tree* T)
{
    if (!T) return;
    free(T->slots);
    free(T->r2);
    free(T->tris);
    free(T);
}

API int64_t restated_num_nodes(const restated_tree* T) { return T->nnodes; }
API int32_t restated_root_ns(const restated_tree* T) { return T->root_ns; }
API int32_t restated_depth(const restated_tree* T) { return T->depth; }
API const uint32_t* restated_slots(const restated_tree* T) { return T->slots; }
API const uint32_t* restated_tris(const restated_tree* T) { return T->tris; }
API const float* restated_r2(const restated_tree* T) { return T->r2; }

/* ------------------------------------------------------------------------------------ */
/* FP32 traversal (restated engine and same-tree check of the GPU engine)                 */
/* ------------------------------------------------------------------------------------ */

/* exact triangle solid angle in FP32, normalised-vector form: with unit a,b,c,
 * Omega = 2 atan2(a.((b-a) x (c-a)), 1 + a.b + a.c + b.c). */
static inline float tri_f32(const float* t, float qx, float qy, float qz)
{
    float ax = t[0] - qx, ay = t[1] - qy, az = t[2] - qz;
    float bx = t[3] - qx, by = t[4] - qy, bz = t[5] - qz;
    float cx = t[6] - qx, cy = t[7] - qy, cz = t[8] - qz;
    float la = sqrtf(fmaf(ax, ax, fmaf(ay, ay, az * az)));
    float lb = sqrtf(fmaf(bx, bx, fmaf(by, by, bz * bz)));
    float lc = sqrtf(fmaf(cx, cx, fmaf(cy, cy, cz * cz)));
    if (la == 0.0f || lb == 0.0f || lc == 0.0f) return 0.0f;
    float ia = 1.0f / la, ib = 1.0f / lb, ic = 1.0f / lc;
    ax *= ia; ay *= ia; az *= ia;
    bx *= ib; by *= ib; bz *= ib;
    cx *= ic; cy *= ic; cz *= ic;
    float ux = bx - ax, uy = by - ay, uz = bz - az;
    float vx = cx - ax, vy = cy - ay, vz = cz - az;
    float nx = fmaf(uy, vz, -uz * vy);
    float ny = fmaf(uz, vx, -ux * vz);
    float nz = fmaf(ux, vy, -uy * vx);
    float num = fmaf(ax, nx, fmaf(ay, ny, az * nz));
    if (num == 0.0f) return 0.0f;
    float ab = fmaf(ax, bx, fmaf(ay, by, az * bz));
    float ac = fmaf(ax, cx, fmaf(ay, cy, az * cz));
    float bc = fmaf(bx, cx, fmaf(by, cy, bz * cz));
    float den = 1.0f + ab + ac + bc;
    return 2.0f * atan2f(num, den);
}

static inline double tri_f64_words(const float* t, float qx, float qy, float qz)
{
    double q[3] = {qx, qy, qz};
    double a[3] = {(double)t[0] - q[0], (double)t[1] - q[1], (double)t[2] - q[2]};
    double b[3] = {(double)t[3] - q[0], (double)t[4] - q[1], (double)t[5] - q[2]};
    double c[3] = {(double)t[6] - q[0], (double)t[7] - q[1], (double)t[8] - q[2]};
    return tri_f64(a, b, c);
}

/* omega + inv3 (o0 + inv2 (q1 + inv2 q2)), o0 = d.N, q1 = d^T A d and q2 = B(d, d, d) homogeneous
 * (tr and C folded into A and B by slot_from_moments), in Horner form: the GPU kernel's operation
 * order (one rounding per FMA), so the same tree gives the same sum up to rsqrt. */
static inline float far_field_acc(const float* w, float dx, float dy, float dz, float d2, float om)
{
    const float inv = 1.0f / sqrtf(d2);
    const float inv2 = inv * inv;
    const float inv3 = inv2 * inv;
    const float* A = w + FWN_W_A;
    const float* Bq = w + FWN_W_B;
    const float o0 = fmaf(dx, w[FWN_W_NX], fmaf(dy, w[FWN_W_NY], dz * w[FWN_W_NZ]));
    const float zz = dz * dz;
    const float a1 = fmaf(A[0], dx, fmaf(A[3], dy, A[5] * dz));
    const float b1 = fmaf(A[1], dy, A[4] * dz);
    const float q1 = fmaf(dy, b1, fmaf(dx, a1, A[2] * zz));
    const float L = fmaf(Bq[7], dx, fmaf(Bq[8], dy, Bq[2] * dz));
    const float P = fmaf(Bq[0], dx, fmaf(Bq[3], dy, Bq[4] * dz));
    const float Qv = fmaf(Bq[5], dy, Bq[9] * dz);
    const float X = fmaf(dy, Qv, dx * P);
    const float R = fmaf(Bq[1], dy, Bq[6] * dz);
    const float q2 = fmaf(dy * dy, R, fmaf(dx, X, zz * L));
    const float t = fmaf(inv2, fmaf(inv2, q2, q1), o0);
    return fmaf(t, inv3, om);
}

typedef struct { int32_t node; int32_t ns; } tstack;

/* One point over a slot tree of the given width.  Children are pushed in slot order and popped
 * last-in-first-out, which is the order the GPU kernel sums a point's FP32 terms in.  A slot is
 * far when |p - q|^2 exceeds its threshold word (beta^2 r2, set_thresholds).  With tau > 0, a
 * near leaf's triangles are taken in FP64 when q lies within tau of the leaf's bounding sphere
 * (|d|^2 < (sqrt(r2) + tau)^2) and, for a single-triangle leaf, within tau of its plane
 * ((d.N)^2 < tau^2 |N|^2), both thresholds read from the leaf's first triangle; each such leaf's
 * FP64 sum is rounded to 2^-40 fixed point and the fixed-point total is added to the FP32 sum
 * once, at the end (the engine's deferred near-field pass does the same). */
static float eval_point(const uint32_t* slots, int width, int32_t root, int32_t root_ns,
                        const uint32_t* tris, float tau, float qx, float qy, float qz, int32_t* st)
{
    tstack stack[1024];
    int sp = 0;
    stack[sp].node = root;
    stack[sp].ns = root_ns;
    sp++;
    float om = 0.0f;
    long long acc = 0;
    int32_t nt = 0, nf = 0, n32 = 0, n64 = 0;
    while (sp > 0) {
        --sp;
        int32_t node = stack[sp].node, ns = stack[sp].ns;
        for (int s = 0; s < ns; ++s) {
            const float* w = (const float*)(slots + ((int64_t)node * width + s) * FWN_SLOT_WORDS);
            const uint32_t link = slots[((int64_t)node * width + s) * FWN_SLOT_WORDS + FWN_W_LINK];
            float dx = w[FWN_W_PX] - qx, dy = w[FWN_W_PY] - qy, dz = w[FWN_W_PZ] - qz;
            float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            nt++;
            if (d2 > w[FWN_W_THR]) {
                om = far_field_acc(w, dx, dy, dz, d2, om);
                nf++;
            } else if (link & FWN_LINK_LEAF) {
                int32_t first = (int32_t)((link & ~FWN_LINK_LEAF) >> 3), cnt = (int32_t)(link & 7u) + 1;
                const float* t0 = (const float*)(tris + (int64_t)first * FWN_TRI_WORDS);
                float nd = fmaf(dx, w[FWN_W_NX], fmaf(dy, w[FWN_W_NY], dz * w[FWN_W_NZ]));
                int use64 = tau > 0.0f && d2 < t0[FWN_T_NEAR2] && (cnt > 1 || nd * nd < t0[FWN_T_PLANE]);
                if (use64) {
                    double s64 = 0.0;
                    for (int32_t t = first; t < first + cnt; ++t)
                        s64 += tri_f64_words((const float*)(tris + (int64_t)t * FWN_TRI_WORDS), qx, qy, qz);
                    acc += llrint(s64 * 1099511627776.0);
                    n64 += cnt;
                } else {
                    for (int32_t t = first; t < first + cnt; ++t) {
                        om += tri_f32((const float*)(tris + (int64_t)t * FWN_TRI_WORDS), qx, qy, qz);
                        n32++;
                    }
                }
            } else {
                if (sp >= 1024) return NAN;
                stack[sp].node = (int32_t)(link >> 4);
                stack[sp].ns = (int32_t)(link & 15u);
                sp++;
            }
        }
    }
    if (acc != 0) om = (float)((double)om + (double)acc * (1.0 / 1099511627776.0));
    if (st) {
        st[0] = nt;
        st[1] = nf;
        st[2] = n32;
        st[3] = n64;
    }
    return om;
}

API void tree_eval(const uint32_t* slots, int32_t width, int32_t root, int32_t root_ns, const uint32_t* tris,
                   const float* Q, int64_t nQ, float tau, float* out, int32_t* stats, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < nQ; ++i)
        out[i] = eval_point(slots, width, root, root_ns, tris, tau, Q[3 * i], Q[3 * i + 1], Q[3 * i + 2],
                            stats ? stats + 4 * i : NULL);
}

/* is_inside exactly as the reference wrapper: float omega, double 4*pi, double compare */
API void inside_from_omega(const float* om, int64_t n, uint8_t* out)
{
    const double pi = 3.14159265358979323846;
    for (int64_t i = 0; i < n; ++i) out[i] = ((double)om[i] / (4.0 * pi) > 0.5) ? 1 : 0;
}

/* ------------------------------------------------------------------------------------ */
/* Morton LBVH triangle order                                                             */
/* ------------------------------------------------------------------------------------ */
static uint32_t expand10(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

static uint32_t quant10(float c, float lo, float inv)
{
    float t = (c - lo) * inv;
    float s = t * 1024.0f;
    if (!(s > 0.0f)) return 0u;
    if (s >= 1023.0f) return 1023u;
    return (uint32_t)s;
}

typedef struct { uint64_t key; int64_t idx; } kv;
static int kv_cmp(const void* a, const void* b)
{
    const kv* x = (const kv*)a;
    const kv* y = (const kv*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

/* Per mesh (f_off has n_meshes+1 entries, facet indices local to v_off ranges): centroid
 * c = ((v0+v1)+v2) * (1/3) in float, quantised to 10 bits per axis inside the mesh's centroid
 * box (inv = 1/extent computed in float, 0 for a flat axis), key = mesh << 32 | morton30.
 * Output: facet ids (global) in sorted order and the keys. */
API void lbvh_sorted_order(int32_t n_meshes, const float* V, const int64_t* v_off, const int32_t* F,
                           const int64_t* f_off, int64_t* order_out, uint64_t* keys_out)
{
    int64_t nF = f_off[n_meshes];
    kv* a = (kv*)malloc(sizeof(kv) * (size_t)(nF > 0 ? nF : 1));
    float* cen = (float*)malloc(sizeof(float) * 3 * (size_t)(nF > 0 ? nF : 1));
    for (int32_t m = 0; m < n_meshes; ++m) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t t = f_off[m]; t < f_off[m + 1]; ++t) {
            for (int d = 0; d < 3; ++d) {
                float a0 = V[3 * (v_off[m] + F[3 * t]) + d];
                float a1 = V[3 * (v_off[m] + F[3 * t + 1]) + d];
                float a2 = V[3 * (v_off[m] + F[3 * t + 2]) + d];
                float c = ((a0 + a1) + a2) * (1.0f / 3.0f);
                cen[3 * t + d] = c;
                if (c < lo[d]) lo[d] = c;
                if (c > hi[d]) hi[d] = c;
            }
        }
        float inv[3];
        for (int d = 0; d < 3; ++d) {
            float e = hi[d] - lo[d];
            inv[d] = (e > 0.0f) ? 1.0f / e : 0.0f;
        }
        for (int64_t t = f_off[m]; t < f_off[m + 1]; ++t) {
            uint32_t x = quant10(cen[3 * t], lo[0], inv[0]);
            uint32_t y = quant10(cen[3 * t + 1], lo[1], inv[1]);
            uint32_t z = quant10(cen[3 * t + 2], lo[2], inv[2]);
            uint32_t mc = (expand10(x) << 2) | (expand10(y) << 1) | expand10(z);
            a[t].key = ((uint64_t)(uint32_t)m << 32) | mc;
            a[t].idx = t;
        }
    }
    qsort(a, (size_t)nF, sizeof(kv), kv_cmp);
    for (int64_t i = 0; i < nF; ++i) {
        order_out[i] = a[i].idx;
        if (keys_out) keys_out[i] = a[i].key;
    }
    free(a);
    free(cen);
}


/* ------------------------------------------------------------------------------------ */
/* closest point (brute force)                                                           */
/* ------------------------------------------------------------------------------------ */

/* squared distance from q to segment (a, b), parameter u of the closest point a + u (b - a) */
static double seg_d2(const double* q, const double* a, const double* b, double* u_out)
{
    double ab[3], aq[3];
    for (int d = 0; d < 3; ++d) {
        ab[d] = b[d] - a[d];
        aq[d] = q[d] - a[d];
    }
    const double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    const double t = aq[0] * ab[0] + aq[1] * ab[1] + aq[2] * ab[2];
    double u = 0.0;
    if (!(t <= 0.0 || l2 == 0.0)) u = t > l2 ? 1.0 : t / l2;
    double r = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double x = q[d] - (a[d] + u * ab[d]);
        r += x * x;
    }
    *u_out = u;
    return r;
}

/* Eberly's minimisation of |v0 + s e0 + t e1 - q|^2 over the triangle, by the region of the
 * unconstrained minimiser (s, t) (point_triangle_squared_distance.h:31-280). */
static double tri_closest(const double* v0, const double* v1, const double* v2, const double* q, double* so,
                          double* to)
{
    double df[3], e0[3], e1[3];
    for (int d = 0; d < 3; ++d) {
        df[d] = v0[d] - q[d];
        e0[d] = v1[d] - v0[d];
        e1[d] = v2[d] - v0[d];
    }
    const double a00 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const double a01 = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    const double a11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double b0 = df[0] * e0[0] + df[1] * e0[1] + df[2] * e0[2];
    const double b1 = df[0] * e1[0] + df[1] * e1[1] + df[2] * e1[2];
    const double c = df[0] * df[0] + df[1] * df[1] + df[2] * df[2];
    const double det = fabs(a00 * a11 - a01 * a01);
    double s = a01 * b1 - a11 * b0, t = a01 * b0 - a00 * b1, r;
    if (det < 1e-30) {
        const double* E[3][2] = {{v0, v1}, {v0, v2}, {v1, v2}};
        r = INFINITY;
        for (int k = 0; k < 3; ++k) {
            double u;
            const double dd = seg_d2(q, E[k][0], E[k][1], &u);
            if (dd < r) {
                r = dd;
                *so = k == 0 ? u : (k == 1 ? 0.0 : 1.0 - u);
                *to = k == 0 ? 0.0 : u;
            }
        }
        return r;
    }
    /* on an edge: minimise the 1D quadratic a s^2 + 2 b s + c with s clamped to [0, 1] */
#define EDGE(ss, aa, bb, rr)                                                                   \
    do {                                                                                       \
        if ((bb) >= 0.0) { ss = 0.0; rr = c; }                                                 \
        else if (-(bb) >= (aa)) { ss = 1.0; rr = (aa) + 2.0 * (bb) + c; }                      \
        else { ss = -(bb) / (aa); rr = (bb) * ss + c; }                                        \
    } while (0)
#define FULLQ(ss, tt) ((ss) * (a00 * (ss) + a01 * (tt) + 2.0 * b0) + (tt) * (a01 * (ss) + a11 * (tt) + 2.0 * b1) + c)
    if (s + t <= det) {
        if (s < 0.0) {
            if (t < 0.0) {                   /* region 4: vertex v0 neighbourhood */
                if (b0 < 0.0) { t = 0.0; EDGE(s, a00, b0, r); }
                else { s = 0.0; EDGE(t, a11, b1, r); }
            } else { s = 0.0; EDGE(t, a11, b1, r); }          /* region 3 */
        } else if (t < 0.0) { t = 0.0; EDGE(s, a00, b0, r); } /* region 5 */
        else {                                                 /* region 0 */
            const double inv = 1.0 / det;
            s = s * inv;
            t = t * inv;
            r = FULLQ(s, t);
        }
    } else {
        const double denom = a00 - 2.0 * a01 + a11;
        if (s < 0.0) {                                          /* region 2 */
            const double tmp0 = a01 + b0, tmp1 = a11 + b1;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0;
                if (numer >= denom) { s = 1.0; t = 0.0; r = a00 + 2.0 * b0 + c; }
                else { s = numer / denom; t = 1.0 - s; r = FULLQ(s, t); }
            } else {
                s = 0.0;
                if (tmp1 <= 0.0) { t = 1.0; r = a11 + 2.0 * b1 + c; }
                else if (b1 >= 0.0) { t = 0.0; r = c; }
                else { t = -b1 / a11; r = b1 * t + c; }
            }
        } else if (t < 0.0) {                                   /* region 6 */
            const double tmp0 = a01 + b1, tmp1 = a00 + b0;
            if (tmp1 > tmp0) {
                const double numer = tmp1 - tmp0;
                if (numer >= denom) { t = 1.0; s = 0.0; r = a11 + 2.0 * b1 + c; }
                else { t = numer / denom; s = 1.0 - t; r = FULLQ(s, t); }
            } else {
                t = 0.0;
                if (tmp1 <= 0.0) { s = 1.0; r = a00 + 2.0 * b0 + c; }
                else if (b0 >= 0.0) { s = 0.0; r = c; }
                else { s = -b0 / a00; r = b0 * s + c; }
            }
        } else {                                                /* region 1 */
            const double numer = a11 + b1 - a01 - b0;
            if (numer <= 0.0) { s = 0.0; t = 1.0; r = a11 + 2.0 * b1 + c; }
            else if (numer >= denom) { s = 1.0; t = 0.0; r = a00 + 2.0 * b0 + c; }
            else { s = numer / denom; t = 1.0 - s; r = FULLQ(s, t); }
        }
    }
#undef EDGE
#undef FULLQ
    *so = s;
    *to = t;
    return r < 0.0 ? 0.0 : r;
}

API double oracle_tri_closest(const double* v0, const double* v1, const double* v2, const double* q, double* st)
{
    return tri_closest(v0, v1, v2, q, st, st + 1);
}

/* For every point: smallest squared distance over all facets (first facet on exact ties), the
 * facet and the closest point. */
API void oracle_closest(const float* V, const int32_t* F, int64_t nF, const float* Q, int64_t nQ, double* d2,
                        int32_t* facet, double* cp, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t i = 0; i < nQ; ++i) {
        const double q[3] = {Q[3 * i], Q[3 * i + 1], Q[3 * i + 2]};
        double best = INFINITY, bs = 0.0, bt = 0.0;
        int32_t bf = -1;
        for (int64_t f = 0; f < nF; ++f) {
            double v[3][3];
            for (int k = 0; k < 3; ++k)
                for (int d = 0; d < 3; ++d) v[k][d] = V[3 * (int64_t)F[3 * f + k] + d];
            double s, t;
            const double r = tri_closest(v[0], v[1], v[2], q, &s, &t);
            if (r < best) {
                best = r;
                bf = (int32_t)f;
                bs = s;
                bt = t;
            }
        }
        d2[i] = best;
        facet[i] = bf;
        for (int d = 0; d < 3; ++d) {
            if (bf < 0) {
                cp[3 * i + d] = NAN;
                continue;
            }
            const double a = V[3 * (int64_t)F[3 * bf] + d], b = V[3 * (int64_t)F[3 * bf + 1] + d],
                         c = V[3 * (int64_t)F[3 * bf + 2] + d];
            cp[3 * i + d] = a + bs * (b - a) + bt * (c - a);
        }
    }
}
