/*
 * CPU oracle for the structured TMOP block-Jacobi smoother.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the reference the HIP path is compared
 * against; nothing in the shipped library links, loads or calls it. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Plain C restatement of the reference algorithm, compiled twice: REAL=double
 * (symbols *_f64, the reference's default CPU build) and REAL=float (*_f32).
 * Everything works in GLOBAL node-id space: halo, share and fan copies of the
 * reference are pure data movement (every copy holds its owner's value when a
 * sweep reads it), so the semantics of one sweep are "every free DOF reads the
 * previous snapshot of the global X and writes its own node".
 *
 * Functions and the reference code they restate (paths relative to the
 * reference repository root):
 *   o_metric2d    src/metric.hpp:38-102 (mu2), :133-142 (untangle), :150-197
 *                 (shape_size); NumPy twin egg/smoothing/batch.py:62-148
 *   o_metric3d    src/metric.hpp:203-243 (mu_cond3, cond3_size, untangle3)
 *   o_patch_eval  src/patch.hpp:84-117 (assemble_vecT), :211-229 (role_Jb),
 *                 :237-320 (accumulate_sample), :325-346 (patch_eval)
 *   o_trial       src/patch.hpp:176-202 (trial_energy_mindet)
 *   o_solve       src/solve.hpp:32-77 (solveNxN<2>, <3>)
 *   o_energy_mindet  src/sweep.hpp:1881-1922 (reduce_energy_mindet); the sum
 *                 and min are accumulated in double (FP64 reductions)
 *   o_jacobi      src/sweep.hpp:1425-1650 (metric + interior update kernels),
 *                 :1938-2099 (jacobi_sweep_once / run_block_jacobi)
 *   o_gauss_seidel  the reference's sequential NumPy node baseline,
 *                 tests/smoothing/test_control_ref.py:95-135 (converged-state
 *                 cross-check of the block-Jacobi path)
 *   dir_*         src/directional.hpp (directional soft-energy terms), NumPy
 *                 twin egg/smoothing/directional.py:304-369; energies and
 *                 gradients pinned by closed forms and finite differences,
 *                 the Gauss-Newton block by finite differences of the energy
 *                 at zero residual (where it is the exact Hessian), the
 *                 stabilisers by their closed forms on symmetric samples
 *   curv_*        src/curvature.hpp (block-interface C2 windows); pinned the
 *                 same way (straight and circular windows have zero residual)
 *   o_entity_eval src/geometry.hpp:288-464 projections and tangents, pinned
 *                 by tests/cpp/golden_geometry.hpp
 *   o_entity_frame, update_constrained  every entity kind and the cold /
 *                 warm constrained update (entity_oracle.c lists the
 *                 reference code and the pins)
 *
 * Pinning: golden vectors from tests/cpp/golden_metric.hpp,
 * golden_metric_3d.hpp and golden_patch.hpp (tests/golden/*.json), finite
 * differences, the delta->0 limit of the untangle surrogate, the analytic
 * minimiser of a Cartesian block, and brute force on tiny inputs (tests/).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
#ifndef SFX
#define SFX _f64
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

#undef real
#define real REAL

/* Per-precision tolerances (reference src/core.hpp:37-51). */
#define IS_F32 (sizeof(real) == 4)
#define TOL_ENERGY ((real)(IS_F32 ? 1e-6 : 1e-12))
#define TOL_ZNORM ((real)(IS_F32 ? 1e-7 : 1e-15))
#define MAX_BACKTRACKS 10 /* reference src/sweep.hpp:43 */

#ifndef EGG_ORACLE_KINDS
#define EGG_ORACLE_KINDS
enum { KIND_SHAPE = 0, KIND_UNTANGLE = 1, KIND_SHAPE_SIZE = 2 };
#endif

/* ------------------------------------------------------------------ 2D metric */

/* d vec(cof)/d vec(T), cof = [d, -c, -b, a]. */
static const real FN(kP2)[16] = {0, 0, 0, 1, 0, 0, -1, 0, 0, -1, 0, 0, 1, 0, 0, 0};

static real FN(untangle_dh)(real D, real delta, real* dh1, real* dh2)
{
    /* Dh = (D + sqrt(D^2 + 4 delta^2)) / 2; Dh' = Dh / r; Dh'' = 2 delta^2 / r^3 */
    const real r = sqrt(D * D + (real)4 * delta * delta);
    const real Dh = (real)0.5 * (D + r);
    if (dh1) *dh1 = Dh / r;
    if (dh2) *dh2 = (real)2 * delta * delta / (r * r * r);
    return Dh;
}

/* value, grad (4) and Hessian (16, row-major) of the 2D metric at vec(T). */
void FN(o_metric2d)(int kind, const real* t, real delta, real* val, real* g, real* H)
{
    const real a = t[0], b = t[1], c = t[2], d = t[3];
    const real D = a * d - b * c;
    const real s = a * a + b * b + c * c + d * d;
    const real cof[4] = {d, -c, -b, a};
    if (kind == KIND_SHAPE_SIZE) {
        /* mu = (s/2D)^2 - 1 + (D-1)^2 */
        const real cond = s / ((real)2 * D);
        if (val) *val = cond * cond - (real)1 + (D - (real)1) * (D - (real)1);
        /* cond' = t/D - s cof/(2 D^2) */
        real cg[4];
        for (int i = 0; i < 4; ++i) cg[i] = t[i] / D - s * cof[i] / ((real)2 * D * D);
        if (g)
            for (int i = 0; i < 4; ++i) g[i] = (real)2 * cond * cg[i] + (real)2 * (D - (real)1) * cof[i];
        if (H) {
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    const real eye = (i == j) ? (real)1 : (real)0;
                    const real ch = eye / D - (t[i] * cof[j] + cof[i] * t[j]) / (D * D) +
                                    s * cof[i] * cof[j] / (D * D * D) -
                                    s * FN(kP2)[i * 4 + j] / ((real)2 * D * D);
                    H[i * 4 + j] = (real)2 * cg[i] * cg[j] + (real)2 * cond * ch +
                                   (real)2 * cof[i] * cof[j] +
                                   (real)2 * (D - (real)1) * FN(kP2)[i * 4 + j];
                }
        }
        return;
    }
    /* shape (phi = D) and untangle (phi = Dh(D)): mu = s / (2 phi) - 1 */
    real p = D, p1 = (real)1, p2 = (real)0;
    if (kind == KIND_UNTANGLE) p = FN(untangle_dh)(D, delta, &p1, &p2);
    if (val) *val = s / ((real)2 * p) - (real)1;
    if (g)
        for (int i = 0; i < 4; ++i) g[i] = t[i] / p - s * p1 * cof[i] / ((real)2 * p * p);
    if (H) {
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) {
                const real eye = (i == j) ? (real)1 : (real)0;
                H[i * 4 + j] = eye / p - p1 * (t[i] * cof[j] + cof[i] * t[j]) / (p * p) +
                               s * p1 * p1 * cof[i] * cof[j] / (p * p * p) -
                               s / ((real)2 * p * p) * (p2 * cof[i] * cof[j] + p1 * FN(kP2)[i * 4 + j]);
            }
    }
}

/* ------------------------------------------------------------------ 3D metric */

static void FN(cof3)(const real* t, real* c)
{
    c[0] = t[4] * t[8] - t[5] * t[7];
    c[1] = t[5] * t[6] - t[3] * t[8];
    c[2] = t[3] * t[7] - t[4] * t[6];
    c[3] = t[2] * t[7] - t[1] * t[8];
    c[4] = t[0] * t[8] - t[2] * t[6];
    c[5] = t[1] * t[6] - t[0] * t[7];
    c[6] = t[1] * t[5] - t[2] * t[4];
    c[7] = t[2] * t[3] - t[0] * t[5];
    c[8] = t[0] * t[4] - t[1] * t[3];
}

/* Symmetric bilinear cofactor form: cof(T) = Q(T,T), Q(A,B)_pq =
 * 1/2 eps_pkl eps_qmn (A_km B_ln + B_km A_ln) / 2 * 2 -> written out as the
 * mixed 2x2 minors. */
static void FN(cofQ)(const real* A, const real* B, real* Q)
{
    static const int r1[3] = {1, 2, 0}, r2[3] = {2, 0, 1};
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) {
            const int k = r1[p], l = r2[p], m = r1[q], n = r2[q];
            Q[p * 3 + q] = (real)0.5 * (A[k * 3 + m] * B[l * 3 + n] - A[k * 3 + n] * B[l * 3 + m] +
                                        B[k * 3 + m] * A[l * 3 + n] - B[k * 3 + n] * A[l * 3 + m]);
        }
}

static real FN(dot9)(const real* a, const real* b)
{
    real s = 0;
    for (int i = 0; i < 9; ++i) s += a[i] * b[i];
    return s;
}

/* value, grad (9) and Hessian (81) of mu_cond3 (kind 0), its untangle
 * surrogate (kind 1) and cond3 + (det-1)^2 (kind 2). Derivatives are taken
 * as second directional derivatives along unit directions e_a, e_b:
 *   f = s q / (9 phi(D)^2), s = |T|^2, q = |cof T|^2, D = det T. */
void FN(o_metric3d)(int kind, const real* t, real delta, real* val, real* g, real* H)
{
    real cof[9];
    FN(cof3)(t, cof);
    const real s = FN(dot9)(t, t);
    const real q = FN(dot9)(cof, cof);
    const real D = t[0] * cof[0] + t[1] * cof[1] + t[2] * cof[2];
    real p = D, p1 = (real)1, p2 = (real)0;
    if (kind == KIND_UNTANGLE) p = FN(untangle_dh)(D, delta, &p1, &p2);
    const real h = p * p;
    const real gq = s * q;
    if (val) {
        *val = gq / ((real)9 * h) - (real)1;
        if (kind == KIND_SHAPE_SIZE) *val += (D - (real)1) * (D - (real)1);
    }
    if (!g && !H) return;
    /* first derivatives along e_a */
    real QT[9][9]; /* Q(T, e_a) */
    real s1[9], D1[9], q1[9], h1[9], g1[9];
    for (int a = 0; a < 9; ++a) {
        real E[9] = {0};
        E[a] = 1;
        FN(cofQ)(t, E, QT[a]);
        s1[a] = (real)2 * t[a];
        D1[a] = cof[a];
        q1[a] = (real)4 * FN(dot9)(cof, QT[a]);
        h1[a] = (real)2 * p * p1 * D1[a];
        g1[a] = s1[a] * q + s * q1[a];
    }
    if (g)
        for (int a = 0; a < 9; ++a) {
            g[a] = (g1[a] * h - gq * h1[a]) / ((real)9 * h * h);
            if (kind == KIND_SHAPE_SIZE) g[a] += (real)2 * (D - (real)1) * D1[a];
        }
    if (H) {
        for (int a = 0; a < 9; ++a)
            for (int b = 0; b < 9; ++b) {
                real Ea[9] = {0}, Eb[9] = {0}, Qab[9];
                Ea[a] = 1;
                Eb[b] = 1;
                FN(cofQ)(Ea, Eb, Qab);
                const real s2 = (a == b) ? (real)2 : (real)0;
                const real D2 = (real)2 * FN(dot9)(QT[a], Eb);
                const real q2 = (real)8 * FN(dot9)(QT[a], QT[b]) + (real)4 * FN(dot9)(cof, Qab);
                const real h2 = (real)2 * (p1 * p1 + p * p2) * D1[a] * D1[b] + (real)2 * p * p1 * D2;
                const real g2 = s2 * q + s1[a] * q1[b] + s1[b] * q1[a] + s * q2;
                real v = (g2 / h - (g1[a] * h1[b] + g1[b] * h1[a]) / (h * h) - gq * h2 / (h * h) +
                          (real)2 * gq * h1[a] * h1[b] / (h * h * h)) /
                         (real)9;
                if (kind == KIND_SHAPE_SIZE) v += (real)2 * D1[a] * D1[b] + (real)2 * (D - (real)1) * D2;
                H[a * 9 + b] = v;
            }
    }
}

static void FN(metric_any)(int dim, int kind, const real* t, real delta, real* val, real* g, real* H)
{
    if (dim == 2)
        FN(o_metric2d)(kind, t, delta, val, g, H);
    else
        FN(o_metric3d)(kind, t, delta, val, g, H);
}

/* ------------------------------------------------------------------ patch */

static real FN(det_n)(int dim, const real* A)
{
    if (dim == 2) return A[0] * A[3] - A[1] * A[2];
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}

/* A[:,k] = s_k (nbr_k - corner); T = A W (reference src/patch.hpp:95-117). */
static void FN(assemble)(int dim, const real* X, int gc, const int* gn, const real* s, const real* W,
                     real* T, real* detA)
{
    real A[9];
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k)
            A[i * dim + k] = s[k] * (X[(size_t)gn[k] * dim + i] - X[(size_t)gc * dim + i]);
    *detA = FN(det_n)(dim, A);
    for (int i = 0; i < dim; ++i)
        for (int k = 0; k < dim; ++k) {
            real acc = 0;
            for (int j = 0; j < dim; ++j) acc += A[i * dim + j] * W[j * dim + k];
            T[i * dim + k] = acc;
        }
}

/* Occurrence arrays for one DOF patch: P occurrences, each with corner gc,
 * neighbours gn[k*P + p], signs s[k*P + p], role, and a W_inv row (w_stride 0
 * = one uniform row). */
typedef struct {
    int P;
    const int* gc;
    const int* gn; /* [dim][P] */
    const real* s; /* [dim][P] */
    const int* role;
    const real* W;
    int w_stride;
    const real* wt; /* [P] per-occurrence energy weight, NULL = 1 (reference src/patch.hpp:45-77) */
} FN(patch_t);

static void FN(occ_fetch)(int dim, const FN(patch_t)* pv, int p, int* gn, real* s)
{
    for (int k = 0; k < dim; ++k) {
        gn[k] = pv->gn[(size_t)k * pv->P + p];
        s[k] = pv->s[(size_t)k * pv->P + p];
    }
}

/* Full patch evaluation: gradient, contracted Hessian, energy and min det A
 * over the patch (reference src/patch.hpp:237-346). */
static void FN(patch_eval_impl)(int dim, int kind, real delta, const FN(patch_t)* pv, const real* X, real* grad,
                            real* hess, real* energy, real* mindet)
{
    const int VT = dim * dim;
    for (int i = 0; i < dim; ++i) grad[i] = 0;
    for (int i = 0; i < dim * dim; ++i) hess[i] = 0;
    *energy = 0;
    *mindet = INFINITY;
    for (int p = 0; p < pv->P; ++p) {
        int gn[3];
        real s[3], T[9], detA, val, g[9], H[81];
        FN(occ_fetch)(dim, pv, p, gn, s);
        const real* W = pv->W + (size_t)pv->w_stride * p;
        FN(assemble)(dim, X, pv->gc[p], gn, s, W, T, &detA);
        if (detA < *mindet) *mindet = detA;
        const int role = pv->role[p];
        /* the weight scales mu's value, gradient and Hessian, not the det barrier
           (reference src/patch.hpp:250-292 accumulate_sample) */
        const real w = pv->wt ? pv->wt[p] : (real)1;
        FN(metric_any)(dim, kind, T, delta, &val, g, role >= 0 ? H : NULL);
        *energy += w * val;
        /* dmu/dA[i][k] = sum_j g[i,j] W[k,j] */
        real dA[3][3];
        for (int i = 0; i < dim; ++i)
            for (int k = 0; k < dim; ++k) {
                real acc = 0;
                for (int j = 0; j < dim; ++j) acc += g[i * dim + j] * W[k * dim + j];
                dA[i][k] = acc;
            }
        if (role == 0) {
            for (int i = 0; i < dim; ++i) {
                real acc = 0;
                for (int k = 0; k < dim; ++k) acc += s[k] * dA[i][k];
                grad[i] += w * -acc;
            }
        } else if (role >= 1) {
            const int ax = role - 1;
            for (int i = 0; i < dim; ++i) grad[i] += w * (s[ax] * dA[i][ax]);
        }
        if (role >= 0) {
            /* Jb (VT x dim): row a=(i,c) has its only nonzero in column i */
            real u[3];
            for (int c = 0; c < dim; ++c) {
                if (role == 0) {
                    real acc = 0;
                    for (int k = 0; k < dim; ++k) acc += s[k] * W[k * dim + c];
                    u[c] = -acc;
                } else {
                    const int ax = role - 1;
                    u[c] = s[ax] * W[ax * dim + c];
                }
            }
            real Jb[81];
            for (int e = 0; e < VT * dim; ++e) Jb[e] = 0;
            for (int a = 0; a < VT; ++a) Jb[a * dim + a / dim] = u[a % dim];
            for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j) {
                    real acc = 0;
                    for (int a = 0; a < VT; ++a) {
                        real hj = 0;
                        for (int b = 0; b < VT; ++b) hj += H[a * VT + b] * Jb[b * dim + j];
                        acc += Jb[a * dim + i] * hj;
                    }
                    hess[i * dim + j] += w * acc;
                }
        }
    }
}

void FN(o_patch_eval)(int dim, int kind, real delta, int P, const int* gc, const int* gn, const real* s,
                      const int* role, const real* W, int w_stride, const real* wt, const real* X, real* grad,
                      real* hess, real* energy, real* mindet)
{
    FN(patch_t) pv = {P, gc, gn, s, role, W, w_stride, wt};
    FN(patch_eval_impl)(dim, kind, delta, &pv, X, grad, hess, energy, mindet);
}

/* Trial energy + min det with node `dof` moved to `trial`
 * (reference src/patch.hpp:176-202). */
static void FN(trial_impl)(int dim, int kind, real delta, const FN(patch_t)* pv, const real* X, int dof,
                       const real* trial, real* e_new, real* mdet)
{
    *e_new = 0;
    *mdet = INFINITY;
    for (int p = 0; p < pv->P; ++p) {
        int gn[3];
        real s[3];
        FN(occ_fetch)(dim, pv, p, gn, s);
        const real* W = pv->W + (size_t)pv->w_stride * p;
        int ids[4] = {pv->gc[p], gn[0], gn[1], dim == 3 ? gn[2] : 0};
        real pts[4][3];
        for (int n = 0; n <= dim; ++n)
            for (int i = 0; i < dim; ++i)
                pts[n][i] = (ids[n] == dof) ? trial[i] : X[(size_t)ids[n] * dim + i];
        real A[9], T[9];
        for (int i = 0; i < dim; ++i)
            for (int k = 0; k < dim; ++k) A[i * dim + k] = s[k] * (pts[k + 1][i] - pts[0][i]);
        const real detA = FN(det_n)(dim, A);
        for (int i = 0; i < dim; ++i)
            for (int k = 0; k < dim; ++k) {
                real acc = 0;
                for (int j = 0; j < dim; ++j) acc += A[i * dim + j] * W[j * dim + k];
                T[i * dim + k] = acc;
            }
        real val;
        FN(metric_any)(dim, kind, T, delta, &val, NULL, NULL);
        *e_new += (pv->wt ? pv->wt[p] : (real)1) * val;
        if (detA < *mdet) *mdet = detA;
    }
}

/* Newton step x = -H^{-1} g with the fallbacks of reference src/solve.hpp:32-77:
 * |g| < znorm -> 0; non-finite -> -0.1 g/|g|. */
void FN(o_solve)(int dim, const real* H, const real* g, real* x)
{
    if (dim == 2) {
        const real gnorm = sqrt(g[0] * g[0] + g[1] * g[1]);
        if (gnorm < TOL_ZNORM) {
            x[0] = x[1] = 0;
            return;
        }
        const real det = H[0] * H[3] - H[1] * H[2];
        const real inv = (real)1 / det;
        x[0] = -inv * (H[3] * g[0] - H[1] * g[1]);
        x[1] = -inv * (-H[2] * g[0] + H[0] * g[1]);
        if (!isfinite(x[0]) || !isfinite(x[1])) {
            const real c = (real)-0.1 / gnorm;
            x[0] = c * g[0];
            x[1] = c * g[1];
        }
        return;
    }
    const real gnorm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (gnorm < TOL_ZNORM) {
        x[0] = x[1] = x[2] = 0;
        return;
    }
    const real c00 = H[4] * H[8] - H[5] * H[7];
    const real c01 = -(H[3] * H[8] - H[5] * H[6]);
    const real c02 = H[3] * H[7] - H[4] * H[6];
    const real det = H[0] * c00 + H[1] * c01 + H[2] * c02;
    const real inv = (real)1 / det;
    const real c10 = -(H[1] * H[8] - H[2] * H[7]);
    const real c11 = H[0] * H[8] - H[2] * H[6];
    const real c12 = -(H[0] * H[7] - H[1] * H[6]);
    const real c20 = H[1] * H[5] - H[2] * H[4];
    const real c21 = -(H[0] * H[5] - H[2] * H[3]);
    const real c22 = H[0] * H[4] - H[1] * H[3];
    x[0] = -inv * (c00 * g[0] + c10 * g[1] + c20 * g[2]);
    x[1] = -inv * (c01 * g[0] + c11 * g[1] + c21 * g[2]);
    x[2] = -inv * (c02 * g[0] + c12 * g[1] + c22 * g[2]);
    if (!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2])) {
        const real c = (real)-0.1 / gnorm;
        for (int i = 0; i < 3; ++i) x[i] = c * g[i];
    }
}

/* ------------------------------------- directional soft-energy terms */

/* Kinds (reference src/directional.hpp:9-27, NumPy twin
 * egg/smoothing/directional.py:317-369): 0 parallel (nodes i, j), E = w c^2
 * with c = (u.ref)/s, u = X[j] - X[i], s = sqrt(|u|^2 + eps^2); 2 stem (nodes
 * n1, c), the same with u = X[n1] - X[c]; 1 line (nodes a, c, b),
 * E = w |q1/s1 + q2/s2|^2 with qk = X[a|b] - X[c], sk = sqrt(|qk|^2 + eps^2).
 * Any other kind contributes nothing. One table type serves the per-DOF
 * windows (K per DOF, slot = the DOF's place, -1 pads) and the whole-sample
 * table of the energy reports (num samples, slot unused). */
#ifndef EGG_ORACLE_DIR_T
#define EGG_ORACLE_DIR_T
typedef struct {
    int K;             /* per-DOF windows per DOF (0 = none) */
    long long num;     /* whole-sample table length */
    const int* nodes;  /* [.. * 4] global node ids, -1 pads */
    const int* slot;   /* [..] the DOF's slot in the sample, -1 = pad */
    const int* kind;   /* [..] */
    const double* ref; /* [.. * dim] frozen reference vector */
    const double* wt;  /* [..] weight */
    double eps;
} dir_t;
static int dir_arity(int kind) { return kind == 1 ? 3 : (kind == 0 || kind == 2) ? 2 : 0; }
#endif

/* Energy of one sample at positions p[0..arity-1] (src/directional.hpp:95-124). */
static real FN(dir_energy)(int dim, int kind, const real p[3][3], const double* ref, real w, real eps)
{
    if (kind == 0 || kind == 2) {
        const int head = kind == 0 ? 1 : 0;
        real ss = eps * eps, dot = 0;
        for (int k = 0; k < dim; ++k) {
            const real u = p[head][k] - p[1 - head][k];
            ss += u * u;
            dot += u * (real)ref[k];
        }
        const real c = dot / (real)sqrt(ss);
        return w * c * c;
    }
    if (kind == 1) {
        real s1 = eps * eps, s2 = eps * eps;
        for (int k = 0; k < dim; ++k) {
            const real a = p[0][k] - p[1][k], b = p[2][k] - p[1][k];
            s1 += a * a;
            s2 += b * b;
        }
        s1 = (real)sqrt(s1);
        s2 = (real)sqrt(s2);
        real e = 0;
        for (int k = 0; k < dim; ++k) {
            const real r = (p[0][k] - p[1][k]) / s1 + (p[2][k] - p[1][k]) / s2;
            e += r * r;
        }
        return w * e;
    }
    return 0;
}

/* Gradient, Gauss-Newton Hessian (row-major dim x dim) and energy of one
 * sample for the node at `slot`, added in, with the coupling stabiliser on
 * the diagonal (src/directional.hpp:126-211): parallel / stem add
 * 2w |dr|^2 I; line adds 2w |J_slot|_F (sum of the other two |J|_F) I. */
static void FN(dir_accumulate)(int dim, int kind, const real p[3][3], const double* ref, real w, real eps, int slot,
                               real* grad, real* hess, real* e0)
{
    if (kind == 0 || kind == 2) {
        const int head = kind == 0 ? 1 : 0;
        real u[3], ss = eps * eps, dot = 0;
        for (int k = 0; k < dim; ++k) {
            u[k] = p[head][k] - p[1 - head][k];
            ss += u[k] * u[k];
            dot += u[k] * (real)ref[k];
        }
        const real sn = (real)sqrt(ss), c = dot / sn;
        real dr[3], dd = 0;
        for (int k = 0; k < dim; ++k) {
            dr[k] = (real)ref[k] / sn - c * u[k] / ss; /* dc/du */
            dd += dr[k] * dr[k];
        }
        *e0 += w * c * c;
        const real gc = 2 * w * c * (slot == head ? (real)1 : (real)-1), hc = 2 * w;
        for (int i = 0; i < dim; ++i) {
            grad[i] += gc * dr[i];
            for (int j = 0; j < dim; ++j) hess[i * dim + j] += hc * dr[i] * dr[j];
            hess[i * dim + i] += hc * dd;
        }
        return;
    }
    if (kind != 1) return;
    /* unit vectors h_k = q_k / s_k and their Jacobians (I - q q^T / s^2) / s */
    real q[2][3], h[2][3], J[2][9], sn[2];
    for (int t = 0; t < 2; ++t) {
        real ss = eps * eps;
        for (int k = 0; k < dim; ++k) {
            q[t][k] = p[2 * t][k] - p[1][k];
            ss += q[t][k] * q[t][k];
        }
        sn[t] = (real)sqrt(ss);
        for (int k = 0; k < dim; ++k) h[t][k] = q[t][k] / sn[t];
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j)
                J[t][i * dim + j] = ((i == j) ? (real)1 / sn[t] : (real)0) - q[t][i] * q[t][j] / (ss * sn[t]);
    }
    real r[3], e = 0;
    for (int k = 0; k < dim; ++k) {
        r[k] = h[0][k] + h[1][k];
        e += r[k] * r[k];
    }
    *e0 += w * e;
    real Js[9], f[3] = {0, 0, 0}; /* |J_a|, |J_b|, |J_a + J_b| */
    for (int k = 0; k < dim * dim; ++k) {
        Js[k] = slot == 0 ? J[0][k] : slot == 2 ? J[1][k] : -(J[0][k] + J[1][k]);
        f[0] += J[0][k] * J[0][k];
        f[1] += J[1][k] * J[1][k];
        f[2] += (J[0][k] + J[1][k]) * (J[0][k] + J[1][k]);
    }
    for (int t = 0; t < 3; ++t) f[t] = (real)sqrt(f[t]);
    const real fs = slot == 0 ? f[0] : slot == 2 ? f[1] : f[2];
    const real gc = 2 * w, stab = gc * fs * (f[0] + f[1] + f[2] - fs);
    for (int i = 0; i < dim; ++i) {
        real gi = 0;
        for (int j = 0; j < dim; ++j) gi += Js[i * dim + j] * r[j];
        grad[i] += gc * gi;
        for (int j = 0; j < dim; ++j) {
            real hij = 0;
            for (int m = 0; m < dim; ++m) hij += Js[i * dim + m] * Js[j * dim + m];
            hess[i * dim + j] += gc * hij;
        }
        hess[i * dim + i] += stab;
    }
}

/* One directional sample's gradient, Hessian and energy for its node
 * `slot` (dir_accumulate, for tests); p holds arity x dim positions. */
void FN(o_dir_eval)(int dim, int kind, const real* p, const double* ref, real w, real eps, int slot, real* grad,
                    real* hess, real* e0)
{
    real q[3][3] = {{0}};
    for (int m = 0; m < dir_arity(kind); ++m)
        for (int k = 0; k < dim; ++k) q[m][k] = p[m * dim + k];
    for (int i = 0; i < dim; ++i) grad[i] = 0;
    for (int i = 0; i < dim * dim; ++i) hess[i] = 0;
    *e0 = 0;
    FN(dir_accumulate)(dim, kind, q, ref, w, eps, slot, grad, hess, e0);
}

/* Sample positions of window j of per-DOF table row d (DOF at `trial` when given). */
static int FN(dir_load)(int dim, const dir_t* dw, long long w, const real* X, int slot, const real* trial,
                        real p[3][3])
{
    const int ar = dir_arity(dw->kind[w]);
    for (int m = 0; m < ar; ++m)
        for (int k = 0; k < dim; ++k)
            p[m][k] = (trial && m == slot) ? trial[k] : X[(size_t)dw->nodes[4 * w + m] * dim + k];
    return ar;
}

/* A DOF's windows into (grad, hess, e0) (src/directional.hpp:225-250). */
static void FN(dir_dof_accumulate)(int dim, const dir_t* dw, int d, const real* X, real* grad, real* hess, real* e0)
{
    for (int j = 0; j < dw->K; ++j) {
        const long long w = (long long)d * dw->K + j;
        const int sl = dw->slot[w];
        real p[3][3];
        if (sl < 0 || !FN(dir_load)(dim, dw, w, X, sl, NULL, p)) continue;
        FN(dir_accumulate)(dim, dw->kind[w], p, dw->ref + (size_t)dim * w, (real)dw->wt[w], (real)dw->eps, sl, grad,
                           hess, e0);
    }
}

/* A DOF's windows' energy with the DOF moved to `trial` (src/directional.hpp:252-277). */
static real FN(dir_dof_trial)(int dim, const dir_t* dw, int d, const real* X, const real* trial)
{
    real e = 0;
    for (int j = 0; j < dw->K; ++j) {
        const long long w = (long long)d * dw->K + j;
        const int sl = dw->slot[w];
        real p[3][3];
        if (sl < 0 || !FN(dir_load)(dim, dw, w, X, sl, trial, p)) continue;
        e += FN(dir_energy)(dim, dw->kind[w], p, dw->ref + (size_t)dim * w, (real)dw->wt[w], (real)dw->eps);
    }
    return e;
}

/* ------------------------------------------------------------ reductions */

/* Sum of mu and min det A over every energy-stencil sample
 * (reference src/sweep.hpp:1881-1922). mu and det are evaluated in double
 * from the REAL positions and accumulated in double. */
void FN(o_energy_mindet)(int dim, int kind, double delta, long long ns, const int* gc, const int* gn,
                         const real* s, const real* W, int w_stride, const real* wt, const real* X,
                         const dir_t* dirs, double* E, double* md)
{
    double e = 0.0, m = INFINITY;
    for (long long i = 0; i < ns; ++i) {
        double A[9], T[9];
        const real* w = W + (size_t)w_stride * i;
        for (int r = 0; r < dim; ++r)
            for (int k = 0; k < dim; ++k) {
                const int nk = gn[(size_t)k * ns + i];
                A[r * dim + k] = (double)s[(size_t)k * ns + i] *
                                 ((double)X[(size_t)nk * dim + r] - (double)X[(size_t)gc[i] * dim + r]);
            }
        for (int r = 0; r < dim; ++r)
            for (int k = 0; k < dim; ++k) {
                double acc = 0;
                for (int j = 0; j < dim; ++j) acc += A[r * dim + j] * (double)w[j * dim + k];
                T[r * dim + k] = acc;
            }
        double val, det;
        if (dim == 2) {
            det = A[0] * A[3] - A[1] * A[2];
            o_metric2d_f64(kind, T, delta, &val, NULL, NULL);
        } else {
            det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                  A[2] * (A[3] * A[7] - A[4] * A[6]);
            o_metric3d_f64(kind, T, delta, &val, NULL, NULL);
        }
        e += (wt ? (double)wt[i] : 1.0) * val;
        if (det < m) m = det;
    }
    /* the whole-sample directional table, evaluated in double from the REAL
       positions, joins the reported energy (src/sweep.hpp:1921) */
    for (long long i = 0; dirs && i < dirs->num; ++i) {
        const int ar = dir_arity(dirs->kind[i]);
        double p[3][3];
        for (int m2 = 0; m2 < ar; ++m2)
            for (int k = 0; k < dim; ++k) p[m2][k] = (double)X[(size_t)dirs->nodes[4 * i + m2] * dim + k];
        e += dir_energy_f64(dim, dirs->kind[i], p, dirs->ref + (size_t)dim * i, dirs->wt[i], dirs->eps);
    }
    *E = e;
    *md = m;
}

/* ---------------------------------------------- constrained boundary DOFs */

/* 2D closed-form entities a boundary DOF slides on (reference
 * src/geometry.hpp:288-320 LineSeg, :322-357 LineRail, :398-431 Circle,
 * :432-464 Ellipse): tag 1 = segment (sx, sy, ex, ey), 2 = circle (cx, cy,
 * r), 3 = ellipse (cx, cy, rx, ry), 15 = rail (sx, sy, ex, ey: a segment
 * that releases the node beyond its ends). */
#define ENT_SEGMENT 1
#define ENT_CIRCLE 2
#define ENT_ELLIPSE 3
#define ENT_RAIL 15
#define TOL_TINY ((real)(IS_F32 ? 1e-20 : 1e-30))

static real FN(rail_foot)(const real* e, const real* p)
{
    const real abx = e[2] - e[0], aby = e[3] - e[1];
    const real ab2 = abx * abx + aby * aby;
    return ((p[0] - e[0]) * abx + (p[1] - e[1]) * aby) / (ab2 > TOL_TINY ? ab2 : TOL_TINY);
}

static void FN(ent_project)(int tag, const real* e, const real* p, real* out)
{
    if (tag == ENT_SEGMENT || tag == ENT_RAIL) {
        const real abx = e[2] - e[0], aby = e[3] - e[1];
        real t = FN(rail_foot)(e, p);
        if (tag == ENT_RAIL && (t < 0 || t > 1)) { /* beyond an end: released */
            out[0] = p[0];
            out[1] = p[1];
            return;
        }
        t = t < 0 ? 0 : (t > 1 ? 1 : t);
        out[0] = e[0] + t * abx;
        out[1] = e[1] + t * aby;
    } else if (tag == ENT_ELLIPSE) {
        const real sx = (p[0] - e[0]) / e[2], sy = (p[1] - e[1]) / e[3];
        const real dist = sqrt(sx * sx + sy * sy);
        const real ux = dist < TOL_ZNORM ? (real)0.70710678118654752440 : sx / dist;
        const real uy = dist < TOL_ZNORM ? (real)0.70710678118654752440 : sy / dist;
        out[0] = e[0] + ux * e[2];
        out[1] = e[1] + uy * e[3];
    } else {
        const real dx = p[0] - e[0], dy = p[1] - e[1];
        const real dist = sqrt(dx * dx + dy * dy);
        if (dist < TOL_ZNORM) {
            out[0] = e[0] + e[2];
            out[1] = e[1];
        } else {
            out[0] = e[0] + e[2] * dx / dist;
            out[1] = e[1] + e[2] * dy / dist;
        }
    }
}

static void FN(ent_tangent)(int tag, const real* e, const real* p, real* t)
{
    if (tag == ENT_ELLIPSE) {
        /* parametric angle of the radially scaled point */
        const real ang = atan2((p[1] - e[1]) / e[3], (p[0] - e[0]) / e[2]);
        const real tx = -e[2] * sin(ang), ty = e[3] * cos(ang);
        const real n = sqrt(tx * tx + ty * ty);
        t[0] = n < TOL_ZNORM ? 1 : tx / n;
        t[1] = n < TOL_ZNORM ? 0 : ty / n;
        return;
    }
    if (tag == ENT_SEGMENT || tag == ENT_RAIL) {
        const real abx = e[2] - e[0], aby = e[3] - e[1];
        const real n = sqrt(abx * abx + aby * aby);
        t[0] = n < TOL_ZNORM ? 1 : abx / n;
        t[1] = n < TOL_ZNORM ? 0 : aby / n;
    } else {
        const real dx = p[0] - e[0], dy = p[1] - e[1];
        const real rn = sqrt(dx * dx + dy * dy);
        const real nx = rn < TOL_ZNORM ? 1 : dx / rn, ny = rn < TOL_ZNORM ? 0 : dy / rn;
        t[0] = -ny;
        t[1] = nx;
    }
}

/* Projection of p onto entity (tag, record) and the unit tangent there, the
 * pair the reference's geometry_project / geometry_tangent bindings return
 * (src/bindings.cpp:1872-1943); tag 0 (free) leaves p and has tangent e_0
 * (Free<2>, src/geometry.hpp). Pinned by tests/cpp/golden_geometry.hpp. */
void FN(o_entity_eval)(int tag, const real* rec, const real* p, real* proj, real* tang)
{
    if (tag == 0) {
        proj[0] = p[0];
        proj[1] = p[1];
        tang[0] = 1;
        tang[1] = 0;
        return;
    }
    FN(ent_project)(tag, rec, p, proj);
    FN(ent_tangent)(tag, rec, p, tang);
}

#include "entity_oracle.c"

/* Tangent-reduced Newton step delta = B y from the orthonormal basis
 * B[0..K-1] (reference src/patch.hpp:348-388): M = B^T H B, r = B^T g, y =
 * solve1x1(M, r) (src/solve.hpp:83-90) or the 2x2 node solve. */
static void FN(reduced_step)(int dim, const real* hess, const real* grad, real B[2][3], int K, real* dl)
{
    real M[4] = {0, 0, 0, 0}, r[2] = {0, 0}, y[2] = {0, 0};
    for (int a = 0; a < K; ++a) {
        real Hb[3];
        for (int i = 0; i < dim; ++i) {
            Hb[i] = 0;
            for (int j = 0; j < dim; ++j) Hb[i] += hess[i * dim + j] * B[a][j];
        }
        for (int c = 0; c < K; ++c) {
            M[a * K + c] = 0;
            for (int i = 0; i < dim; ++i) M[a * K + c] += B[c][i] * Hb[i];
        }
        for (int i = 0; i < dim; ++i) r[a] += B[a][i] * grad[i];
    }
    if (K == 1) {
        const real rn = fabs(r[0]);
        if (rn >= TOL_ZNORM) {
            y[0] = -r[0] / M[0];
            if (!isfinite(y[0])) y[0] = (real)-0.1 * r[0] / rn;
        }
    } else {
        FN(o_solve)(2, M, r, y);
    }
    for (int i = 0; i < dim; ++i) dl[i] = B[0][i] * y[0] + (K == 2 ? B[1][i] * y[1] : 0);
}

/* One constrained DOF's update. Cold (the first sweep of a run; reference
 * src/sweep.hpp:1285-1403): tangent-reduced Newton step at the entity's
 * frame of the DOF (a rail beyond its ends releases the node: full-space
 * step, src/patch.hpp:400-409), damped by omega on the step, trials pos +
 * alpha delta projected onto the entity (the iterative kinds from the
 * DOF's seed, which each projection refreshes). Warm (later sweeps, the
 * parametrized kinds; :1717-1844): project once to the foot q* (seeded),
 * seed = q*, the step reduced to dq = (J^T J)^-1 J^T delta times omega, trials
 * eval(trim clamp(q* + alpha dq)). The free DOFs' accept rule; no post-blend. */
static void FN(update_constrained)(int dim, int kind, real delta, real omega, const FN(patch_t)* pv, const real* X,
                                   int dof, int tag, const real* blob, real* seed, int cold, const dir_t* dw, int d,
                                   real* out)
{
    real grad[3], hess[9], e0, md;
    FN(patch_eval_impl)(dim, kind, delta, pv, X, grad, hess, &e0, &md);
    if (dw) /* directional windows (src/sweep.hpp:1746-1760) */
        FN(dir_dof_accumulate)(dim, dw, d, X, grad, hess, &e0);
    real pos[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) pos[k] = X[(size_t)dof * dim + k];
    const int K = eo_tdim(tag);
    real dl[3] = {0, 0, 0}, qs[2] = {0, 0}, dq[2] = {0, 0}, B[2][3] = {{0, 0, 0}, {0, 0, 0}};
    int pls = 0;
    if (!cold && eo_param(tag)) {
        pls = 1;
        real J[2][3] = {{0, 0, 0}, {0, 0, 0}};
        FN(eo_invert)(tag, blob + 1, pos, seed, isfinite(seed[0]), 1, qs, tag == 14 ? J : NULL);
        if (tag != 14) FN(eo_frame)(tag, blob + 1, qs, J);
        seed[0] = qs[0];
        seed[1] = qs[1];
        memcpy(B, J, sizeof(B));
        FN(eo_gram_schmidt)(B, K, dim);
        real del[3];
        FN(reduced_step)(dim, hess, grad, B, K, del);
        real JtJ[4] = {0, 0, 0, 0}, Jtd[2] = {0, 0};
        for (int a = 0; a < K; ++a) {
            for (int c = 0; c < K; ++c) JtJ[a * 2 + c] = FN(eo_dot)(J[a], J[c], dim);
            Jtd[a] = FN(eo_dot)(J[a], del, dim);
        }
        if (K == 1) {
            dq[0] = fabs(JtJ[0]) > EO_TINY ? Jtd[0] / JtJ[0] : 0;
        } else {
            const real det = JtJ[0] * JtJ[3] - JtJ[1] * JtJ[2];
            if (fabs(det) >= EO_TINY) {
                dq[0] = (JtJ[3] * Jtd[0] - JtJ[1] * Jtd[1]) / det;
                dq[1] = (-JtJ[2] * Jtd[0] + JtJ[0] * Jtd[1]) / det;
            }
        }
        dq[0] *= omega;
        dq[1] *= omega;
    } else {
        int release = 0;
        if (tag == 15 || tag == 16) {
            real ab2 = 0, s = 0;
            const real* e = blob + 1;
            for (int k = 0; k < dim; ++k)
                ab2 += (e[dim + k] - e[k]) * (e[dim + k] - e[k]), s += (pos[k] - e[k]) * (e[dim + k] - e[k]);
            const real ft = s / (ab2 > EO_TINY ? ab2 : EO_TINY);
            release = ft < 0 || ft > 1;
        }
        if (release)
            FN(o_solve)(dim, hess, grad, dl);
        else {
            if (eo_seeded(tag))
                FN(eo_seeded_frame)(tag, blob, pos, seed, 0, NULL, B);
            else
                FN(eo_entity_frame)(tag, blob, pos, NULL, B);
            FN(reduced_step)(dim, hess, grad, B, K, dl);
        }
        for (int k = 0; k < dim; ++k) dl[k] *= omega;
    }
    real alpha = 1, cur[3] = {pos[0], pos[1], pos[2]};
    for (int it = 0; it < MAX_BACKTRACKS; ++it) {
        real trial[3] = {0, 0, 0}, e_new, mdet;
        if (pls) {
            real qt[2] = {qs[0] + alpha * dq[0], qs[1] + alpha * dq[1]};
            FN(eo_clamp_trim)(tag, blob + 1, qt);
            FN(eo_eval)(tag, blob + 1, qt, trial);
        } else {
            real raw[3] = {0, 0, 0};
            for (int k = 0; k < dim; ++k) raw[k] = pos[k] + alpha * dl[k];
            if (eo_seeded(tag))
                FN(eo_seeded_frame)(tag, blob, raw, seed, 0, trial, NULL);
            else
                FN(eo_entity_frame)(tag, blob, raw, trial, NULL);
        }
        FN(trial_impl)(dim, kind, delta, pv, X, dof, trial, &e_new, &mdet);
        if (dw) e_new += FN(dir_dof_trial)(dim, dw, d, X, trial);
        const int ok = isfinite(e_new) && (e_new <= e0 + TOL_ENERGY) && (kind == KIND_UNTANGLE || mdet > 0);
        if (ok) {
            for (int k = 0; k < dim; ++k) cur[k] = trial[k];
            break;
        }
        alpha *= (real)0.5;
    }
    for (int k = 0; k < dim; ++k) out[k] = cur[k];
}

/* ------------------------------------------- block-interface C2 curvature */

/* A window of four nodes p0..p3 along a grid line: residual r = turn(p0, p1,
 * p2) - turn(p1, p2, p3) (signed turning angles), energy w r^2; the DOF is
 * node `slot` of the window (reference src/curvature.hpp). */
static double FN(curv_turn)(const double* a, const double* b, const double* c)
{
    const double e1x = b[0] - a[0], e1y = b[1] - a[1], e2x = c[0] - b[0], e2y = c[1] - b[1];
    return atan2(e1x * e2y - e1y * e2x, e1x * e2x + e1y * e2y);
}

/* r and dr/dp_m (m = 0..3) of one window (reference src/curvature.hpp:57-94):
 * the turning angle's gradient is perp(e)/|e|^2 per edge. */
static double FN(curv_residual_grad)(const double p[4][2], double dr[4][2])
{
    double g[3][2];
    for (int e = 0; e < 3; ++e) {
        const double ex = p[e + 1][0] - p[e][0], ey = p[e + 1][1] - p[e][1];
        const double l = ex * ex + ey * ey;
        const double il = l > 0 ? 1.0 / l : 0.0;
        g[e][0] = -ey * il;
        g[e][1] = ex * il;
    }
    for (int i = 0; i < 2; ++i) {
        dr[0][i] = g[0][i];
        dr[1][i] = -g[0][i] - 2.0 * g[1][i];
        dr[2][i] = 2.0 * g[1][i] + g[2][i];
        dr[3][i] = -g[2][i];
    }
    return FN(curv_turn)(p[0], p[1], p[2]) - FN(curv_turn)(p[1], p[2], p[3]);
}

/* The DOF's windows into grad, hess (2x2) and energy (reference
 * src/curvature.hpp:116-164): gradient 2 w r dr_s, Gauss-Newton Hessian
 * 2 w dr_s dr_s^T plus the diagonal stabiliser 2 w |dr_s| sum_{m != s} |dr_m|,
 * energy w r^2. Windows with slot < 0 are padding. */
static void FN(curv_accumulate)(int K, const int* nodes, const int* slot, const real* wt, const real* X,
                                real* grad, real* hess, real* e0)
{
    for (int j = 0; j < K; ++j) {
        const int s = slot[j];
        if (s < 0) continue;
        double p[4][2], dr[4][2];
        for (int m = 0; m < 4; ++m)
            for (int i = 0; i < 2; ++i) p[m][i] = (double)X[(size_t)nodes[4 * j + m] * 2 + i];
        const double r = FN(curv_residual_grad)(p, dr);
        const double w = (double)wt[j];
        const double gc = 2.0 * w * r, hc = 2.0 * w;
        grad[0] += (real)(gc * dr[s][0]);
        grad[1] += (real)(gc * dr[s][1]);
        double off = 0;
        for (int m = 0; m < 4; ++m)
            if (m != s) off += sqrt(dr[m][0] * dr[m][0] + dr[m][1] * dr[m][1]);
        const double stab = hc * sqrt(dr[s][0] * dr[s][0] + dr[s][1] * dr[s][1]) * off;
        hess[0] += (real)(hc * dr[s][0] * dr[s][0] + stab);
        hess[1] += (real)(hc * dr[s][0] * dr[s][1]);
        hess[2] += (real)(hc * dr[s][1] * dr[s][0]);
        hess[3] += (real)(hc * dr[s][1] * dr[s][1] + stab);
        *e0 += (real)(w * r * r);
    }
}

/* One curvature window's gradient, Hessian and energy for its node `slot`
 * (the per-window body of curv_accumulate, for tests). */
void FN(o_curv_eval)(const real* p8, real w, int slot, real* grad, real* hess, real* e0)
{
    int nodes[4] = {0, 1, 2, 3};
    grad[0] = grad[1] = 0;
    hess[0] = hess[1] = hess[2] = hess[3] = 0;
    *e0 = 0;
    FN(curv_accumulate)(1, nodes, &slot, &w, p8, grad, hess, e0);
}

/* The windows' energy with the DOF (node `slot` of each) moved to `trial`
 * (reference src/curvature.hpp:166-183). */
static real FN(curv_trial)(int K, const int* nodes, const int* slot, const real* wt, const real* X,
                           const real* trial)
{
    double e = 0;
    for (int j = 0; j < K; ++j) {
        const int s = slot[j];
        if (s < 0) continue;
        double p[4][2];
        for (int m = 0; m < 4; ++m)
            for (int i = 0; i < 2; ++i)
                p[m][i] = m == s ? (double)trial[i] : (double)X[(size_t)nodes[4 * j + m] * 2 + i];
        const double r = FN(curv_turn)(p[0], p[1], p[2]) - FN(curv_turn)(p[1], p[2], p[3]);
        e += (double)wt[j] * r * r;
    }
    return (real)e;
}

typedef struct {
    int K;
    const int* nodes; /* [K*4] global node ids */
    const int* slot;  /* [K] the DOF's slot in each window, -1 = padding */
    const real* wt;   /* [K] window weight */
} FN(curv_t);

/* ------------------------------------------------------------ Jacobi sweep */

/* One DOF's block-Jacobi update: metric pass, Newton, <=10 halvings, omega
 * post-blend (reference src/sweep.hpp:1534-1650). X is the frozen snapshot. */
static void FN(update_dof)(int dim, int kind, real delta, real omega, const FN(patch_t)* pv, const real* X, int dof,
                       const real* tau, const FN(curv_t)* cv, const dir_t* dw, int d, real* out)
{
    real grad[3], hess[9], e0, md;
    FN(patch_eval_impl)(dim, kind, delta, pv, X, grad, hess, &e0, &md);
    if (cv) /* curvature windows on top of the shape metric (src/sweep.hpp:1461-1487) */
        FN(curv_accumulate)(cv->K, cv->nodes, cv->slot, cv->wt, X, grad, hess, &e0);
    if (dw) /* directional windows (src/sweep.hpp:1488-1510) */
        FN(dir_dof_accumulate)(dim, dw, d, X, grad, hess, &e0);
    real pos[3];
    for (int k = 0; k < dim; ++k) pos[k] = X[(size_t)dof * dim + k];
    if (tau) {
        for (int k = 0; k < dim; ++k) {
            grad[k] -= tau[(size_t)dof * dim + k];
            e0 -= tau[(size_t)dof * dim + k] * pos[k];
        }
    }
    real dlt[3];
    FN(o_solve)(dim, hess, grad, dlt);
    real alpha = 1, cur[3];
    for (int k = 0; k < dim; ++k) cur[k] = pos[k];
    for (int it = 0; it < MAX_BACKTRACKS; ++it) {
        real trial[3], e_new, mdet;
        for (int k = 0; k < dim; ++k) trial[k] = pos[k] + alpha * dlt[k];
        FN(trial_impl)(dim, kind, delta, pv, X, dof, trial, &e_new, &mdet);
        if (cv) e_new += FN(curv_trial)(cv->K, cv->nodes, cv->slot, cv->wt, X, trial);
        if (dw) e_new += FN(dir_dof_trial)(dim, dw, d, X, trial);
        if (tau)
            for (int k = 0; k < dim; ++k) e_new -= tau[(size_t)dof * dim + k] * trial[k];
        const int ok = isfinite(e_new) && (e_new <= e0 + TOL_ENERGY) && (kind == KIND_UNTANGLE || mdet > 0);
        if (ok) {
            for (int k = 0; k < dim; ++k) cur[k] = trial[k];
            break;
        }
        alpha *= (real)0.5;
    }
    if (omega != (real)1)
        for (int k = 0; k < dim; ++k) cur[k] = pos[k] + omega * (cur[k] - pos[k]);
    for (int k = 0; k < dim; ++k) out[k] = cur[k];
}

/*
 * n_sweeps of double-buffered block-Jacobi over the free DOFs `dof[ndof]`
 * (global ids). DOF d owns occurrences [poff[d], poff[d+1]) of the patch
 * arrays (pgn/ps are [dim][total]). X (n_nodes*dim) is updated in place with
 * the final state. Reports (E, min det) on sweeps (s+1) % k == 0 and the
 * last one, with k = report_every <= 0 ? n_sweeps : report_every.
 * `tau` (nullable, n_nodes*dim) is the FAS linear shift. Returns the number
 * of reports written.
 */
int FN(o_jacobi)(int dim, int kind, double delta, double omega, int n_sweeps, int report_every, int ndof,
                 const int* dof, const long long* poff, const int* pgc, const int* pgn, const real* ps,
                 const int* prole, const real* W, int w_stride, const real* pwt, const int* etag,
                 const long long* eoff, const real* eblob, real* eseed, int curv_k, const int* cnodes, const int* cslot, const real* cwt,
                 const dir_t* dw, const dir_t* dirs, int n_nodes, real* X, const real* tau, long long ns, const int* egc,
                 const int* egn, const real* es, const real* eW, int ew_stride, const real* ewt, real* work,
                 double* energies, double* mindets)
{
    const long long total = poff[ndof];
    const int k = report_every <= 0 ? n_sweeps : report_every;
    int nrep = 0;
    real* xin = X;
    real* xout = work; /* n_nodes*dim */
    memcpy(xout, xin, sizeof(real) * (size_t)n_nodes * dim);
    for (int sw = 0; sw < n_sweeps; ++sw) {
#pragma omp parallel for schedule(static)
        for (int d = 0; d < ndof; ++d) {
            const long long o = poff[d];
            const int P = (int)(poff[d + 1] - o);
            /* per-DOF view into the [dim][total] arrays: gn/s rows have stride total */
            int gn_l[3 * 128];
            real s_l[3 * 128];
            for (int kk = 0; kk < dim; ++kk)
                for (int p = 0; p < P; ++p) {
                    gn_l[kk * P + p] = pgn[(size_t)kk * total + o + p];
                    s_l[kk * P + p] = ps[(size_t)kk * total + o + p];
                }
            FN(patch_t) pv = {P, pgc + o, gn_l, s_l, prole + o, W + (size_t)w_stride * o, w_stride,
                              pwt ? pwt + o : NULL};
            if (etag && etag[d]) /* constrained boundary DOF; the first sweep of a run is cold */
                FN(update_constrained)(dim, kind, (real)delta, (real)omega, &pv, xin, dof[d], etag[d],
                                       eblob + eoff[d], eseed + 2 * (size_t)d, sw == 0, dw && dw->K ? dw : NULL, d,
                                       xout + (size_t)dof[d] * dim);
            else {
                FN(curv_t) cv = {curv_k, cnodes ? cnodes + (size_t)4 * curv_k * d : NULL,
                                 cslot ? cslot + (size_t)curv_k * d : NULL, cwt ? cwt + (size_t)curv_k * d : NULL};
                FN(update_dof)(dim, kind, (real)delta, (real)omega, &pv, xin, dof[d], tau,
                               curv_k > 0 ? &cv : NULL, dw && dw->K ? dw : NULL, d, xout + (size_t)dof[d] * dim);
            }
        }
        if ((sw + 1) % k == 0 || sw == n_sweeps - 1) {
            FN(o_energy_mindet)(dim, kind, delta, ns, egc, egn, es, eW, ew_stride, ewt, xout, dirs, &energies[nrep],
                                &mindets[nrep]);
            ++nrep;
        }
        real* t = xin;
        xin = xout;
        xout = t;
    }
    if (xin != X) memcpy(X, xin, sizeof(real) * (size_t)n_nodes * dim);
    return nrep;
}

/*
 * Sequential Gauss-Seidel restatement of the reference's NumPy node
 * baseline (tests/smoothing/test_control_ref.py:95-135): the DOFs in order,
 * each updated in place from the current X; Newton step -H^{-1} g of its
 * patch (skipped when H is exactly singular, as np.linalg.solve refuses
 * it), alpha = 1, 1/2, ... while alpha >= 1e-4, the first trial with a
 * finite patch energy <= e0 + 1e-12 and patch min det > 0 taken; stop after
 * a sweep that moved the nodes by less than 1e-12 in total. Independent of
 * the block-Jacobi update (no omega, no frozen snapshot, another ladder and
 * solve), it converges to the same minimiser. Returns the sweeps run.
 */
int FN(o_gauss_seidel)(int dim, int kind, double delta, int max_sweeps, int ndof, const int* dof,
                       const long long* poff, const int* pgc, const int* pgn, const real* ps, const int* prole,
                       const real* W, int w_stride, real* X)
{
    const long long total = poff[ndof];
    int sw = 0;
    while (sw < max_sweeps) {
        ++sw;
        double moved = 0;
        for (int d = 0; d < ndof; ++d) {
            const long long o = poff[d];
            const int P = (int)(poff[d + 1] - o);
            int gn_l[3 * 128];
            real s_l[3 * 128];
            for (int kk = 0; kk < dim; ++kk)
                for (int p = 0; p < P; ++p) {
                    gn_l[kk * P + p] = pgn[(size_t)kk * total + o + p];
                    s_l[kk * P + p] = ps[(size_t)kk * total + o + p];
                }
            FN(patch_t) pv = {P, pgc + o, gn_l, s_l, prole + o, W + (size_t)w_stride * o, w_stride, NULL};
            real g[3], H[9], e0, md;
            FN(patch_eval_impl)(dim, kind, (real)delta, &pv, X, g, H, &e0, &md);
            /* -H^{-1} g by the adjugate; an exactly singular H is skipped */
            real adj[9], det;
            if (dim == 2) {
                det = H[0] * H[3] - H[1] * H[2];
                adj[0] = H[3]; adj[1] = -H[1]; adj[2] = -H[2]; adj[3] = H[0];
            } else {
                adj[0] = H[4] * H[8] - H[5] * H[7];
                adj[1] = H[2] * H[7] - H[1] * H[8];
                adj[2] = H[1] * H[5] - H[2] * H[4];
                adj[3] = H[5] * H[6] - H[3] * H[8];
                adj[4] = H[0] * H[8] - H[2] * H[6];
                adj[5] = H[2] * H[3] - H[0] * H[5];
                adj[6] = H[3] * H[7] - H[4] * H[6];
                adj[7] = H[1] * H[6] - H[0] * H[7];
                adj[8] = H[0] * H[4] - H[1] * H[3];
                det = H[0] * adj[0] + H[1] * adj[3] + H[2] * adj[6];
            }
            if (det == 0 || !isfinite(det)) continue;
            real step[3];
            for (int i = 0; i < dim; ++i) {
                real acc = 0;
                for (int j = 0; j < dim; ++j) acc += adj[i * dim + j] * g[j];
                step[i] = -acc / det;
            }
            const int n = dof[d];
            for (real alpha = 1; alpha >= (real)1e-4; alpha *= (real)0.5) {
                real trial[3], e, m;
                for (int i = 0; i < dim; ++i) trial[i] = X[(size_t)n * dim + i] + alpha * step[i];
                FN(trial_impl)(dim, kind, (real)delta, &pv, X, n, trial, &e, &m);
                if (isfinite(e) && e <= e0 + (real)1e-12 && m > 0) {
                    double len = 0;
                    for (int i = 0; i < dim; ++i) {
                        X[(size_t)n * dim + i] = trial[i];
                        len += (double)step[i] * step[i];
                    }
                    moved += alpha * sqrt(len);
                    break;
                }
            }
        }
        if (moved < 1e-12) break;
    }
    return sw;
}

/* Per-DOF patch gradient (minus tau when given) into gout[dof*dim + k]
 * (the FAS restriction input; reference src/sweep.hpp:1425-1514 metric_kernel
 * with HasTau). */
void FN(o_patch_grads)(int dim, int kind, double delta, int ndof, const int* dof, const long long* poff,
                       const int* pgc, const int* pgn, const real* ps, const int* prole, const real* W,
                       int w_stride, const real* pwt, int curv_k, const int* cnodes, const int* cslot,
                       const real* cwt, const dir_t* dw, const real* X, const real* tau, real* gout)
{
    const long long total = poff[ndof];
#pragma omp parallel for schedule(static)
    for (int d = 0; d < ndof; ++d) {
        const long long o = poff[d];
        const int P = (int)(poff[d + 1] - o);
        int gn_l[3 * 128];
        real s_l[3 * 128];
        for (int kk = 0; kk < dim; ++kk)
            for (int p = 0; p < P; ++p) {
                gn_l[kk * P + p] = pgn[(size_t)kk * total + o + p];
                s_l[kk * P + p] = ps[(size_t)kk * total + o + p];
            }
        FN(patch_t) pv = {P, pgc + o, gn_l, s_l, prole + o, W + (size_t)w_stride * o, w_stride,
                          pwt ? pwt + o : NULL};
        real grad[3], hess[9], e0, md;
        FN(patch_eval_impl)(dim, kind, (real)delta, &pv, X, grad, hess, &e0, &md);
        if (curv_k > 0)
            FN(curv_accumulate)(curv_k, cnodes + (size_t)4 * curv_k * d, cslot + (size_t)curv_k * d,
                                cwt + (size_t)curv_k * d, X, grad, hess, &e0);
        if (dw && dw->K) FN(dir_dof_accumulate)(dim, dw, d, X, grad, hess, &e0);
        for (int k = 0; k < dim; ++k)
            gout[(size_t)dof[d] * dim + k] = grad[k] - (tau ? tau[(size_t)dof[d] * dim + k] : (real)0);
    }
}
