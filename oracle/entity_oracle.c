/*
 * Boundary entities of the constrained sweep, CPU restatement (TEST
 * INFRASTRUCTURE ONLY; included by egg_oracle.c once per precision).
 *
 * Records use the compute core's layout (value 0 the tag, then the kind's
 * fields and payloads; the wire conversion in oracle.py builds them from
 * the reference wire). Restated from the reference (paths relative to its
 * root):
 *   closed-form kinds   src/geometry.hpp:288-464 (segment, rail, circle,
 *                       ellipse, 3D rail)
 *   trimmed kinds       :195-258 (invert, period re-representation, trim
 *                       test and clamp, frame), :118-192 (Trim<1>, Trim<2>)
 *   curves              :475-684 (arcs, Beziers), :729-963 (B-spline /
 *                       NURBS; here by the full Cox-de Boor recursion over
 *                       every basis function, started from the clamped span
 *                       so that parameters past the ends extrapolate the
 *                       end piece), :517-553 (seeded foot)
 *   surfaces            :1001-1094 (plane, sphere, cylinder, 3D line),
 *                       :1095-1395 (B-spline surface: Gauss-Newton / exact
 *                       Newton / damped feet, multi-start cold search)
 *   composite           :2580-2612 (nearest segment)
 * Pinned in tests/test_entities.py by closed forms (the NURBS circle is the
 * circle, a degree-1 B-spline is its polyline, a bilinear patch is its
 * plane, sphere and cylinder feet are radial), finite differences of every
 * parametrization's frame, and the golden vectors of the closed-form kinds.
 */

#define EO_TINY ((real)(IS_F32 ? 1e-20 : 1e-30))
#define EO_NEWTON ((real)(IS_F32 ? 1e-5 : 1e-9))
#define EO_TWO_PI ((real)6.28318530717958647692)
#define EO_MAXN 256 /* basis functions per direction in the oracle's tables */

static real FN(eo_dot)(const real* a, const real* b, int n)
{
    real s = 0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
static void FN(eo_normalize)(real* a, int n)
{
    const real nn = sqrt(FN(eo_dot)(a, a, n));
    if (nn < TOL_ZNORM) {
        for (int i = 0; i < n; ++i) a[i] = i == 0 ? 1 : 0;
        return;
    }
    for (int i = 0; i < n; ++i) a[i] = (1 / nn) * a[i];
}
static void FN(eo_gram_schmidt)(real B[2][3], int K, int D)
{
    for (int i = 0; i < K; ++i) {
        for (int j = 0; j < i; ++j) {
            const real c = FN(eo_dot)(B[i], B[j], D);
            for (int k = 0; k < D; ++k) B[i][k] -= c * B[j][k];
        }
        FN(eo_normalize)(B[i], D);
    }
}
static real FN(eo_clamp)(real v, real lo, real hi) { return v < lo ? lo : (v > hi ? hi : v); }

#ifndef EO_KINDS_ONCE
#define EO_KINDS_ONCE
static int eo_dim(int tag) { return (tag == 4 || tag == 5 || tag == 12 || tag == 13 || tag == 14 || tag == 16) ? 3 : 2; }
static int eo_tdim(int tag) { return (tag == 4 || tag == 5 || tag == 12 || tag == 14) ? 2 : 1; }
static int eo_param(int tag) { return tag == 4 || tag == 5 || (tag >= 6 && tag <= 10) || tag == 12 || tag == 13 || tag == 14; }
static int eo_seeded(int tag) { return tag == 10 || tag == 14; }
#endif

/* ----------------------------------------------------------- B-splines */

/* Values (and d, dd when non-null) of every degree-p basis function at u
 * over the n + p + 1 knots U: degree 0 is the indicator of the span the
 * point is assigned to (the last non-empty span at or before u, kept inside
 * [p, n - 1]); then N_{i,k} = (u - U_i)/(U_{i+k} - U_i) N_{i,k-1} +
 * (U_{i+k+1} - u)/(U_{i+k+1} - U_{i+1}) N_{i+1,k-1} for every i, with the
 * derivatives by k N_{i,k-1}/(U_{i+k} - U_i) - k N_{i+1,k-1}/(...). */
static void FN(eo_basis_all)(const real* U, int n, int p, real u, real* N, real* d, real* dd)
{
    const int m = n + p; /* knot intervals */
    int s = p;
    for (int i = p; i <= n - 1; ++i)
        if (U[i] <= u && U[i] < U[i + 1]) s = i;
    real cur[EO_MAXN + 8], prv[EO_MAXN + 8], dcur[EO_MAXN + 8], dprv[EO_MAXN + 8], ddcur[EO_MAXN + 8],
        ddprv[EO_MAXN + 8];
    for (int i = 0; i < m; ++i) {
        cur[i] = i == s ? 1 : 0;
        dcur[i] = 0;
        ddcur[i] = 0;
    }
    for (int k = 1; k <= p; ++k) {
        memcpy(prv, cur, sizeof(real) * (size_t)m);
        memcpy(dprv, dcur, sizeof(real) * (size_t)m);
        memcpy(ddprv, ddcur, sizeof(real) * (size_t)m);
        for (int i = 0; i + k < m; ++i) {
            const real a = U[i + k] - U[i], b = U[i + k + 1] - U[i + 1];
            const real fa = a != 0 ? (u - U[i]) / a : 0, fb = b != 0 ? (U[i + k + 1] - u) / b : 0;
            const real ga = a != 0 ? k / a : 0, gb = b != 0 ? k / b : 0;
            cur[i] = fa * prv[i] + fb * prv[i + 1];
            dcur[i] = ga * prv[i] - gb * prv[i + 1];
            ddcur[i] = ga * dprv[i] - gb * dprv[i + 1];
        }
    }
    for (int i = 0; i < n; ++i) {
        N[i] = cur[i];
        if (d) d[i] = dcur[i];
        if (dd) dd[i] = ddcur[i];
    }
}

/* ------------------------------------------------------------- curves */

/* C, C', C'' of the iterative 2D curves (ellipse arc, Beziers, B-spline) */
static void FN(eo_curve)(int tag, const real* f, real t, real C[3][2])
{
    if (tag == 7) {
        const real cp = cos(f[4]), sp = sin(f[4]), a = f[2], b = f[3];
        const real x[3] = {a * cos(t), -a * sin(t), -a * cos(t)}, y[3] = {b * sin(t), b * cos(t), -b * sin(t)};
        for (int k = 0; k < 3; ++k) {
            C[k][0] = (k == 0 ? f[0] : 0) + cp * x[k] - sp * y[k];
            C[k][1] = (k == 0 ? f[1] : 0) + sp * x[k] + cp * y[k];
        }
        return;
    }
    if (tag == 8 || tag == 9) {
        /* Bernstein polynomials of degree 2 or 3 and their derivatives */
        const int deg = tag == 8 ? 2 : 3;
        const real u = 1 - t;
        real B0[4], B1[4], B2[4];
        if (deg == 2) {
            B0[0] = u * u, B0[1] = 2 * u * t, B0[2] = t * t;
            B1[0] = -2 * u, B1[1] = 2 * (u - t), B1[2] = 2 * t;
            B2[0] = 2, B2[1] = -4, B2[2] = 2;
        } else {
            B0[0] = u * u * u, B0[1] = 3 * u * u * t, B0[2] = 3 * u * t * t, B0[3] = t * t * t;
            B1[0] = -3 * u * u, B1[1] = 3 * u * u - 6 * u * t, B1[2] = 6 * u * t - 3 * t * t, B1[3] = 3 * t * t;
            B2[0] = 6 * u, B2[1] = 6 * t - 12 * u, B2[2] = 6 * u - 12 * t, B2[3] = 6 * t;
        }
        for (int i = 0; i < 2; ++i) {
            C[0][i] = C[1][i] = C[2][i] = 0;
            for (int j = 0; j <= deg; ++j) {
                C[0][i] += B0[j] * f[2 * j + i];
                C[1][i] += B1[j] * f[2 * j + i];
                C[2][i] += B2[j] * f[2 * j + i];
            }
        }
        return;
    }
    /* B-spline / NURBS */
    const int p = (int)f[0], n = (int)f[1];
    const int rat = f[5] != 0;
    const real* U = f + 6;
    const real* P = U + n + p + 1;
    const real* W = P + 2 * n;
    real N[EO_MAXN], d[EO_MAXN], dd[EO_MAXN];
    FN(eo_basis_all)(U, n, p, t, N, d, dd);
    real A[3][2] = {{0, 0}, {0, 0}, {0, 0}}, w[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        const real wi = rat ? W[i] : 1;
        const real b[3] = {N[i] * wi, d[i] * wi, dd[i] * wi};
        for (int k = 0; k < 3; ++k) {
            A[k][0] += b[k] * P[2 * i];
            A[k][1] += b[k] * P[2 * i + 1];
            w[k] += b[k];
        }
    }
    for (int i = 0; i < 2; ++i) {
        if (!rat) {
            C[0][i] = A[0][i], C[1][i] = A[1][i], C[2][i] = A[2][i];
            continue;
        }
        C[0][i] = A[0][i] / w[0];
        C[1][i] = (A[1][i] - w[1] * C[0][i]) / w[0];
        C[2][i] = (A[2][i] - 2 * w[1] * C[1][i] - w[2] * C[0][i]) / w[0];
    }
}

static real FN(eo_curve_foot)(int tag, const real* f, const real* q, real lo, real hi, real seed, int has_seed)
{
    real t = lo;
    if (has_seed) {
        t = FN(eo_clamp)(seed, lo, hi);
    } else {
        real best = (real)INFINITY;
        for (int i = 0; i <= 16; ++i) {
            const real tc = lo + ((hi - lo) * i) / 16;
            real C[3][2];
            FN(eo_curve)(tag, f, tc, C);
            const real dd = (C[0][0] - q[0]) * (C[0][0] - q[0]) + (C[0][1] - q[1]) * (C[0][1] - q[1]);
            if (dd < best) {
                best = dd;
                t = tc;
            }
        }
    }
    for (int it = 0; it < 8; ++it) {
        real C[3][2];
        FN(eo_curve)(tag, f, t, C);
        const real dx = C[0][0] - q[0], dy = C[0][1] - q[1];
        const real fv = dx * C[1][0] + dy * C[1][1];
        const real fp = C[1][0] * C[1][0] + C[1][1] * C[1][1] + dx * C[2][0] + dy * C[2][1];
        if (fabs(fp) < EO_TINY) break;
        t -= fv / fp;
    }
    return t;
}

/* ------------------------------------------------------------ surface */

typedef struct {
    int pu, pv, nu, nv, rat, nl;
    const real *Uu, *Uv, *P, *W, *loops, *verts;
} FN(eo_surf_t);

static FN(eo_surf_t) FN(eo_surf)(const real* f)
{
    FN(eo_surf_t) S;
    S.pu = (int)f[0], S.pv = (int)f[1], S.nu = (int)f[2], S.nv = (int)f[3];
    S.rat = f[4] != 0;
    S.nl = (int)f[5];
    S.Uu = f + 6;
    S.Uv = S.Uu + S.nu + S.pu + 1;
    S.P = S.Uv + S.nv + S.pv + 1;
    S.W = S.P + 3 * S.nu * S.nv;
    S.loops = S.W + (S.rat ? S.nu * S.nv : 0);
    S.verts = S.loops + S.nl;
    return S;
}

/* S, S_u, S_v, S_uv, S_uu, S_vv */
static void FN(eo_surf_eval)(const FN(eo_surf_t) * S, const real* q, real D[6][3])
{
    real Nu[3][EO_MAXN], Nv[3][EO_MAXN];
    FN(eo_basis_all)(S->Uu, S->nu, S->pu, q[0], Nu[0], Nu[1], Nu[2]);
    FN(eo_basis_all)(S->Uv, S->nv, S->pv, q[1], Nv[0], Nv[1], Nv[2]);
    static const int ou[6] = {0, 1, 0, 1, 2, 0}, ov[6] = {0, 0, 1, 1, 0, 2};
    real A[6][3], w[6];
    memset(A, 0, sizeof(A));
    memset(w, 0, sizeof(w));
    for (int a = 0; a < S->nu; ++a)
        for (int b = 0; b < S->nv; ++b) {
            const int c = a * S->nv + b;
            const real wc = S->rat ? S->W[c] : 1;
            for (int m = 0; m < 6; ++m) {
                const real nw = Nu[ou[m]][a] * Nv[ov[m]][b] * wc;
                if (nw == 0) continue;
                for (int k = 0; k < 3; ++k) A[m][k] += nw * S->P[3 * c + k];
                w[m] += nw;
            }
        }
    for (int k = 0; k < 3; ++k) {
        if (!S->rat) {
            for (int m = 0; m < 6; ++m) D[m][k] = A[m][k];
            continue;
        }
        D[0][k] = A[0][k] / w[0];
        D[1][k] = (A[1][k] - w[1] * D[0][k]) / w[0];
        D[2][k] = (A[2][k] - w[2] * D[0][k]) / w[0];
        D[3][k] = (A[3][k] - w[1] * D[2][k] - w[2] * D[1][k] - w[3] * D[0][k]) / w[0];
        D[4][k] = (A[4][k] - 2 * w[1] * D[1][k] - w[4] * D[0][k]) / w[0];
        D[5][k] = (A[5][k] - 2 * w[2] * D[2][k] - w[5] * D[0][k]) / w[0];
    }
}

static real FN(eo_surf_dist2)(const FN(eo_surf_t) * S, const real* q, const real* p)
{
    real D[6][3];
    FN(eo_surf_eval)(S, q, D);
    real s = 0;
    for (int k = 0; k < 3; ++k) s += (D[0][k] - p[k]) * (D[0][k] - p[k]);
    return s;
}

/* Newton on (S - p).S_u = (S - p).S_v = 0, iterates clamped to the domain;
 * exact: with the curvature terms; J: S_u, S_v at the last iterate */
static void FN(eo_surf_newton)(const FN(eo_surf_t) * S, const real* p, real* q, const real* dom, int exact,
                               real J[2][3])
{
    for (int it = 0; it < 12; ++it) {
        real D[6][3], d[3];
        FN(eo_surf_eval)(S, q, D);
        if (J)
            for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
        for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
        const real f1 = FN(eo_dot)(d, D[1], 3), f2 = FN(eo_dot)(d, D[2], 3);
        real j11 = FN(eo_dot)(D[1], D[1], 3), j12 = FN(eo_dot)(D[1], D[2], 3), j22 = FN(eo_dot)(D[2], D[2], 3);
        if (exact) {
            j11 += FN(eo_dot)(d, D[4], 3);
            j12 += FN(eo_dot)(d, D[3], 3);
            j22 += FN(eo_dot)(d, D[5], 3);
        }
        const real det = j11 * j22 - j12 * j12;
        if (fabs(det) < EO_TINY) break;
        const real qu = FN(eo_clamp)(q[0] - (j22 * f1 - j12 * f2) / det, dom[0], dom[1]);
        const real qv = FN(eo_clamp)(q[1] - (-j12 * f1 + j11 * f2) / det, dom[2], dom[3]);
        const int done = fabs(qu - q[0]) + fabs(qv - q[1]) < EO_NEWTON;
        q[0] = qu;
        q[1] = qv;
        if (done) break;
    }
}

static void FN(eo_surf_lm)(const FN(eo_surf_t) * S, const real* p, real* q, const real* dom, real J[2][3])
{
    q[0] = FN(eo_clamp)(q[0], dom[0], dom[1]);
    q[1] = FN(eo_clamp)(q[1], dom[2], dom[3]);
    real D[6][3], d[3];
    FN(eo_surf_eval)(S, q, D);
    for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
    real r = FN(eo_dot)(d, d, 3);
    real lam = (real)1e-3 * (FN(eo_dot)(D[1], D[1], 3) + FN(eo_dot)(D[2], D[2], 3) + EO_TINY);
    for (int it = 0; it < 16; ++it) {
        const real f1 = FN(eo_dot)(d, D[1], 3), f2 = FN(eo_dot)(d, D[2], 3);
        const real j11 = FN(eo_dot)(D[1], D[1], 3) + FN(eo_dot)(d, D[4], 3);
        const real j12 = FN(eo_dot)(D[1], D[2], 3) + FN(eo_dot)(d, D[3], 3);
        const real j22 = FN(eo_dot)(D[2], D[2], 3) + FN(eo_dot)(d, D[5], 3);
        real du = 0, dv = 0;
        int ok = 0;
        for (int bt = 0; bt < 6; ++bt) {
            const real a11 = j11 + lam, a22 = j22 + lam, det = a11 * a22 - j12 * j12;
            if (fabs(det) < EO_TINY) {
                lam *= 4;
                continue;
            }
            du = -((a22 * f1 - j12 * f2) / det);
            dv = -((-j12 * f1 + a11 * f2) / det);
            const real qn[2] = {FN(eo_clamp)(q[0] + du, dom[0], dom[1]), FN(eo_clamp)(q[1] + dv, dom[2], dom[3])};
            const real rn = FN(eo_surf_dist2)(S, qn, p);
            if (rn < r) {
                q[0] = qn[0], q[1] = qn[1];
                r = rn;
                FN(eo_surf_eval)(S, q, D);
                for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
                lam = lam * (real)0.3 > EO_TINY ? lam * (real)0.3 : EO_TINY;
                ok = 1;
                break;
            }
            lam *= 4;
        }
        if (!ok || fabs(du) + fabs(dv) < EO_NEWTON) break;
    }
    if (J)
        for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
}

static void FN(eo_surf_foot)(const FN(eo_surf_t) * S, const real* p, const real* seed, int has_seed, int warm,
                             real* q, real J[2][3])
{
    const real dom[4] = {S->Uu[S->pu], S->Uu[S->nu], S->Uv[S->pv], S->Uv[S->nv]};
    if (warm || has_seed) {
        q[0] = FN(eo_clamp)(seed[0], dom[0], dom[1]);
        q[1] = FN(eo_clamp)(seed[1], dom[2], dom[3]);
        FN(eo_surf_newton)(S, p, q, dom, 0, J);
        return;
    }
    /* the four nearest of a 9 x 9 grid (first found wins ties), each
       polished by the damped Newton; the nearest foot kept */
    real bd[4], bq[4][2];
    for (int s = 0; s < 4; ++s) bd[s] = (real)INFINITY, bq[s][0] = dom[0], bq[s][1] = dom[2];
    for (int i = 0; i <= 8; ++i)
        for (int j = 0; j <= 8; ++j) {
            const real t[2] = {dom[0] + ((dom[1] - dom[0]) * i) / 8, dom[2] + ((dom[3] - dom[2]) * j) / 8};
            const real dd = FN(eo_surf_dist2)(S, t, p);
            int s = 4;
            while (s > 0 && dd < bd[s - 1]) --s;
            if (s == 4) continue;
            for (int r = 3; r > s; --r) bd[r] = bd[r - 1], bq[r][0] = bq[r - 1][0], bq[r][1] = bq[r - 1][1];
            bd[s] = dd, bq[s][0] = t[0], bq[s][1] = t[1];
        }
    q[0] = dom[0], q[1] = dom[2];
    real best = (real)INFINITY;
    for (int s = 0; s < 4; ++s) {
        if (!(bd[s] < (real)INFINITY)) continue;
        real qs[2] = {bq[s][0], bq[s][1]}, Js[2][3];
        FN(eo_surf_lm)(S, p, qs, dom, Js);
        const real dd = FN(eo_surf_dist2)(S, qs, p);
        if (dd < best) {
            best = dd;
            q[0] = qs[0], q[1] = qs[1];
            if (J) memcpy(J, Js, sizeof(Js));
        }
    }
}

static int FN(eo_uv_inside)(const FN(eo_surf_t) * S, const real* uv)
{
    if (S->nl < 2) return 1;
    int in = 0;
    for (int l = 0; l + 1 < S->nl; ++l) {
        const int lo = (int)S->loops[l], hi = (int)S->loops[l + 1];
        for (int i = lo; i < hi; ++i) {
            const int j = i == lo ? hi - 1 : i - 1;
            const real *a = S->verts + 2 * i, *b = S->verts + 2 * j;
            if ((a[1] > uv[1]) != (b[1] > uv[1]) && uv[0] < (b[0] - a[0]) * (uv[1] - a[1]) / (b[1] - a[1]) + a[0])
                in = !in;
        }
    }
    return in;
}

static void FN(eo_uv_edge)(const FN(eo_surf_t) * S, real* uv)
{
    real best = (real)INFINITY, bu = uv[0], bv = uv[1];
    for (int l = 0; l + 1 < S->nl; ++l) {
        const int lo = (int)S->loops[l], hi = (int)S->loops[l + 1];
        for (int i = lo; i < hi; ++i) {
            const int j = i == lo ? hi - 1 : i - 1;
            const real *a = S->verts + 2 * j, *b = S->verts + 2 * i;
            const real ex = b[0] - a[0], ey = b[1] - a[1], e2 = ex * ex + ey * ey;
            real t = e2 > EO_TINY ? ((uv[0] - a[0]) * ex + (uv[1] - a[1]) * ey) / e2 : 0;
            t = FN(eo_clamp)(t, 0, 1);
            const real cx = a[0] + t * ex, cy = a[1] + t * ey;
            const real dd = (cx - uv[0]) * (cx - uv[0]) + (cy - uv[1]) * (cy - uv[1]);
            if (dd < best) best = dd, bu = cx, bv = cy;
        }
    }
    uv[0] = bu, uv[1] = bv;
}

/* ------------------------------------------------- parametrized kinds */

static const real* FN(eo_trim1)(int tag, const real* f)
{
    return tag == 6 ? f + 3 : tag == 7 ? f + 5 : tag == 8 ? f + 6 : tag == 9 ? f + 8 : tag == 10 ? f + 2 : f + 6;
}
static int FN(eo_closed)(int tag, const real* f) { return (tag == 6 || tag == 7 || tag == 10) && FN(eo_trim1)(tag, f)[2] != 0; }

static void FN(eo_eval)(int tag, const real* f, const real* q, real* out)
{
    if (tag == 6) {
        out[0] = f[0] + f[2] * cos(q[0]), out[1] = f[1] + f[2] * sin(q[0]);
    } else if (tag == 13) {
        for (int k = 0; k < 3; ++k) out[k] = f[k] + q[0] * (f[3 + k] - f[k]);
    } else if (tag == 5) {
        for (int k = 0; k < 3; ++k) out[k] = f[k] + q[0] * f[3 + k] + q[1] * f[6 + k];
    } else if (tag == 4) {
        for (int k = 0; k < 3; ++k)
            out[k] = f[k] + f[3] * (cos(q[1]) * cos(q[0]) * f[4 + k] + cos(q[1]) * sin(q[0]) * f[7 + k] +
                                    sin(q[1]) * f[10 + k]);
    } else if (tag == 12) {
        for (int k = 0; k < 3; ++k) out[k] = f[k] + f[12] * (cos(q[0]) * f[3 + k] + sin(q[0]) * f[6 + k]) + q[1] * f[9 + k];
    } else if (tag == 14) {
        const FN(eo_surf_t) S = FN(eo_surf)(f);
        real D[6][3];
        FN(eo_surf_eval)(&S, q, D);
        for (int k = 0; k < 3; ++k) out[k] = D[0][k];
    } else {
        real C[3][2];
        FN(eo_curve)(tag, f, q[0], C);
        out[0] = C[0][0], out[1] = C[0][1];
    }
}

static void FN(eo_frame)(int tag, const real* f, const real* q, real J[2][3])
{
    if (tag == 6) {
        J[0][0] = -f[2] * sin(q[0]), J[0][1] = f[2] * cos(q[0]);
    } else if (tag == 13) {
        for (int k = 0; k < 3; ++k) J[0][k] = f[3 + k] - f[k];
    } else if (tag == 5) {
        for (int k = 0; k < 3; ++k) J[0][k] = f[3 + k], J[1][k] = f[6 + k];
    } else if (tag == 4) {
        const real cu = cos(q[0]), su = sin(q[0]), cv = cos(q[1]), sv = sin(q[1]);
        for (int k = 0; k < 3; ++k) {
            J[0][k] = f[3] * (-cv * su * f[4 + k] + cv * cu * f[7 + k]);
            J[1][k] = f[3] * (-sv * cu * f[4 + k] - sv * su * f[7 + k] + cv * f[10 + k]);
        }
    } else if (tag == 12) {
        for (int k = 0; k < 3; ++k) J[0][k] = f[12] * (-sin(q[0]) * f[3 + k] + cos(q[0]) * f[6 + k]), J[1][k] = f[9 + k];
    } else if (tag == 14) {
        const FN(eo_surf_t) S = FN(eo_surf)(f);
        real D[6][3];
        FN(eo_surf_eval)(&S, q, D);
        for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
    } else {
        real C[3][2];
        FN(eo_curve)(tag, f, q[0], C);
        J[0][0] = C[1][0], J[0][1] = C[1][1];
    }
}

/* the foot parameter (unclamped), seeded for the iterative kinds */
static void FN(eo_invert)(int tag, const real* f, const real* p, const real* seed, int has_seed, int warm, real* q,
                          real J[2][3])
{
    if (tag == 6) {
        q[0] = atan2(p[1] - f[1], p[0] - f[0]);
    } else if (tag == 13) {
        real ab2 = 0, s = 0;
        for (int k = 0; k < 3; ++k) ab2 += (f[3 + k] - f[k]) * (f[3 + k] - f[k]), s += (p[k] - f[k]) * (f[3 + k] - f[k]);
        q[0] = s / (ab2 > EO_TINY ? ab2 : EO_TINY);
    } else if (tag == 5) {
        real d[3];
        for (int k = 0; k < 3; ++k) d[k] = p[k] - f[k];
        q[0] = FN(eo_dot)(d, f + 3, 3), q[1] = FN(eo_dot)(d, f + 6, 3);
    } else if (tag == 4) {
        real m[3];
        for (int k = 0; k < 3; ++k) m[k] = p[k] - f[k];
        FN(eo_normalize)(m, 3);
        q[0] = atan2(FN(eo_dot)(m, f + 7, 3), FN(eo_dot)(m, f + 4, 3));
        q[1] = asin(FN(eo_clamp)(FN(eo_dot)(m, f + 10, 3), -1, 1));
    } else if (tag == 12) {
        real d[3];
        for (int k = 0; k < 3; ++k) d[k] = p[k] - f[k];
        q[0] = atan2(FN(eo_dot)(d, f + 6, 3), FN(eo_dot)(d, f + 3, 3));
        q[1] = FN(eo_dot)(d, f + 9, 3);
    } else if (tag == 14) {
        const FN(eo_surf_t) S = FN(eo_surf)(f);
        FN(eo_surf_foot)(&S, p, seed, has_seed, warm, q, J);
    } else {
        real lo = 0, hi = 1;
        if (tag == 7) hi = EO_TWO_PI;
        if (tag == 10) lo = f[6 + (int)f[0]], hi = f[6 + (int)f[1]];
        q[0] = FN(eo_curve_foot)(tag, f, p, lo, hi, tag == 10 ? seed[0] : 0, tag == 10 && has_seed);
    }
}

static int FN(eo_contains)(int tag, const real* f, const real* q)
{
    if (eo_tdim(tag) == 2) {
        if (tag != 14) return 1;
        const FN(eo_surf_t) S = FN(eo_surf)(f);
        return FN(eo_uv_inside)(&S, q);
    }
    const real* tr = FN(eo_trim1)(tag, f);
    return FN(eo_closed)(tag, f) || (q[0] >= tr[0] && q[0] <= tr[1]);
}

static void FN(eo_clamp_trim)(int tag, const real* f, real* q)
{
    if (eo_tdim(tag) == 2) {
        if (tag == 14) {
            const FN(eo_surf_t) S = FN(eo_surf)(f);
            if (S.nl >= 2) FN(eo_uv_edge)(&S, q);
        }
        return;
    }
    const real* tr = FN(eo_trim1)(tag, f);
    if (FN(eo_closed)(tag, f)) {
        const real L = tr[1] - tr[0];
        if (L > 0) {
            real t = fmod(q[0] - tr[0], L);
            if (t < 0) t += L;
            q[0] = tr[0] + t;
        }
    } else {
        q[0] = FN(eo_clamp)(q[0], tr[0], tr[1]);
    }
}

/* ------------------------------------------------------------ frames */

static void FN(eo_closed_frame)(int tag, const real* e, const real* p, real* pos, real* t)
{
    const int D = tag == 16 ? 3 : 2;
    if (tag == 1 || tag == 15 || tag == 16) {
        real ab2 = 0, s = 0;
        for (int k = 0; k < D; ++k) ab2 += (e[D + k] - e[k]) * (e[D + k] - e[k]), s += (p[k] - e[k]) * (e[D + k] - e[k]);
        real ft = s / (ab2 > EO_TINY ? ab2 : EO_TINY);
        const int released = tag != 1 && (ft < 0 || ft > 1);
        ft = FN(eo_clamp)(ft, 0, 1);
        const real n = sqrt(ab2);
        for (int k = 0; k < D; ++k) {
            if (pos) pos[k] = released ? p[k] : e[k] + ft * (e[D + k] - e[k]);
            if (t) t[k] = n < TOL_ZNORM ? (k == 0 ? 1 : 0) : (e[D + k] - e[k]) / n;
        }
        return;
    }
    if (tag == 3) {
        if (pos) FN(ent_project)(3, e, p, pos);
        if (t) FN(ent_tangent)(3, e, p, t);
        return;
    }
    if (pos) FN(ent_project)(2, e, p, pos);
    if (t) FN(ent_tangent)(2, e, p, t);
}

/* projection of p onto a non-composite kind and its orthonormal basis */
static void FN(eo_simple_frame)(int tag, const real* b, const real* p, real* pos, real B[2][3])
{
    const real* f = b + 1;
    if (!eo_param(tag)) {
        FN(eo_closed_frame)(tag, f, p, pos, B ? B[0] : NULL);
        return;
    }
    real q[2] = {0, 0};
    const real none[2] = {0, 0};
    FN(eo_invert)(tag, f, p, none, 0, 0, q, NULL);
    if ((tag == 6 || tag == 7) && !FN(eo_closed)(tag, f)) {
        const real* tr = FN(eo_trim1)(tag, f);
        const real mid = (real)0.5 * (tr[0] + tr[1]);
        q[0] = mid + remainder(q[0] - mid, EO_TWO_PI);
    }
    if (!FN(eo_contains)(tag, f, q)) FN(eo_clamp_trim)(tag, f, q);
    if (pos) FN(eo_eval)(tag, f, q, pos);
    if (B) {
        FN(eo_frame)(tag, f, q, B);
        FN(eo_gram_schmidt)(B, eo_tdim(tag), eo_dim(tag));
    }
}

static void FN(eo_entity_frame)(int tag, const real* b, const real* p, real* pos, real B[2][3])
{
    if (tag != 11) {
        FN(eo_simple_frame)(tag, b, p, pos, B);
        return;
    }
    const int m = (int)b[1];
    real best = (real)INFINITY;
    if (pos) pos[0] = p[0], pos[1] = p[1];
    if (B) B[0][0] = 1, B[0][1] = 0;
    for (int i = 0; i < m; ++i) {
        const real* sb = b + (long long)b[2 + i];
        real sp[2], sB[2][3];
        FN(eo_simple_frame)((int)sb[0], sb, p, sp, sB);
        const real dd = (sp[0] - p[0]) * (sp[0] - p[0]) + (sp[1] - p[1]) * (sp[1] - p[1]);
        if (dd < best) {
            best = dd;
            if (pos) pos[0] = sp[0], pos[1] = sp[1];
            if (B) B[0][0] = sB[0][0], B[0][1] = sB[0][1];
        }
    }
}

/* seeded projection of the iterative kinds: the unclamped foot goes back
 * into the seed */
static void FN(eo_seeded_frame)(int tag, const real* b, const real* p, real* seed, int warm, real* pos, real B[2][3])
{
    const real* f = b + 1;
    real q[2] = {0, 0};
    FN(eo_invert)(tag, f, p, seed, isfinite(seed[0]), warm, q, NULL);
    seed[0] = q[0];
    if (eo_tdim(tag) == 2) seed[1] = q[1];
    if (!FN(eo_contains)(tag, f, q)) FN(eo_clamp_trim)(tag, f, q);
    if (pos) FN(eo_eval)(tag, f, q, pos);
    if (B) {
        FN(eo_frame)(tag, f, q, B);
        FN(eo_gram_schmidt)(B, eo_tdim(tag), eo_dim(tag));
    }
}

/* Projection of p onto the entity record b and its orthonormal tangent
 * basis (columns B[c][0..dim-1], c < tdim): the reference's cold
 * project_frame. */
void FN(o_entity_frame)(const real* b, const real* p, real* pos, real* basis)
{
    real B[2][3] = {{0, 0, 0}, {0, 0, 0}};
    const int tag = (int)b[0];
    FN(eo_entity_frame)(tag, b, p, pos, B);
    const int D = eo_dim(tag), K = eo_tdim(tag);
    for (int c = 0; c < K; ++c)
        for (int k = 0; k < D; ++k) basis[c * D + k] = B[c][k];
}

/* Raw parametrization access for the tests: mode 0 eval, 1 frame (K
 * columns of D), 2 foot parameter (cold; seed given when seeded kinds have
 * one) */
void FN(o_entity_param)(const real* b, int mode, const real* in, const real* seed, real* out)
{
    const int tag = (int)b[0];
    const real* f = b + 1;
    if (mode == 0) {
        FN(eo_eval)(tag, f, in, out);
    } else if (mode == 1) {
        real J[2][3];
        FN(eo_frame)(tag, f, in, J);
        for (int c = 0; c < eo_tdim(tag); ++c)
            for (int k = 0; k < eo_dim(tag); ++k) out[c * eo_dim(tag) + k] = J[c][k];
    } else {
        const real none[2] = {(real)NAN, (real)NAN};
        const real* s = seed ? seed : none;
        FN(eo_invert)(tag, f, in, s, isfinite(s[0]), 0, out, NULL);
    }
}
