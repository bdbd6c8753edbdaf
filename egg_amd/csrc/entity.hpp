// Boundary entities a constrained DOF slides on (reference src/geometry.hpp,
// tags src/entity_soa.hpp:37-55), shared by the sweep kernels (R = float,
// the FP32 tolerances of src/core.hpp: tiny 1e-20, znorm 1e-7, projection
// Newton break 1e-5) and the host geometry calls (R = double: 1e-30, 1e-15,
// 1e-9).
//
// Every entity is one self-describing record ("blob") of R values: blob[0]
// is the tag, then the tag's fields, then its variable-length payloads:
//    1 segment, 15 rail   sx sy ex ey
//    2 circle             cx cy r
//    3 ellipse            cx cy rx ry              (radial-scaling foot)
//    6 circle arc         cx cy r t0 t1 closed
//    7 ellipse arc        cx cy a b phi t0 t1 closed
//    8 quadratic Bezier   P0 P1 P2 (x y each) t0 t1
//    9 cubic Bezier       P0 P1 P2 P3 t0 t1
//   10 B-spline curve     p n t0 t1 closed has_w | knots[n+p+1] ctrl[2n] (w[n])
//   11 composite path     m off_0 .. off_m-1 | m segment blobs (off_i from
//                         the composite's own blob[0])
//    4 sphere             c(3) r ax(3) ay(3) az(3)
//    5 plane              o(3) ax(3) ay(3)
//   12 cylinder           o(3) ax(3) ay(3) az(3) r
//   13 3D line            p0(3) p1(3) t0 t1
//   16 3D rail            p0(3) p1(3)
//   14 B-spline surface   pu pv nu nv has_w nl | knots_u[nu+pu+1]
//                         knots_v[nv+pv+1] ctrl[3 nu nv] (w[nu nv])
//                         loops[nl] verts[2 loops[nl-1]] (UV trim polygon:
//                         vertex-count offsets 0, n0, n0+n1, ...; nl < 2 =
//                         untrimmed)
// (the host packs them from the reference wire's per-type records and CSR
// payloads, src/geometry.hpp:1717-2520, and checks every size against the
// record before anything reaches the device).
//
// Kinds with a parametrization (arcs, Beziers, B-splines, 3D line, plane,
// sphere, cylinder, B-spline surface; src/geometry.hpp:195-258) project by
// inverting it to the foot parameter, re-representing a periodic angle
// nearest an open arc's interval, clamping into the trim and evaluating;
// their tangent basis is the Gram-Schmidt orthonormalised raw frame there.
// The other kinds have closed-form feet and tangents (:288-464), the
// composite keeps its nearest segment (:2580-2612).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace eggmi
{

enum EntTag : int {
    kEntFree = 0,
    kEntSegment = 1,
    kEntCircle = 2,
    kEntEllipse = 3,
    kEntSphere = 4,
    kEntPlane = 5,
    kEntCircleArc = 6,
    kEntEllipseArc = 7,
    kEntQuadBezier = 8,
    kEntCubicBezier = 9,
    kEntBSpline = 10,
    kEntComposite = 11,
    kEntCylinder = 12,
    kEntLine3 = 13,
    kEntBSplineSurf = 14,
    kEntRail = 15,
    kEntRail3 = 16,
};
constexpr int kEntMaxTag = 16;
// de Boor work sizes (src/geometry.hpp:686-705): curves up to degree 7,
// surfaces up to bicubic
constexpr int kEntCurveCap = 8, kEntSurfCap = 4;

template <class R> struct EntTol;
template <> struct EntTol<float> {
    static constexpr float tiny = 1e-20f, znorm = 1e-7f, newton = 1e-5f, inv_sqrt2 = 0.70710678f;
    static constexpr float two_pi = 6.28318530717958647692f, inf = __builtin_huge_valf();
};
template <> struct EntTol<double> {
    static constexpr double tiny = 1e-30, znorm = 1e-15, newton = 1e-9, inv_sqrt2 = 0.70710678118654752440;
    static constexpr double two_pi = 6.28318530717958647692, inf = __builtin_huge_val();
};

// embedding dimension, tangent dimension, and the two projection families
__host__ __device__ constexpr int ent_dim(int tag)
{
    return (tag == kEntSphere || tag == kEntPlane || tag == kEntCylinder || tag == kEntLine3 ||
            tag == kEntBSplineSurf || tag == kEntRail3)
             ? 3
             : 2;
}
__host__ __device__ constexpr int ent_tdim(int tag)
{
    return (tag == kEntSphere || tag == kEntPlane || tag == kEntCylinder || tag == kEntBSplineSurf) ? 2 : 1;
}
__host__ __device__ constexpr bool ent_is_param(int tag)
{
    return tag == kEntSphere || tag == kEntPlane || (tag >= kEntCircleArc && tag <= kEntBSpline) ||
           tag == kEntCylinder || tag == kEntLine3 || tag == kEntBSplineSurf;
}
// parametrizations whose inverse is an iterative foot that can start from
// the previous sweep's parameter (the per-DOF seed)
__host__ __device__ constexpr bool ent_is_seeded(int tag) { return tag == kEntBSpline || tag == kEntBSplineSurf; }
__host__ __device__ constexpr bool ent_is_rail(int tag) { return tag == kEntRail || tag == kEntRail3; }

template <class R> __host__ __device__ inline R ent_clamp(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class R> __host__ __device__ inline R ent_dot(const R* a, const R* b, int n)
{
    R s = R(0);
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}
// a / |a|, or e_0 below znorm
template <class R> __host__ __device__ inline void ent_normalize(R* a, int n)
{
    const R nn = std::sqrt(ent_dot(a, a, n));
    if (nn < EntTol<R>::znorm) {
        for (int i = 0; i < n; ++i) a[i] = i == 0 ? R(1) : R(0);
        return;
    }
    const R inv = R(1) / nn;
    for (int i = 0; i < n; ++i) a[i] = inv * a[i];
}
// Gram-Schmidt over the K columns B[0..K-1] in D dimensions
template <class R> __host__ __device__ inline void ent_orthonormalize(R (*B)[3], int K, int D)
{
    for (int i = 0; i < K; ++i) {
        for (int j = 0; j < i; ++j) {
            const R c = ent_dot(B[i], B[j], D);
            for (int k = 0; k < D; ++k) B[i][k] = B[i][k] - c * B[j][k];
        }
        ent_normalize(B[i], D);
    }
}

// ------------------------------------------------------------- B-splines

// knot span s with U[s] <= u < U[s+1], clamped into [p, n - 1] (n control
// points; the live domain is [U[p], U[n]])
template <class R> __host__ __device__ inline int bsp_span(const R* U, int n, int p, R u)
{
    if (u >= U[n]) return n - 1;
    if (u <= U[p]) return p;
    int lo = p, hi = n;  // U[lo] <= u < U[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (u < U[mid])
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

// One step of the derivative rule d/du N_{i,k} = k (F_{i,k-1} / (U[i+k] -
// U[i]) - F_{i+1,k-1} / (U[i+k+1] - U[i+1])) over span s: src holds the
// k-1 row (F_{s-k+1+j}, j = 0..k-1), dst receives the k row (j = 0..k);
// a zero knot interval contributes nothing.
template <class R, int CAP> __host__ __device__ inline void bsp_drow(const R* U, int s, int k, const R* src, R* dst)
{
    for (int j = 0; j <= k; ++j) {
        const int i = s - k + j;
        R v = R(0);
        if (j >= 1) {
            const R den = U[i + k] - U[i];
            if (den != R(0)) v += src[j - 1] / den;
        }
        if (j <= k - 1) {
            const R den = U[i + k + 1] - U[i + 1];
            if (den != R(0)) v -= src[j] / den;
        }
        dst[j] = R(k) * v;
    }
}

// The p + 1 nonzero basis functions of degree p at u in span s and their
// first nd <= 2 derivatives: out[k][j] = d^k N_{s-p+j,p}(u). Cox-de Boor
// triangle by degree; the derivatives apply the derivative rule to the
// lower-degree rows.
template <class R, int CAP> __host__ __device__ inline void bsp_basis(const R* U, int p, int s, R u, int nd, R (*out)[CAP])
{
    R N[CAP][CAP];  // N[k][j] = N_{s-k+j,k}(u)
    N[0][0] = R(1);
    for (int k = 1; k <= p; ++k)
        for (int j = 0; j <= k; ++j) {
            const int i = s - k + j;
            R v = R(0);
            if (j >= 1) {
                const R den = U[i + k] - U[i];
                if (den != R(0)) v += (u - U[i]) / den * N[k - 1][j - 1];
            }
            if (j <= k - 1) {
                const R den = U[i + k + 1] - U[i + 1];
                if (den != R(0)) v += (U[i + k + 1] - u) / den * N[k - 1][j];
            }
            N[k][j] = v;
        }
    for (int j = 0; j <= p; ++j) out[0][j] = N[p][j];
    if (nd >= 1) {
        if (p >= 1)
            bsp_drow<R, CAP>(U, s, p, N[p - 1], out[1]);
        else
            out[1][0] = R(0);
    }
    if (nd >= 2) {
        if (p >= 2) {
            R d1[CAP];  // first derivatives of the degree p-1 row
            bsp_drow<R, CAP>(U, s, p - 1, N[p - 2], d1);
            bsp_drow<R, CAP>(U, s, p, d1, out[2]);
        } else {
            for (int j = 0; j <= p; ++j) out[2][j] = R(0);
        }
    }
}

// ---------------------------------------------------------------- curves

// Curve point and its first nd <= 2 derivatives, C[k][0..1], for the
// iterative 2D curves (ellipse arc, Beziers, B-spline; f = blob + 1).
template <class R> __host__ __device__ inline void curve_ders(int tag, const R* f, R t, int nd, R (*C)[2])
{
    if (tag == kEntEllipseArc) {
        const R cp = std::cos(f[4]), sp = std::sin(f[4]);
        const R ct = std::cos(t), st = std::sin(t);
        const R a = f[2], b = f[3];
        // R_phi (a cos t, b sin t) and its t-derivatives
        const R lx[3] = {a * ct, -a * st, -a * ct}, ly[3] = {b * st, b * ct, -b * st};
        for (int k = 0; k <= nd; ++k) {
            C[k][0] = cp * lx[k] - sp * ly[k] + (k == 0 ? f[0] : R(0));
            C[k][1] = sp * lx[k] + cp * ly[k] + (k == 0 ? f[1] : R(0));
        }
        return;
    }
    if (tag == kEntQuadBezier) {
        const R u = R(1) - t;
        for (int i = 0; i < 2; ++i) {
            const R p0 = f[i], p1 = f[2 + i], p2 = f[4 + i];
            C[0][i] = u * u * p0 + R(2) * u * t * p1 + t * t * p2;
            if (nd >= 1) C[1][i] = R(2) * u * (p1 - p0) + R(2) * t * (p2 - p1);
            if (nd >= 2) C[2][i] = R(2) * (p2 - R(2) * p1 + p0);
        }
        return;
    }
    if (tag == kEntCubicBezier) {
        const R u = R(1) - t;
        for (int i = 0; i < 2; ++i) {
            const R p0 = f[i], p1 = f[2 + i], p2 = f[4 + i], p3 = f[6 + i];
            C[0][i] = u * u * u * p0 + R(3) * u * u * t * p1 + R(3) * u * t * t * p2 + t * t * t * p3;
            if (nd >= 1) C[1][i] = R(3) * u * u * (p1 - p0) + R(6) * u * t * (p2 - p1) + R(3) * t * t * (p3 - p2);
            if (nd >= 2) C[2][i] = R(6) * u * (p2 - R(2) * p1 + p0) + R(6) * t * (p3 - R(2) * p2 + p1);
        }
        return;
    }
    // B-spline / NURBS: homogeneous sums A^(k) = sum N^(k) w P, w^(k) =
    // sum N^(k) w, then the quotient rule for C = A / w
    const int p = (int)f[0], n = (int)f[1];
    const bool rat = f[5] != R(0);
    const R* U = f + 6;
    const R* P = U + n + p + 1;
    const R* W = P + 2 * n;
    const int s = bsp_span(U, n, p, t);
    R Nd[3][kEntCurveCap];
    bsp_basis<R, kEntCurveCap>(U, p, s, t, nd, Nd);
    R A[3][2] = {}, w[3] = {};
    for (int j = 0; j <= p; ++j) {
        const int i = s - p + j;
        const R wi = rat ? W[i] : R(1);
        for (int k = 0; k <= nd; ++k) {
            const R nw = Nd[k][j] * wi;
            A[k][0] += nw * P[2 * i];
            A[k][1] += nw * P[2 * i + 1];
            w[k] += nw;
        }
    }
    if (!rat) {
        for (int k = 0; k <= nd; ++k) C[k][0] = A[k][0], C[k][1] = A[k][1];
        return;
    }
    const R iw = R(1) / w[0];
    for (int i = 0; i < 2; ++i) {
        C[0][i] = iw * A[0][i];
        if (nd >= 1) C[1][i] = iw * (A[1][i] - w[1] * C[0][i]);
        if (nd >= 2) C[2][i] = iw * (A[2][i] - R(2) * w[1] * C[1][i] - w[2] * C[0][i]);
    }
}

// Nearest-foot parameter of q on an iterative curve over [lo, hi]
// (src/geometry.hpp:517-553): from the seed, or from the nearest of 17
// evenly spaced samples, 8 Newton steps on (C - q).C' = 0; unclamped.
template <class R>
__host__ __device__ inline R curve_foot(int tag, const R* f, const R* q, R lo, R hi, R seed, bool has_seed)
{
    R t;
    if (has_seed) {
        t = ent_clamp(seed, lo, hi);
    } else {
        t = lo;
        R best = EntTol<R>::inf;
        for (int i = 0; i <= 16; ++i) {
            const R tc = lo + ((hi - lo) * R(i)) / R(16);
            R C[1][2];
            curve_ders(tag, f, tc, 0, C);
            const R dx = C[0][0] - q[0], dy = C[0][1] - q[1];
            const R dd = dx * dx + dy * dy;
            if (dd < best) {
                best = dd;
                t = tc;
            }
        }
    }
    for (int it = 0; it < 8; ++it) {
        R C[3][2];
        curve_ders(tag, f, t, 2, C);
        const R dx = C[0][0] - q[0], dy = C[0][1] - q[1];
        const R fv = dx * C[1][0] + dy * C[1][1];
        const R fp = C[1][0] * C[1][0] + C[1][1] * C[1][1] + dx * C[2][0] + dy * C[2][1];
        if (std::fabs(fp) < EntTol<R>::tiny) break;
        t -= fv / fp;
    }
    return t;
}

// ---------------------------------------------------------- B-spline surface

template <class R> struct SurfRef {
    int pu, pv, nu, nv, nl;
    bool rat;
    const R *Uu, *Uv, *P, *W, *loops, *verts;
};
template <class R> __host__ __device__ inline SurfRef<R> surf_ref(const R* f)
{
    SurfRef<R> S;
    S.pu = (int)f[0];
    S.pv = (int)f[1];
    S.nu = (int)f[2];
    S.nv = (int)f[3];
    S.rat = f[4] != R(0);
    S.nl = (int)f[5];
    S.Uu = f + 6;
    S.Uv = S.Uu + S.nu + S.pu + 1;
    S.P = S.Uv + S.nv + S.pv + 1;
    S.W = S.P + 3 * S.nu * S.nv;
    S.loops = S.W + (S.rat ? S.nu * S.nv : 0);
    S.verts = S.loops + S.nl;
    return S;
}
// surface point and partials: D[0] S, D[1] S_u, D[2] S_v, D[3] S_uv,
// D[4] S_uu, D[5] S_vv (nd = 0, 1 or 2 fills 1, 3 or 6 of them)
template <class R> __host__ __device__ inline void surf_ders(const SurfRef<R>& S, const R* q, int nd, R (*D)[3])
{
    const int su = bsp_span(S.Uu, S.nu, S.pu, q[0]), sv = bsp_span(S.Uv, S.nv, S.pv, q[1]);
    R bu[3][kEntSurfCap], bv[3][kEntSurfCap];
    bsp_basis<R, kEntSurfCap>(S.Uu, S.pu, su, q[0], nd, bu);
    bsp_basis<R, kEntSurfCap>(S.Uv, S.pv, sv, q[1], nd, bv);
    const int nt = nd == 0 ? 1 : (nd == 1 ? 3 : 6);
    // derivative orders (u, v) of the six terms
    const int ou[6] = {0, 1, 0, 1, 2, 0}, ov[6] = {0, 0, 1, 1, 0, 2};
    R A[6][3] = {}, w[6] = {};
    for (int a = 0; a <= S.pu; ++a) {
        const int iu = su - S.pu + a;
        for (int b = 0; b <= S.pv; ++b) {
            const int iv = sv - S.pv + b;
            const int c = iu * S.nv + iv;
            const R wc = S.rat ? S.W[c] : R(1);
            for (int m = 0; m < nt; ++m) {
                const R nw = bu[ou[m]][a] * bv[ov[m]][b] * wc;
                for (int k = 0; k < 3; ++k) A[m][k] += nw * S.P[3 * c + k];
                w[m] += nw;
            }
        }
    }
    if (!S.rat) {
        for (int m = 0; m < nt; ++m)
            for (int k = 0; k < 3; ++k) D[m][k] = A[m][k];
        return;
    }
    // bivariate quotient rule for S = A / w
    const R iw = R(1) / w[0];
    for (int k = 0; k < 3; ++k) {
        D[0][k] = iw * A[0][k];
        if (nd >= 1) {
            D[1][k] = iw * (A[1][k] - w[1] * D[0][k]);
            D[2][k] = iw * (A[2][k] - w[2] * D[0][k]);
        }
        if (nd >= 2) {
            D[3][k] = iw * (A[3][k] - w[1] * D[2][k] - w[2] * D[1][k] - w[3] * D[0][k]);
            D[4][k] = iw * (A[4][k] - R(2) * w[1] * D[1][k] - w[4] * D[0][k]);
            D[5][k] = iw * (A[5][k] - R(2) * w[2] * D[2][k] - w[5] * D[0][k]);
        }
    }
}

// Newton on the foot conditions (S - p).S_u = (S - p).S_v = 0 from q, each
// iterate clamped into the live domain, stopping once it moves less than the
// projection tolerance or the Jacobian is singular (src/geometry.hpp:
// 1214-1244). EXACT: the Jacobian with the curvature terms (d.S_uu ...),
// else Gauss-Newton. J, when given, receives S_u, S_v at the last iterate.
template <class R, bool EXACT>
__host__ __device__ inline void surf_newton(const SurfRef<R>& S, const R* p, R* q, const R* dom, R (*J)[3])
{
    for (int it = 0; it < 12; ++it) {
        R D[6][3];
        surf_ders(S, q, EXACT ? 2 : 1, D);
        if (J)
            for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
        R d[3];
        for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
        const R f1 = ent_dot(d, D[1], 3), f2 = ent_dot(d, D[2], 3);
        R j11 = ent_dot(D[1], D[1], 3), j12 = ent_dot(D[1], D[2], 3), j22 = ent_dot(D[2], D[2], 3);
        if (EXACT) {
            j11 += ent_dot(d, D[4], 3);
            j12 += ent_dot(d, D[3], 3);
            j22 += ent_dot(d, D[5], 3);
        }
        const R det = j11 * j22 - j12 * j12;
        if (std::fabs(det) < EntTol<R>::tiny) break;
        const R qu = ent_clamp(q[0] - (j22 * f1 - j12 * f2) / det, dom[0], dom[1]);
        const R qv = ent_clamp(q[1] - (-j12 * f1 + j11 * f2) / det, dom[2], dom[3]);
        const bool done = std::fabs(qu - q[0]) + std::fabs(qv - q[1]) < EntTol<R>::newton;
        q[0] = qu;
        q[1] = qv;
        if (done) break;
    }
}

// Damped (Levenberg-Marquardt) exact-Newton foot from one start, every
// accepted step lowering |S - p|^2 (src/geometry.hpp:1254-1305).
template <class R>
__host__ __device__ inline void surf_lm(const SurfRef<R>& S, const R* p, R* q, const R* dom, R (*J)[3])
{
    q[0] = ent_clamp(q[0], dom[0], dom[1]);
    q[1] = ent_clamp(q[1], dom[2], dom[3]);
    R D[6][3];
    surf_ders(S, q, 2, D);
    R d[3];
    for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
    R r = ent_dot(d, d, 3);
    R lam = R(1e-3) * (ent_dot(D[1], D[1], 3) + ent_dot(D[2], D[2], 3) + EntTol<R>::tiny);
    for (int it = 0; it < 16; ++it) {
        const R f1 = ent_dot(d, D[1], 3), f2 = ent_dot(d, D[2], 3);
        const R j11 = ent_dot(D[1], D[1], 3) + ent_dot(d, D[4], 3);
        const R j12 = ent_dot(D[1], D[2], 3) + ent_dot(d, D[3], 3);
        const R j22 = ent_dot(D[2], D[2], 3) + ent_dot(d, D[5], 3);
        R du = R(0), dv = R(0);
        bool ok = false;
        for (int bt = 0; bt < 6; ++bt) {
            const R a11 = j11 + lam, a22 = j22 + lam;
            const R det = a11 * a22 - j12 * j12;
            if (std::fabs(det) < EntTol<R>::tiny) {
                lam *= R(4);
                continue;
            }
            du = -((a22 * f1 - j12 * f2) / det);
            dv = -((-j12 * f1 + a11 * f2) / det);
            const R qn[2] = {ent_clamp(q[0] + du, dom[0], dom[1]), ent_clamp(q[1] + dv, dom[2], dom[3])};
            R Dn[1][3];
            surf_ders(S, qn, 0, Dn);
            R dn[3];
            for (int k = 0; k < 3; ++k) dn[k] = Dn[0][k] - p[k];
            const R rn = ent_dot(dn, dn, 3);
            if (rn < r) {
                q[0] = qn[0];
                q[1] = qn[1];
                r = rn;
                surf_ders(S, q, 2, D);
                for (int k = 0; k < 3; ++k) d[k] = D[0][k] - p[k];
                lam = lam * R(0.3) > EntTol<R>::tiny ? lam * R(0.3) : EntTol<R>::tiny;
                ok = true;
                break;
            }
            lam *= R(4);
        }
        if (!ok || std::fabs(du) + std::fabs(dv) < EntTol<R>::newton) break;
    }
    if (J)
        for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
}

// The surface's foot parameter of p (src/geometry.hpp:1322-1395): from a
// seed (warm, or a live seed in the cold pass) Gauss-Newton; else the four
// nearest of a 9 x 9 sample grid each polished by the damped Newton and the
// nearest result kept.
template <class R>
__host__ __device__ inline void surf_foot(const SurfRef<R>& S, const R* p, const R* seed, bool has_seed, bool warm,
                                          R* q, R (*J)[3])
{
    const R dom[4] = {S.Uu[S.pu], S.Uu[S.nu], S.Uv[S.pv], S.Uv[S.nv]};
    if (warm || has_seed) {
        q[0] = ent_clamp(seed[0], dom[0], dom[1]);
        q[1] = ent_clamp(seed[1], dom[2], dom[3]);
        surf_newton<R, false>(S, p, q, dom, J);
        return;
    }
    constexpr int kStarts = 4;
    R bd[kStarts], bq[kStarts][2];
    for (int s = 0; s < kStarts; ++s) {
        bd[s] = EntTol<R>::inf;
        bq[s][0] = dom[0];
        bq[s][1] = dom[2];
    }
    for (int i = 0; i <= 8; ++i)
        for (int j = 0; j <= 8; ++j) {
            const R t[2] = {dom[0] + ((dom[1] - dom[0]) * R(i)) / R(8), dom[2] + ((dom[3] - dom[2]) * R(j)) / R(8)};
            R D[1][3];
            surf_ders(S, t, 0, D);
            R dd = R(0);
            for (int k = 0; k < 3; ++k) dd += (D[0][k] - p[k]) * (D[0][k] - p[k]);
            if (dd < bd[kStarts - 1]) {
                int s = kStarts - 1;
                for (; s > 0 && dd < bd[s - 1]; --s) {
                    bd[s] = bd[s - 1];
                    bq[s][0] = bq[s - 1][0];
                    bq[s][1] = bq[s - 1][1];
                }
                bd[s] = dd;
                bq[s][0] = t[0];
                bq[s][1] = t[1];
            }
        }
    q[0] = dom[0];
    q[1] = dom[2];
    R best = EntTol<R>::inf;
    for (int s = 0; s < kStarts; ++s) {
        if (!(bd[s] < EntTol<R>::inf)) continue;
        R qs[2] = {bq[s][0], bq[s][1]}, Js[2][3];
        surf_lm(S, p, qs, dom, J ? Js : nullptr);
        R D[1][3];
        surf_ders(S, qs, 0, D);
        R dd = R(0);
        for (int k = 0; k < 3; ++k) dd += (D[0][k] - p[k]) * (D[0][k] - p[k]);
        if (dd < best) {
            best = dd;
            q[0] = qs[0];
            q[1] = qs[1];
            if (J)
                for (int k = 0; k < 3; ++k) J[0][k] = Js[0][k], J[1][k] = Js[1][k];
        }
    }
}

// even-odd inside test and nearest point on the loop polylines of a UV trim
// polygon (src/geometry.hpp:144-192)
template <class R> __host__ __device__ inline bool uv_inside(const SurfRef<R>& S, const R* uv)
{
    if (S.nl < 2) return true;
    bool in = false;
    for (int l = 0; l + 1 < S.nl; ++l) {
        const int lo = (int)S.loops[l], hi = (int)S.loops[l + 1];
        for (int i = lo, j = hi - 1; i < hi; j = i++) {
            const R* a = S.verts + 2 * i;
            const R* b = S.verts + 2 * j;
            if ((a[1] > uv[1]) != (b[1] > uv[1]) && uv[0] < (b[0] - a[0]) * (uv[1] - a[1]) / (b[1] - a[1]) + a[0])
                in = !in;
        }
    }
    return in;
}
template <class R> __host__ __device__ inline void uv_nearest_edge(const SurfRef<R>& S, R* uv)
{
    R best = EntTol<R>::inf, bu = uv[0], bv = uv[1];
    for (int l = 0; l + 1 < S.nl; ++l) {
        const int lo = (int)S.loops[l], hi = (int)S.loops[l + 1];
        for (int i = lo, j = hi - 1; i < hi; j = i++) {
            const R* a = S.verts + 2 * j;
            const R* b = S.verts + 2 * i;
            const R ex = b[0] - a[0], ey = b[1] - a[1];
            const R e2 = ex * ex + ey * ey;
            R t = e2 > EntTol<R>::tiny ? ((uv[0] - a[0]) * ex + (uv[1] - a[1]) * ey) / e2 : R(0);
            t = ent_clamp(t, R(0), R(1));
            const R cx = a[0] + t * ex, cy = a[1] + t * ey;
            const R dd = (cx - uv[0]) * (cx - uv[0]) + (cy - uv[1]) * (cy - uv[1]);
            if (dd < best) {
                best = dd;
                bu = cx;
                bv = cy;
            }
        }
    }
    uv[0] = bu;
    uv[1] = bv;
}

// ------------------------------------------------------ parametrizations

// the curve trim (t0, t1, closed) of a 1-parameter kind, in f
template <class R> __host__ __device__ inline const R* curve_trim(int tag, const R* f)
{
    switch (tag) {
    case kEntCircleArc: return f + 3;
    case kEntEllipseArc: return f + 5;
    case kEntQuadBezier: return f + 6;
    case kEntCubicBezier: return f + 8;
    case kEntBSpline: return f + 2;
    default: return f + 6;  // 3D line
    }
}
template <class R> __host__ __device__ inline bool curve_trim_closed(int tag, const R* f)
{
    return (tag == kEntCircleArc || tag == kEntEllipseArc || tag == kEntBSpline) && curve_trim(tag, f)[2] != R(0);
}

// point of a parametrization at q
template <class R> __host__ __device__ inline void par_eval(int tag, const R* f, const R* q, R* out)
{
    switch (tag) {
    case kEntCircleArc:
        out[0] = f[0] + f[2] * std::cos(q[0]);
        out[1] = f[1] + f[2] * std::sin(q[0]);
        return;
    case kEntLine3:
        for (int k = 0; k < 3; ++k) out[k] = f[k] + q[0] * (f[3 + k] - f[k]);
        return;
    case kEntPlane:
        for (int k = 0; k < 3; ++k) out[k] = f[k] + q[0] * f[3 + k] + q[1] * f[6 + k];
        return;
    case kEntSphere: {
        const R cu = std::cos(q[0]), su = std::sin(q[0]), cv = std::cos(q[1]), sv = std::sin(q[1]);
        for (int k = 0; k < 3; ++k) out[k] = f[k] + f[3] * (cv * cu * f[4 + k] + cv * su * f[7 + k] + sv * f[10 + k]);
        return;
    }
    case kEntCylinder: {
        const R cu = std::cos(q[0]), su = std::sin(q[0]);
        for (int k = 0; k < 3; ++k) out[k] = f[k] + f[12] * (cu * f[3 + k] + su * f[6 + k]) + q[1] * f[9 + k];
        return;
    }
    case kEntBSplineSurf: {
        R D[1][3];
        surf_ders(surf_ref(f), q, 0, D);
        for (int k = 0; k < 3; ++k) out[k] = D[0][k];
        return;
    }
    default: {
        R C[1][2];
        curve_ders(tag, f, q[0], 0, C);
        out[0] = C[0][0];
        out[1] = C[0][1];
    }
    }
}

// the raw tangent columns J[0..K-1] of a parametrization at q
template <class R> __host__ __device__ inline void par_frame(int tag, const R* f, const R* q, R (*J)[3])
{
    switch (tag) {
    case kEntCircleArc:
        J[0][0] = -f[2] * std::sin(q[0]);
        J[0][1] = f[2] * std::cos(q[0]);
        return;
    case kEntLine3:
        for (int k = 0; k < 3; ++k) J[0][k] = f[3 + k] - f[k];
        return;
    case kEntPlane:
        for (int k = 0; k < 3; ++k) J[0][k] = f[3 + k], J[1][k] = f[6 + k];
        return;
    case kEntSphere: {
        const R cu = std::cos(q[0]), su = std::sin(q[0]), cv = std::cos(q[1]), sv = std::sin(q[1]);
        for (int k = 0; k < 3; ++k) {
            J[0][k] = f[3] * (cv * -su * f[4 + k] + cv * cu * f[7 + k]);
            J[1][k] = f[3] * (-sv * cu * f[4 + k] - sv * su * f[7 + k] + cv * f[10 + k]);
        }
        return;
    }
    case kEntCylinder: {
        const R cu = std::cos(q[0]), su = std::sin(q[0]);
        for (int k = 0; k < 3; ++k) {
            J[0][k] = f[12] * (-su * f[3 + k] + cu * f[6 + k]);
            J[1][k] = f[9 + k];
        }
        return;
    }
    case kEntBSplineSurf: {
        R D[3][3];
        surf_ders(surf_ref(f), q, 1, D);
        for (int k = 0; k < 3; ++k) J[0][k] = D[1][k], J[1][k] = D[2][k];
        return;
    }
    default: {
        R C[2][2];
        curve_ders(tag, f, q[0], 1, C);
        J[0][0] = C[1][0];
        J[0][1] = C[1][1];
    }
    }
}

// live parameter domain of an iterative curve
template <class R> __host__ __device__ inline void curve_domain(int tag, const R* f, R& lo, R& hi)
{
    if (tag == kEntBSpline) {
        const int p = (int)f[0], n = (int)f[1];
        lo = f[6 + p];
        hi = f[6 + n];
    } else if (tag == kEntEllipseArc) {
        lo = R(0);
        hi = EntTol<R>::two_pi;
    } else {
        lo = R(0);
        hi = R(1);
    }
}

// foot parameter (unclamped) of p; for the iterative kinds from the seed
// when has_seed (warm: the surface skips its cold grid search), J when given
// receives the frame the surface foot ends on (else the caller evaluates it)
template <class R>
__host__ __device__ inline void par_invert(int tag, const R* f, const R* p, const R* seed, bool has_seed, bool warm,
                                           R* q, R (*J)[3] = nullptr)
{
    switch (tag) {
    case kEntCircleArc: q[0] = std::atan2(p[1] - f[1], p[0] - f[0]); return;
    case kEntLine3: {
        R ab[3], d[3];
        for (int k = 0; k < 3; ++k) ab[k] = f[3 + k] - f[k], d[k] = p[k] - f[k];
        const R ab2 = ent_dot(ab, ab, 3);
        q[0] = ent_dot(d, ab, 3) / (ab2 > EntTol<R>::tiny ? ab2 : EntTol<R>::tiny);
        return;
    }
    case kEntPlane: {
        R d[3];
        for (int k = 0; k < 3; ++k) d[k] = p[k] - f[k];
        q[0] = ent_dot(d, f + 3, 3);
        q[1] = ent_dot(d, f + 6, 3);
        return;
    }
    case kEntSphere: {
        R m[3];
        for (int k = 0; k < 3; ++k) m[k] = p[k] - f[k];
        ent_normalize(m, 3);
        q[0] = std::atan2(ent_dot(m, f + 7, 3), ent_dot(m, f + 4, 3));
        q[1] = std::asin(ent_clamp(ent_dot(m, f + 10, 3), R(-1), R(1)));
        return;
    }
    case kEntCylinder: {
        R d[3];
        for (int k = 0; k < 3; ++k) d[k] = p[k] - f[k];
        q[0] = std::atan2(ent_dot(d, f + 6, 3), ent_dot(d, f + 3, 3));
        q[1] = ent_dot(d, f + 9, 3);
        return;
    }
    case kEntBSplineSurf: surf_foot(surf_ref(f), p, seed, has_seed, warm, q, J); return;
    default: {
        R lo, hi;
        curve_domain(tag, f, lo, hi);
        q[0] = curve_foot(tag, f, p, lo, hi, tag == kEntBSpline ? seed[0] : R(0), tag == kEntBSpline && has_seed);
    }
    }
}

// trim of a parametrization: inside test and clamp (a curve interval,
// wrapped when closed; a UV polygon for the B-spline surface; the analytic
// surfaces are untrimmed)
template <class R> __host__ __device__ inline bool par_contains(int tag, const R* f, const R* q)
{
    if (ent_tdim(tag) == 2) return tag != kEntBSplineSurf || uv_inside(surf_ref(f), q);
    const R* tr = curve_trim(tag, f);
    return curve_trim_closed(tag, f) || (q[0] >= tr[0] && q[0] <= tr[1]);
}
template <class R> __host__ __device__ inline void par_clamp(int tag, const R* f, R* q)
{
    if (ent_tdim(tag) == 2) {
        if (tag == kEntBSplineSurf) {
            const SurfRef<R> S = surf_ref(f);
            if (S.nl >= 2) uv_nearest_edge(S, q);
        }
        return;
    }
    const R* tr = curve_trim(tag, f);
    if (curve_trim_closed(tag, f)) {
        const R L = tr[1] - tr[0];
        if (L > R(0)) {
            R t = std::fmod(q[0] - tr[0], L);
            if (t < R(0)) t += L;
            q[0] = tr[0] + t;
        }
    } else {
        q[0] = ent_clamp(q[0], tr[0], tr[1]);
    }
}

// ------------------------------------------------------ closed-form kinds

// foot parameter of p on the line through a segment (0 at its start, 1 at
// its end), D = 2 or 3
template <class R> __host__ __device__ inline R rail_foot(const R* e, const R* p, int D)
{
    R ab2 = R(0), s = R(0);
    for (int k = 0; k < D; ++k) {
        const R ab = e[D + k] - e[k];
        ab2 += ab * ab;
        s += (p[k] - e[k]) * ab;
    }
    return s / (ab2 > EntTol<R>::tiny ? ab2 : EntTol<R>::tiny);
}

template <class R> __host__ __device__ inline void closed_project(int tag, const R* e, const R* p, R* out)
{
    using T = EntTol<R>;
    if (tag == kEntSegment || tag == kEntRail || tag == kEntRail3) {
        const int D = tag == kEntRail3 ? 3 : 2;
        R t = rail_foot(e, p, D);
        const bool released = tag != kEntSegment && (t < R(0) || t > R(1));
        t = ent_clamp(t, R(0), R(1));
        for (int k = 0; k < D; ++k) out[k] = released ? p[k] : e[k] + t * (e[D + k] - e[k]);
    } else if (tag == kEntEllipse) {
        const R sx = (p[0] - e[0]) / e[2], sy = (p[1] - e[1]) / e[3];
        const R dist = std::sqrt(sx * sx + sy * sy);
        out[0] = e[0] + (dist < T::znorm ? T::inv_sqrt2 : sx / dist) * e[2];
        out[1] = e[1] + (dist < T::znorm ? T::inv_sqrt2 : sy / dist) * e[3];
    } else {  // circle
        const R dx = p[0] - e[0], dy = p[1] - e[1];
        const R dist = std::sqrt(dx * dx + dy * dy);
        out[0] = dist < T::znorm ? e[0] + e[2] : e[0] + e[2] * dx / dist;
        out[1] = dist < T::znorm ? e[1] : e[1] + e[2] * dy / dist;
    }
}

template <class R> __host__ __device__ inline void closed_tangent(int tag, const R* e, const R* p, R* t)
{
    using T = EntTol<R>;
    if (tag == kEntEllipse) {
        // (-rx sin a, ry cos a) at the parametric angle a of the radially
        // scaled point: sin a, cos a as its coordinates over its length
        const R sx = (p[0] - e[0]) / e[2], sy = (p[1] - e[1]) / e[3];
        const R dist = std::sqrt(sx * sx + sy * sy);
        const R sa = dist < T::znorm ? R(0) : sy / dist, ca = dist < T::znorm ? R(1) : sx / dist;
        const R tx = -e[2] * sa, ty = e[3] * ca;
        const R n = std::sqrt(tx * tx + ty * ty);
        t[0] = n < T::znorm ? R(1) : tx / n;
        t[1] = n < T::znorm ? R(0) : ty / n;
    } else if (tag == kEntSegment || tag == kEntRail || tag == kEntRail3) {
        const int D = tag == kEntRail3 ? 3 : 2;
        for (int k = 0; k < D; ++k) t[k] = e[D + k] - e[k];
        const R n = std::sqrt(ent_dot(t, t, D));
        for (int k = 0; k < D; ++k) t[k] = n < T::znorm ? (k == 0 ? R(1) : R(0)) : t[k] / n;
    } else {  // circle
        const R dx = p[0] - e[0], dy = p[1] - e[1];
        const R rn = std::sqrt(dx * dx + dy * dy);
        t[0] = rn < T::znorm ? -R(0) : -dy / rn;
        t[1] = rn < T::znorm ? R(1) : dx / rn;
    }
}

// ------------------------------------------------------------- frames

// Projection of p onto a non-composite entity and, when B is given, its
// orthonormal tangent basis there (src/geometry.hpp:202-224 for the
// parametrizations; the closed-form kinds' project and tangent).
template <class R> __host__ __device__ inline void simple_frame(int tag, const R* b, const R* p, R* pos, R (*B)[3])
{
    const R* f = b + 1;
    const int D = ent_dim(tag), K = ent_tdim(tag);
    if (tag == kEntFree) {
        for (int k = 0; k < D; ++k) pos[k] = p[k];
        if (B)
            for (int i = 0; i < D; ++i)
                for (int k = 0; k < D; ++k) B[i][k] = i == k ? R(1) : R(0);
        return;
    }
    if (!ent_is_param(tag)) {
        closed_project(tag, f, p, pos);
        if (B) closed_tangent(tag, f, p, B[0]);
        return;
    }
    R q[2] = {R(0), R(0)};
    const R none[2] = {R(0), R(0)};
    par_invert(tag, f, p, none, false, false, q);
    // a periodic angle is re-represented nearest an open arc's interval, so
    // a foot across the seam clamps to the near end
    if ((tag == kEntCircleArc || tag == kEntEllipseArc) && !curve_trim_closed(tag, f)) {
        const R* tr = curve_trim(tag, f);
        const R mid = R(0.5) * (tr[0] + tr[1]);
        q[0] = mid + std::remainder(q[0] - mid, EntTol<R>::two_pi);
    }
    if (!par_contains(tag, f, q)) par_clamp(tag, f, q);
    par_eval(tag, f, q, pos);
    if (B) {
        par_frame(tag, f, q, B);
        ent_orthonormalize(B, K, D);
    }
}

// Any entity: a composite keeps the nearest of its segments (strictly
// nearer wins; the first on ties; the tangent is that segment's).
template <class R> __host__ __device__ inline void ent_frame(int tag, const R* b, const R* p, R* pos, R (*B)[3])
{
    if (tag != kEntComposite) {
        simple_frame(tag, b, p, pos, B);
        return;
    }
    const int m = (int)b[1];
    R best = EntTol<R>::inf;
    pos[0] = p[0];
    pos[1] = p[1];
    if (B) B[0][0] = R(1), B[0][1] = R(0);
    for (int i = 0; i < m; ++i) {
        const R* sb = b + (long long)b[2 + i];
        const int st = (int)sb[0];
        R sp[2], sB[1][3];
        simple_frame(st, sb, p, sp, B ? sB : nullptr);
        const R dd = (sp[0] - p[0]) * (sp[0] - p[0]) + (sp[1] - p[1]) * (sp[1] - p[1]);
        if (dd < best) {
            best = dd;
            pos[0] = sp[0];
            pos[1] = sp[1];
            if (B) B[0][0] = sB[0][0], B[0][1] = sB[0][1];
        }
    }
}

// Projection onto a seeded (iterative) parametrization: the unclamped foot
// from the seed goes back into the seed, then the trim clamp and the point
// (and, when B is given, the orthonormal frame there) follow
// (src/geometry.hpp:226-258).
template <class R>
__host__ __device__ inline void seeded_frame(int tag, const R* b, const R* p, R* seed, bool warm, R* pos, R (*B)[3])
{
    const R* f = b + 1;
    const bool has = std::isfinite(seed[0]);
    R q[2] = {R(0), R(0)};
    par_invert(tag, f, p, seed, has, warm, q);
    seed[0] = q[0];
    if (ent_tdim(tag) == 2) seed[1] = q[1];
    if (!par_contains(tag, f, q)) par_clamp(tag, f, q);
    if (pos) par_eval(tag, f, q, pos);
    if (B) {
        par_frame(tag, f, q, B);
        ent_orthonormalize(B, ent_tdim(tag), ent_dim(tag));
    }
}

}  // namespace eggmi

namespace eggmi
{
// Host: checks the blob at b (at most `avail` values; entity.hpp layout) for
// the session's dimension and returns its length, throwing InvalidArgument
// on a bad tag, a size or count that does not fit, a degree above the caps,
// non-monotone knots, non-positive weights, a bad trim polygon, a nested
// composite or a non-finite value (entity_host.cpp).
long long entity_blob_check(const double* b, long long avail, int dim);
}  // namespace eggmi
