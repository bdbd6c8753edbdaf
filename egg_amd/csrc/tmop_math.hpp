// Per-occurrence TMOP math for the HIP sweep kernels (gfx950).
//
// Every metric sample ("occurrence") of a node's patch is a corner plus its D
// axis neighbours: A[:,k] = s_k (nbr_k - corner), T = A W (reference
// src/patch.hpp:95-117). The moving node enters T through one column pattern
// only: dT/dx_i = e_i u^T, where u depends on the node's role in the sample
// (corner: u_c = -sum_k s_k W[k][c]; axis-k neighbour: u_c = s_k W[k][c];
// reference src/patch.hpp:211-229). The node's gradient and contracted
// Hessian J_b^T H J_b are therefore the first and second derivatives of
// mu along the D directions E_i = e_i u^T, and the second derivative of det T
// along any pair of them is zero (det is linear in each row, and two rows
// equal to u make it vanish). That turns the reference's 4x4 (2D) / 9x9
// (3D) Hessian contraction into a few dot products per sample.
//
// Metrics (identity-target shape family, reference src/metric.hpp):
//   2D shape      mu = s/(2 phi) - 1,          phi = det T            (:38-102)
//   2D untangle   same with phi = (D + sqrt(D^2 + 4 delta^2))/2        (:133-142)
//   2D shape_size mu = (s/2D)^2 - 1 + (D-1)^2                          (:150-197)
//   3D shape      mu = s q/(9 phi^2) - 1, q = |cof T|^2                 (:203-243)
//   3D untangle   phi = Dh(D) as above                                 (:247-261)
//   3D shape_size mu_cond3 + (D-1)^2                                   (:226-231)
#pragma once

#include <hip/hip_runtime.h>

namespace eggmi
{

enum Kind : int { kShape = 0, kUntangle = 1, kShapeSize = 2 };

template <class R> __host__ __device__ __forceinline__ R rcp(R x) { return R(1) / x; }
// Hardware reciprocal (v_rcp_f32, 1 ulp) for the FP32 hot loops on the device.
template <> __host__ __device__ __forceinline__ float rcp<float>(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;  // host instantiation (CPU-side checks of this header)
#endif
}

// phi(D) and its first two derivatives for the barrier (phi = D) or the
// untangle surrogate phi = (D + r)/2, r = sqrt(D^2 + 4 delta^2).
template <int KIND, class R> struct Phi {
    R p, p1, p2;
    __host__ __device__ __forceinline__ Phi(R D, R delta)
    {
        if constexpr (KIND == kUntangle) {
            const R r = sqrt(D * D + R(4) * delta * delta);
            const R ir = rcp(r);
            p = R(0.5) * (D + r);
            p1 = p * ir;
            p2 = R(2) * delta * delta * ir * ir * ir;
        } else {
            p = D;
            p1 = R(1);
            p2 = R(0);
        }
    }
};

// ---------------------------------------------------------------- 2D

// Value only (line-search trials, reductions). t = [a, b, c, d] = vec(T).
template <int KIND, class R> __host__ __device__ __forceinline__ R mu2_value(R a, R b, R c, R d, R delta)
{
    const R D = a * d - b * c;
    const R s = a * a + b * b + c * c + d * d;
    if constexpr (KIND == kShapeSize) {
        const R cond = R(0.5) * s * rcp(D);
        return cond * cond - R(1) + (D - R(1)) * (D - R(1));
    } else {
        const Phi<KIND, R> ph(D, delta);
        return R(0.5) * s * rcp(ph.p) - R(1);  // same operation order as mu2_grad_hess
    }
}

// Value + node gradient (2) + contracted Hessian (h00, h01, h11) of one
// sample along u (u0, u1). Adds into g/h and returns mu.
template <int KIND, class R>
__host__ __device__ __forceinline__ R mu2_grad_hess(R a, R b, R c, R d, R u0, R u1, R delta, R& g0, R& g1, R& h00,
                                           R& h01, R& h11)
{
    const R D = a * d - b * c;
    const R s = a * a + b * b + c * c + d * d;
    // s'_i = 2 (T u)_i ; D'_i = (cof(T) u)_i, cof = [[d, -c], [-b, a]]
    const R s0 = R(2) * (a * u0 + b * u1);
    const R s1 = R(2) * (c * u0 + d * u1);
    const R D0 = d * u0 - c * u1;
    const R D1 = -b * u0 + a * u1;
    const R uu = R(2) * (u0 * u0 + u1 * u1);  // s''_ii
    if constexpr (KIND == kShapeSize) {
        const R iD = rcp(D);
        const R cnd = R(0.5) * s * iD;  // s / 2D
        // c_i = s_i/(2D) - s D_i/(2D^2); c_ij = uu d_ij/(2D) - (s_i D_j + s_j D_i)/(2D^2) + s D_i D_j / D^3
        const R ih = R(0.5) * iD;
        const R c0 = s0 * ih - cnd * D0 * iD;
        const R c1 = s1 * ih - cnd * D1 * iD;
        const R iD2 = iD * iD;
        const R c00 = uu * ih - s0 * D0 * iD2 + R(2) * cnd * D0 * D0 * iD2;
        const R c01 = -R(0.5) * (s0 * D1 + s1 * D0) * iD2 + R(2) * cnd * D0 * D1 * iD2;
        const R c11 = uu * ih - s1 * D1 * iD2 + R(2) * cnd * D1 * D1 * iD2;
        const R dm1 = D - R(1);
        g0 += R(2) * cnd * c0 + R(2) * dm1 * D0;
        g1 += R(2) * cnd * c1 + R(2) * dm1 * D1;
        h00 += R(2) * c0 * c0 + R(2) * cnd * c00 + R(2) * D0 * D0;
        h01 += R(2) * c0 * c1 + R(2) * cnd * c01 + R(2) * D0 * D1;
        h11 += R(2) * c1 * c1 + R(2) * cnd * c11 + R(2) * D1 * D1;
        return cnd * cnd - R(1) + dm1 * dm1;
    } else {
        const Phi<KIND, R> ph(D, delta);
        const R ip = rcp(ph.p);
        const R hs = R(0.5) * s * ip;  // s/(2 phi)
        // h_i = phi' D_i; f_i = s_i/(2 phi) - s h_i/(2 phi^2)
        const R k0 = ph.p1 * D0, k1 = ph.p1 * D1;
        g0 += R(0.5) * s0 * ip - hs * k0 * ip;
        g1 += R(0.5) * s1 * ip - hs * k1 * ip;
        // f_ij = s_ij/(2phi) - (s_i h_j + s_j h_i)/(2 phi^2) - s h_ij/(2 phi^2) + s h_i h_j / phi^3
        const R ip2 = ip * ip;
        const R q2 = hs * ph.p2 * ip;  // s phi''/(2 phi^2)
        const R q3 = R(2) * hs * ip2;  // s / phi^3
        h00 += R(0.5) * uu * ip - s0 * k0 * ip2 - q2 * D0 * D0 + q3 * k0 * k0;
        h01 += -R(0.5) * (s0 * k1 + s1 * k0) * ip2 - q2 * D0 * D1 + q3 * k0 * k1;
        h11 += R(0.5) * uu * ip - s1 * k1 * ip2 - q2 * D1 * D1 + q3 * k1 * k1;
        return hs - R(1);
    }
}

// ---------------------------------------------------------------- 3D

template <class R> struct M3 {
    R t[9];
};

template <class R> __host__ __device__ __forceinline__ R det3(const R* t)
{
    return t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
           t[2] * (t[3] * t[7] - t[4] * t[6]);
}

template <class R> __host__ __device__ __forceinline__ void cof3(const R* t, R* c)
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

template <int KIND, class R> __host__ __device__ __forceinline__ R mu3_value(const R* t, R delta)
{
    R c[9];
    cof3(t, c);
    R s = R(0), q = R(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += t[i] * t[i];
        q += c[i] * c[i];
    }
    const R D = t[0] * c[0] + t[1] * c[1] + t[2] * c[2];
    const Phi<KIND, R> ph(D, delta);
    const R gq = s * q;
    const R ih = rcp(ph.p * ph.p);
    R v = gq * (R(1) / R(9)) * ih - R(1);  // same operation order as mu3_grad_hess
    if constexpr (KIND == kShapeSize) v += (D - R(1)) * (D - R(1));
    return v;
}

// Value + node gradient (3) + contracted Hessian (6 unique, row-major upper:
// h00 h01 h02 h11 h12 h22) of one 3D sample along u.
//   v = T u, w = T^T v, z = T w: s'_i = 2 v_i, q'_i = 2 (s v_i - z_i),
//   D'_i = (cof u)_i, s''_ij = 2|u|^2 d_ij, D''_ij = 0,
//   q''_ij = 2 (d_ij sum_k |R_k|^2 - <R_i, R_j>), R_k = (row k of T) x u.
template <int KIND, class R>
__host__ __device__ __forceinline__ R mu3_grad_hess(const R* t, const R* u, R delta, R* g, R* h)
{
    R c[9];
    cof3(t, c);
    R s = R(0), q = R(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += t[i] * t[i];
        q += c[i] * c[i];
    }
    const R D = t[0] * c[0] + t[1] * c[1] + t[2] * c[2];
    R v[3], w[3], z[3], Dd[3], Rk[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v[i] = t[3 * i] * u[0] + t[3 * i + 1] * u[1] + t[3 * i + 2] * u[2];
        Dd[i] = c[3 * i] * u[0] + c[3 * i + 1] * u[1] + c[3 * i + 2] * u[2];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) w[j] = t[j] * v[0] + t[3 + j] * v[1] + t[6 + j] * v[2];
#pragma unroll
    for (int i = 0; i < 3; ++i) z[i] = t[3 * i] * w[0] + t[3 * i + 1] * w[1] + t[3 * i + 2] * w[2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const R* r = t + 3 * k;
        Rk[k][0] = r[1] * u[2] - r[2] * u[1];
        Rk[k][1] = r[2] * u[0] - r[0] * u[2];
        Rk[k][2] = r[0] * u[1] - r[1] * u[0];
    }
    R G[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) G[i][j] = Rk[i][0] * Rk[j][0] + Rk[i][1] * Rk[j][1] + Rk[i][2] * Rk[j][2];
    const R trG = G[0][0] + G[1][1] + G[2][2];
    const R uu = R(2) * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);

    const Phi<KIND, R> ph(D, delta);
    const R hh = ph.p * ph.p;  // h = phi^2
    const R ih = rcp(hh);
    const R gq = s * q;
    const R ninth = R(1) / R(9);
    R s1[3], q1[3], g1[3], h1[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        s1[i] = R(2) * v[i];
        q1[i] = R(2) * (s * v[i] - z[i]);
        g1[i] = s1[i] * q + s * q1[i];
        h1[i] = R(2) * ph.p * ph.p1 * Dd[i];
        g[i] += ninth * (g1[i] - gq * h1[i] * ih) * ih;
    }
    const R hc = R(2) * (ph.p1 * ph.p1 + ph.p * ph.p2);  // h_ij = hc D_i D_j
    const R dm1 = D - R(1);
    int e = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 3; ++j) {
            const R sij = (i == j) ? uu : R(0);
            const R qij = R(2) * (((i == j) ? trG : R(0)) - G[i][j]);
            const R gij = sij * q + s1[i] * q1[j] + s1[j] * q1[i] + s * qij;
            const R hij = hc * Dd[i] * Dd[j];
            R val = ninth * ih *
                    (gij - (g1[i] * h1[j] + g1[j] * h1[i]) * ih - gq * hij * ih + R(2) * gq * h1[i] * h1[j] * ih * ih);
            if constexpr (KIND == kShapeSize) val += R(2) * Dd[i] * Dd[j];
            h[e++] += val;
        }
    if constexpr (KIND == kShapeSize) {
#pragma unroll
        for (int i = 0; i < 3; ++i) g[i] += R(2) * dm1 * Dd[i];
    }
    R val = gq * ninth * ih - R(1);
    if constexpr (KIND == kShapeSize) val += dm1 * dm1;
    return val;
}

// ---------------------------------------------------------------- solves

// x = -H^{-1} g with the reference fallbacks (src/solve.hpp:32-77).
template <class R> struct Tol;
template <> struct Tol<float> {
    static constexpr float energy = 1e-6f, znorm = 1e-7f;
};
template <> struct Tol<double> {
    static constexpr double energy = 1e-12, znorm = 1e-15;
};

template <class R>
__host__ __device__ __forceinline__ void solve2(R h00, R h01, R h10, R h11, R g0, R g1, R& x0, R& x1)
{
    const R gn = sqrt(g0 * g0 + g1 * g1);
    if (gn < Tol<R>::znorm) {
        x0 = x1 = R(0);
        return;
    }
    const R inv = rcp(h00 * h11 - h01 * h10);
    x0 = -inv * (h11 * g0 - h01 * g1);
    x1 = -inv * (-h10 * g0 + h00 * g1);
    if (!isfinite(x0) || !isfinite(x1)) {
        const R c = R(-0.1) * rcp(gn);
        x0 = c * g0;
        x1 = c * g1;
    }
}

template <class R> __host__ __device__ __forceinline__ void solve3(const R* H, const R* g, R* x)
{
    const R gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (gn < Tol<R>::znorm) {
        x[0] = x[1] = x[2] = R(0);
        return;
    }
    const R c00 = H[4] * H[8] - H[5] * H[7];
    const R c01 = -(H[3] * H[8] - H[5] * H[6]);
    const R c02 = H[3] * H[7] - H[4] * H[6];
    const R inv = rcp(H[0] * c00 + H[1] * c01 + H[2] * c02);
    const R c10 = -(H[1] * H[8] - H[2] * H[7]);
    const R c11 = H[0] * H[8] - H[2] * H[6];
    const R c12 = -(H[0] * H[7] - H[1] * H[6]);
    const R c20 = H[1] * H[5] - H[2] * H[4];
    const R c21 = -(H[0] * H[5] - H[2] * H[3]);
    const R c22 = H[0] * H[4] - H[1] * H[3];
    x[0] = -inv * (c00 * g[0] + c10 * g[1] + c20 * g[2]);
    x[1] = -inv * (c01 * g[0] + c11 * g[1] + c21 * g[2]);
    x[2] = -inv * (c02 * g[0] + c12 * g[1] + c22 * g[2]);
    if (!isfinite(x[0]) || !isfinite(x[1]) || !isfinite(x[2])) {
        const R c = R(-0.1) * rcp(gn);
        x[0] = c * g[0];
        x[1] = c * g[1];
        x[2] = c * g[2];
    }
}

}  // namespace eggmi
