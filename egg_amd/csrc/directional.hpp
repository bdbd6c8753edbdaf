// Directional soft-energy samples (reference src/directional.hpp,
// NumPy twin egg/smoothing/directional.py:304-369). Kinds:
//   0 parallel (nodes i, j): E = w c^2, c = (u . ref) / s, u = p_j - p_i,
//     s = sqrt(|u|^2 + eps^2): a chain segment kept parallel to a wall;
//   2 stem (nodes n1, c): the same with u = p_n1 - p_c: a leg kept
//     perpendicular to a through rail;
//   1 line (nodes a, c, b): E = w |q1 / s1 + q2 / s2|^2, qk = p_{a|b} - p_c:
//     two legs kept opposed.
// Any other kind adds nothing. The per-DOF update takes the sample's
// gradient, its Gauss-Newton Hessian block and a diagonal stabiliser
// (coupling to the frozen neighbours); the det barrier is untouched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace eggmi {

__host__ __device__ constexpr int dir_arity(int kind) { return kind == 1 ? 3 : (kind == 0 || kind == 2) ? 2 : 0; }

// Packed index of H(i, j), i <= j, in the upper-triangle layout the Newton
// solve reads ({h00, h01, h11} in 2D, {h00, h01, h02, h11, h12, h22} in 3D).
template <int D> __host__ __device__ constexpr int dir_hidx(int i, int j) { return i * D - i * (i - 1) / 2 + (j - i); }

// A sample's energy at positions p[0 .. arity-1] (src/directional.hpp:95-124);
// T = float in the sweeps, double in the FP64 reductions.
template <int D, typename T>
__device__ __forceinline__ T dir_energy(int kind, const T (&p)[3][D], const float* ref, T w, T eps)
{
    if (kind == 0 || kind == 2) {
        const int hd = kind == 0 ? 1 : 0;
        T ss = eps * eps, dot = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T u = p[hd][k] - p[1 - hd][k];
            ss += u * u;
            dot += u * T(ref[k]);
        }
        const T c = dot / sqrt(ss);
        return w * c * c;
    }
    if (kind == 1) {
        T sa = eps * eps, sb = eps * eps;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T a = p[0][k] - p[1][k], b = p[2][k] - p[1][k];
            sa += a * a;
            sb += b * b;
        }
        sa = sqrt(sa);
        sb = sqrt(sb);
        T e = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const T r = (p[0][k] - p[1][k]) / sa + (p[2][k] - p[1][k]) / sb;
            e += r * r;
        }
        return w * e;
    }
    return T(0);
}

// Gradient, packed Hessian and energy of one sample for the node at `slot`,
// added in (src/directional.hpp:126-211).
template <int D>
__device__ __forceinline__ void dir_accumulate(int kind, const float (&p)[3][D], const float* ref, float w,
                                               float eps, int slot, float (&g)[D], float (&h)[D * (D + 1) / 2],
                                               float& e)
{
    if (kind == 0 || kind == 2) {
        const int hd = kind == 0 ? 1 : 0;
        float u[D], ss = eps * eps, dot = 0.0f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            u[k] = p[hd][k] - p[1 - hd][k];
            ss += u[k] * u[k];
            dot += u[k] * ref[k];
        }
        const float sn = sqrtf(ss), c = dot / sn;
        float dr[D], dd = 0.0f;  // dc/du and the other endpoint's coupling |dr|^2
#pragma unroll
        for (int k = 0; k < D; ++k) {
            dr[k] = ref[k] / sn - c * u[k] / ss;
            dd += dr[k] * dr[k];
        }
        const float gc = 2.0f * w * c * (slot == hd ? 1.0f : -1.0f), hc = 2.0f * w;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            g[i] += gc * dr[i];
#pragma unroll
            for (int j = i; j < D; ++j) h[dir_hidx<D>(i, j)] += hc * dr[i] * dr[j] + (i == j ? hc * dd : 0.0f);
        }
        e += w * c * c;
        return;
    }
    if (kind != 1) return;
    // unit legs hk = qk / sk and their Jacobians Jk = (I - qk qk^T / sk^2) / sk
    float J[2][D][D], hv[2][D];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        float q[D], ss = eps * eps;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            q[k] = p[2 * t][k] - p[1][k];
            ss += q[k] * q[k];
        }
        const float sn = sqrtf(ss), i3 = 1.0f / (ss * sn);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            hv[t][i] = q[i] / sn;
#pragma unroll
            for (int j = 0; j < D; ++j) J[t][i][j] = (i == j ? 1.0f / sn : 0.0f) - q[i] * q[j] * i3;
        }
    }
    float r[D], Js[D][D], f[3] = {0.0f, 0.0f, 0.0f}, rr = 0.0f;  // f: |J_a|_F, |J_b|_F, |J_a + J_b|_F
#pragma unroll
    for (int i = 0; i < D; ++i) {
        r[i] = hv[0][i] + hv[1][i];
        rr += r[i] * r[i];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const float a = J[0][i][j], b = J[1][i][j];
            Js[i][j] = slot == 0 ? a : slot == 2 ? b : -(a + b);
            f[0] += a * a;
            f[1] += b * b;
            f[2] += (a + b) * (a + b);
        }
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) f[t] = sqrtf(f[t]);
    const float fs = slot == 0 ? f[0] : slot == 2 ? f[1] : f[2];
    const float gc = 2.0f * w, stab = gc * fs * (f[0] + f[1] + f[2] - fs);
#pragma unroll
    for (int i = 0; i < D; ++i) {
        float gi = 0.0f;
#pragma unroll
        for (int j = 0; j < D; ++j) gi += Js[i][j] * r[j];
        g[i] += gc * gi;
#pragma unroll
        for (int j = i; j < D; ++j) {
            float hij = 0.0f;
#pragma unroll
            for (int m = 0; m < D; ++m) hij += Js[i][m] * Js[j][m];
            h[dir_hidx<D>(i, j)] += gc * hij + (i == j ? stab : 0.0f);
        }
    }
    e += w * rr;
}

}  // namespace eggmi
