// Directional soft-energy samples (2D; reference src/directional.hpp,
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

// A sample's energy at positions p[0 .. arity-1] (src/directional.hpp:95-124);
// T = float in the sweeps, double in the FP64 reductions.
template <typename T>
__device__ __forceinline__ T dir_energy(int kind, const T (&p)[3][2], const float* ref, T w, T eps)
{
    if (kind == 0 || kind == 2) {
        const int hd = kind == 0 ? 1 : 0;
        const T ux = p[hd][0] - p[1 - hd][0], uy = p[hd][1] - p[1 - hd][1];
        const T c = (ux * T(ref[0]) + uy * T(ref[1])) / sqrt(eps * eps + ux * ux + uy * uy);
        return w * c * c;
    }
    if (kind == 1) {
        const T ax = p[0][0] - p[1][0], ay = p[0][1] - p[1][1], bx = p[2][0] - p[1][0], by = p[2][1] - p[1][1];
        const T sa = sqrt(eps * eps + ax * ax + ay * ay), sb = sqrt(eps * eps + bx * bx + by * by);
        const T rx = ax / sa + bx / sb, ry = ay / sa + by / sb;
        return w * (rx * rx + ry * ry);
    }
    return T(0);
}

// Gradient, Hessian (h00, h01, h11) and energy of one sample for the node at
// `slot`, added in (src/directional.hpp:126-211).
__device__ __forceinline__ void dir_accumulate(int kind, const float (&p)[3][2], const float* ref, float w,
                                               float eps, int slot, float (&g)[2], float (&h)[3], float& e)
{
    if (kind == 0 || kind == 2) {
        const int hd = kind == 0 ? 1 : 0;
        const float ux = p[hd][0] - p[1 - hd][0], uy = p[hd][1] - p[1 - hd][1];
        const float ss = eps * eps + ux * ux + uy * uy, sn = sqrtf(ss);
        const float c = (ux * ref[0] + uy * ref[1]) / sn;
        const float drx = ref[0] / sn - c * ux / ss, dry = ref[1] / sn - c * uy / ss;  // dc/du
        const float gc = 2.0f * w * c * (slot == hd ? 1.0f : -1.0f), hc = 2.0f * w;
        const float dd = drx * drx + dry * dry;  // the other endpoint's coupling
        g[0] += gc * drx;
        g[1] += gc * dry;
        h[0] += hc * drx * drx + hc * dd;
        h[1] += hc * drx * dry;
        h[2] += hc * dry * dry + hc * dd;
        e += w * c * c;
        return;
    }
    if (kind != 1) return;
    // unit legs hk = qk / sk and their Jacobians Jk = (I - qk qk^T / sk^2) / sk
    float J[2][3], hx[2], hy[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const float qx = p[2 * t][0] - p[1][0], qy = p[2 * t][1] - p[1][1];
        const float ss = eps * eps + qx * qx + qy * qy, sn = sqrtf(ss), i3 = 1.0f / (ss * sn);
        hx[t] = qx / sn;
        hy[t] = qy / sn;
        J[t][0] = 1.0f / sn - qx * qx * i3;
        J[t][1] = -qx * qy * i3;
        J[t][2] = 1.0f / sn - qy * qy * i3;
    }
    const float rx = hx[0] + hx[1], ry = hy[0] + hy[1];
    float Js[3], f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) Js[k] = slot == 0 ? J[0][k] : slot == 2 ? J[1][k] : -(J[0][k] + J[1][k]);
    // Frobenius norms of J_a, J_b, J_a + J_b (the off-diagonal counts twice)
    auto frob = [](float a, float b, float c) { return sqrtf(a * a + 2.0f * b * b + c * c); };
    f[0] = frob(J[0][0], J[0][1], J[0][2]);
    f[1] = frob(J[1][0], J[1][1], J[1][2]);
    f[2] = frob(J[0][0] + J[1][0], J[0][1] + J[1][1], J[0][2] + J[1][2]);
    const float fs = slot == 0 ? f[0] : slot == 2 ? f[1] : f[2];
    const float gc = 2.0f * w, stab = gc * fs * (f[0] + f[1] + f[2] - fs);
    g[0] += gc * (Js[0] * rx + Js[1] * ry);
    g[1] += gc * (Js[1] * rx + Js[2] * ry);
    h[0] += gc * (Js[0] * Js[0] + Js[1] * Js[1]) + stab;
    h[1] += gc * (Js[0] * Js[1] + Js[1] * Js[2]);
    h[2] += gc * (Js[1] * Js[1] + Js[2] * Js[2]) + stab;
    e += w * (rx * rx + ry * ry);
}

}  // namespace eggmi
