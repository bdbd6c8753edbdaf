// Shared pieces of the sweep kernels: compile-time loops, the interior
// patch occurrence tables and the line-search accept rule.
#pragma once

#include "kernels.hpp"
#include "tmop_math.hpp"

#include <utility>

namespace eggmi
{

template <int N, class F> __device__ __forceinline__ void static_for(F&& f)
{
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (f(std::integral_constant<int, I> {}), ...);
    }(std::make_integer_sequence<int, N> {});
}

constexpr float kInf = __builtin_huge_valf();

// ------------------------------------------------------------------ 2D tables

// Occurrence `occ` of an interior 2D patch in local 3x3 coordinates (the DOF
// sits at (1,1)): corner (cr, cc), axis-0 neighbour (ar, ac), axis-1
// neighbour (br, bc), signs and the DOF's role.
struct Occ2 {
    int cr, cc, ar, ac, br, bc, s0, s1, role;
};

__host__ __device__ constexpr Occ2 occ2(int occ)
{
    const int cell = occ >> 2, o = occ & 3;
    const int o0 = o >> 1, o1 = o & 1;
    const int cr = (cell >> 1) + o0, cc = (cell & 1) + o1;
    const int ar = cr + (o0 ? -1 : 1), ac = cc;
    const int br = cr, bc = cc + (o1 ? -1 : 1);
    const int role = (cr == 1 && cc == 1) ? 0 : (ar == 1 && ac == 1) ? 1 : (br == 1 && bc == 1) ? 2 : -1;
    return {cr, cc, ar, ac, br, bc, o0 ? -1 : 1, o1 ? -1 : 1, role};
}

__host__ __device__ constexpr int occ2_m1_index(int occ)
{
    int n = 0;
    for (int o = 0; o < occ; ++o) n += occ2(o).role < 0 ? 1 : 0;
    return n;
}

// ------------------------------------------------------------------ 3D tables

struct Occ3 {
    int c;     // corner local index a0*9 + a1*3 + a2 (DOF = 13)
    int n[3];  // axis-k neighbour local index
    int s[3];
    int role;
};

__host__ __device__ constexpr Occ3 occ3(int occ)
{
    const int cell = occ >> 3, o = occ & 7;
    int corner[3] = {0, 0, 0}, ob[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        ob[k] = (o >> (2 - k)) & 1;
        corner[k] = ((cell >> (2 - k)) & 1) + ob[k];
    }
    Occ3 r {};
    r.c = corner[0] * 9 + corner[1] * 3 + corner[2];
    r.role = (r.c == 13) ? 0 : -1;
    for (int k = 0; k < 3; ++k) {
        int nb[3] = {corner[0], corner[1], corner[2]};
        nb[k] += ob[k] ? -1 : 1;
        r.n[k] = nb[0] * 9 + nb[1] * 3 + nb[2];
        r.s[k] = ob[k] ? -1 : 1;
        if (r.role < 0 && r.n[k] == 13) r.role = k + 1;
    }
    return r;
}

__host__ __device__ constexpr int occ3_m1_index(int occ)
{
    int n = 0;
    for (int o = 0; o < occ; ++o) n += occ3(o).role < 0 ? 1 : 0;
    return n;
}

// ------------------------------------------------------------------ common

template <int KIND> __device__ __forceinline__ bool accept(float e, float e0, float md)
{
    return isfinite(e) && (e <= e0 + Tol<float>::energy) && (KIND == kUntangle || md > 0.0f);
}

}  // namespace eggmi
