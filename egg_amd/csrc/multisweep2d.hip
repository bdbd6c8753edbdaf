// Multi-sweep kernel of the 2D structured smoother: k block-Jacobi sweeps per
// launch over LDS-staged dependency closures (tables: multisweep.cpp; data
// layout: kernels.hpp, MsArgs).
//
// Semantics per sweep are the reference's frozen-snapshot Jacobi sweep
// (src/sweep.hpp:1938-1984: every DOF of sweep s reads the state of sweep
// s-1); a sub-sweep computes each DOF with the one-sweep kernel's own device
// functions (synth_solve, stored_group). The stored DOFs' updates come out
// bit for bit; the compiler fuses some multiply-adds of the synthesized
// DOFs' update differently here, which moves about 1 % of them by one float
// ulp per sweep against the one-sweep kernel.
//
// Correctness of the redundant rings: sub-sweep s computes exactly the DOFs
// within distance k-1-s of the core (the tile's entry lists are sorted by
// distance, so they are a prefix). Such a DOF reads only nodes within
// distance k-s, which sub-sweep s-1 computed (or, for fixed nodes and the
// first sub-sweep, staged); ghosts of a region take their roots' new state
// before the next sub-sweep. Every LDS entry the core's k-th sweep depends
// on therefore holds the right state when it is read.
#include "sweep_impl.hpp"

#include <stdexcept>
#include <string>

namespace eggmi
{

// Stored entries of a sub-sweep: positions from the LDS buffer `src`, the
// result into `dst` (or, in the last sub-sweep, the store and its replicas).
struct MsIO {
    const MsArgs* m;
    const float* src;
    float* dst;
    int S;       // LDS plane stride
    bool fin;    // last sub-sweep: the entries are the core
    int core_end;  // stored entries below it are the core (counted)
    __device__ const StoredView& st(const SweepArgs&) const { return m->st; }
    __device__ const float* X(const SweepArgs&) const { return src; }
    __device__ long long PL(const SweepArgs&) const { return S; }
    __device__ long long slot(int d, int) const { return m->st_slot[d]; }
    template <int D> __device__ void finish(const SweepArgs& a, int d, int dof, const float* cur, int ntr) const
    {
        if (fin) {
            store_node<D>(a, m->st_slot[d], cur);
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) dst[k * S + dof] = cur[k];
        }
        if (d < core_end) count_update(a, ntr);
    }
};

template <int KIND, bool TAU, int CPL, int NT>
__global__ __launch_bounds__(NT) void multi_sweep_kernel(const SweepArgs a, const MsArgs m)
{
    constexpr int D = 2, NC = 4, NL = NC / CPL, G = stored_lanes(2);
    static_assert(NT % 64 == 0 && NT % G == 0, "whole waves and stored lane groups");
    extern __shared__ float lds[];
    const int S = m.lds_stride;
    float* const b0 = lds;
    float* const b1 = lds + D * S;
    const MsTile& T = m.tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const long long PL = a.plane;

    // ---- stage the closure in both buffers (rows of each region coalesced)
    for (int ri = 0; ri < T.nreg; ++ri) {
        const MsRegion& R = m.regions[T.reg0 + ri];
        const int n = R.h * R.w;
        for (int e = tid; e < n; e += NT) {
            const int row = e / R.w, col = e - row * R.w;
            const long long slot = R.base + (long long)row * R.st0 + col;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float v = a.xin[k * PL + slot];
                b0[k * S + R.lds + e] = v;
                b1[k * S + R.lds + e] = v;
            }
        }
    }
    const int e0 = T.n_lds - T.next;
    for (int e = tid; e < T.next; e += NT) {
        const long long slot = m.ext[T.ext0 + e];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float v = a.xin[k * PL + slot];
            b0[k * S + e0 + e] = v;
            b1[k * S + e0 + e] = v;
        }
    }
    __syncthreads();

    for (int s = 0; s < m.k; ++s) {
        const int r = m.k - 1 - s;  // the farthest distance computed in this sub-sweep
        const bool fin = r == 0;
        const float* src = (s & 1) ? b1 : b0;
        float* dst = (s & 1) ? b0 : b1;

        // ---- synthesized DOFs within distance r (3x3 row addressing in their region)
        const int nsyn = T.syn_cnt[r];
        for (int it = tid / NL; it < nsyn; it += NT / NL) {
            const int2 en = m.syn[T.syn0 + it];
            const int li = en.x & 0xFFFF, w = en.x >> 16;
            const int gl = tid % NL;
            float rho[CPL][D];
            float P[CPL][NC][D];
#pragma unroll
            for (int j = 0; j < CPL; ++j)
                synth_corners<D>(gl * CPL + j, rho[j], [&](int op, const int* loc) {
                    const int q = li + (loc[0] - 1) * w + (loc[1] - 1);
#pragma unroll
                    for (int k = 0; k < D; ++k) P[j][op][k] = src[k * S + q];
                });
            float cur[D];
            const int ntr = synth_solve<D, KIND, TAU, false, CPL>(a, P, rho, gl, en.y, cur);
            if (gl == 0) {
                if (fin) {
                    store_node<D>(a, en.y, cur);
                } else {
#pragma unroll
                    for (int k = 0; k < D; ++k) dst[k * S + li] = cur[k];
                }
                if (it < T.syn_cnt[0]) count_update(a, ntr);  // the core, once per sweep
            }
        }

        // ---- stored DOFs within distance r (entries sorted by distance)
        const int ns = T.st_cnt[r];
        const MsIO io {&m, src, dst, S, fin, T.st0 + T.st_cnt[0]};
        for (int it = tid / G; it < ns; it += NT / G)
            stored_group<D, KIND, TAU, false, G, false, false, MsIO>(a, T.st0 + it, tid % G, true, io);
        if (fin) break;
        __syncthreads();
        // ---- ghosts staged in a region take their roots' new state
        if (T.npush > 0) {
            for (int p = tid; p < T.npush; p += NT) {
                const int ps = m.push_src[T.push0 + p], pd = m.push_dst[T.push0 + p];
#pragma unroll
                for (int k = 0; k < D; ++k) dst[k * S + pd] = dst[k * S + ps];
            }
            __syncthreads();
        }
    }
}

namespace
{
constexpr MsVariant kMsVariants[] = {
    {1, 256},  // one lane per cell
    {2, 256},  // two cells per lane
    {4, 256},  // one lane per node
    {2, 128},
    {4, 128},
    {1, 512},
};
constexpr int kNumMsVariants = sizeof(kMsVariants) / sizeof(kMsVariants[0]);

template <int KIND, bool TAU, int CPL, int NT> void ms_launch(const SweepArgs& a, const MsArgs& m, hipStream_t s)
{
    const size_t bytes = ms_lds_bytes(2, m.lds_stride);
    multi_sweep_kernel<KIND, TAU, CPL, NT><<<dim3(m.n_tiles), dim3(NT), bytes, s>>>(a, m);
}

template <int KIND, bool TAU> void ms_variant_launch(int v, const SweepArgs& a, const MsArgs& m, hipStream_t s)
{
    switch (v) {
    case 0: return ms_launch<KIND, TAU, 1, 256>(a, m, s);
    case 1: return ms_launch<KIND, TAU, 2, 256>(a, m, s);
    case 2: return ms_launch<KIND, TAU, 4, 256>(a, m, s);
    case 3: return ms_launch<KIND, TAU, 2, 128>(a, m, s);
    case 4: return ms_launch<KIND, TAU, 4, 128>(a, m, s);
    case 5: return ms_launch<KIND, TAU, 1, 512>(a, m, s);
    default: throw std::invalid_argument("multi-sweep variant " + std::to_string(v) + " out of range");
    }
}
}  // namespace

int num_ms_variants() { return kNumMsVariants; }

MsVariant ms_variant(int v)
{
    if (v < 0 || v >= kNumMsVariants) throw std::invalid_argument("multi-sweep variant out of range");
    return kMsVariants[v];
}

void launch_multi_sweep(int kind, int variant, const SweepArgs& a, const MsArgs& m, hipStream_t s)
{
    if (m.n_tiles == 0 || m.k <= 0) return;
    if (m.k > kMsMaxK) throw std::invalid_argument("multi-sweep: k above the table depth");
    if (ms_lds_bytes(2, m.lds_stride) > kMsMaxLdsBytes) throw std::invalid_argument("multi-sweep: LDS stride too large");
    if (a.tau) {
        if (kind == kShapeSize)
            ms_variant_launch<kShapeSize, true>(variant, a, m, s);
        else
            ms_variant_launch<kShape, true>(variant, a, m, s);
        return;
    }
    if (kind == kUntangle)
        ms_variant_launch<kUntangle, false>(variant, a, m, s);
    else if (kind == kShapeSize)
        ms_variant_launch<kShapeSize, false>(variant, a, m, s);
    else
        ms_variant_launch<kShape, false>(variant, a, m, s);
}

}  // namespace eggmi
