// Device descriptors and launchers of the FAS transfer kernels (mg_kernels.hip)
// and the line-search pair reduction (halo_reduce.hip).
#pragma once

#include "kernels.hpp"

namespace eggmi
{

// One fine/coarse level pair: per-block descriptors of both levels, the
// per-block per-axis coarsening factor (2 or 1, three ints per block) and the
// prefix of the enumerated node counts (coarse interior nodes for
// restriction, fine interior nodes for prolongation).
struct TransferArgs {
    const BlockDesc* fine = nullptr;
    const BlockDesc* coarse = nullptr;
    const int* factor = nullptr;
    const long long* off = nullptr;
    int nb = 0;
    long long n = 0;
    long long fplane = 0, cplane = 0;
};

// The restriction's optional second and third coarse state buffers and the
// child level's all-roots push table: with off set, the injected state goes
// to cx, x1 and rx at every interior node and at every replica slot it is
// the root of (no halo or copy launch after it).
struct RestrictPush {
    float* x1 = nullptr;
    float* rx = nullptr;
    const int* off = nullptr;
    const int* dst = nullptr;
    const uint8_t* is_dst = nullptr;
};
// Injection of fx into cx (null: none) and full weighting of fgrad into ctau
// (null: none) at the child's nodes.
void launch_restrict(int dim, const TransferArgs& a, const float* fx, float* cx, const float* fgrad, float* ctau,
                     const uint8_t* cfree, hipStream_t s, const RestrictPush& push = RestrictPush {});
// Multilinear prolongation of the coarse change cx - crx over the fine
// level's real nodes: add = true adds it to fout at free fine slots and
// writes the sums into their replica slots (push table cp_off / cp_dst of the
// fine level, null = none); add = false writes it to fout at the free slots
// and zero at the other real nodes, and with a push table also into the free
// slots' replicas, skipping every replica slot (is_dst) in its own thread.
void launch_prolong(int dim, const TransferArgs& a, const float* cx, const float* crx, float* fout,
                    const uint8_t* ffree, bool add, const int* cp_off, const int* cp_dst, hipStream_t s,
                    const uint8_t* is_dst = nullptr);
void launch_masked_axpy(float* x, const float* c, float alpha, const uint8_t* free, long long n, long long plane,
                        int dim, hipStream_t s);

// Fine energy and min det of x + a0 c and x + a1 c in one pass over the
// structured cells (FP64 accumulation): out = {E(a0), md(a0), E(a1), md(a1)}.
// The FAS correction's step-length ladder: kLadder trials alpha = 2^-(t-1),
// t = 1..kLadder (reference src/fas.hpp:549-581, ten trials).
constexpr int kLadder = 10;
__host__ __device__ inline float ladder_alpha(int t) { return 1.0f / float(1 << (t - 1)); }
// All trials (and alpha = 0) in one pass plus the pick on the device:
// partials needs 2 (kLadder + 1) kReduceBlocks doubles, out 2 (kLadder + 1);
// *alpha receives the accepted step (0: none). `dir` (nullable): whole-sample
// directional terms added to every slot's energy.
void launch_linesearch_ladder(int dim, int kind, const CellReduceArgs& a, const float* corr, double* partials,
                              double tol, double* out, float* alpha, hipStream_t s, const DirView* dir = nullptr);
// a[i] = b[i] = src[i], i < n (a kernel: no copy-engine node in captured cycles)
void launch_copy2(float* a, float* b, const float* src, long long n, hipStream_t s);
// x += (*alpha) c at free slots and their replica slots (push table
// cp_off / cp_dst, null = none); nothing when *alpha == 0.
void launch_masked_axpy_dev(float* x, const float* c, const float* alpha, const uint8_t* free, long long n,
                            long long plane, int dim, hipStream_t s, const int* cp_off = nullptr,
                            const int* cp_dst = nullptr);
// partials: 4 kReduceBlocks doubles.
void launch_linesearch_pair(int dim, int kind, const CellReduceArgs& a, const float* corr, float a0, float a1,
                            double* partials, double* out, hipStream_t s, const DirView* dir = nullptr);

}  // namespace eggmi
