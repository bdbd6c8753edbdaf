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

void launch_restrict(int dim, const TransferArgs& a, const float* fx, float* cx, const float* fgrad, float* ctau,
                     const uint8_t* cfree, hipStream_t s);
void launch_prolong(int dim, const TransferArgs& a, const float* ccorr, float* fcorr, hipStream_t s);
void launch_masked_axpy(float* x, const float* c, float alpha, const uint8_t* free, long long n, long long plane,
                        int dim, hipStream_t s);
void launch_zero_unmasked(float* c, const uint8_t* free, long long n, long long plane, int dim, hipStream_t s);
void launch_vec_sub(float* out, const float* a, const float* b, long long n, hipStream_t s);

// Fine energy and min det of x + a0 c and x + a1 c in one pass over the
// structured cells (FP64 accumulation): out = {E(a0), md(a0), E(a1), md(a1)}.
void launch_linesearch_pair(int dim, int kind, const CellReduceArgs& a, const float* corr, float a0, float a1,
                            double* partials, double* out, hipStream_t s);

}  // namespace eggmi
