// Jacobi-preconditioned CG (FP64) on the device: operator view, work buffers
// and the driver (pcg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace eggmi
{

void hip_check(hipError_t e, const char* what);

struct CsrView {
    long long n = 0;
    const long long* rowptr = nullptr;
    const int* col = nullptr;
    const double* val = nullptr;
};

struct PcgWork {
    double *r = nullptr, *z = nullptr, *p = nullptr, *ap = nullptr, *dinv = nullptr, *part = nullptr;
    void* state = nullptr;
};

struct PcgResult {
    int iters = 0;
    double relres = 0.0;
    int status = 0;  // 0 max_iter reached, 1 converged, 2 non-positive curvature, 3 b = 0
};

size_t pcg_state_bytes();
PcgResult pcg_solve_device(const CsrView& A, const double* b, double* x, double rtol, int max_iter, PcgWork& w,
                           hipStream_t s);

}  // namespace eggmi
