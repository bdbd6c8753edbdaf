// Jacobi-preconditioned CG (FP64) on the device: operator views, the work
// area and the drivers (pcg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace eggmi
{

void hip_check(hipError_t e, const char* what);

// General sparse operator, CSR (rowptr[n + 1], col / val[nnz]).
struct CsrView {
    long long n = 0;
    const long long* rowptr = nullptr;
    const int* col = nullptr;
    const double* val = nullptr;
};

// Matrix-free (2 dim + 1)-point lattice operator on shape[0] x .. x
// shape[dim-1] unknowns (row-major, last axis fastest): `diag` on the
// diagonal, `off` for every neighbour one step along an axis that lies in
// the lattice (Dirichlet closure: neighbours outside add nothing).
struct LatticeView {
    int dim = 3;
    int shape[3] = {1, 1, 1};
    long long n = 0;
    double diag = 0.0, off = 0.0;
};

// Device work area of one solve. Every vector holds n doubles; `part`
// holds 7 * pcg_parts(n) doubles (the p.Ap partials, then three per
// iteration parity); `state` kPcgStateDoubles. The standard recurrence
// uses u[0] for z, w[0] for Ap and s[0] / s[1] for p.
struct PcgWork {
    double *r = nullptr, *dinv = nullptr, *p = nullptr, *part = nullptr;
    double* u[2] = {nullptr, nullptr};
    double* w[2] = {nullptr, nullptr};
    double* s[2] = {nullptr, nullptr};
    void* state = nullptr;
};

// kStandard: the reference's recurrence, two launches per iteration;
// kOneReduction: Chronopoulos-Gear, one launch per iteration (pcg.hip).
enum class PcgRecurrence { kStandard = 0, kOneReduction = 1 };

struct PcgResult {
    int iters = 0;
    double relres = 0.0;
    int status = 0;  // 0 max_iter reached, 1 converged, 2 non-positive curvature, 3 b = 0
};

constexpr int kPcgThreads = 256;
constexpr long long pcg_parts(long long n) { return (n + kPcgThreads - 1) / kPcgThreads; }
constexpr size_t kPcgStateDoubles = 32;  // room for the solver state (checked in pcg.hip)
constexpr size_t pcg_align256(size_t doubles) { return (doubles + 31) & ~size_t(31); }
// doubles of one contiguous work area for n unknowns: 9 vectors, the
// partials and the state, each 256-byte aligned
constexpr size_t pcg_work_doubles(long long n)
{
    const long long m = n > 1 ? n : 1, np = pcg_parts(n) > 1 ? pcg_parts(n) : 1;
    return 9 * pcg_align256((size_t)m) + pcg_align256(7 * (size_t)np) + kPcgStateDoubles;
}
// lay the work area out over `base` (pcg_work_doubles(n) doubles)
void pcg_carve(double* base, long long n, PcgWork& w);

// How the host follows the solve: kQueued reads the device state
// asynchronously through pinned memory (two slots and two events per device
// and host thread, made on first use and kept until the process ends);
// kNoAlloc reads it synchronously and allocates nothing (no pinned memory,
// no events).
enum class PcgLaunch { kQueued, kNoAlloc };
constexpr int kPcgChunk = 16;

// x0 = 0; x receives the solution.
PcgResult pcg_solve_csr(const CsrView& A, const double* b, double* x, double rtol, int max_iter, PcgWork& w,
                        hipStream_t s, PcgLaunch mode, PcgRecurrence rec);
PcgResult pcg_solve_lattice(const LatticeView& A, const double* b, double* x, double rtol, int max_iter,
                            PcgWork& w, hipStream_t s, PcgLaunch mode, PcgRecurrence rec);

// The same CSR solve with its work area in a file-scope device array
// (pcg_noalloc.hip) instead of an allocation: operands in device memory, at
// most kPcgNoAllocMaxN unknowns, nothing allocated (kNoAlloc launching).
constexpr long long kPcgNoAllocMaxN = 1ll << 18;
PcgResult pcg_solve_csr_noalloc(const CsrView& A, const double* b, double* x, double rtol, int max_iter,
                                hipStream_t s, PcgRecurrence rec);

}  // namespace eggmi
