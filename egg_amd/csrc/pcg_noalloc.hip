// Allocation-free PCG: the solver of pcg.hip on operands the caller already
// holds in device memory, with its work area (vectors, partials, state) in
// a file-scope device array instead of an allocation. For callers that may
// not allocate device memory (tight loops, harnesses that forbid
// hipMalloc). Each GPU holds its own copy of the array until the process
// ends; every call on one GPU shares it, so calls that overlap on one GPU
// race. The solver initialises every entry it reads, so nothing is cleared
// between calls.
#include "pcg.hpp"

#include <mutex>
#include <stdexcept>

namespace eggmi
{

namespace
{

// kPcgNoAllocMaxN unknowns: 9 vectors + partials + state, ~18 MiB.
__device__ __align__(256) double scratch[pcg_work_doubles(kPcgNoAllocMaxN)];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// The array's address on the current device, looked up once per device. In
// host code the array's name is only a stand-in for the symbol, never an
// address a kernel may use.
double* scratch_on_current_device()
{
    int dev = -1;
    hip_check(hipGetDevice(&dev), "pcg no-alloc: hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("pcg no-alloc: device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "pcg no-alloc: hipGetSymbolAddress");
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    return g_scratch[dev];
}

}  // namespace

PcgResult pcg_solve_csr_noalloc(const CsrView& A, const double* b, double* x, double rtol, int max_iter,
                                hipStream_t s, PcgRecurrence rec)
{
    PcgWork w;
    pcg_carve(scratch_on_current_device(), A.n, w);
    return pcg_solve_csr(A, b, x, rtol, max_iter, w, s, PcgLaunch::kNoAlloc, rec);
}

}  // namespace eggmi
