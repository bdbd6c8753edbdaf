// Jacobi-preconditioned conjugate gradients on the MI355X in FP64
// (vectors and inner products), for a CSR operator: the solver skeleton of
// the reference's control-net Gauss-Newton step (src/control_solve.hpp:
// 761-818): x0 = 0, r = b, z = M^-1 r, p = z; per iteration Ap, alpha =
// rz / pAp (stop on non-positive curvature), x += alpha p, r -= alpha Ap,
// stop when |r|^2 <= rtol^2 |b|^2, beta = rz'/rz, p = z + beta p.
//
// The scalars never leave the device: each reduction ends in the last
// workgroup to finish (a device-scope ticket), which folds the per-workgroup
// partials in workgroup order (deterministic) and updates alpha / beta and
// the stop flag. Three launches per iteration, no host round trip; the host
// polls the flag every few iterations.
#include "pcg.hpp"

#include <cmath>

namespace eggmi
{

constexpr int kPcgThreads = 256;

struct PcgState {
    double b2, stop2, rz, rr, pap, alpha, beta;
    int done;  // 1 converged, 2 non-positive curvature, 3 zero right-hand side
    int iters;
    unsigned ticket;
};

__device__ __forceinline__ double wg_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ double red[kPcgThreads / 64];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < kPcgThreads / 64; ++i) t += red[i];
    __syncthreads();
    return t;  // valid in thread 0
}

// Publish this workgroup's partial(s); true in every thread of the last workgroup.
__device__ __forceinline__ bool last_block(PcgState* st)
{
    __shared__ bool last;
    __threadfence();
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(&st->ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    return last;
}

__device__ __forceinline__ double fold(const double* part, int n)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) v += part[i];
    // fixed-order fold: strided partial sums, then the workgroup reduction
    return wg_sum(v);
}

__global__ __launch_bounds__(kPcgThreads) void pcg_init(const CsrView A, const double* __restrict__ dinv,
                                                         const double* __restrict__ b, double* x, double* r,
                                                         double* z, double* p, double* part, PcgState* st,
                                                         double rtol2)
{
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < A.n) {
        const double bi = b[i];
        x[i] = 0.0;
        r[i] = bi;
        const double zi = dinv[i] * bi;
        z[i] = zi;
        p[i] = zi;
        rr = bi * bi;
        rz = bi * zi;
    }
    const double s_rr = wg_sum(rr), s_rz = wg_sum(rz);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s_rr;
        part[gridDim.x + blockIdx.x] = s_rz;
    }
    if (!last_block(st)) return;
    const double b2 = fold(part, gridDim.x);
    const double rz0 = fold(part + gridDim.x, gridDim.x);
    if (threadIdx.x == 0) {
        st->b2 = b2;
        st->rr = b2;
        st->rz = rz0;
        st->stop2 = rtol2 * b2;
        st->done = (b2 == 0.0) ? 3 : 0;
        st->iters = 0;
        st->ticket = 0;
    }
}

__global__ __launch_bounds__(kPcgThreads) void pcg_matvec(const CsrView A, const double* __restrict__ p, double* ap,
                                                          double* part, PcgState* st)
{
    if (st->done) return;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    double d = 0.0;
    if (i < A.n) {
        double acc = 0.0;
        for (long long q = A.rowptr[i]; q < A.rowptr[i + 1]; ++q) acc += A.val[q] * p[A.col[q]];
        ap[i] = acc;
        d = p[i] * acc;
    }
    const double s = wg_sum(d);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    if (!last_block(st)) return;
    const double pap = fold(part, gridDim.x);
    if (threadIdx.x == 0) {
        st->pap = pap;
        if (!(pap > 0.0))
            st->done = 2;
        else
            st->alpha = st->rz / pap;
        st->ticket = 0;
    }
}

__global__ __launch_bounds__(kPcgThreads) void pcg_update(const CsrView A, const double* __restrict__ dinv,
                                                          double* x, double* r, double* z, const double* p,
                                                          const double* ap, double* part, PcgState* st)
{
    if (st->done) return;
    const double a = st->alpha;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < A.n) {
        x[i] += a * p[i];
        const double ri = r[i] - a * ap[i];
        r[i] = ri;
        const double zi = dinv[i] * ri;
        z[i] = zi;
        rr = ri * ri;
        rz = ri * zi;
    }
    const double s_rr = wg_sum(rr), s_rz = wg_sum(rz);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s_rr;
        part[gridDim.x + blockIdx.x] = s_rz;
    }
    if (!last_block(st)) return;
    const double rr_all = fold(part, gridDim.x);
    const double rz_new = fold(part + gridDim.x, gridDim.x);
    if (threadIdx.x == 0) {
        st->rr = rr_all;
        st->iters += 1;
        if (rr_all <= st->stop2) {
            st->done = 1;
        } else {
            st->beta = rz_new / st->rz;
            st->rz = rz_new;
        }
        st->ticket = 0;
    }
}

__global__ __launch_bounds__(kPcgThreads) void pcg_direction(long long n, const double* z, double* p,
                                                             const PcgState* st)
{
    if (st->done) return;
    const double beta = st->beta;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(kPcgThreads) void jacobi_diag(const CsrView A, double* dinv)
{
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    if (i >= A.n) return;
    double d = 0.0;
    for (long long q = A.rowptr[i]; q < A.rowptr[i + 1]; ++q)
        if (A.col[q] == i) d += A.val[q];
    dinv[i] = 1.0 / d;
}

PcgResult pcg_solve_device(const CsrView& A, const double* b, double* x, double rtol, int max_iter,
                           PcgWork& w, hipStream_t s)
{
    const unsigned grid = (unsigned)((A.n + kPcgThreads - 1) / kPcgThreads);
    jacobi_diag<<<grid, kPcgThreads, 0, s>>>(A, w.dinv);
    PcgState* st = reinterpret_cast<PcgState*>(w.state);
    hip_check(hipMemsetAsync(st, 0, sizeof(PcgState), s), "pcg state");
    pcg_init<<<grid, kPcgThreads, 0, s>>>(A, w.dinv, b, x, w.r, w.z, w.p, w.part, st, rtol * rtol);
    PcgState h {};
    constexpr int kPoll = 16;
    for (int it = 0; it < max_iter; ++it) {
        pcg_matvec<<<grid, kPcgThreads, 0, s>>>(A, w.p, w.ap, w.part, st);
        pcg_update<<<grid, kPcgThreads, 0, s>>>(A, w.dinv, x, w.r, w.z, w.p, w.ap, w.part, st);
        pcg_direction<<<grid, kPcgThreads, 0, s>>>(A.n, w.z, w.p, st);
        if ((it + 1) % kPoll == 0 || it + 1 == max_iter) {
            hip_check(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s), "pcg poll");
            hip_check(hipStreamSynchronize(s), "pcg poll");
            if (h.done) break;
        }
    }
    hip_check(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s), "pcg result");
    hip_check(hipStreamSynchronize(s), "pcg result");
    PcgResult res;
    res.iters = h.iters;
    res.relres = h.b2 > 0.0 ? std::sqrt(h.rr / h.b2) : 0.0;
    res.status = h.done;
    return res;
}

size_t pcg_state_bytes() { return sizeof(PcgState); }

}  // namespace eggmi
