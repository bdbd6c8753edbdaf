// Jacobi-preconditioned conjugate gradients on the MI355X in FP64
// (vectors and inner products), for a CSR operator: the solver skeleton of
// the reference's control-net Gauss-Newton step (src/control_solve.hpp:
// 761-818): x0 = 0, r = b, z = M^-1 r, p = z; per iteration Ap, alpha =
// rz / pAp (stop on non-positive curvature), x += alpha p, r -= alpha Ap,
// stop when |r|^2 <= rtol^2 |b|^2, beta = rz'/rz, p = z + beta p.
//
// The scalars never leave the device: each reduction writes one FP64
// partial per workgroup and a one-workgroup kernel folds them in workgroup
// order (deterministic) and updates alpha / beta and the stop flag. A kernel
// boundary, not an in-kernel device-scope fence, makes the partials of all
// XCDs visible (a device-scope release per workgroup writes back its XCD's
// L2 and costs far more than a launch). Five launches per iteration, no host
// round trip; the host polls the flag every few iterations.
#include "pcg.hpp"

#include <cmath>

namespace eggmi
{

constexpr int kPcgThreads = 256;

struct PcgState {
    double b2, stop2, rz, rr, pap, alpha, beta;
    int done;  // 1 converged, 2 non-positive curvature, 3 zero right-hand side
    int iters;
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

// fixed-order fold of n partials by one workgroup (valid in thread 0)
__device__ __forceinline__ double fold(const double* part, int n)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) v += part[i];
    return wg_sum(v);
}

__global__ __launch_bounds__(kPcgThreads) void pcg_init(const CsrView A, const double* __restrict__ dinv,
                                                         const double* __restrict__ b, double* x, double* r,
                                                         double* z, double* p, double* part)
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
}

__global__ __launch_bounds__(kPcgThreads) void pcg_init_finish(const double* part, int nparts, PcgState* st,
                                                                double rtol2)
{
    const double b2 = fold(part, nparts);
    const double rz0 = fold(part + nparts, nparts);
    if (threadIdx.x == 0) {
        st->b2 = b2;
        st->rr = b2;
        st->rz = rz0;
        st->stop2 = rtol2 * b2;
        st->done = (b2 == 0.0) ? 3 : 0;
        st->iters = 0;
    }
}

__global__ __launch_bounds__(kPcgThreads) void pcg_matvec(const CsrView A, const double* __restrict__ p, double* ap,
                                                          double* part, const PcgState* st)
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
}

__global__ __launch_bounds__(kPcgThreads) void pcg_alpha(const double* part, int nparts, PcgState* st)
{
    if (st->done) return;
    const double pap = fold(part, nparts);
    if (threadIdx.x == 0) {
        st->pap = pap;
        if (!(pap > 0.0))
            st->done = 2;
        else
            st->alpha = st->rz / pap;
    }
}

__global__ __launch_bounds__(kPcgThreads) void pcg_update(const CsrView A, const double* __restrict__ dinv,
                                                          double* x, double* r, double* z, const double* p,
                                                          const double* ap, double* part, const PcgState* st)
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
}

__global__ __launch_bounds__(kPcgThreads) void pcg_beta(const double* part, int nparts, PcgState* st)
{
    if (st->done) return;
    const double rr_all = fold(part, nparts);
    const double rz_new = fold(part + nparts, nparts);
    if (threadIdx.x == 0) {
        st->rr = rr_all;
        st->iters += 1;
        if (rr_all <= st->stop2) {
            st->done = 1;
        } else {
            st->beta = rz_new / st->rz;
            st->rz = rz_new;
        }
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
    const int np = (int)grid;
    jacobi_diag<<<grid, kPcgThreads, 0, s>>>(A, w.dinv);
    PcgState* st = reinterpret_cast<PcgState*>(w.state);
    pcg_init<<<grid, kPcgThreads, 0, s>>>(A, w.dinv, b, x, w.r, w.z, w.p, w.part);
    pcg_init_finish<<<1, kPcgThreads, 0, s>>>(w.part, np, st, rtol * rtol);
    PcgState h {};
    constexpr int kPoll = 16;
    for (int it = 0; it < max_iter; ++it) {
        pcg_matvec<<<grid, kPcgThreads, 0, s>>>(A, w.p, w.ap, w.part, st);
        pcg_alpha<<<1, kPcgThreads, 0, s>>>(w.part, np, st);
        pcg_update<<<grid, kPcgThreads, 0, s>>>(A, w.dinv, x, w.r, w.z, w.p, w.ap, w.part, st);
        pcg_beta<<<1, kPcgThreads, 0, s>>>(w.part, np, st);
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
