// Jacobi-preconditioned conjugate gradients on the MI355X in FP64
// (vectors and inner products): the solver skeleton of the reference's
// control-net Gauss-Newton step (src/control_solve.hpp:761-818). x0 = 0,
// stop when |r|^2 <= rtol^2 |b|^2 (iterations counted with the converging
// one), stop on non-positive curvature p.Ap <= 0 (not counted), nothing
// checked before the first iteration. Two recurrences, the same iterates in
// exact arithmetic:
//
// kStandard, the reference's: r = b, z = M^-1 r, p = z; per iteration Ap,
//   alpha = rz / pAp, x += alpha p, r -= alpha Ap, check |r|, z = M^-1 r,
//   beta = rz' / rz, p = z + beta p. Two launches per iteration:
//   step_a(t): fold the r.r / r.z partials of iteration t-1, decide
//     convergence, beta; p_t = z + beta p_(t-1) and A p_t = A z + beta A
//     p_(t-1) (the operator's loads do not wait for beta, so they overlap
//     the fold), the p.Ap partial of every workgroup;
//   step_b(t): fold p.Ap, alpha, update x, r, z, the r.r / r.z partials.
//
// kOneReduction (Chronopoulos and Gear, 1989): u = M^-1 r and w = A u are
//   carried as well, so the three inner products an iteration needs,
//   (r, u), (w, u) and (r, r), are reduced together: gamma = (r, u),
//   beta = gamma / gamma_prev, pAp = (w, u) - beta gamma / alpha_prev,
//   alpha = gamma / pAp; p = u + beta p, s = w + beta s (= A p),
//   x += alpha p, r -= alpha s, u -= alpha M^-1 s, w = A u. One launch per
//   iteration: A u_(i+1) is formed from the iteration-i values at the
//   neighbours, A u - alpha A M^-1 (w + beta s), so no kernel reads a vector
//   entry another workgroup of the same kernel writes.
//
// Every workgroup folds the partials itself in the same fixed order, so all
// of them take the same decisions, and the grid never synchronises inside a
// kernel: a kernel boundary is the visibility point (a device-scope release
// per workgroup writes back its XCD's L2, which costs more than a launch).
// Scalars written by block 0 live in parity-indexed slots (iteration t uses
// slot t & 1), so no kernel reads a slot that a block of the same kernel
// writes; vectors read at neighbours alternate between two buffers.
#include "pcg.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace eggmi
{

namespace
{

struct PcgState {
    double rtol2, b2, stop2;
    double rz[2];     // standard: r.z entering iteration t, slot t & 1
    double rr[2];     // r.r entering iteration t (b.b at t = 0), slot t & 1
    double alpha[2];  // one-reduction: alpha_t, slot t & 1
    double gamma[2];  // one-reduction: (r, u) entering iteration t
    int t[2];         // iteration index of the last iteration kernel of each parity
    int halt_a[2], halt_b[2];  // 1: the solve has ended (written by that step)
    int max_iter;
    int status;  // -1 while running, then the PcgResult status
    int iters;
    double rr_end;
};

// A DPP lane move of a double (both halves through the same control; every
// source lane of the controls used here lies in the row, so `old` is never
// taken).
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// 64-lane sum in register moves (no LDS round trips). Each DPP step adds to
// every lane the value of a partner lane holding the same partial as the
// other members of its group, and a + b == b + a exactly, so after the four
// steps all 16 lanes of a row hold the same row sum; the four row sums are
// then combined in one fixed order. Every lane, and every workgroup folding
// the same inputs, receives the same bits.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
    v += dpp_mov<0x141>(v);  // row_half_mirror: the other quad of the eight
    v += dpp_mov<0x140>(v);  // row_mirror: the other eight of the row
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// 256-thread workgroup sums; every thread receives the same bits.
__device__ __forceinline__ double wg_sum(double v)
{
    v = wave_sum(v);
    __shared__ double red[kPcgThreads / 64];
    __syncthreads();  // the previous call's readers are done with red
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

template <int K> __device__ __forceinline__ void wg_sums(double (&v)[K])
{
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __shared__ double red[K][kPcgThreads / 64];
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < K; ++k) red[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (red[k][0] + red[k][1]) + (red[k][2] + red[k][3]);
}

// this thread's share of a fixed-order fold of n partials; a workgroup sum
// of it gives the same bits in every workgroup
__device__ __forceinline__ double part_sum(const double* part, long long n)
{
    double v = 0.0;
    for (long long i = threadIdx.x; i < n; i += kPcgThreads) v += part[i];
    return v;
}

// partial k (0..2) of parity par: 6 arrays of np partials after the p.Ap ones
__device__ __forceinline__ double* part_of(double* part, long long np, int par, int k)
{
    return part + np * (1 + 3 * par + k);
}

struct CsrOp {
    CsrView A;
    static constexpr bool kDiagTable = true;  // M^-1 varies by row: kept in dinv
    __device__ long long n() const { return A.n; }
    __device__ double dinv_at(const double* dinv, long long i) const { return dinv[i]; }
    __device__ double inv_diag(long long i) const
    {
        double d = 0.0;
        for (long long q = A.rowptr[i]; q < A.rowptr[i + 1]; ++q)
            if (A.col[q] == i) d += A.val[q];
        return 1.0 / d;
    }
    // (A u)_i and (A v)_i in one walk over the row
    template <bool TWO> __device__ void row(long long i, const double* u, const double* v, double& au, double& av) const
    {
        double su = 0.0, sv = 0.0;
        for (long long q = A.rowptr[i]; q < A.rowptr[i + 1]; ++q) {
            const int j = A.col[q];
            const double a = A.val[q];
            su += a * u[j];
            if (TWO) sv += a * v[j];
        }
        au = su;
        av = sv;
    }
    // (A u)_i, (A D w)_i and (A D s)_i, D = M^-1 = diag(dinv)
    __device__ void row3(long long i, const double* u, const double* w, const double* s, const double* dinv,
                         double& au, double& adw, double& ads) const
    {
        double su = 0.0, sw = 0.0, ss = 0.0;
        for (long long q = A.rowptr[i]; q < A.rowptr[i + 1]; ++q) {
            const int j = A.col[q];
            const double a = A.val[q], ad = a * dinv[j];
            su += a * u[j];
            sw += ad * w[j];
            ss += ad * s[j];
        }
        au = su;
        adw = sw;
        ads = ss;
    }
};

// Unknowns are fewer than 2^31 (checked by the entry point), so the lattice
// coordinates come from 32-bit divisions, once per row.
struct LatticeOp {
    LatticeView A;
    static constexpr bool kDiagTable = false;  // M^-1 = 1 / diag on every row
    __device__ long long n() const { return A.n; }
    __device__ double inv_diag(long long) const { return 1.0 / A.diag; }
    __device__ double dinv_at(const double*, long long) const { return 1.0 / A.diag; }
    // the in-lattice neighbours of row i: their count and indices
    __device__ int neighbours(long long i, int (&nb)[6]) const
    {
        unsigned rem = (unsigned)i, stride = 1;
        int m = 0;
        for (int k = A.dim - 1; k >= 0; --k) {
            const unsigned e = (unsigned)A.shape[k], c = rem % e;
            rem /= e;
            if (c > 0) nb[m++] = (int)(i - stride);
            if (c + 1 < e) nb[m++] = (int)(i + stride);
            stride *= e;
        }
        return m;
    }
    template <bool TWO> __device__ void row(long long i, const double* u, const double* v, double& au, double& av) const
    {
        int nb[6];
        const int m = neighbours(i, nb);
        double su = 0.0, sv = 0.0;
        for (int k = 0; k < m; ++k) {
            su += u[nb[k]];
            if (TWO) sv += v[nb[k]];
        }
        au = A.diag * u[i] + A.off * su;
        av = TWO ? A.diag * v[i] + A.off * sv : 0.0;
    }
    // (A u)_i, (A D w)_i and (A D s)_i, D = diag^-1 I
    __device__ void row3(long long i, const double* u, const double* w, const double* s, const double*, double& au,
                         double& adw, double& ads) const
    {
        int nb[6];
        const int m = neighbours(i, nb);
        double su = 0.0, sw = 0.0, ss = 0.0;
        for (int k = 0; k < m; ++k) {
            su += u[nb[k]];
            sw += w[nb[k]];
            ss += s[nb[k]];
        }
        const double d = 1.0 / A.diag;
        au = A.diag * u[i] + A.off * su;
        adw = d * (A.diag * w[i] + A.off * sw);
        ads = d * (A.diag * s[i] + A.off * ss);
    }
};

__device__ __forceinline__ void finish(PcgState* st, int status, int iters, double rr)
{
    st->status = status;
    st->iters = iters;
    st->rr_end = rr;
}

__device__ __forceinline__ void reset_state(PcgState* st, double rtol2, int max_iter)
{
    st->rtol2 = rtol2;
    st->t[0] = -1;  // no iteration kernel of this solve has run yet
    st->t[1] = -1;
    st->halt_a[1] = 0;
    st->halt_b[1] = 0;
    st->max_iter = max_iter;
    st->status = -1;
    st->iters = 0;
    st->rr_end = 0.0;
}

// ------------------------------------------------------------ kStandard

template <class Op>
__global__ __launch_bounds__(kPcgThreads) void pcg_start(const Op op, const double* __restrict__ b, double* x,
                                                         double* r, double* z, double* dinv, double* part,
                                                         PcgState* st, double rtol2, int max_iter)
{
    const long long n = op.n(), np = gridDim.x;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        const double bi = b[i], di = op.inv_diag(i);
        dinv[i] = di;
        x[i] = 0.0;
        r[i] = bi;
        const double zi = di * bi;
        z[i] = zi;
        v[0] = bi * bi;
        v[1] = bi * zi;
    }
    wg_sums<2>(v);
    if (threadIdx.x == 0) {
        // the partials a step_b of parity 1 would leave: step_a(0) folds them
        part_of(part, np, 1, 0)[blockIdx.x] = v[0];
        part_of(part, np, 1, 1)[blockIdx.x] = v[1];
        if (blockIdx.x == 0) reset_state(st, rtol2, max_iter);
    }
}

// Both steps issue every load they need (the state slots, the partials and
// the vector entries) before anything waits on one, so the latencies of the
// fold, the state and the vectors overlap; the decisions come after.
template <class Op>
__global__ __launch_bounds__(kPcgThreads) void pcg_step_a(const Op op, const double* __restrict__ z, double* p_new,
                                                          const double* __restrict__ p_old, double* ap, double* part,
                                                          PcgState* st, int par)
{
    const int q = par ^ 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const long long n = op.n(), np = gridDim.x;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    const long long ii = i < n ? i : n - 1;  // loads stay in range; stores are guarded
    const int halted = st->halt_a[q] | st->halt_b[q];
    const int t = st->t[q] + 1;
    const double rz_old = st->rz[q], stop2 = st->stop2;
    const int max_iter = st->max_iter;
    double v[2] = {part_sum(part_of(part, np, q, 0), np), part_sum(part_of(part, np, q, 1), np)};
    // A z and A p_(t-1) do not depend on beta (p_(t-1) is not yet written at t = 0: dropped below)
    double az, aq;
    op.template row<true>(ii, z, p_old, az, aq);
    const double zi = z[ii];
    double pi = p_old[ii];
    if (halted) {
        if (lead) st->halt_a[par] = 1;
        return;
    }
    wg_sums<2>(v);
    const double rr = v[0], rz = v[1];
    double beta = 0.0;
    if (t == 0) {
        aq = 0.0;
        pi = 0.0;
        if (rr == 0.0 || max_iter == 0) {  // b = 0, or no iteration allowed
            if (lead) {
                st->b2 = rr;
                finish(st, rr == 0.0 ? 3 : 0, 0, rr);
                st->halt_a[par] = 1;
            }
            return;
        }
    } else {
        const bool conv = rr <= stop2;
        if (conv || t >= max_iter) {
            if (lead) {
                finish(st, conv ? 1 : 0, t, rr);
                st->halt_a[par] = 1;
            }
            return;
        }
        beta = rz / rz_old;
    }
    const double pn = __builtin_fma(beta, pi, zi);
    const double apn = __builtin_fma(beta, aq, az);
    if (i < n) {
        p_new[i] = pn;
        ap[i] = apn;
    }
    const double pap = wg_sum(i < n ? pn * apn : 0.0);
    if (threadIdx.x == 0) part[blockIdx.x] = pap;
    if (lead) {
        st->t[par] = t;
        st->rz[par] = rz;
        st->rr[par] = rr;
        if (t == 0) {
            st->b2 = rr;
            st->stop2 = st->rtol2 * rr;
        }
        st->halt_a[par] = 0;
    }
}

template <class Op>
__global__ __launch_bounds__(kPcgThreads) void pcg_step_b(const Op op, const double* __restrict__ dinv, double* x,
                                                          double* r, double* z, const double* __restrict__ p,
                                                          const double* __restrict__ ap, double* part, PcgState* st,
                                                          int par)
{
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const long long n = op.n(), np = gridDim.x;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    const long long ii = i < n ? i : n - 1;
    const int halted = st->halt_a[par];
    const double rz_t = st->rz[par];
    double pap = part_sum(part, np);
    const double xi = x[ii], pi = p[ii], ri = r[ii], api = ap[ii], di = dinv[ii];
    if (halted) {
        if (lead) st->halt_b[par] = 1;
        return;
    }
    pap = wg_sum(pap);
    if (!(pap > 0.0)) {  // non-positive curvature: stop at the last iterate
        if (lead) {
            finish(st, 2, st->t[par], st->rr[par]);
            st->halt_b[par] = 1;
        }
        return;
    }
    const double alpha = rz_t / pap;
    double v[2] = {0.0, 0.0};
    if (i < n) {
        x[i] = xi + alpha * pi;
        const double rn = ri - alpha * api;
        r[i] = rn;
        const double zn = di * rn;
        z[i] = zn;
        v[0] = rn * rn;
        v[1] = rn * zn;
    }
    wg_sums<2>(v);
    if (threadIdx.x == 0) {
        part_of(part, np, par, 0)[blockIdx.x] = v[0];
        part_of(part, np, par, 1)[blockIdx.x] = v[1];
    }
    if (lead) st->halt_b[par] = 0;
}

// ------------------------------------------------------------ kOneReduction

template <class Op>
__global__ __launch_bounds__(kPcgThreads) void cg1_diag(const Op op, double* dinv)
{
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    if (i < op.n()) dinv[i] = op.inv_diag(i);
}

// x = 0, r = b, u = M^-1 b, w = A u, p = s = 0; the partials of iteration 0
// in the parity-1 slots
template <class Op>
__global__ __launch_bounds__(kPcgThreads) void cg1_start(const Op op, const double* __restrict__ b, double* x,
                                                         double* r, const double* __restrict__ dinv, double* p,
                                                         double* u, double* w, double* s, double* part, PcgState* st,
                                                         double rtol2, int max_iter)
{
    const long long n = op.n(), np = gridDim.x;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    const long long ii = i < n ? i : n - 1;
    double au, adw, ads;  // A D b: w_0 = A u_0 with u_0 = D b
    op.row3(ii, b, b, b, dinv, au, adw, ads);
    double v[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const double bi = b[i], ui = op.dinv_at(dinv, i) * bi;
        x[i] = 0.0;
        r[i] = bi;
        p[i] = 0.0;
        s[i] = 0.0;
        u[i] = ui;
        w[i] = adw;
        v[0] = bi * ui;
        v[1] = adw * ui;
        v[2] = bi * bi;
    }
    wg_sums<3>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) part_of(part, np, 1, k)[blockIdx.x] = v[k];
        if (blockIdx.x == 0) reset_state(st, rtol2, max_iter);
    }
}

template <class Op>
__global__ __launch_bounds__(kPcgThreads) void cg1_iter(const Op op, double* x, double* r,
                                                        const double* __restrict__ dinv, double* p,
                                                        const double* __restrict__ u_old,
                                                        const double* __restrict__ w_old,
                                                        const double* __restrict__ s_old, double* u_new, double* w_new,
                                                        double* s_new, double* part, PcgState* st, int par)
{
    const int q = par ^ 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const long long n = op.n(), np = gridDim.x;
    const long long i = (long long)blockIdx.x * kPcgThreads + threadIdx.x;
    const long long ii = i < n ? i : n - 1;
    const int halted = st->halt_a[q];
    const int t = st->t[q] + 1;
    const double alpha_prev = st->alpha[q], gamma_prev = st->gamma[q], stop2 = st->stop2;
    const int max_iter = st->max_iter;
    double v[3] = {part_sum(part_of(part, np, q, 0), np), part_sum(part_of(part, np, q, 1), np),
                   part_sum(part_of(part, np, q, 2), np)};
    // A u, A D w and A D s at the iteration-t values: A u_(t+1) = A u - alpha (A D w + beta A D s)
    double au, adw, ads;
    op.row3(ii, u_old, w_old, s_old, dinv, au, adw, ads);
    const double xi = x[ii], ri = r[ii], pi = p[ii], ui = u_old[ii], wi = w_old[ii], si = s_old[ii], di = op.dinv_at(dinv, ii);
    if (halted) {
        if (lead) st->halt_a[par] = 1;
        return;
    }
    wg_sums<3>(v);
    const double gamma = v[0], delta = v[1], rr = v[2];
    double beta = 0.0, pap = delta;
    if (t == 0) {
        if (rr == 0.0 || max_iter == 0) {  // b = 0, or no iteration allowed
            if (lead) {
                st->b2 = rr;
                finish(st, rr == 0.0 ? 3 : 0, 0, rr);
                st->halt_a[par] = 1;
            }
            return;
        }
    } else {
        const bool conv = rr <= stop2;
        if (conv || t >= max_iter) {
            if (lead) {
                finish(st, conv ? 1 : 0, t, rr);
                st->halt_a[par] = 1;
            }
            return;
        }
        beta = gamma / gamma_prev;
        pap = delta - beta * gamma / alpha_prev;
    }
    if (!(pap > 0.0)) {  // non-positive curvature: stop at the last iterate
        if (lead) {
            if (t == 0) st->b2 = rr;
            finish(st, 2, t, rr);
            st->halt_a[par] = 1;
        }
        return;
    }
    const double alpha = gamma / pap;
    const double pn = __builtin_fma(beta, pi, ui);
    const double sn = __builtin_fma(beta, si, wi);
    const double un = __builtin_fma(-alpha, di * sn, ui);
    const double rn = __builtin_fma(-alpha, sn, ri);
    const double wn = __builtin_fma(-alpha, __builtin_fma(beta, ads, adw), au);
    double o[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        x[i] = __builtin_fma(alpha, pn, xi);
        r[i] = rn;
        p[i] = pn;
        s_new[i] = sn;
        u_new[i] = un;
        w_new[i] = wn;
        o[0] = rn * un;
        o[1] = wn * un;
        o[2] = rn * rn;
    }
    wg_sums<3>(o);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 3; ++k) part_of(part, np, par, k)[blockIdx.x] = o[k];
    if (lead) {
        st->t[par] = t;
        st->rr[par] = rr;  // r.r entering iteration t (the host sizes the last chunks from it)
        st->alpha[par] = alpha;
        st->gamma[par] = gamma;
        if (t == 0) {
            st->b2 = rr;
            st->stop2 = st->rtol2 * rr;
        }
        st->halt_a[par] = 0;
    }
}

// ------------------------------------------------------------ drivers

template <class Op> struct Launcher {
    const Op& op;
    unsigned grid;
    double* x;
    PcgWork& w;
    PcgRecurrence rec;
    PcgState* st() const { return reinterpret_cast<PcgState*>(w.state); }
    void start(const double* b, double rtol2, int max_iter, hipStream_t s) const
    {
        if (rec == PcgRecurrence::kStandard) {
            pcg_start<Op><<<grid, kPcgThreads, 0, s>>>(op, b, x, w.r, w.u[0], w.dinv, w.part, st(), rtol2, max_iter);
        } else {
            if (Op::kDiagTable) cg1_diag<Op><<<grid, kPcgThreads, 0, s>>>(op, w.dinv);
            cg1_start<Op><<<grid, kPcgThreads, 0, s>>>(op, b, x, w.r, w.dinv, w.p, w.u[1], w.w[1], w.s[1], w.part,
                                                       st(), rtol2, max_iter);
        }
    }
    // one iteration of parity par; check_only: the test that follows the last one
    void iteration(int par, hipStream_t s, bool check_only = false) const
    {
        const int q = par ^ 1;
        if (rec == PcgRecurrence::kStandard) {
            // z = u[0], Ap = w[0], p alternates over s[0] / s[1]
            pcg_step_a<Op><<<grid, kPcgThreads, 0, s>>>(op, w.u[0], w.s[par], w.s[q], w.w[0], w.part, st(), par);
            if (!check_only)
                pcg_step_b<Op><<<grid, kPcgThreads, 0, s>>>(op, w.dinv, x, w.r, w.u[0], w.s[par], w.w[0], w.part,
                                                            st(), par);
        } else {
            cg1_iter<Op><<<grid, kPcgThreads, 0, s>>>(op, x, w.r, w.dinv, w.p, w.u[q], w.w[q], w.s[q], w.u[par],
                                                      w.w[par], w.s[par], w.part, st(), par);
        }
    }
};

// The queued driver's pinned read-back slots (two states) and poll events:
// made once per device and host thread on first use and kept until the
// process ends (a solve would otherwise pay a pinned allocation, two event
// creations and their release, tens of microseconds, on every call). A
// solve returns only once its stream has drained, so solves made one after
// another on a thread share them safely.
struct Poller {
    PcgState* pin = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

Poller& poller()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "pcg device");
    thread_local std::vector<Poller*> by_dev;
    if ((int)by_dev.size() <= dev) by_dev.resize(dev + 1, nullptr);
    if (!by_dev[dev]) {
        Poller p;
        try {
            hip_check(hipHostMalloc((void**)&p.pin, 2 * sizeof(PcgState), hipHostMallocDefault), "pcg pinned status");
            for (auto& e : p.ev) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "pcg event");
        } catch (...) {
            for (auto e : p.ev)
                if (e) (void)hipEventDestroy(e);
            if (p.pin) (void)hipHostFree(p.pin);
            throw;
        }
        by_dev[dev] = new Poller(p);
    }
    return *by_dev[dev];
}

template <class Op>
PcgResult solve(const Op& op, long long n, const double* b, double* x, double rtol, int max_iter, PcgWork& w,
                hipStream_t s, PcgLaunch mode, PcgRecurrence rec)
{
    PcgResult res;
    if (n <= 0) {
        res.status = 3;
        return res;
    }
    const Launcher<Op> L {op, (unsigned)pcg_parts(n), x, w, rec};
    PcgState* st = L.st();
    // iterations are launched kernel by kernel, kPcgChunk at a time
    // (parities 0, 1, 0, ...); after each chunk the host reads the device
    // state: synchronously when nothing may be allocated, else into pinned
    // memory with one chunk of look-ahead (chunk k's state is read once
    // chunk k + 1 is queued, so the device never idles on the poll). Kernels
    // launched after the solve ended exit at once but still cost a launch,
    // so once two polls show the residual's decay rate, the next chunk is
    // cut to the iterations the rate predicts are left (plus two), the
    // queued ones counted. Replaying the chunks as captured hipGraphs
    // measured slower on the MI355X (33^3 lattice, 126 iterations: 1.27
    // against 1.09 ms).
    PcgState* pin = nullptr;
    hipEvent_t* ev = nullptr;
    if (mode == PcgLaunch::kQueued) {
        Poller& P = poller();
        pin = P.pin;
        ev = P.ev;
    }
    bool halted = false;
    PcgState h {};
    {
        L.start(b, rtol * rtol, max_iter, s);
        int t = 0, k = 0, status = -1, m = kPcgChunk;
        int t_a = -1;  // the last two polls' newest iteration and its r.r
        double rr_a = 0.0;
        while (!halted && t < max_iter) {
            m = std::min(m, max_iter - t);
            for (int c = 0; c < m; ++c) L.iteration((t + c) & 1, s);
            t += m;
            if (mode == PcgLaunch::kNoAlloc) {
                hip_check(hipMemcpyAsync(&status, &st->status, sizeof(int), hipMemcpyDeviceToHost, s), "pcg poll");
                hip_check(hipStreamSynchronize(s), "pcg poll");
                halted = status >= 0;
                continue;
            }
            hip_check(hipMemcpyAsync(&pin[k & 1], st, sizeof(PcgState), hipMemcpyDeviceToHost, s), "pcg poll");
            hip_check(hipEventRecord(ev[k & 1], s), "pcg poll");
            m = kPcgChunk;
            if (k > 0) {
                hip_check(hipEventSynchronize(ev[(k - 1) & 1]), "pcg poll");
                h = pin[(k - 1) & 1];
                halted = h.status >= 0;
                const int tb = std::max(h.t[0], h.t[1]);
                const double rr_b = tb >= 0 ? h.rr[tb & 1] : 0.0;
                if (!halted && t_a >= 0 && tb > t_a && rr_b > 0.0 && rr_b < rr_a && h.stop2 > 0.0) {
                    const double rate = std::log(rr_b / rr_a) / (double)(tb - t_a);  // < 0
                    const double left = std::log(h.stop2 / rr_b) / rate;  // iterations after tb
                    const int queued = t - (tb + 1);
                    const int want = (int)std::ceil(std::max(left, 0.0)) + 2 - queued;
                    m = std::max(1, std::min(kPcgChunk, want));
                }
                if (tb > t_a) {
                    t_a = tb;
                    rr_a = rr_b;
                }
            }
            ++k;
        }
        hip_check(hipStreamSynchronize(s), "pcg poll");
        if (!halted && k > 0) {
            h = pin[(k - 1) & 1];
            halted = h.status >= 0;
        }
    }
    // A halted state no longer changes (later iteration kernels only mark
    // their parity halted), so the last poll's copy is the result; else the
    // check that follows the last iteration (or the b = 0 / max_iter = 0
    // exits) runs and the state is read once more.
    if (!halted || mode == PcgLaunch::kNoAlloc) {
        if (!halted) L.iteration(max_iter & 1, s, true);
        hip_check(hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s), "pcg result");
        hip_check(hipStreamSynchronize(s), "pcg result");
    }
    res.iters = h.iters;
    res.status = h.status;
    res.relres = h.b2 > 0.0 ? std::sqrt(h.rr_end / h.b2) : 0.0;
    return res;
}

static_assert(sizeof(PcgState) <= kPcgStateDoubles * sizeof(double), "PcgState outgrew its slot");

}  // namespace

void pcg_carve(double* base, long long n, PcgWork& w)
{
    const size_t v = pcg_align256((size_t)std::max<long long>(n, 1));
    double* const vec[9] = {base, base + v, base + 2 * v, base + 3 * v, base + 4 * v,
                                       base + 5 * v, base + 6 * v, base + 7 * v, base + 8 * v};
    w.r = vec[0];
    w.dinv = vec[1];
    w.p = vec[2];
    w.u[0] = vec[3];
    w.u[1] = vec[4];
    w.w[0] = vec[5];
    w.w[1] = vec[6];
    w.s[0] = vec[7];
    w.s[1] = vec[8];
    w.part = base + 9 * v;
    w.state = w.part + pcg_align256(7 * (size_t)std::max<long long>(pcg_parts(n), 1));
}

PcgResult pcg_solve_csr(const CsrView& A, const double* b, double* x, double rtol, int max_iter, PcgWork& w,
                        hipStream_t s, PcgLaunch mode, PcgRecurrence rec)
{
    return solve(CsrOp {A}, A.n, b, x, rtol, max_iter, w, s, mode, rec);
}

PcgResult pcg_solve_lattice(const LatticeView& A, const double* b, double* x, double rtol, int max_iter,
                            PcgWork& w, hipStream_t s, PcgLaunch mode, PcgRecurrence rec)
{
    return solve(LatticeOp {A}, A.n, b, x, rtol, max_iter, w, s, mode, rec);
}

}  // namespace eggmi
