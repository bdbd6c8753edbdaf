// Inter-grid transfer kernels of the FAS multigrid (gfx950), one launch per
// operation over all blocks of a level (the block of a thread's node is
// found by binary search over the per-block node-count prefix):
//   restrict_inject      coarse(I) = fine(I * f) (exact copy)
//   restrict_full_weight tensor (1/4, 1/2, 1/4) stencil on coarsened axes of
//                        the ghost-inclusive fine gradient, at free coarse
//                        nodes (zero elsewhere)
//   prolong_correction   multilinear interpolation of the coarse correction
//                        x_c - rx_c, added at a coarse level's free slots or
//                        written as the fine correction (zero where not free)
//   masked_axpy          x += a c at free slots
//   linesearch_pair      fine energy / min det at x + a0 c and x + a1 c
// (reference src/mg_transfer.hpp:116-341 and src/fas.hpp:174-234).
#include "mg.hpp"

namespace eggmi
{

__device__ __forceinline__ int find_block(const long long* off, int nb, long long i)
{
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// logical index of the r-th node of an n[0] x n[1] (x n[2]) block, row-major
template <int D> __device__ __forceinline__ void unravel(long long r, const int* n, int* I)
{
    if constexpr (D == 2) {
        I[0] = int(r / n[1]);
        I[1] = int(r - (long long)I[0] * n[1]);
    } else {
        const long long s12 = (long long)n[1] * n[2];
        I[0] = int(r / s12);
        r -= (long long)I[0] * s12;
        I[1] = int(r / n[2]);
        I[2] = int(r - (long long)I[1] * n[2]);
    }
}

template <int D> __device__ __forceinline__ long long padded_slot(const BlockDesc& b, const int* P)
{
    long long s = b.base;
#pragma unroll
    for (int k = 0; k < D; ++k) s += (long long)P[k] * b.st[k];
    return s;
}

// Threads enumerate the coarse interior nodes of every block. With a push
// table (p.off), the injected state goes to cx and the two extra coarse
// buffers p.x1 and p.rx, at the node and at every replica slot it is the
// root of; a replica slot's own thread writes nothing there.
template <int D> __global__ __launch_bounds__(256) void restrict_kernel(const TransferArgs a, const float* fx,
                                                                       float* cx, const float* fgrad,
                                                                       float* ctau, const uint8_t* cfree,
                                                                       const RestrictPush p)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int b = find_block(a.off, a.nb, i);
    const BlockDesc fb = a.fine[b], cb = a.coarse[b];
    int I[3], Pc[3], Pf[3];
    unravel<D>(i - a.off[b], cb.n, I);
    int f[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        f[k] = a.factor[b * 3 + k];
        Pc[k] = I[k] + 1;
        Pf[k] = I[k] * f[k] + 1;
    }
    const long long cs = padded_slot<D>(cb, Pc);
    const long long fs = padded_slot<D>(fb, Pf);
    if (cx && p.off) {
        if (!p.is_dst[cs]) {
            float v[D];
#pragma unroll
            for (int k = 0; k < D; ++k) v[k] = fx[k * a.fplane + fs];
            for (int q = p.off[cs] - 1; q < p.off[cs + 1]; ++q) {
                const long long d = q < p.off[cs] ? cs : (long long)p.dst[q];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    cx[k * a.cplane + d] = v[k];
                    p.x1[k * a.cplane + d] = v[k];
                    p.rx[k * a.cplane + d] = v[k];
                }
            }
        }
    } else if (cx) {
#pragma unroll
        for (int k = 0; k < D; ++k) cx[k * a.cplane + cs] = fx[k * a.fplane + fs];
    }
    if (ctau) {
        float acc[D] = {};
        if (cfree[cs]) {
            constexpr int NO = (D == 2) ? 9 : 27;
            for (int e = 0; e < NO; ++e) {
                int rem = e, P[3];
                float w = 1.0f;
                bool ok = true;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const int o = rem % 3 - 1;
                    rem /= 3;
                    if (f[k] == 1) {
                        if (o != 0) ok = false;
                        P[k] = Pf[k];
                    } else {
                        w *= (o == 0) ? 0.5f : 0.25f;
                        P[k] = Pf[k] + o;
                    }
                }
                if (!ok) continue;
                const long long s = padded_slot<D>(fb, P);
#pragma unroll
                for (int k = 0; k < D; ++k) acc[k] += w * fgrad[k * a.fplane + s];
            }
        }
#pragma unroll
        for (int k = 0; k < D; ++k) ctau[k * a.cplane + cs] = acc[k];
    }
}

// Threads enumerate the fine interior nodes of every block. The coarse
// correction is the coarse solve's change cx - crx, interpolated
// multilinearly. ADD: added in place to the fine state at its free slots (a
// coarse level's correction); otherwise written as the fine correction, zero
// at the nodes that are not free (the step the line search scales).
template <int D, bool ADD>
__global__ __launch_bounds__(256) void prolong_kernel(const TransferArgs a, const float* cx, const float* crx,
                                                      float* fout, const uint8_t* ffree, const int* cp_off,
                                                      const int* cp_dst, const uint8_t* is_dst)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const int b = find_block(a.off, a.nb, i);
    const BlockDesc fb = a.fine[b], cb = a.coarse[b];
    int J[3], c[3], r[3], P[3];
    unravel<D>(i - a.off[b], fb.n, J);
    float w0 = 1.0f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int f = a.factor[b * 3 + k];
        c[k] = J[k] / f;
        r[k] = J[k] - c[k] * f;
        if (r[k] != 0) w0 *= 0.5f;
        P[k] = J[k] + 1;
    }
    float acc[D] = {};
    for (int e = 0; e < (1 << D); ++e) {
        int C[3];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            C[k] = c[k] + 1;
            if ((e >> k) & 1) {
                if (r[k] == 0) ok = false;
                C[k] += 1;
            }
        }
        if (!ok) continue;
        const long long s = padded_slot<D>(cb, C);
#pragma unroll
        for (int k = 0; k < D; ++k) acc[k] += w0 * (cx[k * a.cplane + s] - crx[k * a.cplane + s]);
    }
    const long long fs = padded_slot<D>(fb, P);
    const bool fr = ffree[fs] != 0;
    if constexpr (ADD) {
        if (!fr) return;
        float v[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            v[k] = fout[k * a.fplane + fs] + acc[k];
            fout[k * a.fplane + fs] = v[k];
        }
        if (cp_off)  // the DOF's replicas (only free slots are written, so no other thread writes them)
            for (int q = cp_off[fs]; q < cp_off[fs + 1]; ++q) {
                const long long d = cp_dst[q];
#pragma unroll
                for (int k = 0; k < D; ++k) fout[k * a.fplane + d] = v[k];
            }
    } else {
        // with a push table, a replica slot is its root's to write (a relaxed
        // root pushes its value; the others' replicas keep their zero)
        if (is_dst && is_dst[fs]) return;
#pragma unroll
        for (int k = 0; k < D; ++k) fout[k * a.fplane + fs] = fr ? acc[k] : 0.0f;
        if (fr && cp_off)
            for (int q = cp_off[fs]; q < cp_off[fs + 1]; ++q) {
                const long long d = cp_dst[q];
#pragma unroll
                for (int k = 0; k < D; ++k) fout[k * a.fplane + d] = acc[k];
            }
    }
}

__global__ __launch_bounds__(256) void masked_axpy_kernel(float* x, const float* c, float alpha, const uint8_t* free,
                                                          long long n, long long plane, int dim)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !free[i]) return;
    for (int k = 0; k < dim; ++k) x[k * plane + i] = x[k * plane + i] + alpha * c[k * plane + i];
}

// x += (*alpha) c at free slots, the sums pushed into their replica slots
// (push table cp_off / cp_dst, null = none)
__global__ __launch_bounds__(256) void masked_axpy_dev_kernel(float* x, const float* c, const float* alpha,
                                                              const uint8_t* free, long long n, long long plane,
                                                              int dim, const int* cp_off, const int* cp_dst)
{
    const float al = *alpha;
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (al == 0.0f || i >= n || !free[i]) return;
    for (int k = 0; k < dim; ++k) {
        const float v = x[k * plane + i] + al * c[k * plane + i];
        x[k * plane + i] = v;
        if (cp_off)
            for (int q = cp_off[i]; q < cp_off[i + 1]; ++q) x[k * plane + cp_dst[q]] = v;
    }
}

__global__ __launch_bounds__(256) void copy2_kernel(float* __restrict__ a, float* __restrict__ b,
                                                    const float* __restrict__ src, long long n)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v = src[i];
    a[i] = v;
    b[i] = v;
}

static unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

void launch_restrict(int dim, const TransferArgs& a, const float* fx, float* cx, const float* fgrad, float* ctau,
                     const uint8_t* cfree, hipStream_t s, const RestrictPush& push)
{
    if (a.n <= 0) return;
    if (dim == 2)
        restrict_kernel<2><<<nblk(a.n), 256, 0, s>>>(a, fx, cx, fgrad, ctau, cfree, push);
    else
        restrict_kernel<3><<<nblk(a.n), 256, 0, s>>>(a, fx, cx, fgrad, ctau, cfree, push);
}

void launch_prolong(int dim, const TransferArgs& a, const float* cx, const float* crx, float* fout,
                    const uint8_t* ffree, bool add, const int* cp_off, const int* cp_dst, hipStream_t s,
                    const uint8_t* is_dst)
{
    if (a.n <= 0) return;
    if (!add && !is_dst) cp_off = cp_dst = nullptr;  // a pushed correction needs the replica mask
    if (dim == 2) {
        if (add)
            prolong_kernel<2, true><<<nblk(a.n), 256, 0, s>>>(a, cx, crx, fout, ffree, cp_off, cp_dst, nullptr);
        else
            prolong_kernel<2, false><<<nblk(a.n), 256, 0, s>>>(a, cx, crx, fout, ffree, cp_off, cp_dst, is_dst);
    } else {
        if (add)
            prolong_kernel<3, true><<<nblk(a.n), 256, 0, s>>>(a, cx, crx, fout, ffree, cp_off, cp_dst, nullptr);
        else
            prolong_kernel<3, false><<<nblk(a.n), 256, 0, s>>>(a, cx, crx, fout, ffree, cp_off, cp_dst, is_dst);
    }
}

void launch_copy2(float* a, float* b, const float* src, long long n, hipStream_t s)
{
    if (n > 0) copy2_kernel<<<nblk(n), 256, 0, s>>>(a, b, src, n);
}

void launch_masked_axpy_dev(float* x, const float* c, const float* alpha, const uint8_t* free, long long n,
                            long long plane, int dim, hipStream_t s, const int* cp_off, const int* cp_dst)
{
    if (n > 0) masked_axpy_dev_kernel<<<nblk(n), 256, 0, s>>>(x, c, alpha, free, n, plane, dim, cp_off, cp_dst);
}

void launch_masked_axpy(float* x, const float* c, float alpha, const uint8_t* free, long long n, long long plane,
                        int dim, hipStream_t s)
{
    if (n > 0) masked_axpy_kernel<<<nblk(n), 256, 0, s>>>(x, c, alpha, free, n, plane, dim);
}

}  // namespace eggmi
