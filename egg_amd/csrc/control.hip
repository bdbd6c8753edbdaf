// Kernels of the control-net reduced Gauss-Newton smoother (gfx950): map
// evaluation, per-sample metric derivatives, node gathers, pullbacks and the
// FP64 vector algebra of its PCG (reference src/control_net.hpp:486-828,
// :1037-1090). Every gather runs in a fixed order (deterministic results).
#include "control.hpp"
#include "tmop_math.hpp"

#include <algorithm>

namespace eggmi
{

namespace
{

constexpr int kThreads = 256;
unsigned blocks_for(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Hyper-dual number v + a e1 + b e2 + ab e1 e2 (e1^2 = e2^2 = 0): the e1 e2
// part of f(x + e1 u + e2 w) is the exact second derivative of f along u
// and w, the e1 part the first derivative along u.
struct HD {
    double v, a, b, ab;
    __host__ __device__ HD(double x = 0.0) : v(x), a(0.0), b(0.0), ab(0.0) {}
    __host__ __device__ HD(double x, double p, double q, double r) : v(x), a(p), b(q), ab(r) {}
};
__host__ __device__ inline HD operator+(HD x, HD y) { return {x.v + y.v, x.a + y.a, x.b + y.b, x.ab + y.ab}; }
__host__ __device__ inline HD operator-(HD x, HD y) { return {x.v - y.v, x.a - y.a, x.b - y.b, x.ab - y.ab}; }
__host__ __device__ inline HD operator-(HD x) { return {-x.v, -x.a, -x.b, -x.ab}; }
__host__ __device__ inline HD operator*(HD x, HD y)
{
    return {x.v * y.v, x.v * y.a + x.a * y.v, x.v * y.b + x.b * y.v, x.v * y.ab + x.a * y.b + x.b * y.a + x.ab * y.v};
}
__host__ __device__ inline HD operator/(HD x, HD y)
{
    // 1/y: f' = -1/v^2, f'' = 2/v^3
    const double r = 1.0 / y.v, r2 = r * r;
    const HD inv {r, -y.a * r2, -y.b * r2, -y.ab * r2 + 2.0 * y.a * y.b * r2 * r};
    return x * inv;
}
__host__ __device__ inline HD& operator+=(HD& x, HD y) { return x = x + y; }
__host__ __device__ inline HD& operator-=(HD& x, HD y) { return x = x - y; }

template <int D> constexpr int kV = D * D;
template <int D> constexpr int kNHs = kV<D> * (kV<D> + 1) / 2;

// index of (q, r), q <= r, in the upper triangle (row-major) of a V x V matrix
template <int D> __device__ __forceinline__ int sym_index(int q, int r)
{
    constexpr int V = kV<D>;
    if (q > r) {
        const int t = q;
        q = r;
        r = t;
    }
    return q * V - q * (q - 1) / 2 + (r - q);
}

template <int D, int KIND> __device__ __forceinline__ HD metric_value(const HD* t)
{
    if constexpr (D == 2)
        return mu2_value<KIND, HD>(t[0], t[1], t[2], t[3], HD(0.0));
    else
        return mu3_value<KIND, HD>(t, HD(0.0));
}

// logical index of block-local node n (row-major, last axis fastest)
template <int D> __device__ __forceinline__ void unravel_node(const CtrlGeom& g, long long n, int* I)
{
    if constexpr (D == 2) {
        I[0] = int(n / g.n[1]);
        I[1] = int(n - (long long)I[0] * g.n[1]);
        I[2] = 0;
    } else {
        const long long s12 = (long long)g.n[1] * g.n[2];
        I[0] = int(n / s12);
        const long long r = n - (long long)I[0] * s12;
        I[1] = int(r / g.n[2]);
        I[2] = int(r - (long long)I[1] * g.n[2]);
    }
}

template <int D> __device__ __forceinline__ long long node_slot(const CtrlGeom& g, const int* I)
{
    long long s = g.bd.base;
#pragma unroll
    for (int k = 0; k < D; ++k) s += (long long)(I[k] + 1) * g.bd.st[k];
    return s;
}

template <int D> __device__ __forceinline__ long long node_local(const CtrlGeom& g, const int* I)
{
    return D == 2 ? (long long)I[0] * g.n[1] + I[1] : ((long long)I[0] * g.n[1] + I[1]) * g.n[2] + I[2];
}

// ---------------------------------------------------------------- map

template <int D, bool DIR>
__global__ __launch_bounds__(kThreads) void ctrl_eval_kernel(const CtrlGeom g, const double* src, const double* b,
                                                             double* out_dir, float* out_plane)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= g.n_nodes) return;
    int I[3];
    unravel_node<D>(g, n, I);
    int a0[3];
#pragma unroll
    for (int k = 0; k < D; ++k) a0[k] = g.off[k][I[k]];
    double acc[D] = {};
    for (int t0 = 0; t0 < g.ks; ++t0) {
        const double w0 = g.wt[0][(long long)I[0] * g.ks + t0];
        if (w0 == 0.0) continue;
        for (int t1 = 0; t1 < g.ks; ++t1) {
            const double w1 = w0 * g.wt[1][(long long)I[1] * g.ks + t1];
            if (w1 == 0.0) continue;
            for (int t2 = 0; t2 < (D == 3 ? g.ks : 1); ++t2) {
                double w = w1;
                long long a = (long long)(a0[0] + t0) * g.c[1] + (a0[1] + t1);
                if constexpr (D == 3) {
                    w *= g.wt[2][(long long)I[2] * g.ks + t2];
                    a = a * g.c[2] + (a0[2] + t2);
                }
                if (w == 0.0) continue;
                if constexpr (DIR) {
                    const int f = g.l2f[a];
                    if (f < 0) continue;
#pragma unroll
                    for (int c = 0; c < D; ++c) acc[c] += w * src[(long long)f * D + c];
                } else {
#pragma unroll
                    for (int c = 0; c < D; ++c) acc[c] += w * src[a * D + c];
                }
            }
        }
    }
    if constexpr (!DIR) {
#pragma unroll
        for (int c = 0; c < D; ++c) acc[c] += b ? b[n * D + c] : 0.0;
    }
    if (out_dir)
#pragma unroll
        for (int c = 0; c < D; ++c) out_dir[n * D + c] = acc[c];
    if (out_plane) {
        const long long s = node_slot<D>(g, I);
#pragma unroll
        for (int c = 0; c < D; ++c) out_plane[c * g.plane + s] = (float)acc[c];
    }
}

// ---------------------------------------------------------------- samples

// corner b of cell I0: logical index of corner `bits` (axis 0 the MSB)
template <int D> __device__ __forceinline__ void corner_of(const int* I0, int bits, int* I)
{
#pragma unroll
    for (int k = 0; k < D; ++k) I[k] = I0[k] + ((bits >> (D - 1 - k)) & 1);
    if constexpr (D == 2) I[2] = 0;
}

// cell origin of block-local cell `cell` (row-major over n - 1)
template <int D> __device__ __forceinline__ void cell_origin(const CtrlGeom& g, long long cell, int* I0)
{
    I0[2] = 0;
    if constexpr (D == 2) {
        I0[0] = int(cell / (g.n[1] - 1));
        I0[1] = int(cell - (long long)I0[0] * (g.n[1] - 1));
    } else {
        const long long s12 = (long long)(g.n[1] - 1) * (g.n[2] - 1);
        I0[0] = int(cell / s12);
        const long long r = cell - (long long)I0[0] * s12;
        I0[1] = int(r / (g.n[2] - 1));
        I0[2] = int(r - (long long)I0[1] * (g.n[2] - 1));
    }
}

// A of sample (cell I0, corner b) at xb + a dir (dir null: xb), row-major
template <int D>
__device__ __forceinline__ void sample_A(const CtrlGeom& g, const int* I0, int b, const double* xb, const double* dir,
                                         double a, double* t)
{
    int Ic[3];
    corner_of<D>(I0, b, Ic);
    const long long nc = node_local<D>(g, Ic);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        int In[3];
        corner_of<D>(I0, b ^ (1 << (D - 1 - k)), In);
        const long long nn = node_local<D>(g, In);
        const double sk = ((b >> (D - 1 - k)) & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int i = 0; i < D; ++i) {
            double pn = xb[nn * D + i], pc = xb[nc * D + i];
            if (dir) {
                pn += a * dir[nn * D + i];
                pc += a * dir[nc * D + i];
            }
            t[i * D + k] = sk * (pn - pc);
        }
    }
}

template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void ctrl_samples_kernel(const CtrlGeom g, const double* xb, double* gs,
                                                                double* hs)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= g.n_cells * NC) return;
    const long long cell = p / NC;
    const int b = int(p - cell * NC);
    int I0[3];
    cell_origin<D>(g, cell, I0);
    double t[V];
    sample_A<D>(g, I0, b, xb, nullptr, 0.0, t);
    // component-major (entry e of sample p at e * ns + p): coalesced
    const long long ns = g.n_cells * NC;
    int e = 0;
    for (int q = 0; q < V; ++q)
        for (int r = q; r < V; ++r) {
            HD T[V];
#pragma unroll
            for (int j = 0; j < V; ++j) T[j] = HD(t[j]);
            T[q].a = 1.0;
            T[r].b = 1.0;
            const HD m = metric_value<D, KIND>(T);
            hs[(long long)(e++) * ns + p] = m.ab;
            if (q == r) gs[(long long)q * ns + p] = m.a;
        }
}

// z_p = H_p dA_p per sample, dA_p from the block-local direction field
template <int D>
__global__ __launch_bounds__(kThreads) void ctrl_sample_hv_kernel(const CtrlGeom g, const double* hs,
                                                                  const double* dir, double* zs)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= g.n_cells * NC) return;
    const long long cell = p / NC;
    const int b = int(p - cell * NC);
    int I0[3];
    cell_origin<D>(g, cell, I0);
    double dA[V];
    sample_A<D>(g, I0, b, dir, nullptr, 0.0, dA);
    const long long ns = g.n_cells * NC;
    double H[kNHs<D>];
#pragma unroll
    for (int e = 0; e < kNHs<D>; ++e) H[e] = hs[(long long)e * ns + p];
#pragma unroll
    for (int q = 0; q < V; ++q) {
        double y = 0.0;
#pragma unroll
        for (int r = 0; r < V; ++r) y += H[sym_index<D>(q, r)] * dA[r];
        zs[(long long)q * ns + p] = y;
    }
}

// FP64 cell energy and min det of the block at two step lengths: each
// workgroup walks a contiguous run of samples with stride kThreads, then a
// fixed tree per workgroup and a one-workgroup fold over the partials
template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void ctrl_energy_kernel(const CtrlGeom g, const double* xb, const double* dir,
                                                               double a0, double a1, double* partials)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long ns = g.n_cells * NC;
    const long long per = (ns + gridDim.x - 1) / gridDim.x;
    const long long lo = (long long)blockIdx.x * per, hi = min(ns, lo + per);
    double e0 = 0.0, e1 = 0.0, m0 = __builtin_huge_val(), m1 = __builtin_huge_val();
    for (long long p = lo + threadIdx.x; p < hi; p += kThreads) {
        const long long cell = p / NC;
        const int b = int(p - cell * NC);
        int I0[3];
        cell_origin<D>(g, cell, I0);
        double t[V];
        sample_A<D>(g, I0, b, xb, dir, a0, t);
        if constexpr (D == 2) {
            e0 += mu2_value<KIND, double>(t[0], t[1], t[2], t[3], 0.0);
            m0 = fmin(m0, t[0] * t[3] - t[1] * t[2]);
        } else {
            e0 += mu3_value<KIND, double>(t, 0.0);
            m0 = fmin(m0, det3(t));
        }
        sample_A<D>(g, I0, b, xb, dir, a1, t);
        if constexpr (D == 2) {
            e1 += mu2_value<KIND, double>(t[0], t[1], t[2], t[3], 0.0);
            m1 = fmin(m1, t[0] * t[3] - t[1] * t[2]);
        } else {
            e1 += mu3_value<KIND, double>(t, 0.0);
            m1 = fmin(m1, det3(t));
        }
    }
    __shared__ double sh[4][kThreads];
    sh[0][threadIdx.x] = e0;
    sh[1][threadIdx.x] = m0;
    sh[2][threadIdx.x] = e1;
    sh[3][threadIdx.x] = m1;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] = fmin(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
            sh[3][threadIdx.x] = fmin(sh[3][threadIdx.x], sh[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) partials[blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(kThreads) void ctrl_energy_finish(const double* partials, int n, double* out)
{
    // n <= kCtrlParts = kThreads: one partial per thread, a fixed tree
    __shared__ double sh[4][kThreads];
    const bool in = (int)threadIdx.x < n;
    sh[0][threadIdx.x] = in ? partials[4 * threadIdx.x] : 0.0;
    sh[1][threadIdx.x] = in ? partials[4 * threadIdx.x + 1] : __builtin_huge_val();
    sh[2][threadIdx.x] = in ? partials[4 * threadIdx.x + 2] : 0.0;
    sh[3][threadIdx.x] = in ? partials[4 * threadIdx.x + 3] : __builtin_huge_val();
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] = fmin(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
            sh[2][threadIdx.x] += sh[2][threadIdx.x + w];
            sh[3][threadIdx.x] = fmin(sh[3][threadIdx.x], sh[3][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------- node gathers

template <int D, int MODE>
__global__ __launch_bounds__(kThreads) void ctrl_node_kernel(const CtrlGeom g, const double* gs, const double* hs,
                                                             const double* dir, double* out)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= g.n_nodes) return;
    int I[3];
    unravel_node<D>(g, n, I);
    const long long ns = g.n_cells * NC;
    double acc[MODE == 2 ? V : D] = {};
    // incident cells: origin I - o for o in {0, 1}^D, the node at corner o
    for (int o = 0; o < NC; ++o) {
        int I0[3] = {0, 0, 0};
        bool ok = true;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            I0[k] = I[k] - ((o >> (D - 1 - k)) & 1);
            ok = ok && I0[k] >= 0 && I0[k] <= g.n[k] - 2;
        }
        if (!ok) continue;
        const long long cell = D == 2 ? (long long)I0[0] * (g.n[1] - 1) + I0[1]
                                      : ((long long)I0[0] * (g.n[1] - 1) + I0[1]) * (g.n[2] - 1) + I0[2];
        for (int b = 0; b < NC; ++b) {
            // the node's role in sample (cell, b): corner, axis-k neighbour or none
            double u[D];
            const int x = o ^ b;
            if (x == 0) {
#pragma unroll
                for (int k = 0; k < D; ++k) u[k] = ((b >> (D - 1 - k)) & 1) ? 1.0 : -1.0;  // -s_k
            } else if ((x & (x - 1)) == 0) {
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    const int bit = 1 << (D - 1 - k);
                    u[k] = (x == bit) ? (((b & bit) ? -1.0 : 1.0)) : 0.0;  // s_k on axis k
                }
            } else {
                continue;
            }
            const long long p = cell * NC + b;
            if constexpr (MODE == 0) {
#pragma unroll
                for (int i = 0; i < D; ++i)
#pragma unroll
                    for (int k = 0; k < D; ++k) acc[i] += u[k] * gs[(long long)(i * D + k) * ns + p];
            } else {

#pragma unroll
                for (int i = 0; i < D; ++i)
#pragma unroll
                    for (int j = 0; j < D; ++j) {
                        double h = 0.0;
#pragma unroll
                        for (int k = 0; k < D; ++k)
#pragma unroll
                            for (int l = 0; l < D; ++l)
                                h += u[k] * u[l] * hs[(long long)sym_index<D>(i * D + k, j * D + l) * ns + p];
                        acc[i * D + j] += h;
                    }
            }
        }
    }
    constexpr int NO = MODE == 2 ? V : D;
#pragma unroll
    for (int c = 0; c < NO; ++c) out[n * NO + c] = acc[c];
}

// ---------------------------------------------------------------- pullback

// one workgroup per free control: its threads split the tensor product of
// the per-axis support lists, then fold in a fixed tree
template <int D, int MODE>
__global__ __launch_bounds__(kThreads) void ctrl_pullback_kernel(const CtrlGeom g, const double* y, double* out)
{
    constexpr int NO = MODE == 1 ? D * D : D;
    const int f = blockIdx.x;
    long long rem = g.free_idx[f];
    int a[3] = {0, 0, 0};
    for (int k = D - 1; k >= 0; --k) {
        a[k] = int(rem % g.c[k]);
        rem /= g.c[k];
    }
    int lo[3] = {0, 0, 0}, len[3] = {1, 1, 1};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        lo[k] = g.sup_off[k][a[k]];
        len[k] = g.sup_off[k][a[k] + 1] - lo[k];
    }
    const int total = len[0] * len[1] * len[2];
    double acc[NO] = {};
    for (int t = threadIdx.x; t < total; t += kThreads) {
        const int q2 = t % len[2], q1 = (t / len[2]) % len[1], q0 = t / (len[2] * len[1]);
        double w = g.sup_w[0][lo[0] + q0] * g.sup_w[1][lo[1] + q1];
        long long n = (long long)g.sup_node[0][lo[0] + q0] * g.n[1] + g.sup_node[1][lo[1] + q1];
        if constexpr (D == 3) {
            w *= g.sup_w[2][lo[2] + q2];
            n = n * g.n[2] + g.sup_node[2][lo[2] + q2];
        }
        const double ww = MODE == 1 ? w * w : w;
#pragma unroll
        for (int c = 0; c < NO; ++c) acc[c] += ww * y[n * NO + c];
    }
    __shared__ double sh[NO][kThreads];
#pragma unroll
    for (int c = 0; c < NO; ++c) sh[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int c = 0; c < NO; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < NO) out[(long long)f * NO + threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------- vectors

template <int D>
__global__ __launch_bounds__(kThreads) void ctrl_precond_kernel(int n_free, const double* P, double lam,
                                                                const double* r, double* z)
{
    const int f = blockIdx.x * kThreads + threadIdx.x;
    if (f >= n_free) return;
    double A[D * D], rhs[D];
#pragma unroll
    for (int i = 0; i < D * D; ++i) A[i] = P[(long long)f * D * D + i];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        A[i * D + i] += lam;
        rhs[i] = r[(long long)f * D + i];
    }
    // Gaussian elimination without pivoting (the blocks are sums of w^2
    // times node Hessians plus lam I; a zero pivot leaves the component as is)
    double x[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int j = i + 1; j < D; ++j) {
            const double m = A[i * D + i] != 0.0 ? A[j * D + i] / A[i * D + i] : 0.0;
#pragma unroll
            for (int k = i; k < D; ++k) A[j * D + k] -= m * A[i * D + k];
            rhs[j] -= m * rhs[i];
        }
    }
#pragma unroll
    for (int i = D - 1; i >= 0; --i) {
        double s = rhs[i];
#pragma unroll
        for (int k = i + 1; k < D; ++k) s -= A[i * D + k] * x[k];
        x[i] = A[i * D + i] != 0.0 ? s / A[i * D + i] : s;
    }
#pragma unroll
    for (int i = 0; i < D; ++i) z[(long long)f * D + i] = x[i];
}

// one workgroup of 1024 threads: strided partial sums, then a fixed tree
template <int OP>
__global__ __launch_bounds__(1024) void ctrl_reduce_kernel(const double* x, const double* y, long long n, int dim,
                                                           double* out)
{
    __shared__ double sh[1024];
    double acc = 0.0;
    for (long long i = threadIdx.x; i < n; i += 1024) {
        if constexpr (OP == 0)
            acc += x[i] * y[i];
        else if constexpr (OP == 1)
            acc = fmax(acc, fabs(x[i]));
        else
            for (int k = 0; k < dim; ++k) acc += x[i * dim * dim + k * dim + k];
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            sh[threadIdx.x] = OP == 1 ? fmax(sh[threadIdx.x], sh[threadIdx.x + w]) : sh[threadIdx.x] + sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

__global__ __launch_bounds__(kThreads) void ctrl_axpy_kernel(long long n, double a, const double* x, double* y,
                                                             bool xpay)
{
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    y[i] = xpay ? x[i] + a * y[i] : y[i] + a * x[i];
}

__global__ __launch_bounds__(kThreads) void ctrl_neg_kernel(long long n, const double* x, double* y)
{
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) y[i] = -x[i];
}

__global__ __launch_bounds__(kThreads) void ctrl_scatter_free_kernel(int n_free, int dim, const int* free_idx,
                                                                     double a, const double* v, double* C)
{
    const int f = blockIdx.x * kThreads + threadIdx.x;
    if (f >= n_free) return;
    for (int c = 0; c < dim; ++c) C[(long long)free_idx[f] * dim + c] += a * v[(long long)f * dim + c];
}

}  // namespace

void launch_ctrl_eval_state(const CtrlGeom& g, const double* C, const double* b, double* out, float* out_plane,
                            hipStream_t s)
{
    if (g.n_nodes <= 0) return;
    if (g.dim == 2)
        ctrl_eval_kernel<2, false><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, C, b, out, out_plane);
    else
        ctrl_eval_kernel<3, false><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, C, b, out, out_plane);
}

void launch_ctrl_energy(const CtrlGeom& g, int kind, const double* xb, const double* dir, double a0, double a1,
                        double* partials, double* out, hipStream_t s)
{
    const long long ns = g.n_cells * (1LL << g.dim);
    const int nb = (int)std::max<long long>(1, std::min<long long>(kCtrlParts, (ns + kThreads - 1) / kThreads));
#define EGGMI_CE(D, K) ctrl_energy_kernel<D, K><<<nb, kThreads, 0, s>>>(g, xb, dir, a0, a1, partials)
    if (g.dim == 2) {
        if (kind == kShapeSize)
            EGGMI_CE(2, kShapeSize);
        else
            EGGMI_CE(2, kShape);
    } else {
        if (kind == kShapeSize)
            EGGMI_CE(3, kShapeSize);
        else
            EGGMI_CE(3, kShape);
    }
#undef EGGMI_CE
    ctrl_energy_finish<<<1, kThreads, 0, s>>>(partials, nb, out);
}

void launch_ctrl_eval_dir(const CtrlGeom& g, const double* v, double* out_dir, float* out_plane, hipStream_t s)
{
    if (g.n_nodes <= 0) return;
    if (g.dim == 2)
        ctrl_eval_kernel<2, true><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, v, nullptr, out_dir, out_plane);
    else
        ctrl_eval_kernel<3, true><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, v, nullptr, out_dir, out_plane);
}

void launch_ctrl_samples(const CtrlGeom& g, int kind, const double* x, double* gs, double* hs, hipStream_t s)
{
    const long long n = g.n_cells * (1LL << g.dim);
    if (n <= 0) return;
    if (g.dim == 2) {
        if (kind == kShapeSize)
            ctrl_samples_kernel<2, kShapeSize><<<blocks_for(n), kThreads, 0, s>>>(g, x, gs, hs);
        else
            ctrl_samples_kernel<2, kShape><<<blocks_for(n), kThreads, 0, s>>>(g, x, gs, hs);
    } else {
        if (kind == kShapeSize)
            ctrl_samples_kernel<3, kShapeSize><<<blocks_for(n), kThreads, 0, s>>>(g, x, gs, hs);
        else
            ctrl_samples_kernel<3, kShape><<<blocks_for(n), kThreads, 0, s>>>(g, x, gs, hs);
    }
}

void launch_ctrl_sample_hv(const CtrlGeom& g, const double* hs, const double* dir, double* zs, hipStream_t s)
{
    const long long n = g.n_cells * (1LL << g.dim);
    if (n <= 0) return;
    if (g.dim == 2)
        ctrl_sample_hv_kernel<2><<<blocks_for(n), kThreads, 0, s>>>(g, hs, dir, zs);
    else
        ctrl_sample_hv_kernel<3><<<blocks_for(n), kThreads, 0, s>>>(g, hs, dir, zs);
}

void launch_ctrl_node(const CtrlGeom& g, int mode, const double* gs, const double* hs, const double* dir,
                      double* out, hipStream_t s)
{
    if (g.n_nodes <= 0) return;
    const unsigned nb = blocks_for(g.n_nodes);
#define EGGMI_CN(D, M) ctrl_node_kernel<D, M><<<nb, kThreads, 0, s>>>(g, gs, hs, dir, out)
    if (g.dim == 2) {
        if (mode == 0)
            EGGMI_CN(2, 0);
        else
            EGGMI_CN(2, 2);
    } else {
        if (mode == 0)
            EGGMI_CN(3, 0);
        else
            EGGMI_CN(3, 2);
    }
#undef EGGMI_CN
}

void launch_ctrl_pullback(const CtrlGeom& g, int mode, const double* y, double* out, hipStream_t s)
{
    if (g.n_free <= 0) return;
    const unsigned nb = (unsigned)g.n_free;  // one workgroup per free control
    if (g.dim == 2) {
        if (mode == 0)
            ctrl_pullback_kernel<2, 0><<<nb, kThreads, 0, s>>>(g, y, out);
        else
            ctrl_pullback_kernel<2, 1><<<nb, kThreads, 0, s>>>(g, y, out);
    } else {
        if (mode == 0)
            ctrl_pullback_kernel<3, 0><<<nb, kThreads, 0, s>>>(g, y, out);
        else
            ctrl_pullback_kernel<3, 1><<<nb, kThreads, 0, s>>>(g, y, out);
    }
}

void launch_ctrl_precond(const CtrlGeom& g, const double* P, double lam, const double* r, double* z, hipStream_t s)
{
    if (g.n_free <= 0) return;
    if (g.dim == 2)
        ctrl_precond_kernel<2><<<blocks_for(g.n_free), kThreads, 0, s>>>(g.n_free, P, lam, r, z);
    else
        ctrl_precond_kernel<3><<<blocks_for(g.n_free), kThreads, 0, s>>>(g.n_free, P, lam, r, z);
}

void launch_ctrl_reduce(int op, const double* x, const double* y, long long n, int dim, double* out, hipStream_t s)
{
    if (op == 0)
        ctrl_reduce_kernel<0><<<1, 1024, 0, s>>>(x, y, n, dim, out);
    else if (op == 1)
        ctrl_reduce_kernel<1><<<1, 1024, 0, s>>>(x, y, n, dim, out);
    else
        ctrl_reduce_kernel<2><<<1, 1024, 0, s>>>(x, y, n, dim, out);
}

void launch_ctrl_axpy(long long n, double a, const double* x, double* y, bool xpay, hipStream_t s)
{
    if (n > 0) ctrl_axpy_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, a, x, y, xpay);
}

void launch_ctrl_neg(long long n, const double* x, double* y, hipStream_t s)
{
    if (n > 0) ctrl_neg_kernel<<<blocks_for(n), kThreads, 0, s>>>(n, x, y);
}

void launch_ctrl_scatter_free(const CtrlGeom& g, double a, const double* v, double* C, hipStream_t s)
{
    if (g.n_free > 0)
        ctrl_scatter_free_kernel<<<blocks_for(g.n_free), kThreads, 0, s>>>(g.n_free, g.dim, g.free_idx, a, v, C);
}

}  // namespace eggmi
