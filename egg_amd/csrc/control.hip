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

// M src at block-local node I (dim values into acc): FREE_MAP reads a
// free-control vector through l2f (fixed controls 0), otherwise src is the
// whole lattice
template <int D, bool FREE_MAP>
__device__ __forceinline__ void eval_at(const CtrlGeom& g, const int* I, const double* src, double* acc)
{
    int a0[3];
#pragma unroll
    for (int k = 0; k < D; ++k) a0[k] = g.off[k][I[k]];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0.0;
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
                if constexpr (FREE_MAP) {
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
}

template <int D, bool DIR>
__global__ __launch_bounds__(kThreads) void ctrl_eval_kernel(const CtrlGeom g, const double* src, const double* b,
                                                             double* out_dir, float* out_plane)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= g.n_nodes || (g.halt && *g.halt != 0.0)) return;
    int I[3];
    unravel_node<D>(g, n, I);
    double acc[D];
    eval_at<D, DIR>(g, I, src, acc);
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

// sign of eps_(i k m), m the third index of a 3-permutation (i != k)
__device__ __forceinline__ double lc_sign(int i, int k) { return ((k - i + 3) % 3 == 1) ? 1.0 : -1.0; }

// d^2 det / dT_x dT_z of a 3 x 3 row-major matrix m (x = 3 i + j, z = 3 k + l):
// eps_ikm' eps_jln' m_m'n', zero on a shared row or column
__device__ __forceinline__ double d2det3(const double* m, int x, int z)
{
    const int i = x / 3, j = x % 3, k = z / 3, l = z % 3;
    if (i == k || j == l) return 0.0;
    return lc_sign(i, k) * lc_sign(j, l) * m[(3 - i - k) * 3 + (3 - j - l)];
}

// The 3D barrier-phase metrics' gradient and Hessian in closed form:
// mu = s q / (9 D^2) - 1 (+ (D - 1)^2 for shape_size), s = |T|^2, q =
// |cof T|^2, D = det T. With K = d^2 D (symmetric, K_xz = d2det3(T)),
// grad D = cof T = c, grad q = 2 K c and hess q = 2 (K K + M), M_xz =
// d2det3(c) (the third derivatives of det contracted with c); the product
// and quotient rules give the rest. Stored as sample_derivs stores the
// hyper-dual results (gs: 9 entries, hs: the 45-entry upper triangle).
template <int KIND>
__device__ __forceinline__ void mu3_derivs_closed(const double* t, double* gs, double* hs, long long ns, long long q)
{
    double c[9];
    cof3(t, c);
    double s = 0.0, qq = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += t[i] * t[i];
        qq += c[i] * c[i];
    }
    const double D = t[0] * c[0] + t[1] * c[1] + t[2] * c[2];
    double gq[9];
#pragma unroll
    for (int x = 0; x < 9; ++x) {
        double acc = 0.0;
#pragma unroll
        for (int z = 0; z < 9; ++z) acc += d2det3(t, x, z) * c[z];
        gq[x] = 2.0 * acc;
    }
    const double a = s * qq, iD = 1.0 / D, w = iD * iD * (1.0 / 9.0);
    const double w1 = -(2.0 / 9.0) * iD * iD * iD;      // grad w = w1 c
    const double w2 = (6.0 / 9.0) * iD * iD * iD * iD;  // hess w = w2 c c^T + w1 K
    double ga[9];
#pragma unroll
    for (int x = 0; x < 9; ++x) {
        ga[x] = qq * 2.0 * t[x] + s * gq[x];
        double gx = w * ga[x] + a * w1 * c[x];
        if constexpr (KIND == kShapeSize) gx += 2.0 * (D - 1.0) * c[x];
        gs[(long long)x * ns + q] = gx;
    }
    int e = 0;
#pragma unroll
    for (int x = 0; x < 9; ++x)
#pragma unroll
        for (int y = x; y < 9; ++y) {
            double kk = 0.0;
#pragma unroll
            for (int z = 0; z < 9; ++z) kk += d2det3(t, x, z) * d2det3(t, z, y);
            const double Kxy = d2det3(t, x, y);
            const double Hq = 2.0 * (kk + d2det3(c, x, y));
            const double Ha = (x == y ? 2.0 * qq : 0.0) + s * Hq + 2.0 * t[x] * gq[y] + 2.0 * gq[x] * t[y];
            double hv = w * Ha + w1 * (ga[x] * c[y] + c[x] * ga[y]) + a * (w2 * c[x] * c[y] + w1 * Kxy);
            if constexpr (KIND == kShapeSize) hv += 2.0 * c[x] * c[y] + 2.0 * (D - 1.0) * Kxy;
            hs[(long long)(e++) * ns + q] = hv;
        }
}

// metric gradient and symmetric Hessian of block-local sample p (positions
// xb of the block), stored component-major at index q of an ns-sample space
template <int D, int KIND>
__device__ __forceinline__ void sample_derivs(const CtrlGeom& g, long long p, const double* xb, double* gs,
                                              double* hs, long long ns, long long q)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long cell = p / NC;
    const int b = int(p - cell * NC);
    int I0[3];
    cell_origin<D>(g, cell, I0);
    double t[V];
    sample_A<D>(g, I0, b, xb, nullptr, 0.0, t);
    if constexpr (D == 3) {
        // the closed form: 45 hyper-dual metric evaluations per sample were
        // 20x slower, compute-bound, where this one streams its 54 outputs
        mu3_derivs_closed<KIND>(t, gs, hs, ns, q);
        return;
    }
    int e = 0;
    for (int i = 0; i < V; ++i)
        for (int r = i; r < V; ++r) {
            HD T[V];
#pragma unroll
            for (int j = 0; j < V; ++j) T[j] = HD(t[j]);
            T[i].a = 1.0;
            T[r].b = 1.0;
            const HD m = metric_value<D, KIND>(T);
            hs[(long long)(e++) * ns + q] = m.ab;
            if (i == r) gs[(long long)i * ns + q] = m.a;
        }
}

template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void ctrl_samples_kernel(const CtrlGeom g, const double* xb, double* gs,
                                                                double* hs)
{
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    const long long ns = g.n_cells * (1 << D);
    // component-major (entry e of sample p at e * ns + p): coalesced
    if (p < ns) sample_derivs<D, KIND>(g, p, xb, gs, hs, ns, p);
}

// energy and det of block-local sample p at xb + a dir
template <int D, int KIND>
__device__ __forceinline__ void sample_energy(const CtrlGeom& g, long long p, const double* xb, const double* dir,
                                              double a, double& e, double& m)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const long long cell = p / NC;
    const int b = int(p - cell * NC);
    int I0[3];
    cell_origin<D>(g, cell, I0);
    double t[V];
    sample_A<D>(g, I0, b, xb, dir, a, t);
    if constexpr (D == 2) {
        e += mu2_value<KIND, double>(t[0], t[1], t[2], t[3], 0.0);
        m = fmin(m, t[0] * t[3] - t[1] * t[2]);
    } else {
        e += mu3_value<KIND, double>(t, 0.0);
        m = fmin(m, det3(t));
    }
}

// fixed-tree fold of the four per-thread partials of a workgroup into
// partials[4 * blockIdx.x ..]
__device__ __forceinline__ void energy_fold(double e0, double m0, double e1, double m1, double* partials)
{
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

// FP64 cell energy and min det of the block at two step lengths: each
// workgroup walks a contiguous run of samples with stride kThreads, then a
// fixed tree per workgroup and a one-workgroup fold over the partials
template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void ctrl_energy_kernel(const CtrlGeom g, const double* xb, const double* dir,
                                                               double a0, double a1, double* partials)
{
    const long long ns = g.n_cells * (1 << D);
    const long long per = (ns + gridDim.x - 1) / gridDim.x;
    const long long lo = (long long)blockIdx.x * per, hi = min(ns, lo + per);
    double e0 = 0.0, e1 = 0.0, m0 = __builtin_huge_val(), m1 = __builtin_huge_val();
    for (long long p = lo + threadIdx.x; p < hi; p += kThreads) {
        sample_energy<D, KIND>(g, p, xb, dir, a0, e0, m0);
        sample_energy<D, KIND>(g, p, xb, dir, a1, e1, m1);
    }
    energy_fold(e0, m0, e1, m1, partials);
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

// ---------------------------------------------------------------- pullback

// one workgroup per control (lattice index a): its threads split the
// tensor product of the per-axis support lists, then fold in a fixed tree
// into out[0 .. NO)
template <int D, int MODE>
__device__ __forceinline__ void pullback_wg(const CtrlGeom& g, long long a_flat, const double* y, double* out)
{
    constexpr int NO = MODE == 1 ? D * D : D;
    long long rem = a_flat;
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
    if ((int)threadIdx.x < NO) out[threadIdx.x] = sh[threadIdx.x][0];
}

template <int D, int MODE>
__global__ __launch_bounds__(kThreads) void ctrl_pullback_kernel(const CtrlGeom g, const double* y, double* out)
{
    constexpr int NO = MODE == 1 ? D * D : D;
    const int f = blockIdx.x;
    if (g.halt && *g.halt != 0.0) return;  // uniform over the workgroup
    pullback_wg<D, MODE>(g, g.free_idx[f], y, out + (long long)f * NO);
}

// ---------------------------------------------------------------- several blocks

// the segment s of prefix table pfx[0 .. n] with pfx[s] <= i < pfx[s + 1]
__device__ __forceinline__ int find_seg(const long long* pfx, int n, long long i)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pfx[mid] <= i)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// positions of every block node from the FP32 store (its own slot)
template <int D>
__global__ __launch_bounds__(kThreads) void topo_load_kernel(const TopoView t, const float* x, double* out)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= t.n_all) return;
    const TopoBlock& B = t.blk[find_seg(t.npfx, t.nb, n)];
    int I[3];
    unravel_node<D>(B.g, n - B.node0, I);
    const long long s = node_slot<D>(B.g, I);
#pragma unroll
    for (int c = 0; c < D; ++c) out[n * D + c] = (double)x[c * B.g.plane + s];
}

// every block node takes its owner's net value: state M C + b (nodes whose
// owner block has no net keep their value), direction M v (0 there); the
// store receives the state at each node's own slot
template <int D, bool DIR>
__global__ __launch_bounds__(kThreads) void topo_eval_kernel(const TopoView t, const double* src, double* out,
                                                             float* out_plane)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= t.n_all || (t.halt && *t.halt != 0.0)) return;
    const long long o = t.owner[n];
    const TopoBlock& Bo = t.blk[find_seg(t.npfx, t.nb, o)];
    double acc[D];
    if (Bo.net < 0) {
        if constexpr (DIR) {
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] = 0.0;
        } else {
            return;
        }
    } else {
        int I[3];
        const long long lo = o - Bo.node0;
        unravel_node<D>(Bo.g, lo, I);
        eval_at<D, false>(Bo.g, I, src + Bo.ctrl0 * D, acc);
        if constexpr (!DIR) {
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c] += Bo.b ? Bo.b[lo * D + c] : 0.0;
        }
    }
#pragma unroll
    for (int c = 0; c < D; ++c) out[n * D + c] = acc[c];
    if (out_plane) {
        const TopoBlock& B = t.blk[find_seg(t.npfx, t.nb, n)];
        int I[3];
        unravel_node<D>(B.g, n - B.node0, I);
        const long long s = node_slot<D>(B.g, I);
#pragma unroll
        for (int c = 0; c < D; ++c) out_plane[c * B.g.plane + s] = (float)acc[c];
    }
}

template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void topo_samples_kernel(const TopoView t, const double* xb, double* gs,
                                                                double* hs)
{
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= t.ns_all) return;
    const TopoBlock& B = t.blk[find_seg(t.spfx, t.nb, p)];
    sample_derivs<D, KIND>(B.g, p - B.samp0, xb + B.node0 * D, gs, hs, t.ns_all, p);
}

// Per sample, its contributions to the cell's corner nodes, summed per
// corner inside the cell's 2^D consecutive lanes: sample b gives its
// corner node sum_k u_k z_(i,k) (u_k = -s_k) and its axis-k neighbour
// (corner b ^ bit k) s_k z_(i,k); lane b adds its own term and what the
// lanes b ^ bit k give their axis-k neighbour. z is the sample's Hessian
// action H dA (HV, dA from the block-local field dir) or its metric
// gradient; the sample is pl of block geometry g, at index p of the
// ns-sample component-major arrays. cc[p D + i]: corner b of the sample's
// cell. Every lane of a cell calls this together (the sample count and
// kThreads are multiples of 2^D, and a block's first sample is too).
static_assert(kThreads % 8 == 0, "a cell's lanes share a workgroup");
template <int D, bool HV>
__device__ __forceinline__ void cell_corners(const CtrlGeom& g, long long pl, const double* gs, const double* hs,
                                             const double* dir, long long ns, long long p, double* cc)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const int b = int(pl & (NC - 1));
    double z[V];
    if constexpr (HV) {
        int I0[3];
        cell_origin<D>(g, pl / NC, I0);
        double dA[V];
        sample_A<D>(g, I0, b, dir, nullptr, 0.0, dA);
        double H[kNHs<D>];
#pragma unroll
        for (int e = 0; e < kNHs<D>; ++e) H[e] = hs[(long long)e * ns + p];
#pragma unroll
        for (int i = 0; i < V; ++i) {
            double y = 0.0;
#pragma unroll
            for (int r = 0; r < V; ++r) y += H[sym_index<D>(i, r)] * dA[r];
            z[i] = y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < V; ++i) z[i] = gs[(long long)i * ns + p];
    }
    double tot[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double own = 0.0;
#pragma unroll
        for (int k = 0; k < D; ++k) own += (((b >> (D - 1 - k)) & 1) ? 1.0 : -1.0) * z[i * D + k];
        tot[i] = own;
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int bit = 1 << (D - 1 - k);
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const double nb = ((b & bit) ? -1.0 : 1.0) * z[i * D + k];  // to corner b ^ bit
            tot[i] += __shfl_xor(nb, bit);
        }
    }
#pragma unroll
    for (int i = 0; i < D; ++i) cc[p * D + i] = tot[i];
}

// The contracted node Hessians the same way: sample b gives its corner
// sum_kl u_k u_l H_(ik,jl) and its axis-k neighbour H_(ik,jk) (u_k^2 = 1),
// D x D values each; cc[p D^2 + i D + j]: corner b of the sample's cell.
template <int D>
__device__ __forceinline__ void cell_corner_hessians(long long pl, const double* hs, long long ns, long long p,
                                                     double* cc)
{
    constexpr int NC = 1 << D, V = kV<D>;
    const int b = int(pl & (NC - 1));
    double H[kNHs<D>];
#pragma unroll
    for (int e = 0; e < kNHs<D>; ++e) H[e] = hs[(long long)e * ns + p];
    double u[D];
#pragma unroll
    for (int k = 0; k < D; ++k) u[k] = ((b >> (D - 1 - k)) & 1) ? 1.0 : -1.0;
    double tot[V];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double h = 0.0;
#pragma unroll
            for (int k = 0; k < D; ++k)
#pragma unroll
                for (int l = 0; l < D; ++l) h += u[k] * u[l] * H[sym_index<D>(i * D + k, j * D + l)];
            tot[i * D + j] = h;
        }
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int bit = 1 << (D - 1 - k);
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) tot[i * D + j] += __shfl_xor(H[sym_index<D>(i * D + k, j * D + k)], bit);
    }
#pragma unroll
    for (int e = 0; e < V; ++e) cc[p * V + e] = tot[e];
}

// add the corner entries of block-local node I's incident cells (the
// block's samples from index s0 on) to acc
template <int D, int NO = D>
__device__ __forceinline__ void node_corners(const CtrlGeom& g, const int* I, const double* cc, long long s0,
                                             double* acc)
{
    constexpr int NC = 1 << D;
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
        const long long q = s0 + cell * NC + o;
#pragma unroll
        for (int i = 0; i < NO; ++i) acc[i] += cc[q * NO + i];
    }
}

template <int D, bool HV>
__global__ __launch_bounds__(kThreads) void ctrl_cell_kernel(const CtrlGeom g, const double* gs, const double* hs,
                                                             const double* dir, double* cc)
{
    if (g.halt && *g.halt != 0.0) return;  // uniform over the launch
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    const long long ns = g.n_cells * (1 << D);
    if (p < ns) cell_corners<D, HV>(g, p, gs, hs, dir, ns, p, cc);  // whole cells in or out
}

template <int D>
__global__ __launch_bounds__(kThreads) void ctrl_cell_hess_kernel(const CtrlGeom g, const double* hs, double* cc)
{
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    const long long ns = g.n_cells * (1 << D);
    if (p < ns) cell_corner_hessians<D>(p, hs, ns, p, cc);  // whole cells in or out
}

template <int D, int NO>
__global__ __launch_bounds__(kThreads) void ctrl_node_cc_kernel(const CtrlGeom g, const double* cc, double* out)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= g.n_nodes || (g.halt && *g.halt != 0.0)) return;
    int I[3];
    unravel_node<D>(g, n, I);
    double acc[NO] = {};
    node_corners<D, NO>(g, I, cc, 0, acc);
#pragma unroll
    for (int i = 0; i < NO; ++i) out[n * NO + i] = acc[i];
}

template <int D, bool HV>
__global__ __launch_bounds__(kThreads) void topo_cell_kernel(const TopoView t, const double* gs, const double* hs,
                                                             const double* dir, double* cc)
{
    if (t.halt && *t.halt != 0.0) return;  // uniform over the launch
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= t.ns_all) return;  // whole cells in or out
    const TopoBlock& B = t.blk[find_seg(t.spfx, t.nb, p)];
    cell_corners<D, HV>(B.g, p - B.samp0, gs, hs, dir ? dir + B.node0 * D : nullptr, t.ns_all, p, cc);
}

template <int D>
__global__ __launch_bounds__(kThreads) void topo_cell_hess_kernel(const TopoView t, const double* hs, double* cc)
{
    const long long p = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (p >= t.ns_all) return;  // whole cells in or out
    const TopoBlock& B = t.blk[find_seg(t.spfx, t.nb, p)];
    cell_corner_hessians<D>(p - B.samp0, hs, t.ns_all, p, cc);
}

// per block node: an owner sums the corner entries (NO values each) of its
// incident cells and of its copies' incident cells; a copy receives 0
template <int D, int NO>
__global__ __launch_bounds__(kThreads) void topo_node_cc_kernel(const TopoView t, const double* cc, double* out)
{
    const long long n = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (n >= t.n_all || (t.halt && *t.halt != 0.0)) return;
    double acc[NO] = {};
    if (t.owner[n] == n) {
        for (int e = -1; e < t.grp_off[n + 1] - t.grp_off[n]; ++e) {
            const long long m = e < 0 ? n : (long long)t.grp_mem[t.grp_off[n] + e];
            const TopoBlock& B = t.blk[find_seg(t.npfx, t.nb, m)];
            int I[3];
            unravel_node<D>(B.g, m - B.node0, I);
            node_corners<D, NO>(B.g, I, cc, B.samp0, acc);
        }
    }
#pragma unroll
    for (int i = 0; i < NO; ++i) out[n * NO + i] = acc[i];
}

// one workgroup per stacked control
template <int D, int MODE>
__global__ __launch_bounds__(kThreads) void topo_pullback_kernel(const TopoView t, const double* y, double* out)
{
    constexpr int NO = MODE == 1 ? D * D : D;
    const long long f = blockIdx.x;
    if (t.halt && *t.halt != 0.0) return;  // uniform over the workgroup
    const TopoBlock& B = t.blk[t.net_blk[find_seg(t.cpfx, t.n_nets, f)]];
    pullback_wg<D, MODE>(B.g, f - B.ctrl0, y + B.node0 * NO, out + f * NO);
}

template <int D, int KIND>
__global__ __launch_bounds__(kThreads) void topo_energy_kernel(const TopoView t, const double* xb, const double* dir,
                                                               double a0, double a1, double* partials)
{
    const long long ns = t.ns_all;
    const long long per = (ns + gridDim.x - 1) / gridDim.x;
    const long long lo = (long long)blockIdx.x * per, hi = min(ns, lo + per);
    double e0 = 0.0, e1 = 0.0, m0 = __builtin_huge_val(), m1 = __builtin_huge_val();
    for (long long p = lo + threadIdx.x; p < hi; p += kThreads) {
        const TopoBlock& B = t.blk[find_seg(t.spfx, t.nb, p)];
        const double* xl = xb + B.node0 * D;
        const double* dl = dir ? dir + B.node0 * D : nullptr;
        sample_energy<D, KIND>(B.g, p - B.samp0, xl, dl, a0, e0, m0);
        sample_energy<D, KIND>(B.g, p - B.samp0, xl, dl, a1, e1, m1);
    }
    energy_fold(e0, m0, e1, m1, partials);
}

// y = A x (+ lam z) over n_rows CSR rows
__global__ __launch_bounds__(kThreads) void csr_apply_kernel(long long n_rows, const int* off, const int* col,
                                                             const double* coef, const double* x, double lam,
                                                             const double* z, double* y)
{
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n_rows) return;
    double acc = 0.0;
    for (int e = off[i]; e < off[i + 1]; ++e) acc += coef[e] * x[col[e]];
    if (z) acc += lam * z[i];
    y[i] = acc;
}

// y_i += sum over the transpose row i of coef * w_j * r_j
__global__ __launch_bounds__(kThreads) void csrT_accum_kernel(long long n_rows, const int* off, const int* row,
                                                              const double* coef, const double* w, const double* r,
                                                              double* y)
{
    const long long i = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n_rows) return;
    double acc = 0.0;
    for (int e = off[i]; e < off[i + 1]; ++e) acc += coef[e] * w[row[e]] * r[row[e]];
    y[i] += acc;
}

// diagonal of R^T P R for block-diagonal P (one D x D block per control):
// the entries of column c (rows ascending) grouped by control, each group's
// coefficient vector v adding v^T P_control v
template <int D>
__global__ __launch_bounds__(kThreads) void qdiag_build_kernel(long long nq, const int* g_off, const int* g_row,
                                                               const double* g_coef, const double* P, double* diag)
{
    const long long c = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (c >= nq) return;
    double d = 0.0;
    int e = g_off[c];
    const int end = g_off[c + 1];
    while (e < end) {
        const int ctrl = g_row[e] / D;
        double v[D] = {};
        for (; e < end && g_row[e] / D == ctrl; ++e) v[g_row[e] - ctrl * D] = g_coef[e];
        const double* Pc = P + (long long)ctrl * D * D;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int j = 0; j < D; ++j) d += v[i] * Pc[i * D + j] * v[j];
    }
    diag[c] = d;
}

__global__ __launch_bounds__(kThreads) void qdiag_apply_kernel(long long nq, const double* diag, double lam,
                                                               const double* r, double* z)
{
    const long long c = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (c >= nq) return;
    const double d = diag[c] + lam;
    z[c] = d > 0.0 ? r[c] / d : r[c];
}

// ---------------------------------------------------------------- device-resident reduced PCG

// The reduced-space PCG's recurrences with its scalars in device memory
// (s: 0 r.z, 1 r.r, 2 p.Ap, 3 |b|^2, 4 stop threshold, 5 iterations, 6
// state: 0 running, 1 converged, 2 negative curvature, 3 iteration cap),
// one workgroup of 1024 threads each: strided partial sums and the same
// fixed tree as ctrl_reduce_kernel, so the iterates are those of the
// host-driven loop.
__device__ __forceinline__ double wg1024_sum(double v, double* sh)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// z = M^-1 r by the workgroup, then its threads' partials of r.z in the
// element order of ctrl_reduce_kernel: BLOCK false, M = diag + lam I (a
// scalar per element); BLOCK true, M = one D x D block plus lam I per
// control (Gaussian elimination as ctrl_precond_kernel)
template <int D, bool BLOCK>
__device__ __forceinline__ double wg_precond(long long nq, const double* P, double lam, const double* r, double* z)
{
    if constexpr (!BLOCK) {
        double rz = 0.0;
        for (long long i = threadIdx.x; i < nq; i += 1024) {
            const double d = P[i] + lam, ri = r[i];
            const double zi = d > 0.0 ? ri / d : ri;
            z[i] = zi;
            rz += ri * zi;
        }
        return rz;
    } else {
        const long long nf = nq / D;
        for (long long f = threadIdx.x; f < nf; f += 1024) {
            double A[D * D], rhs[D], x[D];
#pragma unroll
            for (int i = 0; i < D * D; ++i) A[i] = P[f * D * D + i];
#pragma unroll
            for (int i = 0; i < D; ++i) {
                A[i * D + i] += lam;
                rhs[i] = r[f * D + i];
            }
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
                double t = rhs[i];
#pragma unroll
                for (int k = i + 1; k < D; ++k) t -= A[i * D + k] * x[k];
                x[i] = A[i * D + i] != 0.0 ? t / A[i * D + i] : t;
            }
#pragma unroll
            for (int i = 0; i < D; ++i) z[f * D + i] = x[i];
        }
        __syncthreads();  // z of other threads' controls, read below
        double rz = 0.0;
        for (long long i = threadIdx.x; i < nq; i += 1024) rz += r[i] * z[i];
        return rz;
    }
}

template <int D, bool BLOCK>
__global__ __launch_bounds__(1024) void qpcg_init_kernel(long long nq, const double* G, const double* P, double lam,
                                                         double rtol, double* x, double* r, double* z, double* p,
                                                         double* s)
{
    __shared__ double sh[1024];
    double rr = 0.0;
    for (long long i = threadIdx.x; i < nq; i += 1024) {
        x[i] = 0.0;
        const double ri = -G[i];
        r[i] = ri;
        rr += ri * ri;
    }
    rr = wg1024_sum(rr, sh);  // its barrier publishes r
    double rz = wg_precond<D, BLOCK>(nq, P, lam, r, z);
    rz = wg1024_sum(rz, sh);
    for (long long i = threadIdx.x; i < nq; i += 1024) p[i] = z[i];
    if (threadIdx.x == 0) {
        s[0] = rz;
        s[1] = rr;
        s[3] = rr;
        s[4] = rtol * rtol * rr;
        s[5] = 0.0;
        s[6] = rr == 0.0 ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(1024) void qpcg_pap_kernel(long long nq, const double* p, const double* Ap, double* s)
{
    __shared__ double sh[1024];
    if (s[6] != 0.0) return;
    double acc = 0.0;
    for (long long i = threadIdx.x; i < nq; i += 1024) acc += p[i] * Ap[i];
    acc = wg1024_sum(acc, sh);
    if (threadIdx.x == 0) s[2] = acc;
}

template <int D, bool BLOCK>
__global__ __launch_bounds__(1024) void qpcg_update_kernel(long long nq, const double* P, double lam, int max_iter,
                                                           double* x, double* r, double* z, double* p,
                                                           const double* Ap, double* s)
{
    __shared__ double sh[1024];
    if (s[6] != 0.0) return;
    const double pap = s[2], rz_old = s[0], stop2 = s[4];
    const int it = (int)s[5] + 1;
    if (!(pap > 0.0)) {  // negative curvature: keep the last iterate
        __syncthreads();
        if (threadIdx.x == 0) s[6] = 2.0;
        return;
    }
    const double a = rz_old / pap;
    double rr = 0.0;
    for (long long i = threadIdx.x; i < nq; i += 1024) {
        x[i] = x[i] + a * p[i];
        const double ri = r[i] + (-a) * Ap[i];
        r[i] = ri;
        rr += ri * ri;
    }
    rr = wg1024_sum(rr, sh);  // its barrier publishes r
    if (rr <= stop2) {
        if (threadIdx.x == 0) {
            s[1] = rr;
            s[5] = it;
            s[6] = 1.0;
        }
        return;
    }
    double rz = wg_precond<D, BLOCK>(nq, P, lam, r, z);
    rz = wg1024_sum(rz, sh);
    const double beta = rz / rz_old;
    for (long long i = threadIdx.x; i < nq; i += 1024) p[i] = z[i] + beta * p[i];
    if (threadIdx.x == 0) {
        s[0] = rz;
        s[1] = rr;
        s[5] = it;
        if (it >= max_iter) s[6] = 3.0;
    }
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

void launch_ctrl_cell(const CtrlGeom& g, bool hv, const double* gs, const double* hs, const double* dir, double* cc,
                      hipStream_t s)
{
    const long long n = g.n_cells * (1LL << g.dim);
    if (n <= 0) return;
    const unsigned nb = blocks_for(n);
    if (g.dim == 2) {
        if (hv)
            ctrl_cell_kernel<2, true><<<nb, kThreads, 0, s>>>(g, gs, hs, dir, cc);
        else
            ctrl_cell_kernel<2, false><<<nb, kThreads, 0, s>>>(g, gs, hs, dir, cc);
    } else {
        if (hv)
            ctrl_cell_kernel<3, true><<<nb, kThreads, 0, s>>>(g, gs, hs, dir, cc);
        else
            ctrl_cell_kernel<3, false><<<nb, kThreads, 0, s>>>(g, gs, hs, dir, cc);
    }
}

void launch_ctrl_node_cc(const CtrlGeom& g, const double* cc, double* out, hipStream_t s)
{
    if (g.n_nodes <= 0) return;
    if (g.dim == 2)
        ctrl_node_cc_kernel<2, 2><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, cc, out);
    else
        ctrl_node_cc_kernel<3, 3><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, cc, out);
}

void launch_ctrl_node_hessians(const CtrlGeom& g, const double* hs, double* cc, double* out, hipStream_t s)
{
    const long long ns = g.n_cells * (1LL << g.dim);
    if (g.n_nodes <= 0) return;
    if (g.dim == 2) {
        if (ns > 0) ctrl_cell_hess_kernel<2><<<blocks_for(ns), kThreads, 0, s>>>(g, hs, cc);
        ctrl_node_cc_kernel<2, 4><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, cc, out);
    } else {
        if (ns > 0) ctrl_cell_hess_kernel<3><<<blocks_for(ns), kThreads, 0, s>>>(g, hs, cc);
        ctrl_node_cc_kernel<3, 9><<<blocks_for(g.n_nodes), kThreads, 0, s>>>(g, cc, out);
    }
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

// ---------------------------------------------------------------- several blocks

void launch_topo_load(const TopoView& t, const float* x, double* out, hipStream_t s)
{
    if (t.n_all <= 0) return;
    if (t.dim == 2)
        topo_load_kernel<2><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, x, out);
    else
        topo_load_kernel<3><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, x, out);
}

void launch_topo_eval(const TopoView& t, bool dir, const double* src, double* out, float* out_plane, hipStream_t s)
{
    if (t.n_all <= 0) return;
    const unsigned nb = blocks_for(t.n_all);
    if (t.dim == 2) {
        if (dir)
            topo_eval_kernel<2, true><<<nb, kThreads, 0, s>>>(t, src, out, out_plane);
        else
            topo_eval_kernel<2, false><<<nb, kThreads, 0, s>>>(t, src, out, out_plane);
    } else {
        if (dir)
            topo_eval_kernel<3, true><<<nb, kThreads, 0, s>>>(t, src, out, out_plane);
        else
            topo_eval_kernel<3, false><<<nb, kThreads, 0, s>>>(t, src, out, out_plane);
    }
}

void launch_topo_samples(const TopoView& t, int kind, const double* xb, double* gs, double* hs, hipStream_t s)
{
    if (t.ns_all <= 0) return;
    const unsigned nb = blocks_for(t.ns_all);
#define EGGMI_TS(D, K) topo_samples_kernel<D, K><<<nb, kThreads, 0, s>>>(t, xb, gs, hs)
    if (t.dim == 2) {
        if (kind == kShapeSize)
            EGGMI_TS(2, kShapeSize);
        else
            EGGMI_TS(2, kShape);
    } else {
        if (kind == kShapeSize)
            EGGMI_TS(3, kShapeSize);
        else
            EGGMI_TS(3, kShape);
    }
#undef EGGMI_TS
}

void launch_topo_pullback(const TopoView& t, int mode, const double* y, double* out, hipStream_t s)
{
    if (t.n_stack <= 0) return;
    const unsigned nb = (unsigned)t.n_stack;  // one workgroup per stacked control
    if (t.dim == 2) {
        if (mode == 0)
            topo_pullback_kernel<2, 0><<<nb, kThreads, 0, s>>>(t, y, out);
        else
            topo_pullback_kernel<2, 1><<<nb, kThreads, 0, s>>>(t, y, out);
    } else {
        if (mode == 0)
            topo_pullback_kernel<3, 0><<<nb, kThreads, 0, s>>>(t, y, out);
        else
            topo_pullback_kernel<3, 1><<<nb, kThreads, 0, s>>>(t, y, out);
    }
}

void launch_topo_energy(const TopoView& t, int kind, const double* xb, const double* dir, double a0, double a1,
                        double* partials, double* out, hipStream_t s)
{
    const int nb = (int)std::max<long long>(1, std::min<long long>(kCtrlParts, (t.ns_all + kThreads - 1) / kThreads));
#define EGGMI_TE(D, K) topo_energy_kernel<D, K><<<nb, kThreads, 0, s>>>(t, xb, dir, a0, a1, partials)
    if (t.dim == 2) {
        if (kind == kShapeSize)
            EGGMI_TE(2, kShapeSize);
        else
            EGGMI_TE(2, kShape);
    } else {
        if (kind == kShapeSize)
            EGGMI_TE(3, kShapeSize);
        else
            EGGMI_TE(3, kShape);
    }
#undef EGGMI_TE
    ctrl_energy_finish<<<1, kThreads, 0, s>>>(partials, nb, out);
}

void launch_csr_apply(long long n_rows, const int* off, const int* col, const double* coef, const double* x,
                      double lam, const double* z, double* y, hipStream_t s)
{
    if (n_rows > 0) csr_apply_kernel<<<blocks_for(n_rows), kThreads, 0, s>>>(n_rows, off, col, coef, x, lam, z, y);
}

void launch_csrT_accum(long long n_rows, const int* off, const int* row, const double* coef, const double* w,
                       const double* r, double* y, hipStream_t s)
{
    if (n_rows > 0) csrT_accum_kernel<<<blocks_for(n_rows), kThreads, 0, s>>>(n_rows, off, row, coef, w, r, y);
}

void launch_qdiag_build(int dim, long long nq, const int* g_off, const int* g_row, const double* g_coef,
                        const double* P, double* diag, hipStream_t s)
{
    if (nq <= 0) return;
    if (dim == 2)
        qdiag_build_kernel<2><<<blocks_for(nq), kThreads, 0, s>>>(nq, g_off, g_row, g_coef, P, diag);
    else
        qdiag_build_kernel<3><<<blocks_for(nq), kThreads, 0, s>>>(nq, g_off, g_row, g_coef, P, diag);
}

void launch_qdiag_apply(long long nq, const double* diag, double lam, const double* r, double* z, hipStream_t s)
{
    if (nq > 0) qdiag_apply_kernel<<<blocks_for(nq), kThreads, 0, s>>>(nq, diag, lam, r, z);
}

void launch_qpcg_init(int dim, bool block, long long nq, const double* G, const double* P, double lam, double rtol,
                      double* x, double* r, double* z, double* p, double* s, hipStream_t st)
{
    if (!block)
        qpcg_init_kernel<2, false><<<1, 1024, 0, st>>>(nq, G, P, lam, rtol, x, r, z, p, s);
    else if (dim == 2)
        qpcg_init_kernel<2, true><<<1, 1024, 0, st>>>(nq, G, P, lam, rtol, x, r, z, p, s);
    else
        qpcg_init_kernel<3, true><<<1, 1024, 0, st>>>(nq, G, P, lam, rtol, x, r, z, p, s);
}

void launch_qpcg_pap(long long nq, const double* p, const double* Ap, double* s, hipStream_t st)
{
    qpcg_pap_kernel<<<1, 1024, 0, st>>>(nq, p, Ap, s);
}

void launch_qpcg_update(int dim, bool block, long long nq, const double* P, double lam, int max_iter, double* x,
                        double* r, double* z, double* p, const double* Ap, double* s, hipStream_t st)
{
    if (!block)
        qpcg_update_kernel<2, false><<<1, 1024, 0, st>>>(nq, P, lam, max_iter, x, r, z, p, Ap, s);
    else if (dim == 2)
        qpcg_update_kernel<2, true><<<1, 1024, 0, st>>>(nq, P, lam, max_iter, x, r, z, p, Ap, s);
    else
        qpcg_update_kernel<3, true><<<1, 1024, 0, st>>>(nq, P, lam, max_iter, x, r, z, p, Ap, s);
}

void launch_topo_cell(const TopoView& t, bool hv, const double* gs, const double* hs, const double* dir, double* cc,
                      hipStream_t s)
{
    if (t.ns_all <= 0) return;
    const unsigned nb = blocks_for(t.ns_all);
    if (t.dim == 2) {
        if (hv)
            topo_cell_kernel<2, true><<<nb, kThreads, 0, s>>>(t, gs, hs, dir, cc);
        else
            topo_cell_kernel<2, false><<<nb, kThreads, 0, s>>>(t, gs, hs, dir, cc);
    } else {
        if (hv)
            topo_cell_kernel<3, true><<<nb, kThreads, 0, s>>>(t, gs, hs, dir, cc);
        else
            topo_cell_kernel<3, false><<<nb, kThreads, 0, s>>>(t, gs, hs, dir, cc);
    }
}

void launch_topo_node_cc(const TopoView& t, const double* cc, double* out, hipStream_t s)
{
    if (t.n_all <= 0) return;
    if (t.dim == 2)
        topo_node_cc_kernel<2, 2><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, cc, out);
    else
        topo_node_cc_kernel<3, 3><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, cc, out);
}

void launch_topo_node_hessians(const TopoView& t, const double* hs, double* cc, double* out, hipStream_t s)
{
    if (t.n_all <= 0) return;
    if (t.dim == 2) {
        if (t.ns_all > 0) topo_cell_hess_kernel<2><<<blocks_for(t.ns_all), kThreads, 0, s>>>(t, hs, cc);
        topo_node_cc_kernel<2, 4><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, cc, out);
    } else {
        if (t.ns_all > 0) topo_cell_hess_kernel<3><<<blocks_for(t.ns_all), kThreads, 0, s>>>(t, hs, cc);
        topo_node_cc_kernel<3, 9><<<blocks_for(t.n_all), kThreads, 0, s>>>(t, cc, out);
    }
}

}  // namespace eggmi
