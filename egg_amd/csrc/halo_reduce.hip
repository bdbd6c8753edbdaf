// HIP kernels around the sweep: the fused halo / fan / share copy and the
// FP64 energy + min-det reductions, plus the dimension dispatch of the
// sweep launchers.
#include "directional.hpp"
#include "mg.hpp"
#include "sweep_common.hpp"

#include <stdexcept>

namespace eggmi
{

void launch_sweep2d(int kind, int v, const SweepArgs& a, hipStream_t s);
void launch_grad2d(int kind, int v, const SweepArgs& a, hipStream_t s);
void launch_sweep3d(int kind, int v, const SweepArgs& a, hipStream_t s);
void launch_grad3d(int kind, int v, const SweepArgs& a, hipStream_t s);
void launch_level_smooth2d(int kind, int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s);
void launch_level_smooth3d(int kind, int v, const SweepArgs& a, float* x0, float* x1, int n, hipStream_t s);

// node tiles (tx = last axis); every node is worked on by 2^D lanes
static constexpr TileShape kTiles2[] = {{32, 2, 1}, {16, 4, 1}, {64, 1, 1}, {32, 4, 1},
                                        {16, 1, 1}, {8, 2, 1},  // one lane per sample
                                        {64, 1, 1}, {32, 2, 1},   // no LDS staging
                                        {32, 4, 1}, {64, 4, 1},   // 2 / 4 cells per lane
                                        {32, 4, 1}, {64, 2, 1},   // 2 cells per lane, without / with LDS
                                        {32, 2, 1, 4}, {32, 2, 1, 8}};  // 4 / 8 lanes per stored DOF
static constexpr TileShape kTiles3[] = {{8, 4, 2}, {4, 4, 4}, {8, 4, 2} /* no LDS staging */,
                                        {8, 4, 2} /* 2 cells per lane */,
                                        {8, 4, 2, 16} /* no LDS staging, 16 lanes per stored DOF */};

int num_tile_variants(int dim) { return dim == 2 ? 14 : 5; }
bool tile_variant_has_proposal(int dim, int v) { return dim == 2 ? v < 4 : v < 2; }
int tile_variant_proposal_partner(int dim, int v)
{
    if (tile_variant_has_proposal(dim, v)) return v;
    const TileShape t = tile_variant_shape(dim, v);
    for (int p = 0; p < num_tile_variants(dim); ++p) {
        if (!tile_variant_has_proposal(dim, p)) continue;
        const TileShape q = tile_variant_shape(dim, p);
        if (q.tx == t.tx && q.ty == t.ty && q.tz == t.tz &&
            tile_variant_stored_lanes(dim, p) == tile_variant_stored_lanes(dim, v))
            return p;
    }
    return -1;
}
TileShape tile_variant_shape(int dim, int v) { return dim == 2 ? kTiles2[v] : kTiles3[v]; }
int tile_variant_stored_lanes(int dim, int v)
{
    const TileShape t = tile_variant_shape(dim, v);
    return t.sg ? t.sg : stored_lanes(dim);
}

void launch_sweep(int dim, int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (dim == 2)
        launch_sweep2d(kind, v, a, s);
    else
        launch_sweep3d(kind, v, a, s);
}

void launch_level_smooth(int dim, int kind, int v, const SweepArgs& a, float* x0, float* x1, int n_sweeps,
                         hipStream_t s)
{
    if (a.st.n > 0 || a.con_n > 0 || !a.tau)
        throw std::invalid_argument("one-workgroup smoothing takes tau-shifted levels of synthesized DOFs only");
    if (dim == 2)
        launch_level_smooth2d(kind, v, a, x0, x1, n_sweeps, s);
    else
        launch_level_smooth3d(kind, v, a, x0, x1, n_sweeps, s);
}

void launch_grad(int dim, int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (dim == 2)
        launch_grad2d(kind, v, a, s);
    else
        launch_grad3d(kind, v, a, s);
}

// ------------------------------------------------------------------ copies

// Fused ghost-halo + fan-mirror + shared-node copy (reference
// src/structured_halo.hpp:198-237): x[dst] = x[src] for every coordinate.
__global__ __launch_bounds__(256) void copy_nodes(float* x, long long plane, int dim, const int* __restrict__ src,
                                                  const int* __restrict__ dst, long long n)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int s = src[e], d = dst[e];
    for (int k = 0; k < dim; ++k) x[k * plane + d] = x[k * plane + s];
}

// ------------------------------------------------------------------ reductions

__device__ __forceinline__ void block_reduce_store(double e, double m, double* out)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        m = fmin(m, __shfl_xor(m, off, 64));
    }
    __shared__ double se[kReduceThreads / 64], sm[kReduceThreads / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) {
        se[w] = e;
        sm[w] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double te = 0.0, tm = __builtin_huge_val();
        for (int i = 0; i < kReduceThreads / 64; ++i) {
            te += se[i];
            tm = fmin(tm, sm[i]);
        }
        out[0] = te;
        out[1] = tm;
    }
}

// Cells of the structured blocks (ids in block order, prefix a.cell_off) are
// cut into one contiguous chunk per workgroup, walked by its threads with
// stride kReduceThreads (coalesced rows); a thread finds its first cell's
// block once and then only steps forward as its ids grow.
__device__ __forceinline__ void cell_chunk(long long ncells, long long& c0, long long& c1)
{
    const long long g = gridDim.x;
    long long per = (ncells + g - 1) / g;
    per = (per + kReduceThreads - 1) / kReduceThreads * kReduceThreads;
    c0 = (long long)blockIdx.x * per;
    c1 = c0 + per < ncells ? c0 + per : ncells;
}
__device__ __forceinline__ int cell_block(const long long* off, int nblocks, long long cid)
{
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= cid)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int cell_block_step(const long long* off, int nblocks, int lo, long long cid)
{
    while (lo + 1 < nblocks && off[lo + 1] <= cid) ++lo;
    return lo;
}

template <int KIND> __device__ __forceinline__ double mu2_d(double a, double b, double c, double d, double delta)
{
    return mu2_value<KIND, double>(a, b, c, d, delta);
}

// Sum of mu and min det A over every (cell, corner) sample of the structured
// blocks, evaluated and accumulated in FP64 from the FP32 store
// (reference src/sweep.hpp:1881-1922; energy stencil of
// egg/smoothing/flat_context.py:64-133).
template <int DIM, int KIND>
__global__ __launch_bounds__(kReduceThreads) void reduce_cells(const CellReduceArgs a, double* partials)
{
    double esum = 0.0, emin = __builtin_huge_val();
    long long c0, c1;
    cell_chunk(a.ncells, c0, c1);
    int lo = c0 + threadIdx.x < c1 ? cell_block(a.cell_off, a.nblocks, c0 + threadIdx.x) : 0;
    for (long long cid = c0 + threadIdx.x; cid < c1; cid += kReduceThreads) {
        lo = cell_block_step(a.cell_off, a.nblocks, lo, cid);
        const BlockDesc bd = a.blocks[lo];
        int r = int(cid - a.cell_off[lo]);  // cells of one block < 2^31
        if constexpr (DIM == 2) {
            const int w1 = bd.n[1] - 1;
            const int ci = r / w1, cj = r - ci * w1;
            double px[2][2], py[2][2];
#pragma unroll
            for (int o0 = 0; o0 < 2; ++o0)
#pragma unroll
                for (int o1 = 0; o1 < 2; ++o1) {
                    const long long node = bd.base + (long long)(ci + o0 + 1) * bd.st[0] + (cj + o1 + 1);
                    px[o0][o1] = a.x[node];
                    py[o0][o1] = a.x[a.plane + node];
                }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int o0 = b >> 1, o1 = b & 1;
                const double s0 = o0 ? -1.0 : 1.0, s1 = o1 ? -1.0 : 1.0;
                const double A00 = s0 * (px[1 - o0][o1] - px[o0][o1]);
                const double A01 = s1 * (px[o0][1 - o1] - px[o0][o1]);
                const double A10 = s0 * (py[1 - o0][o1] - py[o0][o1]);
                const double A11 = s1 * (py[o0][1 - o1] - py[o0][o1]);
                esum += mu2_d<KIND>(A00, A01, A10, A11, a.delta);
                emin = fmin(emin, A00 * A11 - A01 * A10);
            }
        } else {
            const int w2 = bd.n[2] - 1, w1 = bd.n[1] - 1;
            const int ci = r / (w1 * w2);
            r -= ci * w1 * w2;
            const int cj = r / w2, ck = r - cj * w2;
            double p[3][8];
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int o0 = o >> 2, o1 = (o >> 1) & 1, o2 = o & 1;
                const long long node =
                  bd.base + (long long)(ci + o0 + 1) * bd.st[0] + (long long)(cj + o1 + 1) * bd.st[1] + (ck + o2 + 1);
#pragma unroll
                for (int c = 0; c < 3; ++c) p[c][o] = a.x[c * a.plane + node];
            }
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                double A[9];
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const int bit = 1 << (2 - kk);
                    const double s = (b & bit) ? -1.0 : 1.0;
#pragma unroll
                    for (int ii = 0; ii < 3; ++ii) A[ii * 3 + kk] = s * (p[ii][b ^ bit] - p[ii][b]);
                }
                esum += mu3_value<KIND, double>(A, a.delta);
                emin = fmin(emin, det3(A));
            }
        }
    }
    block_reduce_store(esum, emin, partials + 2 * blockIdx.x);
}

template <int DIM, int KIND>
__global__ __launch_bounds__(kReduceThreads) void reduce_samples(const SampleReduceArgs a, double* partials)
{
    double esum = 0.0, emin = __builtin_huge_val();
    const long long stride = (long long)gridDim.x * kReduceThreads;
    for (long long i = (long long)blockIdx.x * kReduceThreads + threadIdx.x; i < a.ns; i += stride) {
        const int ic = a.gc[i];
        double A[9], T[9];
#pragma unroll
        for (int kk = 0; kk < DIM; ++kk) {
            const int in = a.gn[kk * a.ns + i];
            const double s = a.s[kk * a.ns + i];
#pragma unroll
            for (int ii = 0; ii < DIM; ++ii)
                A[ii * DIM + kk] = s * ((double)a.x[ii * a.plane + in] - (double)a.x[ii * a.plane + ic]);
        }
#pragma unroll
        for (int ii = 0; ii < DIM; ++ii)
#pragma unroll
            for (int kk = 0; kk < DIM; ++kk) {
                double acc = 0.0;
#pragma unroll
                for (int jj = 0; jj < DIM; ++jj)
                    acc += A[ii * DIM + jj] * (a.W ? (double)a.W[i * DIM * DIM + jj * DIM + kk] : a.Wu[jj * DIM + kk]);
                T[ii * DIM + kk] = acc;
            }
        // a sample's weight scales mu, not det (reference src/patch.hpp:162)
        const double w = a.wt ? (double)a.wt[i] : 1.0;
        if constexpr (DIM == 2) {
            esum += w * mu2_d<KIND>(T[0], T[1], T[2], T[3], a.delta);
            emin = fmin(emin, A[0] * A[3] - A[1] * A[2]);
        } else {
            esum += w * mu3_value<KIND, double>(T, a.delta);
            emin = fmin(emin, det3(A));
        }
    }
    block_reduce_store(esum, emin, partials + 2 * blockIdx.x);
}

__global__ __launch_bounds__(kReduceThreads) void reduce_finish(const double* partials, int nparts, double* out_e,
                                                                double* out_m, double* rep, int cap, int* count)
{
    double e = 0.0, m = __builtin_huge_val();
    for (int i = threadIdx.x; i < nparts; i += kReduceThreads) {
        e += partials[2 * i];
        m = fmin(m, partials[2 * i + 1]);
    }
    __shared__ double tmp[2];
    block_reduce_store(e, m, tmp);
    __syncthreads();
    if (threadIdx.x == 0) {
        *out_e = tmp[0];
        *out_m = tmp[1];
        if (rep) {
            const int i = *count;
            if (i < cap) {
                rep[i] = tmp[0];
                rep[cap + i] = tmp[1];
            }
            *count = i + 1;
        }
    }
}

// Corner slots of structured cell `cid` (found from the block walk `lo`).
template <int DIM>
__device__ __forceinline__ void cell_nodes(const BlockDesc& bd, int r, long long* nodes)
{
    if constexpr (DIM == 2) {
        const int w1 = bd.n[1] - 1;
        const int ci = r / w1, cj = r - ci * w1;
#pragma unroll
        for (int o = 0; o < 4; ++o) nodes[o] = bd.base + (long long)(ci + (o >> 1) + 1) * bd.st[0] + (cj + (o & 1) + 1);
    } else {
        const int w2 = bd.n[2] - 1, w1 = bd.n[1] - 1;
        const int ci = r / (w1 * w2);
        r -= ci * w1 * w2;
        const int cj = r / w2, ck = r - cj * w2;
#pragma unroll
        for (int o = 0; o < 8; ++o)
            nodes[o] = bd.base + (long long)(ci + (o >> 2) + 1) * bd.st[0] + (long long)(cj + ((o >> 1) & 1) + 1) * bd.st[1] +
                       (ck + (o & 1) + 1);
    }
}

// Energy and min det of one cell's 2^D corner samples at x + al c (positions
// formed in FP32 exactly as masked_axpy forms them), in FP64.
template <int DIM, int KIND>
__device__ __forceinline__ void cell_trial(const CellReduceArgs& a, const float* corr, const long long* nodes, float al,
                                           double& es, double& ms)
{
    constexpr int NC = 1 << DIM;
    double p[DIM][NC];
#pragma unroll
    for (int o = 0; o < NC; ++o)
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            const long long q = c * a.plane + nodes[o];
            p[c][o] = (double)(a.x[q] + al * corr[q]);
        }
    es = 0.0;
    ms = __builtin_huge_val();
#pragma unroll
    for (int b = 0; b < NC; ++b) {
        double A[DIM * DIM];
#pragma unroll
        for (int kk = 0; kk < DIM; ++kk) {
            const int bit = 1 << (DIM - 1 - kk);
            const double sg = (b & bit) ? -1.0 : 1.0;
#pragma unroll
            for (int ii = 0; ii < DIM; ++ii) A[ii * DIM + kk] = sg * (p[ii][b ^ bit] - p[ii][b]);
        }
        if constexpr (DIM == 2) {
            es += mu2_value<KIND, double>(A[0], A[1], A[2], A[3], a.delta);
            ms = fmin(ms, A[0] * A[3] - A[1] * A[2]);
        } else {
            es += mu3_value<KIND, double>(A, a.delta);
            ms = fmin(ms, det3(A));
        }
    }
}

// Two line-search trials in one pass (reference src/fas.hpp:174-234): every
// cell sample is evaluated at x + a0 c and x + a1 c, accumulated in FP64.
template <int DIM, int KIND>
__global__ __launch_bounds__(kReduceThreads) void linesearch_pair(const CellReduceArgs a, const float* corr, float a0,
                                                                  float a1, double* partials)
{
    double e0 = 0.0, e1 = 0.0, m0 = __builtin_huge_val(), m1 = __builtin_huge_val();
    long long c0, c1;
    cell_chunk(a.ncells, c0, c1);
    int lo = c0 + threadIdx.x < c1 ? cell_block(a.cell_off, a.nblocks, c0 + threadIdx.x) : 0;
    for (long long cid = c0 + threadIdx.x; cid < c1; cid += kReduceThreads) {
        lo = cell_block_step(a.cell_off, a.nblocks, lo, cid);
        const BlockDesc bd = a.blocks[lo];
        long long nodes[1 << DIM];
        cell_nodes<DIM>(bd, int(cid - a.cell_off[lo]), nodes);  // cells of one block < 2^31
        double es, ms;
        cell_trial<DIM, KIND>(a, corr, nodes, a0, es, ms);
        e0 += es;
        m0 = fmin(m0, ms);
        cell_trial<DIM, KIND>(a, corr, nodes, a1, es, ms);
        e1 += es;
        m1 = fmin(m1, ms);
    }
    block_reduce_store(e0, m0, partials + 4 * blockIdx.x);
    __syncthreads();
    block_reduce_store(e1, m1, partials + 4 * blockIdx.x + 2);
}

// The whole safeguarded ladder in one pass: slot 0 at alpha = 0 (the energy
// before the correction), slot k = 1..kLadder at alpha = 2^-(k-1), each
// accumulated exactly as linesearch_pair accumulates its trials (same cell
// walk, same per-thread and per-block folds), so a slot equals the pair
// kernel's value for the same alpha bit for bit.
template <int DIM, int KIND>
__global__ __launch_bounds__(kReduceThreads) void linesearch_ladder(const CellReduceArgs a, const float* corr,
                                                                    double* partials)
{
    constexpr int NA = kLadder + 1;
    double e[NA], m[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) {
        e[t] = 0.0;
        m[t] = __builtin_huge_val();
    }
    long long c0, c1;
    cell_chunk(a.ncells, c0, c1);
    int lo = c0 + threadIdx.x < c1 ? cell_block(a.cell_off, a.nblocks, c0 + threadIdx.x) : 0;
    for (long long cid = c0 + threadIdx.x; cid < c1; cid += kReduceThreads) {
        lo = cell_block_step(a.cell_off, a.nblocks, lo, cid);
        const BlockDesc bd = a.blocks[lo];
        long long nodes[1 << DIM];
        cell_nodes<DIM>(bd, int(cid - a.cell_off[lo]), nodes);
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            double es, ms;
            cell_trial<DIM, KIND>(a, corr, nodes, t == 0 ? 0.0f : ladder_alpha(t), es, ms);
            e[t] += es;
            m[t] = fmin(m[t], ms);
        }
    }
    for (int t = 0; t < NA; ++t) {
        block_reduce_store(e[t], m[t], partials + 2 * NA * blockIdx.x + 2 * t);
        __syncthreads();
    }
}

// Fold the ladder's partials per slot (in pair_finish's order) and pick the
// step: the first alpha = 1, 1/2, ... whose energy is finite, at most
// e(0) + tol, with every det > 0 (reference src/fas.hpp:549-581); 0 if none.
// out[2t], out[2t+1]: slot t's energy and min det.
__global__ __launch_bounds__(kReduceThreads) void ladder_finish(const double* partials, int nparts, double tol,
                                                                double* out, float* alpha)
{
    constexpr int NA = kLadder + 1;
    __shared__ double tmp[2 * NA];
    for (int t = 0; t < NA; ++t) {
        double e = 0.0, m = __builtin_huge_val();
        for (int i = threadIdx.x; i < nparts; i += kReduceThreads) {
            e += partials[2 * NA * i + 2 * t];
            m = fmin(m, partials[2 * NA * i + 2 * t + 1]);
        }
        block_reduce_store(e, m, tmp + 2 * t);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float al = 0.0f;
        for (int t = 1; t < NA; ++t)
            if (isfinite(tmp[2 * t]) && tmp[2 * t] <= tmp[0] + tol && tmp[2 * t + 1] > 0.0) {
                al = ladder_alpha(t);
                break;
            }
        *alpha = al;
        for (int i = 0; i < 2 * NA; ++i) out[i] = tmp[i];
    }
}

__global__ __launch_bounds__(kReduceThreads) void pair_finish(const double* partials, int nparts, double* out)
{
    double e0 = 0.0, e1 = 0.0, m0 = __builtin_huge_val(), m1 = __builtin_huge_val();
    for (int i = threadIdx.x; i < nparts; i += kReduceThreads) {
        e0 += partials[4 * i];
        m0 = fmin(m0, partials[4 * i + 1]);
        e1 += partials[4 * i + 2];
        m1 = fmin(m1, partials[4 * i + 3]);
    }
    __shared__ double tmp[4];
    block_reduce_store(e0, m0, tmp);
    __syncthreads();
    block_reduce_store(e1, m1, tmp + 2);
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; ++i) out[i] = tmp[i];
}

void launch_linesearch_ladder(int dim, int kind, const CellReduceArgs& a, const float* corr, double* partials,
                              double tol, double* out, float* alpha, hipStream_t s, const DirView* dir)
{
    const int nb = kReduceBlocks / 2;  // the pair kernel's grid: the same cell chunks
#define EGGMI_LL(D, K) linesearch_ladder<D, K><<<nb, kReduceThreads, 0, s>>>(a, corr, partials)
    if (dim == 2) {
        if (kind == kShapeSize)
            EGGMI_LL(2, kShapeSize);
        else
            EGGMI_LL(2, kShape);
    } else {
        if (kind == kShapeSize)
            EGGMI_LL(3, kShapeSize);
        else
            EGGMI_LL(3, kShape);
    }
#undef EGGMI_LL
    const int nd = dir ? launch_dir_reduce(dim, 2, *dir, a.x, a.plane, corr, 0.0f, 0.0f, partials, nb, s) : 0;
    ladder_finish<<<1, kReduceThreads, 0, s>>>(partials, nb + nd, tol, out, alpha);
}

void launch_linesearch_pair(int dim, int kind, const CellReduceArgs& a, const float* corr, float a0, float a1,
                            double* partials, double* out, hipStream_t s, const DirView* dir)
{
    const int nb = kReduceBlocks / 2;  // 4 partials per block fit the 2*kReduceBlocks buffer
#define EGGMI_LP(D, K) linesearch_pair<D, K><<<nb, kReduceThreads, 0, s>>>(a, corr, a0, a1, partials)
    if (dim == 2) {
        if (kind == kShapeSize)
            EGGMI_LP(2, kShapeSize);
        else
            EGGMI_LP(2, kShape);
    } else {
        if (kind == kShapeSize)
            EGGMI_LP(3, kShapeSize);
        else
            EGGMI_LP(3, kShape);
    }
#undef EGGMI_LP
    const int nd = dir ? launch_dir_reduce(dim, 1, *dir, a.x, a.plane, corr, a0, a1, partials, nb, s) : 0;
    pair_finish<<<1, kReduceThreads, 0, s>>>(partials, nb + nd, out);
}

__global__ void set_int(int* p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

void launch_set_int(int* p, int v, hipStream_t s) { set_int<<<1, 64, 0, s>>>(p, v); }

void launch_copy(float* x, long long plane, int dim, const int* src, const int* dst, long long n, hipStream_t s)
{
    if (n <= 0) return;
    copy_nodes<<<dim3((unsigned)((n + 255) / 256)), 256, 0, s>>>(x, plane, dim, src, dst, n);
}

static int reduce_grid(long long n)
{
    long long b = (n + kReduceThreads - 1) / kReduceThreads;
    if (b > kReduceBlocks) b = kReduceBlocks;
    return b < 1 ? 1 : int(b);
}

void launch_reduce_cells(int dim, int kind, const CellReduceArgs& a, double* partials, hipStream_t s)
{
    const int nb = kReduceBlocks;  // fixed partial count; finish reads all of them
#define EGGMI_RC(D, K) reduce_cells<D, K><<<nb, kReduceThreads, 0, s>>>(a, partials)
    if (dim == 2) {
        if (kind == kUntangle)
            EGGMI_RC(2, kUntangle);
        else if (kind == kShapeSize)
            EGGMI_RC(2, kShapeSize);
        else
            EGGMI_RC(2, kShape);
    } else {
        if (kind == kUntangle)
            EGGMI_RC(3, kUntangle);
        else if (kind == kShapeSize)
            EGGMI_RC(3, kShapeSize);
        else
            EGGMI_RC(3, kShape);
    }
#undef EGGMI_RC
    (void)reduce_grid;
}

void launch_reduce_samples(int dim, int kind, const SampleReduceArgs& a, double* partials, hipStream_t s)
{
    const int nb = kReduceBlocks;
#define EGGMI_RS(D, K) reduce_samples<D, K><<<nb, kReduceThreads, 0, s>>>(a, partials)
    if (dim == 2) {
        if (kind == kUntangle)
            EGGMI_RS(2, kUntangle);
        else if (kind == kShapeSize)
            EGGMI_RS(2, kShapeSize);
        else
            EGGMI_RS(2, kShape);
    } else {
        if (kind == kUntangle)
            EGGMI_RS(3, kUntangle);
        else if (kind == kShapeSize)
            EGGMI_RS(3, kShapeSize);
        else
            EGGMI_RS(3, kShape);
    }
#undef EGGMI_RS
}

void launch_reduce_finish(const double* partials, int nparts, double* out_e, double* out_m, hipStream_t s,
                          double* rep, int cap, int* count)
{
    reduce_finish<<<1, kReduceThreads, 0, s>>>(partials, nparts, out_e, out_m, rep, cap, count);
}

// Whole-sample directional energies at NA step lengths (reference
// src/directional.hpp:279-345, added to the reported energy after the shape
// total, src/sweep.hpp:1921, and to the FAS line-search energies,
// src/fas.hpp:233). Positions are formed in FP32 as cell_trial forms them,
// the energy evaluated and accumulated in FP64; the min det slot is +inf.
// MODE 0: alpha 0; 1: a0, a1; 2: 0 then the ladder's alphas.
template <int D, int MODE>
__global__ __launch_bounds__(kReduceThreads) void dir_reduce(const DirView v, const float* x, long long plane,
                                                             const float* corr, float a0, float a1, double* partials)
{
    constexpr int NA = MODE == 0 ? 1 : MODE == 1 ? 2 : kLadder + 1;
    double e[NA];
#pragma unroll
    for (int t = 0; t < NA; ++t) e[t] = 0.0;
    for (int i = blockIdx.x * kReduceThreads + threadIdx.x; i < v.n; i += gridDim.x * kReduceThreads) {
        const int kind = v.kind[i], ar = dir_arity(kind);
        if (!ar) continue;
#pragma unroll
        for (int t = 0; t < NA; ++t) {
            const float al = MODE == 0 ? 0.0f : MODE == 1 ? (t == 0 ? a0 : a1) : (t == 0 ? 0.0f : ladder_alpha(t));
            double p[3][D] = {};
            for (int m = 0; m < ar; ++m)
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const long long q = c * plane + v.nodes[4 * i + m];
                    p[m][c] = (double)(corr ? x[q] + al * corr[q] : x[q]);
                }
            e[t] += dir_energy<D, double>(kind, p, v.ref + D * i, (double)v.w[i], (double)v.eps);
        }
    }
    for (int t = 0; t < NA; ++t) {
        block_reduce_store(e[t], __builtin_huge_val(), partials + 2 * NA * blockIdx.x + 2 * t);
        __syncthreads();
    }
}

template <int D>
static void dir_reduce_d(int mode, int nb, const DirView& v, const float* x, long long plane, const float* corr,
                         float a0, float a1, double* partials, int pbase, hipStream_t s)
{
    if (mode == 0)
        dir_reduce<D, 0><<<nb, kReduceThreads, 0, s>>>(v, x, plane, corr, a0, a1, partials + 2 * pbase);
    else if (mode == 1)
        dir_reduce<D, 1><<<nb, kReduceThreads, 0, s>>>(v, x, plane, corr, a0, a1, partials + 4 * pbase);
    else
        dir_reduce<D, 2><<<nb, kReduceThreads, 0, s>>>(v, x, plane, corr, a0, a1,
                                                       partials + 2 * (kLadder + 1) * pbase);
}

int launch_dir_reduce(int dim, int mode, const DirView& v, const float* x, long long plane, const float* corr,
                      float a0, float a1, double* partials, int pbase, hipStream_t s)
{
    if (v.n <= 0) return 0;
    int nb = (v.n + kReduceThreads - 1) / kReduceThreads;
    if (nb > kReduceBlocks / 2) nb = kReduceBlocks / 2;
    if (dim == 2)
        dir_reduce_d<2>(mode, nb, v, x, plane, corr, a0, a1, partials, pbase, s);
    else
        dir_reduce_d<3>(mode, nb, v, x, plane, corr, a0, a1, partials, pbase, s);
    return nb;
}

}  // namespace eggmi
