// Dimension-generic HIP sweep kernel of the structured TMOP smoother (gfx950).
//
// Work decomposition. One sweep is one launch. Workgroups [0, n_tiles) each
// own a tile of nodes of one block and stage the tile plus a one-node halo of
// every coordinate plane in LDS (coalesced row loads; the halo reaches into
// the ghost layer of the padded block, so face DOFs of coarse levels work as
// well as interior ones). Each node is worked on by 2^D lanes, one per cell
// of its patch: a lane evaluates the 2^D corner samples of its cell, and the
// node's gradient, contracted Hessian, patch energy and min det are reduced
// across the lane group with wavefront shuffles. Newton step, <= 10
// halvings and the omega post-blend then run redundantly on every lane of
// the group (identical values, so the accept decision is uniform), and lane
// 0 writes the node and pushes it to every ghost / copy / mirror slot that
// replicates it (the reference's per-sweep halo exchange, done by the
// writer). The remaining workgroups give each stored-stencil DOF (block
// faces, interfaces, singular corners) a group of G lanes, one per
// occurrence, with the same shuffle reductions.
//
// Semantics per DOF follow the reference's metric + interior update
// kernels (src/sweep.hpp:1425-1650): frozen snapshot `xin`, accept iff the
// trial energy is finite, <= e0 + 1e-6 and (barrier phases) min det > 0.
// Sample order inside the reductions differs from the reference's
// occurrence order (floating-point summation order only).
#pragma once

#include "directional.hpp"
#include "entity.hpp"
#include "mlp.hpp"
#include "sweep_common.hpp"

namespace eggmi
{

template <int D> struct Vec {
    float v[D];
};

// Reductions inside aligned groups of W lanes (W = 4, 8, 16 or 64) with DPP
// lane moves (a VALU modifier, no LDS round trip). Every step pairs lanes
// symmetrically (lane a takes b's value exactly when b takes a's), so by
// commutativity every lane of the group ends with the bitwise-same result:
//   W >= 16: row_ror 8 (= xor 8), then row_ror 4 (= xor 4 once the values
//            repeat with period 8); then quad_perm xor 1 and xor 2;
//   W == 8:  the two quad sums swapped by row_half_mirror;
//   W == 64: the four 16-lane row sums combined by xor 16 and xor 32.
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
constexpr int kDppRor8 = 0x128, kDppRor4 = 0x124, kDppQuadXor1 = 0xB1, kDppQuadXor2 = 0x4E,
              kDppHalfMirror = 0x141;

template <int W, class Op> __device__ __forceinline__ float group_reduce(float v, Op op)
{
    static_assert(W == 1 || W == 2 || W == 4 || W == 8 || W == 16 || W == 64, "group width");
    if constexpr (W == 1) return v;
    if constexpr (W == 2) return op(v, dpp_mov<kDppQuadXor1>(v));
    if constexpr (W >= 16) {
        v = op(v, dpp_mov<kDppRor8>(v));
        v = op(v, dpp_mov<kDppRor4>(v));
    }
    v = op(v, dpp_mov<kDppQuadXor1>(v));
    v = op(v, dpp_mov<kDppQuadXor2>(v));
    if constexpr (W == 8) v = op(v, dpp_mov<kDppHalfMirror>(v));
    if constexpr (W == 64) {
        v = op(v, __shfl_xor(v, 16, 64));
        v = op(v, __shfl_xor(v, 32, 64));
    }
    return v;
}
template <int W> __device__ __forceinline__ float group_sum(float v)
{
    return group_reduce<W>(v, [](float x, float y) { return x + y; });
}
template <int W> __device__ __forceinline__ float group_min(float v)
{
    return group_reduce<W>(v, [](float x, float y) { return fminf(x, y); });
}

// Metric value, gradient and contracted Hessian (upper triangle, row-major)
// of one sample with matrix T along u; returns mu.
template <int D, int KIND>
__device__ __forceinline__ float sample_grad_hess(const float* T, const float* u, float delta, float* g, float* h)
{
    if constexpr (D == 2) return mu2_grad_hess<KIND, float>(T[0], T[1], T[2], T[3], u[0], u[1], delta, g[0], g[1], h[0], h[1], h[2]);
    else return mu3_grad_hess<KIND, float>(T, u, delta, g, h);
}
template <int D, int KIND> __device__ __forceinline__ float sample_value(const float* T, float delta)
{
    if constexpr (D == 2) return mu2_value<KIND, float>(T[0], T[1], T[2], T[3], delta);
    else return mu3_value<KIND, float>(T, delta);
}
template <int D> __device__ __forceinline__ float detD(const float* A)
{
    if constexpr (D == 2) return A[0] * A[3] - A[1] * A[2];
    else return det3(A);
}

template <int D> constexpr int kNH = D * (D + 1) / 2;

// Newton step with the reference fallbacks from the reduced gradient/Hessian.
template <int D> __device__ __forceinline__ void newton(const float* g, const float* h, float* dl)
{
    if constexpr (D == 2) {
        solve2<float>(h[0], h[1], h[1], h[2], g[0], g[1], dl[0], dl[1]);
    } else {
        const float H[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
        solve3<float>(H, g, dl);
    }
}

// Tangent-reduced Newton step delta = B y from the orthonormal basis B[0..K-1]
// (K = 1 or 2; src/patch.hpp:348-388): M = B^T H B, r = B^T g, y the
// reference's solve of M y = -r (1x1: src/solve.hpp:83-90; 2x2 with the
// same fallbacks as the node solve).
template <int D>
__device__ __forceinline__ void reduced_step(const float* g, const float* h, const float (*B)[3], int K, float* dl)
{
    float H[D][D];
    if constexpr (D == 2) {
        H[0][0] = h[0], H[0][1] = H[1][0] = h[1], H[1][1] = h[2];
    } else {
        H[0][0] = h[0], H[0][1] = H[1][0] = h[1], H[0][2] = H[2][0] = h[2];
        H[1][1] = h[3], H[1][2] = H[2][1] = h[4], H[2][2] = h[5];
    }
    float M[2][2] = {}, r[2] = {};
    for (int c = 0; c < K; ++c) {
        float Hb[D];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < D; ++j) s += H[i][j] * B[c][j];
            Hb[i] = s;
        }
        for (int e = 0; e < K; ++e) {
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < D; ++i) s += B[e][i] * Hb[i];
            M[c][e] = s;
        }
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < D; ++i) s += B[c][i] * g[i];
        r[c] = s;
    }
    float y[2] = {0.0f, 0.0f};
    if (K == 1) {
        const float rn = fabsf(r[0]);
        if (rn >= EntTol<float>::znorm) {
            y[0] = -r[0] / M[0][0];
            if (!isfinite(y[0])) y[0] = -0.1f * r[0] / rn;
        }
    } else {
        solve2<float>(M[0][0], M[0][1], M[1][0], M[1][1], r[0], r[1], y[0], y[1]);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) dl[i] = B[0][i] * y[0] + (K == 2 ? B[1][i] * y[1] : 0.0f);
}

// Write v into the replica slots of a node (its ghost, fan and share
// copies whose root is that node), when the level has a push table.
template <int D>
__device__ __forceinline__ void push_replicas(const SweepArgs& a, float* out, long long node, const float* v)
{
    if (!a.cp_off) return;
    const int q0 = a.cp_off[node], q1 = a.cp_off[node + 1];
    for (int q = q0; q < q1; ++q) {
        const int dst = a.cp_dst[q];
#pragma unroll
        for (int k = 0; k < D; ++k) out[k * a.plane + dst] = v[k];
    }
}

// Write the accepted node and push it to its replicas.
template <int D>
__device__ __forceinline__ void store_node(const SweepArgs& a, long long node, const float* cur)
{
#pragma unroll
    for (int k = 0; k < D; ++k) a.xout[k * a.plane + node] = cur[k];
    push_replicas<D>(a, a.xout, node, cur);
}

// Diagnostic counters (off when null): spread over kCounterSlots cache lines
// by workgroup so that the atomics do not serialise on one address.
__device__ __forceinline__ unsigned long long* counter_slot(const SweepArgs& a)
{
    return a.counters + (blockIdx.x & (kCounterSlots - 1)) * kCounterStride;
}
__device__ __forceinline__ void count_update(const SweepArgs& a, int ntr, bool proposal = false)
{
    if (a.counters) {
        unsigned long long* c = counter_slot(a);
        atomicAdd(c, (unsigned long long)ntr);
        atomicAdd(c + 1, 1ull);
        if (proposal) atomicAdd(c + 2, 1ull);
    }
}

// ------------------------------------------------------------------ stored

// Stored-stencil DOF d handled by a group of G lanes (lane l owns
// occurrences l, l+G, ... of the DOF's patch).
// The entities a constrained boundary DOF slides on: entity.hpp.

// Block-interface C2 curvature window p0..p3 (reference src/curvature.hpp):
// residual r = turn(p0, p1, p2) - turn(p1, p2, p3) of the signed turning
// angles and, with grad, dr/dp_m = perp(e)/|e|^2 combinations per edge.
__device__ __forceinline__ float curv_turn(const float* a, const float* b, const float* c)
{
    const float e1x = b[0] - a[0], e1y = b[1] - a[1], e2x = c[0] - b[0], e2y = c[1] - b[1];
    return atan2f(e1x * e2y - e1y * e2x, e1x * e2x + e1y * e2y);
}
__device__ __forceinline__ float curv_residual(const float (&p)[4][2], float (&dr)[4][2])
{
    float gx[3], gy[3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const float ex = p[e + 1][0] - p[e][0], ey = p[e + 1][1] - p[e][1];
        const float l = ex * ex + ey * ey;
        const float il = l > 0.0f ? 1.0f / l : 0.0f;
        gx[e] = -ey * il;
        gy[e] = ex * il;
    }
    dr[0][0] = gx[0];
    dr[0][1] = gy[0];
    dr[1][0] = -gx[0] - 2.0f * gx[1];
    dr[1][1] = -gy[0] - 2.0f * gy[1];
    dr[2][0] = 2.0f * gx[1] + gx[2];
    dr[2][1] = 2.0f * gy[1] + gy[2];
    dr[3][0] = -gx[2];
    dr[3][1] = -gy[2];
    return curv_turn(p[0], p[1], p[2]) - curv_turn(p[1], p[2], p[3]);
}

// WT: per-occurrence energy weights (st.wt); a separate instantiation, so the
// unweighted sums keep their exact operation order. ENT: d is a constrained
// boundary DOF (2D; stored index st.n + i slides on entity i): tangent-reduced
// Newton step damped by omega, trials projected onto the entity, no
// post-blend (reference src/sweep.hpp:1285-1403, :1848-1871).
// Where a stored DOF's positions come from and where its result goes. The
// one-sweep kernels read the snapshot `xin` and write the node (and its
// replicas) into `xout`; the multi-sweep kernel substitutes its LDS staging.
struct StoreIO {
    __device__ const StoredView& st(const SweepArgs& a) const { return a.st; }
    __device__ const float* X(const SweepArgs& a) const { return a.xin; }
    __device__ long long PL(const SweepArgs& a) const { return a.plane; }
    // store slot of stored DOF d (at position `dof` of X)
    __device__ long long slot(int, int dof) const { return dof; }
    template <int D> __device__ void finish(const SweepArgs& a, int, int dof, const float* cur, int ntr) const
    {
        store_node<D>(a, dof, cur);
        count_update(a, ntr);
    }
};

template <int D, int KIND, bool TAU, bool GRAD, int G, bool WT, bool ENT = false, class IO = StoreIO>
__device__ void stored_group(const SweepArgs& a, int d, int l, bool valid, const IO& io = IO {})
{
    static_assert(!ENT || !TAU, "special DOFs: fine level");
    static_assert(G >= 4 && G <= stored_lanes(D), "stored lane group size");
    const StoredView& st = io.st(a);
    const float* X = io.X(a);
    const long long PL = io.PL(a);
    const int R = st.rounds;
    const int dof = valid ? st.dof[d] : 0;
    float pos[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pos[k] = valid ? X[k * PL + dof] : 0.0f;

    // One occurrence: its corner and neighbour positions and sign flips,
    // which of them is the DOF itself, role and W row. The lane's first
    // occurrence (r = 0) is loaded once and kept in registers for the
    // line-search trials.
    struct Occ {
        float cn[D], nb[D][D], s[D];
        bool cd, nd[D], live;
        int role;
        const float* w;
        float wt;  // energy weight: scales mu's value, gradient and Hessian, not det
    };
    auto load = [&](int r, Occ& o) {
        const long long q = ((long long)d * R + r) * G + l;
        const int m = valid ? st.meta[q] : kStoredPad;
        o.live = !(m & kStoredPad);
        o.role = (m & 7) - 1;
        o.w = st.W ? st.W + (long long)(D * D) * q : a.Wu;
        o.wt = 1.0f;
        if (!o.live) return;
        if constexpr (WT) o.wt = st.wt[q];
        const int ic = st.gc[q];
        o.cd = ic == dof;
#pragma unroll
        for (int i = 0; i < D; ++i) o.cn[i] = X[i * PL + ic];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int in = st.gn[k * st.total + q];
            o.nd[k] = in == dof;
            o.s[k] = (m & (8 << k)) ? -1.0f : 1.0f;
#pragma unroll
            for (int i = 0; i < D; ++i) o.nb[k][i] = X[i * PL + in];
        }
    };
    // builds T (and det A) of an occurrence, with the DOF moved to `sub` when given
    auto build = [&](const Occ& o, const float* sub, float* T, float& det) {
        float cn[D], nb[D][D];
#pragma unroll
        for (int i = 0; i < D; ++i) cn[i] = (sub && o.cd) ? sub[i] : o.cn[i];
#pragma unroll
        for (int k = 0; k < D; ++k)
#pragma unroll
            for (int i = 0; i < D; ++i) nb[k][i] = (sub && o.nd[k]) ? sub[i] : o.nb[k][i];
        float A[D * D];
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) A[i * D + k] = o.s[k] * (nb[k][i] - cn[i]);
        det = detD<D>(A);
        const float* w = o.w;
#pragma unroll
        for (int i = 0; i < D; ++i)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float acc = 0.0f;
#pragma unroll
                for (int j = 0; j < D; ++j) acc += A[i * D + j] * w[j * D + k];
                T[i * D + k] = acc;
            }
    };

    // special DOFs (ENT): tag != 0 slides on an entity (its blob at
    // con_blob + con_off); tag 0 is free and carries curvature windows, lane
    // l taking windows l, l + G, ...
    int etag = 0;
    const float* eb = nullptr;
    const int ci = d - st.n;
    if constexpr (ENT) {
        if (valid) {
            etag = a.con_tag[ci];
            if (etag) eb = a.con_blob + a.con_off[ci];
        }
    }
    const bool curv = D == 2 && ENT && valid && etag == 0 && a.curv_k > 0;  // entities and curvature: 2D
    // directional windows, on free and constrained special DOFs alike
    // (reference src/sweep.hpp:1488-1510, :1746-1760), lane l taking l, l + G, ...
    const bool dirw = ENT && valid && a.dir.n > 0;
    // positions of window w's nodes, the DOF's own at `sub` when given
    auto dir_load = [&](long long w, int sl, const float* sub, float (&p)[3][D]) {
        const int ar = dir_arity(a.dir.kind[w]);
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
            for (int i = 0; i < D; ++i)
                p[m][i] = m >= ar ? 0.0f : (sub && m == sl) ? sub[i] : X[i * PL + a.dir.nodes[4 * w + m]];
        return ar;
    };
    Occ o0;
    load(0, o0);
    float g[D] = {}, h[kNH<D>] = {}, e = 0.0f, md = kInf;
    for (int r = 0; r < R; ++r) {
        Occ o = o0;
        if (r > 0) load(r, o);
        if (!o.live) break;  // padding (a DOF's occurrences come first)
        float T[D * D], det;
        build(o, nullptr, T, det);
        md = fminf(md, det);
        const int role = o.role;
        if (role >= 0) {
            const float* w = o.w;
            float sr = o.s[0];  // sign of the role's axis (selected, not indexed)
#pragma unroll
            for (int k = 1; k < D; ++k)
                if (role - 1 == k) sr = o.s[k];
            float u[D];
#pragma unroll
            for (int c = 0; c < D; ++c) {
                if (role == 0) {
                    float acc = 0.0f;
#pragma unroll
                    for (int k = 0; k < D; ++k) acc += o.s[k] * w[k * D + c];
                    u[c] = -acc;
                } else {
                    u[c] = sr * w[(role - 1) * D + c];
                }
            }
            if constexpr (WT) {  // weighted occurrences (reference src/patch.hpp:250-292)
                float gw[D] = {}, hw[kNH<D>] = {};
                const float v = sample_grad_hess<D, KIND>(T, u, a.delta, gw, hw);
                e += o.wt * v;
#pragma unroll
                for (int k = 0; k < D; ++k) g[k] += o.wt * gw[k];
#pragma unroll
                for (int k = 0; k < kNH<D>; ++k) h[k] += o.wt * hw[k];
            } else {
                e += sample_grad_hess<D, KIND>(T, u, a.delta, g, h);
            }
        } else {
            if constexpr (WT)
                e += o.wt * sample_value<D, KIND>(T, a.delta);
            else
                e += sample_value<D, KIND>(T, a.delta);
        }
    }
    if constexpr (ENT) {
        if constexpr (D == 2)
        if (curv)  // curvature on top of the shape metric (reference src/sweep.hpp:1461-1487)
            for (int j = l; j < a.curv_k; j += G) {
                const long long w = (long long)ci * a.curv_k + j;
                const int sl = a.curv_slot[w];
                if (sl < 0) continue;
                float p[4][2], dr[4][2];
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int i = 0; i < 2; ++i) p[m][i] = X[i * PL + a.curv_nodes[4 * w + m]];
                const float r = curv_residual(p, dr);
                const float wt = a.curv_w[w], gc = 2.0f * wt * r, hc = 2.0f * wt;
                float dsx = 0.0f, dsy = 0.0f, off = 0.0f;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const float nm = sqrtf(dr[m][0] * dr[m][0] + dr[m][1] * dr[m][1]);
                    if (m == sl) {
                        dsx = dr[m][0];
                        dsy = dr[m][1];
                    } else {
                        off += nm;
                    }
                }
                // Gauss-Newton Hessian plus the diagonal stabiliser 2 w |dr_s| sum |dr_m|
                const float stab = hc * sqrtf(dsx * dsx + dsy * dsy) * off;
                g[0] += gc * dsx;
                g[1] += gc * dsy;
                h[0] += hc * dsx * dsx + stab;
                h[1] += hc * dsx * dsy;
                h[2] += hc * dsy * dsy + stab;
                e += wt * r * r;
            }
        if (dirw)
            for (int j = l; j < a.dir.n; j += G) {
                const long long w = (long long)ci * a.dir.n + j;
                const int sl = a.dir.slot[w];
                float p[3][D];
                if (sl < 0 || !dir_load(w, sl, nullptr, p)) continue;
                dir_accumulate<D>(a.dir.kind[w], p, a.dir.ref + D * w, a.dir.w[w], a.dir.eps, sl, g, h, e);
            }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) g[k] = group_sum<G>(g[k]);
    float e0 = group_sum<G>(e);
    if (!valid) return;
    float tau[D] = {};
    if constexpr (TAU) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            tau[k] = a.tau[k * a.plane + io.slot(d, dof)];
            g[k] -= tau[k];
            e0 -= tau[k] * pos[k];
        }
    }
    if constexpr (GRAD) {
        if (l == 0)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float& o = a.gout[k * a.plane + dof];
                o = a.gsub ? g[k] - o : g[k];
            }
        if (l == 0 && a.gpush) push_replicas<D>(a, a.gout, dof, g);
        return;
    }
#pragma unroll
    for (int k = 0; k < kNH<D>; ++k) h[k] = group_sum<G>(h[k]);
    float dl[D];
    // constrained DOF state: the foot parameter the last projection left
    // (the warm start of the iterative kinds) in registers, identical in
    // every lane of the group; pls: a warm sweep on a parametrization
    // line-searches in its parameter from the foot qs along dq
    float sd[2] = {0.0f, 0.0f}, qs[2] = {0.0f, 0.0f}, dq[2] = {0.0f, 0.0f};
    bool pls = false;
    if constexpr (ENT) {
        if (etag != 0) {
            sd[0] = a.con_seed[2 * ci];
            sd[1] = a.con_seed[2 * ci + 1];
            // the first update of a DOF in a run is cold (world-space line
            // search, reprojected trials); later ones are warm
            // (src/sweep.hpp:1938-1984)
            const bool cold = a.con_epoch[ci] != *a.run_epoch;
            const int K = ent_tdim(etag);
            float B[2][3] = {};
            bool release = false;
            if (!cold && ent_is_param(etag)) {
                // warm: project once to the foot, reduce the tangent step to
                // the parameter, dq = (J^T J)^-1 J^T delta, damped by omega
                // (src/sweep.hpp:1786-1820)
                pls = true;
                const float* f = eb + 1;
                float J[2][3] = {};
                par_invert(etag, f, pos, sd, std::isfinite(sd[0]), true, qs,
                           etag == kEntBSplineSurf ? J : nullptr);
                if (etag != kEntBSplineSurf) par_frame(etag, f, qs, J);
                sd[0] = qs[0];
                sd[1] = qs[1];
#pragma unroll
                for (int c = 0; c < 2; ++c)
#pragma unroll
                    for (int k = 0; k < 3; ++k) B[c][k] = J[c][k];
                ent_orthonormalize(B, K, D);
                float del[D];
                reduced_step<D>(g, h, B, K, del);
                float JtJ[4] = {}, Jtd[2] = {};
                for (int c = 0; c < K; ++c) {
                    for (int e = 0; e < K; ++e) JtJ[c * 2 + e] = ent_dot(J[c], J[e], D);
                    Jtd[c] = ent_dot(J[c], del, D);
                }
                if (K == 1) {
                    dq[0] = fabsf(JtJ[0]) > EntTol<float>::tiny ? Jtd[0] / JtJ[0] : 0.0f;
                } else {
                    const float det = JtJ[0] * JtJ[3] - JtJ[1] * JtJ[2];
                    if (fabsf(det) >= EntTol<float>::tiny) {
                        const float inv = 1.0f / det;
                        dq[0] = inv * (JtJ[3] * Jtd[0] - JtJ[1] * Jtd[1]);
                        dq[1] = inv * (-JtJ[2] * Jtd[0] + JtJ[0] * Jtd[1]);
                    }
                }
                dq[0] *= a.omega;
                dq[1] *= a.omega;
            } else if (ent_is_rail(etag)) {
                // a rail beyond its ends releases the node: full-space step
                // (src/patch.hpp:400-409), still damped on the step
                const float ft = rail_foot(eb + 1, pos, D);
                release = ft < 0.0f || ft > 1.0f;
                if (!release) closed_tangent(etag, eb + 1, pos, B[0]);
            } else if (ent_is_seeded(etag)) {
                seeded_frame(etag, eb, pos, sd, false, (float*)nullptr, B);
            } else {
                float foot[3];
                ent_frame(etag, eb, pos, foot, B);
            }
            if (!pls) {
                if (release)
                    newton<D>(g, h, dl);
                else
                    reduced_step<D>(g, h, B, K, dl);
#pragma unroll
                for (int k = 0; k < D; ++k) dl[k] *= a.omega;
            }
        } else {
            newton<D>(g, h, dl);
        }
    } else {
        newton<D>(g, h, dl);
    }
    float alpha = 1.0f, cur[D];
    int ntr = 10;  // trials evaluated (10 = all rejected)
#pragma unroll
    for (int k = 0; k < D; ++k) cur[k] = pos[k];
    for (int it = 0; it < 10; ++it) {
        float tr[D];
#pragma unroll
        for (int k = 0; k < D; ++k) tr[k] = pos[k] + alpha * dl[k];
        if constexpr (ENT) {
            if (etag != 0) {
                if (pls) {  // the trial is the point at the clamped parameter
                    float qt[2] = {qs[0] + alpha * dq[0], qs[1] + alpha * dq[1]};
                    par_clamp(etag, eb + 1, qt);
                    par_eval(etag, eb + 1, qt, tr);
                } else {
                    float raw[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) raw[k] = tr[k];
                    if (ent_is_seeded(etag))
                        seeded_frame(etag, eb, raw, sd, false, tr, (float(*)[3]) nullptr);
                    else
                        ent_frame(etag, eb, raw, tr, (float(*)[3]) nullptr);
                }
            }
        }
        float et = 0.0f, mt = kInf;
        if constexpr (ENT) {
            if constexpr (D == 2)
            if (curv)  // the windows' energy with the DOF moved to the trial
                for (int j = l; j < a.curv_k; j += G) {
                    const long long w = (long long)ci * a.curv_k + j;
                    const int sl = a.curv_slot[w];
                    if (sl < 0) continue;
                    float p[4][2];
#pragma unroll
                    for (int m = 0; m < 4; ++m)
#pragma unroll
                        for (int i = 0; i < 2; ++i) p[m][i] = m == sl ? tr[i] : X[i * PL + a.curv_nodes[4 * w + m]];
                    const float r = curv_turn(p[0], p[1], p[2]) - curv_turn(p[1], p[2], p[3]);
                    et += a.curv_w[w] * r * r;
                }
            if (dirw)  // the directional windows' energy with the DOF at the trial
                for (int j = l; j < a.dir.n; j += G) {
                    const long long w = (long long)ci * a.dir.n + j;
                    const int sl = a.dir.slot[w];
                    float p[3][D];
                    if (sl < 0 || !dir_load(w, sl, tr, p)) continue;
                    et += dir_energy<D, float>(a.dir.kind[w], p, a.dir.ref + D * w, a.dir.w[w], a.dir.eps);
                }
        }
        for (int r = 0; r < R; ++r) {
            Occ o = o0;
            if (r > 0) load(r, o);
            if (!o.live) break;
            float T[D * D], det;
            build(o, tr, T, det);
            if constexpr (WT)
                et += o.wt * sample_value<D, KIND>(T, a.delta);
            else
                et += sample_value<D, KIND>(T, a.delta);
            mt = fminf(mt, det);
        }
        et = group_sum<G>(et);
        mt = group_min<G>(mt);
        if constexpr (TAU) {
#pragma unroll
            for (int k = 0; k < D; ++k) et -= tau[k] * tr[k];
        }
        if (accept<KIND>(et, e0, mt)) {
#pragma unroll
            for (int k = 0; k < D; ++k) cur[k] = tr[k];
            ntr = it + 1;
            break;
        }
        alpha *= 0.5f;
    }
    if (etag == 0 && a.omega != 1.0f)  // constrained DOFs take omega on the step instead
#pragma unroll
        for (int k = 0; k < D; ++k) cur[k] = pos[k] + a.omega * (cur[k] - pos[k]);
    if (l == 0) {
        io.template finish<D>(a, d, dof, cur, ntr);
        if constexpr (ENT) {
            if (etag != 0) {
                a.con_seed[2 * ci] = sd[0];
                a.con_seed[2 * ci + 1] = sd[1];
                a.con_epoch[ci] = *a.run_epoch;
            }
        }
    }
}

// The special stored DOFs of a fine-level sweep or gradient pass
// (constrained ones and those with curvature windows, 2D, or directional
// windows), in a launch of their own beside the sweep kernel: both read xin
// and write disjoint DOF slots.
template <int D, int KIND, int G, bool WT, bool GRAD>
__global__ __launch_bounds__(256) void constrained_kernel(const SweepArgs a)
{
    const int gid = int(blockIdx.x) * (256 / G) + int(threadIdx.x) / G;
    stored_group<D, KIND, false, GRAD, G, WT, true>(a, a.st.n + gid, int(threadIdx.x) % G, gid < a.con_n);
}

template <int D, int KIND, int G, bool GRAD> static void constrained_g(const SweepArgs& a, hipStream_t s)
{
    const unsigned grid = (unsigned)((a.con_n + 256 / G - 1) / (256 / G));
    if (a.st.wt)
        constrained_kernel<D, KIND, G, true, GRAD><<<dim3(grid), dim3(256), 0, s>>>(a);
    else
        constrained_kernel<D, KIND, G, false, GRAD><<<dim3(grid), dim3(256), 0, s>>>(a);
}

// ------------------------------------------------------------------ tiles

// Matrix A of permuted sample OP of a synthesized DOF's cell whose corners
// Pj are permuted so that corner 0 is the DOF (rhoj: the frame's axis sign
// flips); the DOF moved to x0 when given.
template <int D, class OPT>
__device__ __forceinline__ void synth_A(const float (&Pj)[1 << D][D], const float (&rhoj)[D], OPT, const float* x0,
                                       float* A)
{
    constexpr int op = OPT::value;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int nb = op ^ (1 << (D - 1 - k));
        const float sk = (((op >> (D - 1 - k)) & 1) ? -1.0f : 1.0f) * rhoj[k];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const float pn = (nb == 0 && x0) ? x0[i] : Pj[nb][i];
            const float pc = (op == 0 && x0) ? x0[i] : Pj[op][i];
            A[i * D + k] = sk * (pn - pc);
        }
    }
}

// Newton update of one synthesized interior DOF by its lane group (NL =
// 2^D / CPL lanes; lane gl holds cells gl * CPL .. of the patch, corners in
// P, P[0][0] = the DOF): gradient, contracted Hessian and e0 reduced over the
// group, Newton step, <= 10 halvings, omega post-blend. `node` is the DOF's
// store slot (FAS shift tau, gradient output). Returns the trials evaluated
// (10 = all rejected) with the result in cur; GRAD: writes the gradient
// (lane 0) and returns -1.
template <int D, int KIND, bool TAU, bool GRAD, int CPL>
__device__ __forceinline__ int synth_solve(const SweepArgs& a, const float (&P)[CPL][1 << D][D],
                                           const float (&rho)[CPL][D], int gl, long long node, float* cur)
{
    constexpr int NC = 1 << D;
    constexpr int NL = NC / CPL;
    const long long PL = a.plane;
    const float delta = a.delta;
    // sum over the node's cells in the order of the lane-group reduction
    // ((c0 + c1) + (c2 + c3)) + ...: a lane folds its CPL cells the same way,
    // then the group folds the lanes (the fold order of every CPL is the
    // same; the compiler's multiply-add fusion inside a cell may differ)
    auto cell_sum = [&](const float* x) -> float {
        if constexpr (CPL == 1) return group_sum<NL>(x[0]);
        else if constexpr (CPL == 2) return group_sum<NL>(x[0] + x[1]);
        else if constexpr (CPL == 4) return group_sum<NL>((x[0] + x[1]) + (x[2] + x[3]));
        else return group_sum<NL>(((x[0] + x[1]) + (x[2] + x[3])) + ((x[4] + x[5]) + (x[6] + x[7])));
    };
    // per cell: gradient and Hessian of the samples touching the DOF (em),
    // value of the frozen ones (ec) and their min det (mc)
    float gs[CPL][D] = {}, hs[CPL][kNH<D>] = {}, em[CPL], ec[CPL], mc = kInf;
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
        em[j] = 0.0f;
        ec[j] = 0.0f;
        static_for<NC>([&](auto OP) {
            constexpr int op = decltype(OP)::value;
            constexpr int pc = __builtin_popcount(op);
            float A[D * D];
            synth_A<D>(P[j], rho[j], OP, nullptr, A);
            const float det = detD<D>(A);
            if constexpr (pc <= 1) {  // the DOF is the corner (op = 0) or an axis neighbour
                float u[D];
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const float sc = (((op >> (D - 1 - c)) & 1) ? -1.0f : 1.0f) * rho[j][c];
                    if constexpr (op == 0)
                        u[c] = -sc;
                    else
                        u[c] = (op == (1 << (D - 1 - c))) ? sc : 0.0f;
                }
                em[j] += sample_grad_hess<D, KIND>(A, u, delta, gs[j], hs[j]);
            } else {
                ec[j] += sample_value<D, KIND>(A, delta);
                mc = fminf(mc, det);
            }
        });
    }
    float g[D], h[kNH<D>];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        float x[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) x[j] = gs[j][k];
        g[k] = cell_sum(x);
    }
    float e0;
    {
        float x[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) x[j] = ec[j] + em[j];
        e0 = cell_sum(x);
    }
    float pos[D];
#pragma unroll
    for (int k = 0; k < D; ++k) pos[k] = P[0][0][k];
    float tau[D] = {};
    if constexpr (TAU) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            tau[k] = a.tau[k * PL + node];
            g[k] -= tau[k];
            e0 -= tau[k] * pos[k];
        }
    }
    if constexpr (GRAD) {
        if (gl == 0)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float& o = a.gout[k * PL + node];
                o = a.gsub ? g[k] - o : g[k];
            }
        if (gl == 0 && a.gpush) push_replicas<D>(a, a.gout, node, g);
        return -1;
    }
#pragma unroll
    for (int k = 0; k < kNH<D>; ++k) {
        float x[CPL];
#pragma unroll
        for (int j = 0; j < CPL; ++j) x[j] = hs[j][k];
        h[k] = cell_sum(x);
    }
    float dl[D];
    newton<D>(g, h, dl);
    float alpha = 1.0f;
    int ntr = 10;  // trials evaluated (10 = all rejected)
#pragma unroll
    for (int k = 0; k < D; ++k) cur[k] = pos[k];
    for (int it = 0; it < 10; ++it) {
        float tr[D];
#pragma unroll
        for (int k = 0; k < D; ++k) tr[k] = pos[k] + alpha * dl[k];
        float ets[CPL], mt = mc;
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
            float et = 0.0f;
            static_for<NC>([&](auto OP) {
                constexpr int op = decltype(OP)::value;
                if constexpr (__builtin_popcount(op) <= 1) {
                    float A[D * D];
                    synth_A<D>(P[j], rho[j], OP, tr, A);
                    et += sample_value<D, KIND>(A, delta);
                    mt = fminf(mt, detD<D>(A));
                }
            });
            ets[j] = ec[j] + et;
        }
        float et = cell_sum(ets);
        mt = group_min<NL>(mt);
        if constexpr (TAU) {
#pragma unroll
            for (int k = 0; k < D; ++k) et -= tau[k] * tr[k];
        }
        if (accept<KIND>(et, e0, mt)) {
#pragma unroll
            for (int k = 0; k < D; ++k) cur[k] = tr[k];
            ntr = it + 1;
            break;
        }
        alpha *= 0.5f;
    }
    if (a.omega != 1.0f)
#pragma unroll
        for (int k = 0; k < D; ++k) cur[k] = pos[k] + a.omega * (cur[k] - pos[k]);
    return ntr;
}

// Corner permutation of cell cj of a synthesized DOF's patch: the frame's
// axis sign flips rho and, per permuted corner op, its offset loc[k] in
// {0, 1, 2} from the 3^D neighbourhood's low corner (the DOF at 1).
template <int D, class F> __device__ __forceinline__ void synth_corners(int cj, float (&rho)[D], F&& corner)
{
    constexpr int NC = 1 << D;
    int cb[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < D; ++k) cb[k] = (cj >> (D - 1 - k)) & 1;
#pragma unroll
    for (int k = 0; k < D; ++k) rho[k] = cb[k] ? 1.0f : -1.0f;  // r_k = 1 - cb_k; rho = r_k ? -1 : 1
#pragma unroll
    for (int op = 0; op < NC; ++op) {
        int loc[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const int obit = (op >> (D - 1 - k)) & 1;
            const int o = obit ^ (1 - cb[k]);  // actual corner bit
            loc[k] = cb[k] + o;
        }
        corner(op, loc);
    }
}

// Tile of TX (last axis) x TY x TZ nodes; 2^D lanes per node.
// LDS = false: no staging; each lane reads its cell's corners straight from
// the store (cached), with no workgroup barrier (an autotuner candidate for
// latency-bound grids).
template <int D, int KIND, int TX, int TY, int TZ, bool TAU, bool GRAD, bool PROP, bool LDS, int CPL>
__device__ void tile_sweep(const SweepArgs& a, int tile)
{
    static_assert(LDS || !PROP, "the proposal reads the one-ring from LDS");
    static_assert(CPL == 1 || !PROP, "the proposal's matrix-core tiles take one lane per cell");
    constexpr int NC = 1 << D;
    constexpr int NL = NC / CPL;  // lanes per node
    constexpr int NT = NL * TX * TY * TZ;
    constexpr int LX = TX + 2, LY = TY + 2, LZ = (D == 3) ? TZ + 2 : 1;
    __shared__ float sh[D][LDS ? LZ * LY * LX : 1];
    const TileDesc td = a.tiles[tile];
    const BlockDesc& bd = td.b;
    const long long PL = a.plane;
    const int tid = threadIdx.x;

    // this lane's node and its mask first: the mask load then overlaps the
    // tile staging instead of following the barrier
    const int gl = tid & (NL - 1);  // lane within the node's group: cells gl * CPL + j
    const int cell = gl;            // (the proposal block runs with CPL == 1)
    const int nid = tid / NL;
    const int tx = nid % TX, ty = (nid / TX) % TY, tz = nid / (TX * TY);
    int lg[3];
    long long node;
    bool inside;
    if constexpr (D == 2) {
        lg[0] = td.i0 + ty;
        lg[1] = td.j0 + tx;
        lg[2] = 0;
        inside = !(lg[0] > bd.n[0] - 1 || lg[1] > bd.n[1] - 1);
        node = bd.base + (long long)(min(lg[0], bd.n[0] - 1) + 1) * bd.st[0] + (min(lg[1], bd.n[1] - 1) + 1);
    } else {
        lg[0] = td.i0 + tz;
        lg[1] = td.j0 + ty;
        lg[2] = td.k0 + tx;
        inside = !(lg[0] > bd.n[0] - 1 || lg[1] > bd.n[1] - 1 || lg[2] > bd.n[2] - 1);
        node = bd.base + (long long)(min(lg[0], bd.n[0] - 1) + 1) * bd.st[0] +
               (long long)(min(lg[1], bd.n[1] - 1) + 1) * bd.st[1] + (min(lg[2], bd.n[2] - 1) + 1);
    }
    const bool active = inside && a.mask[node];

    // ---- stage the tile + one-node halo (logical -1 .. n: padded ghosts included)
    if constexpr (LDS)
    for (int e = tid; e < LZ * LY * LX; e += NT) {
        const int c = e % LX, q = (e / LX) % LY, r = e / (LX * LY);
        long long node;
        if constexpr (D == 2) {
            const int li = min(td.i0 - 1 + q, bd.n[0]);
            const int lj = min(td.j0 - 1 + c, bd.n[1]);
            node = bd.base + (long long)(li + 1) * bd.st[0] + (lj + 1);
        } else {
            const int l0 = min(td.i0 - 1 + r, bd.n[0]);
            const int l1 = min(td.j0 - 1 + q, bd.n[1]);
            const int l2 = min(td.k0 - 1 + c, bd.n[2]);
            node = bd.base + (long long)(l0 + 1) * bd.st[0] + (long long)(l1 + 1) * bd.st[1] + (l2 + 1);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) sh[k][e] = a.xin[k * PL + node];
    }
    if constexpr (LDS) __syncthreads();

    // whole lane groups leave together; the proposal variant keeps every lane
    // until its matrix-core forward pass is done (converged wave)
    if constexpr (!PROP) {
        if (!active) return;
    }

    // ---- the lane's cells: corners permuted so that corner 0 is the DOF
    float rho[CPL][D];  // axis sign flips of each cell's permuted frame
    float P[CPL][NC][D];
#pragma unroll
    for (int j = 0; j < CPL; ++j)
        synth_corners<D>(gl * CPL + j, rho[j], [&](int op, const int* loc) {
            if constexpr (LDS) {
                int li;
                if constexpr (D == 2)
                    li = (ty + loc[0]) * LX + (tx + loc[1]);
                else
                    li = ((tz + loc[0]) * LY + (ty + loc[1])) * LX + (tx + loc[2]);
#pragma unroll
                for (int k = 0; k < D; ++k) P[j][op][k] = sh[k][li];
            } else {
                long long g = node;
#pragma unroll
                for (int k = 0; k < D; ++k) g += (long long)(loc[k] - 1) * (k == D - 1 ? 1 : bd.st[k]);
#pragma unroll
                for (int k = 0; k < D; ++k) P[j][op][k] = a.xin[k * PL + g];
            }
        });

    const float delta = a.delta;
    if constexpr (PROP) {
        // ---- learned proposal: features of the node's one-ring (LDS), the
        // wave's matrix-core forward pass over its 16 (2D) / 8 (3D) nodes,
        // accepted iff the patch energy drops and every det stays > 0
        using ML = MlpLayout<D>;
        constexpr int NODES = TX * TY * TZ;
        constexpr int PER_LANE = ML::INP / NC;
        __shared__ __bf16 feat[NODES][ML::INP];
        auto ring = [&](int r, int k) -> float {
            int li;
            if constexpr (D == 2)
                li = (ty + r / 3) * LX + (tx + r % 3);
            else
                li = ((tz + r / 9) * LY + (ty + (r / 3) % 3)) * LX + (tx + r % 3);
            return sh[k][li];
        };
        const Frame fr = local_frame<D>(ring);
        float dprev[D];
#pragma unroll
        for (int k = 0; k < D; ++k) dprev[k] = P[0][0][k] - a.xout[k * PL + node];
        // each lane of the group computes and stores its PER_LANE-wide slice
        const float rh = fr.ok ? __builtin_amdgcn_rcpf(fr.h) : 0.0f;
#pragma unroll
        for (int q = 0; q < PER_LANE; ++q) {
            const int j = cell * PER_LANE + q;
            feat[nid][j] = (__bf16)feature_rt<D>(ring, fr, dprev, rh, j);
        }
        __syncthreads();
        constexpr int WN = 64 / NC;  // nodes per wave
        const int lane = tid & 63, wave_node0 = (tid >> 6) * WN;
        const int col_node = wave_node0 + ((lane & 15) % WN);
        bf16x8 xin[ML::KS1];
#pragma unroll
        for (int ks = 0; ks < ML::KS1; ++ks)
#pragma unroll
            for (int i = 0; i < 8; ++i) xin[ks][i] = feat[col_node][ks * 32 + 8 * (lane >> 4) + i];
        const f32x4 y = mlp_forward<D, false>(a.mlp, xin);
        if (!active) return;
        float pl[D], pw[D];
#pragma unroll
        for (int k = 0; k < D; ++k) pl[k] = __shfl(y[k], nid - wave_node0, 64);
        to_world<D>(fr, pl, pw);
        float tr[D];
        bool finite_p = fr.ok;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            tr[k] = P[0][0][k] + a.prop_scale * pw[k];
            finite_p = finite_p && isfinite(tr[k]);
        }
        float e_now = 0.0f, e_try = 0.0f, m_try = kInf;
        static_for<NC>([&](auto OP) {
            float A[D * D];
            synth_A<D>(P[0], rho[0], OP, nullptr, A);
            e_now += sample_value<D, KIND>(A, delta);
            if constexpr (__builtin_popcount(decltype(OP)::value) <= 1) {
                if (finite_p) synth_A<D>(P[0], rho[0], OP, tr, A);
            }
            e_try += sample_value<D, KIND>(A, delta);
            m_try = fminf(m_try, detD<D>(A));
        });
        e_now = group_sum<NC>(e_now);
        e_try = group_sum<NC>(e_try);
        m_try = group_min<NC>(m_try);
        // "drops": by more than the FP32 energy slack plus the rounding noise
        // of a 2^D x 2^D-sample sum (32 ulp), so that a (near-)zero proposal
        // never counts as progress and never displaces the Newton step
        const float noise = Tol<float>::energy + 32.0f * 1.1920929e-7f * fabsf(e_now);
        const bool ok = finite_p && isfinite(e_try) && e_try < e_now - noise && m_try > 0.0f;
        if (ok) {
            if (a.prop_mode == 0) {  // replace: the proposal is the update
                if (cell == 0) {
                    store_node<D>(a, node, tr);
                    count_update(a, 0, true);
                }
                return;
            }
            // predict-correct: the plain Newton update continues from the
            // accepted proposal (corner 0 of every lane's cell is the DOF)
            if (cell == 0 && a.counters) atomicAdd(counter_slot(a) + 2, 1ull);
#pragma unroll
            for (int k = 0; k < D; ++k) P[0][0][k] = tr[k];
        }
    }

    float cur[D];
    const int ntr = synth_solve<D, KIND, TAU, GRAD, CPL>(a, P, rho, gl, node, cur);
    if constexpr (GRAD) return;
    if (gl == 0) {
        store_node<D>(a, node, cur);
        count_update(a, ntr);
    }
}

// ------------------------------------------------------------------ 2D samples

// 2D tile of TX x TY nodes with one lane per (cell, corner) sample: 16 lanes
// per node, so a wave holds 4 nodes and each lane's dependency chain is one
// sample long (the cell-per-lane tile runs four samples per lane). Same
// semantics and sample set as tile_sweep: lane q = 4 * cell + op evaluates
// the permuted corner op of its cell (corner 0 = the DOF); samples touching
// the DOF give value, gradient and contracted Hessian, the fourth is frozen.
template <int KIND, int TX, int TY, bool TAU, bool GRAD>
__device__ void sample_sweep(const SweepArgs& a, int tile)
{
    constexpr int D = 2, NL = 16;
    constexpr int NT = NL * TX * TY;
    constexpr int LX = TX + 2, LY = TY + 2;
    __shared__ float sh[D][LY * LX];
    const TileDesc td = a.tiles[tile];
    const BlockDesc& bd = td.b;
    const long long PL = a.plane;
    const int tid = threadIdx.x;
    for (int e = tid; e < LY * LX; e += NT) {
        const int c = e % LX, q = e / LX;
        const int li = min(td.i0 - 1 + q, bd.n[0]);
        const int lj = min(td.j0 - 1 + c, bd.n[1]);
        const long long nd = bd.base + (long long)(li + 1) * bd.st[0] + (lj + 1);
#pragma unroll
        for (int k = 0; k < D; ++k) sh[k][e] = a.xin[k * PL + nd];
    }
    __syncthreads();

    const int q = tid & (NL - 1);
    const int nid = tid / NL;
    const int tx = nid % TX, ty = nid / TX;
    const int lg0 = td.i0 + ty, lg1 = td.j0 + tx;
    if (lg0 > bd.n[0] - 1 || lg1 > bd.n[1] - 1) return;
    const long long node = bd.base + (long long)(lg0 + 1) * bd.st[0] + (lg1 + 1);
    if (!a.mask[node]) return;  // whole lane group leaves together

    const int cell = q >> 2, op = q & 3;
    const int cb0 = (cell >> 1) & 1, cb1 = cell & 1;
    const float rho0 = cb0 ? 1.0f : -1.0f, rho1 = cb1 ? 1.0f : -1.0f;
    auto lds_of = [&](int o) {
        const int o0 = ((o >> 1) & 1) ^ (1 - cb0), o1 = (o & 1) ^ (1 - cb1);
        return (ty + cb0 + o0) * LX + (tx + cb1 + o1);
    };
    const int ic = lds_of(op), i0 = lds_of(op ^ 2), i1 = lds_of(op ^ 1);
    const int inode = (ty + 1) * LX + (tx + 1);
    float pc[D], pn0[D], pn1[D], pos[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        pc[k] = sh[k][ic];
        pn0[k] = sh[k][i0];
        pn1[k] = sh[k][i1];
        pos[k] = sh[k][inode];
    }
    const float s0 = ((op >> 1) & 1 ? -1.0f : 1.0f) * rho0;
    const float s1 = ((op & 1) ? -1.0f : 1.0f) * rho1;
    const bool moves = op != 3;  // the sample has the DOF as corner or axis neighbour
    const float delta = a.delta;
    // A with the DOF moved to x0 (null: as is)
    auto build = [&](const float* x0, float* A) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const float c = (x0 && op == 0) ? x0[i] : pc[i];
            const float n0 = (x0 && op == 2) ? x0[i] : pn0[i];
            const float n1 = (x0 && op == 1) ? x0[i] : pn1[i];
            A[i * D + 0] = s0 * (n0 - c);
            A[i * D + 1] = s1 * (n1 - c);
        }
    };
    float A[D * D];
    build(nullptr, A);
    float g[D] = {}, h[kNH<D>] = {}, e_l, frozen_e = 0.0f, frozen_m = kInf;
    if (moves) {
        float u[D];
        u[0] = op == 0 ? -s0 : (op == 2 ? s0 : 0.0f);
        u[1] = op == 0 ? -s1 : (op == 1 ? s1 : 0.0f);
        e_l = sample_grad_hess<D, KIND>(A, u, delta, g, h);
    } else {
        e_l = sample_value<D, KIND>(A, delta);
        frozen_e = e_l;
        frozen_m = detD<D>(A);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) g[k] = group_sum<NL>(g[k]);
    float e0 = group_sum<NL>(e_l);
    float tau[D] = {};
    if constexpr (TAU) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            tau[k] = a.tau[k * PL + node];
            g[k] -= tau[k];
            e0 -= tau[k] * pos[k];
        }
    }
    if constexpr (GRAD) {
        if (q == 0)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float& o = a.gout[k * PL + node];
                o = a.gsub ? g[k] - o : g[k];
            }
        if (q == 0 && a.gpush) push_replicas<D>(a, a.gout, node, g);
        return;
    }
#pragma unroll
    for (int k = 0; k < kNH<D>; ++k) h[k] = group_sum<NL>(h[k]);
    float dl[D];
    newton<D>(g, h, dl);
    float alpha = 1.0f, cur[D];
    int ntr = 10;
#pragma unroll
    for (int k = 0; k < D; ++k) cur[k] = pos[k];
    for (int it = 0; it < 10; ++it) {
        float tr[D];
#pragma unroll
        for (int k = 0; k < D; ++k) tr[k] = pos[k] + alpha * dl[k];
        float et = frozen_e, mt = frozen_m;
        if (moves) {
            float At[D * D];
            build(tr, At);
            et = sample_value<D, KIND>(At, delta);
            mt = detD<D>(At);
        }
        et = group_sum<NL>(et);
        mt = group_min<NL>(mt);
        if constexpr (TAU) {
#pragma unroll
            for (int k = 0; k < D; ++k) et -= tau[k] * tr[k];
        }
        if (accept<KIND>(et, e0, mt)) {
#pragma unroll
            for (int k = 0; k < D; ++k) cur[k] = tr[k];
            ntr = it + 1;
            break;
        }
        alpha *= 0.5f;
    }
    if (a.omega != 1.0f)
#pragma unroll
        for (int k = 0; k < D; ++k) cur[k] = pos[k] + a.omega * (cur[k] - pos[k]);
    if (q == 0) {
        store_node<D>(a, node, cur);
        count_update(a, ntr);
    }
}

template <int KIND, int TX, int TY, bool TAU, bool GRAD>
__global__ __launch_bounds__(16 * TX * TY) void sample_kernel(const SweepArgs a)
{
    constexpr int NT = 16 * TX * TY, G = 16;
    const int n_st = int(gridDim.x) - a.n_tiles;
    if (int(blockIdx.x) < n_st) {
        const int gid = int(blockIdx.x) * (NT / G) + int(threadIdx.x) / G;
        if (a.st.wt)
            stored_group<2, KIND, TAU, GRAD, G, true>(a, gid, int(threadIdx.x) % G, gid < a.st.n);
        else
            stored_group<2, KIND, TAU, GRAD, G, false>(a, gid, int(threadIdx.x) % G, gid < a.st.n);
        return;
    }
    sample_sweep<KIND, TX, TY, TAU, GRAD>(a, int(blockIdx.x) - n_st);
}

template <int KIND, int TX, int TY, bool TAU, bool GRAD>
void launch_sample_variant(const SweepArgs& a, hipStream_t s)
{
    constexpr int NT = 16 * TX * TY;
    const int per_wg = NT / 16;
    const int nst = (a.st.n + per_wg - 1) / per_wg;
    const unsigned grid = (unsigned)(a.n_tiles + nst);
    if (grid == 0) return;
    sample_kernel<KIND, TX, TY, TAU, GRAD><<<dim3(grid), dim3(NT), 0, s>>>(a);
}

// One launch per sweep: the leading workgroups serve the stored DOFs (one
// lane group of SG lanes per DOF), the rest one tile each with 2^D / CPL
// lanes per node (CPL cells per lane).
template <int D, int KIND, int TX, int TY, int TZ, bool TAU, bool GRAD, bool PROP, bool LDS, int CPL, int SG>
__device__ __forceinline__ void sweep_body(const SweepArgs& a)
{
    constexpr int NT = ((1 << D) / CPL) * TX * TY * TZ;
    constexpr int G = SG;
    static_assert(NT % G == 0, "a workgroup holds whole stored lane groups");
    // stored-DOF workgroups first: their gather chains are the longest, so
    // they start while the tiles fill the machine
    const int n_st = int(gridDim.x) - a.n_tiles;
    if (int(blockIdx.x) < n_st) {
        const int gid = int(blockIdx.x) * (NT / G) + int(threadIdx.x) / G;
        if (a.st.wt)
            stored_group<D, KIND, TAU, GRAD, G, true>(a, gid, int(threadIdx.x) % G, gid < a.st.n);
        else
            stored_group<D, KIND, TAU, GRAD, G, false>(a, gid, int(threadIdx.x) % G, gid < a.st.n);
        return;
    }
    tile_sweep<D, KIND, TX, TY, TZ, TAU, GRAD, PROP, LDS, CPL>(a, int(blockIdx.x) - n_st);
}

template <int D, int KIND, int TX, int TY, int TZ, bool TAU, bool GRAD, bool PROP, bool LDS,
          int SG = stored_lanes(D)>
__global__ __launch_bounds__((1 << D) * TX * TY * TZ) void sweep_kernel(const SweepArgs a)
{
    sweep_body<D, KIND, TX, TY, TZ, TAU, GRAD, PROP, LDS, 1, SG>(a);
}

// several cells per lane: the compiler interleaves a lane's cells for ILP;
// MINW waves per SIMD bounds the registers that may take (set where the
// bound costs no spills)
template <int D, int KIND, int TX, int TY, int TZ, bool TAU, bool GRAD, bool LDS, int CPL, int MINW>
__global__ __attribute__((amdgpu_flat_work_group_size(1, ((1 << D) / CPL) * TX * TY * TZ),
                          amdgpu_waves_per_eu(MINW))) void sweep_kernel_cpl(const SweepArgs a)
{
    sweep_body<D, KIND, TX, TY, TZ, TAU, GRAD, false, LDS, CPL, stored_lanes(D)>(a);
}

template <int D, int KIND, int TX, int TY, int TZ, bool TAU, bool GRAD, bool PROP = false, bool LDS = true,
          int CPL = 1, int SG = stored_lanes(D)>
void launch_variant(const SweepArgs& a, hipStream_t s)
{
    static_assert(CPL == 1 || SG == stored_lanes(D), "cells-per-lane kernels take the default stored groups");
    constexpr int NT = ((1 << D) / CPL) * TX * TY * TZ;
    constexpr int G = SG;
    const int per_wg = NT / G;
    const int nst = (a.st.n + per_wg - 1) / per_wg;
    const unsigned grid = (unsigned)(a.n_tiles + nst);
    if (grid == 0) return;
    if constexpr (CPL == 1)
        sweep_kernel<D, KIND, TX, TY, TZ, TAU, GRAD, PROP, LDS, SG><<<dim3(grid), dim3(NT), 0, s>>>(a);
    else
        sweep_kernel_cpl<D, KIND, TX, TY, TZ, TAU, GRAD, LDS, CPL, (D == 2 && CPL == 2 ? 4 : 2)>
            <<<dim3(grid), dim3(NT), 0, s>>>(a);
}

// A whole FAS coarse level smoothed by one workgroup: n_sweeps tau-shifted
// sweeps, each walking the level's tiles in turn. A workgroup barrier
// follows every tile (the next tile restages the LDS tile), so the barrier
// after a sweep's last tile also orders its writes, replicas included,
// before the next sweep's reads: the part a kernel boundary plays between
// one-sweep launches (all waves of a workgroup share one CU and its L1).
// x0 holds the state on entry; sweep s reads x_(s mod 2) and writes the
// other buffer. Per DOF the arithmetic is the one-sweep kernel's.
template <int D, int KIND, int TX, int TY, int TZ, bool LDS, int CPL, bool SAMPLE>
__global__ __launch_bounds__(SAMPLE ? 16 * TX * TY : ((1 << D) / CPL) * TX * TY * TZ) void level_smooth_kernel(
    const SweepArgs a0, float* x0, float* x1, int n_sweeps)
{
    SweepArgs a = a0;
    for (int s = 0; s < n_sweeps; ++s) {
        a.xin = (s & 1) ? x1 : x0;
        a.xout = (s & 1) ? x0 : x1;
        for (int t = 0; t < a.n_tiles; ++t) {
            if constexpr (SAMPLE)
                sample_sweep<KIND, TX, TY, true, false>(a, t);
            else
                tile_sweep<D, KIND, TX, TY, TZ, true, false, false, LDS, CPL>(a, t);
            __syncthreads();
        }
    }
}

template <int D, int KIND, int TX, int TY, int TZ, bool LDS = true, int CPL = 1, bool SAMPLE = false>
void launch_level_variant(const SweepArgs& a, float* x0, float* x1, int n_sweeps, hipStream_t s)
{
    constexpr int NT = SAMPLE ? 16 * TX * TY : ((1 << D) / CPL) * TX * TY * TZ;
    if (a.n_tiles == 0 || n_sweeps <= 0) return;
    level_smooth_kernel<D, KIND, TX, TY, TZ, LDS, CPL, SAMPLE><<<dim3(1), dim3(NT), 0, s>>>(a, x0, x1, n_sweeps);
}

}  // namespace eggmi
