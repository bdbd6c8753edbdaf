// Learned Newton-step proposal: a small MLP evaluated and trained with the
// gfx950 BF16 matrix cores (FP32 accumulation).
//
// Network (per dimension D):
//   2D: 22 inputs (3x3 one-ring relative coordinates, 18; two target
//       parameters; the node's previous displacement, 2) -> 64 -> 64 -> 2
//   3D: 85 inputs (3x3x3 one-ring, 81; one constant; previous
//       displacement, 3) -> 128 -> 128 -> 3
// GELU (tanh form) after both hidden layers. Inputs are expressed in the
// node's local frame (axis-0 direction, Gram-Schmidt on axis 1, right-handed
// completion) and scaled by the local edge length h; the output is the
// proposed displacement in that frame divided by h.
//
// Matrix-core formulation. Everything runs transposed: a layer computes
// Z^T[out x 16 nodes] = W^T[out x in] . X^T[in x 16 nodes] with
// v_mfma_f32_16x16x32_bf16 (A = weights, B = activations). Lane l of the
// wave holds B rows 8*(l/16)..+7 of column l%16 and receives D rows
// 4*(l/16)..+3 of column l%16 (verified on the device). Two adjacent
// 16-row output tiles of one layer therefore ARE the next layer's 32-deep B
// fragment, up to a fixed permutation of the reduction index, which the
// weight packing absorbs: slot 8g+i of a 32-chunk is feature 4g+i (i < 4)
// or 16+4g+(i-4) (i >= 4) of that chunk. No activation leaves registers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace eggmi
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int D> struct MlpDims;
template <> struct MlpDims<2> {
    static constexpr int IN = 22, INP = 32, H = 64, OUT = 2;
};
template <> struct MlpDims<3> {
    static constexpr int IN = 85, INP = 96, H = 128, OUT = 3;
};

// Sizes of the FP32 master parameter vector (natural layout):
//   W1[IN][H], b1[H], W2[H][H], b2[H], W3[H][OUT], b3[OUT]
template <int D> struct MlpLayout {
    using M = MlpDims<D>;
    static constexpr int IN = M::IN, INP = M::INP, H = M::H, OUT = M::OUT;
    static constexpr long oW1 = 0, ob1 = oW1 + (long)IN * H, oW2 = ob1 + H, ob2 = oW2 + (long)H * H,
                          oW3 = ob2 + H, ob3 = oW3 + (long)H * OUT, n_params = ob3 + OUT;
    // forward A fragments (64 lanes x 8 bf16 each), order: layer 1 [mt][ks],
    // layer 2 [mt][ks], layer 3 [ks]
    static constexpr int KS1 = INP / 32, MT1 = H / 16, KS2 = H / 32, MT2 = H / 16, KS3 = H / 32;
    static constexpr int F1 = 0, F2 = F1 + MT1 * KS1, F3 = F2 + MT2 * KS2, n_fwd_frags = F3 + KS3;
    // backward A fragments: W3 for dH2 = W3 . dY (16x16x16, 4 bf16 per lane,
    // [mt] over H), W2 for dH1 = W2 . dZ2 (16x16x32, [mt][ks])
    static constexpr int n_bw3_frags = H / 16, n_bw2_frags = (H / 16) * (H / 32);
};

// Feature index held by reduction slot `slot` of a chained (D-layout) 32-chunk.
__host__ __device__ inline int chain_feature(int slot)
{
    const int c = slot & ~31, r = slot & 31, g = r >> 3, i = r & 7;
    return c + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// tanh(u) = 1 - 2 / (1 + e^{2u}) with the hardware reciprocal (saturates
// correctly: e^{2u} = inf gives 1, e^{2u} = 0 gives -1)
__device__ __forceinline__ float tanh_fast(float u)
{
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(1.0f + __expf(2.0f * u));
}
__device__ __forceinline__ float gelu(float x)
{
    const float t = tanh_fast(0.7978845608f * (x + 0.044715f * x * x * x));
    return 0.5f * x * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad(float x)
{
    const float t = tanh_fast(0.7978845608f * (x + 0.044715f * x * x * x));
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608f * (1.0f + 0.134145f * x * x);
}

// gelu(x) and gelu'(x) from one tanh (the training forward keeps only the
// derivative for the backward pass)
__device__ __forceinline__ float gelu_and_grad(float x, float& dgelu)
{
    const float t = tanh_fast(0.7978845608f * (x + 0.044715f * x * x * x));
    dgelu = 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608f * (1.0f + 0.134145f * x * x);
    return 0.5f * x * (1.0f + t);
}

__device__ __forceinline__ f32x4 mfma32(const bf16x8& a, const bf16x8& b, const f32x4& c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Two adjacent D-layout output tiles (rows 16t.. and 16(t+1)..) -> one
// chained B fragment (bf16, round to nearest even).
__device__ __forceinline__ bf16x8 chain(const f32x4& lo, const f32x4& hi)
{
    bf16x8 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b[i] = (__bf16)lo[i];
        b[i + 4] = (__bf16)hi[i];
    }
    return b;
}

// Device view of a packed network.
struct MlpView {
    const bf16x8* fwd = nullptr;  // [n_fwd_frags][64]
    const float* params = nullptr;  // FP32 master (biases are read from here)
};

// Forward pass of one wave over 16 nodes (columns). xin[ks] is the lane's B
// fragment of input chunk ks (identity slot order). Returns the lane's
// D-layout output tile (rows 4*(l/16)+i = output component, column l%16).
// Optionally keeps the pre-activations of both hidden layers.
template <int D, bool KEEP>
__device__ __forceinline__ f32x4 mlp_forward(const MlpView& w, const bf16x8* xin, f32x4* z1 = nullptr,
                                             f32x4* z2 = nullptr)
{
    using L = MlpLayout<D>;
    const int lane = threadIdx.x & 63;
    const int rg = 4 * (lane >> 4);
    f32x4 h1[L::MT1];
#pragma unroll
    for (int mt = 0; mt < L::MT1; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS1; ++ks) acc = mfma32(w.fwd[(L::F1 + mt * L::KS1 + ks) * 64 + lane], xin[ks], acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += w.params[L::ob1 + mt * 16 + rg + i];
            if constexpr (KEEP) z1[mt][i] = acc[i];
            acc[i] = gelu(acc[i]);
        }
        h1[mt] = acc;
    }
    f32x4 h2[L::MT2];
#pragma unroll
    for (int mt = 0; mt < L::MT2; ++mt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS2; ++ks)
            acc = mfma32(w.fwd[(L::F2 + mt * L::KS2 + ks) * 64 + lane], chain(h1[2 * ks], h1[2 * ks + 1]), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += w.params[L::ob2 + mt * 16 + rg + i];
            if constexpr (KEEP) z2[mt][i] = acc[i];
            acc[i] = gelu(acc[i]);
        }
        h2[mt] = acc;
    }
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < L::KS3; ++ks) y = mfma32(w.fwd[(L::F3 + ks) * 64 + lane], chain(h2[2 * ks], h2[2 * ks + 1]), y);
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (rg + i < L::OUT) y[i] += w.params[L::ob3 + rg + i];
    return y;
}

// ---- local frame and features of a node from its 3^D one-ring
// ring(a, k) returns coordinate k of ring node a (a = row-major index over
// offsets {-1,0,1}^D, axis 0 most significant; a = (3^D-1)/2 is the node).
struct Frame {
    float e[3][3];  // rows: frame axes (world coordinates)
    float h;        // local length scale
    bool ok;
};

template <int D, class Ring> __device__ __forceinline__ Frame local_frame(const Ring& ring)
{
    Frame f {};
    constexpr int C = D == 2 ? 4 : 13;
    constexpr int S0 = D == 2 ? 3 : 9, S1 = D == 2 ? 1 : 3;
    float a0[3] = {}, a1[3] = {}, a2[3] = {};
#pragma unroll
    for (int k = 0; k < D; ++k) {
        a0[k] = ring(C + S0, k) - ring(C - S0, k);
        a1[k] = ring(C + S1, k) - ring(C - S1, k);
        if constexpr (D == 3) a2[k] = ring(C + 1, k) - ring(C - 1, k);
    }
    float n0 = 0.f, n1 = 0.f, n2 = 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        n0 += a0[k] * a0[k];
        n1 += a1[k] * a1[k];
        n2 += a2[k] * a2[k];
    }
    n0 = sqrtf(n0);
    n1 = sqrtf(n1);
    n2 = sqrtf(n2);
    f.h = D == 2 ? 0.25f * (n0 + n1) : (n0 + n1 + n2) * (1.0f / 6.0f);
    f.ok = n0 > 0.f && f.h > 0.f;
    const float r0 = n0 > 0.f ? 1.0f / n0 : 0.f;
    if constexpr (D == 2) {
        f.e[0][0] = a0[0] * r0;
        f.e[0][1] = a0[1] * r0;
        f.e[1][0] = -f.e[0][1];
        f.e[1][1] = f.e[0][0];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) f.e[0][k] = a0[k] * r0;
        const float p = a1[0] * f.e[0][0] + a1[1] * f.e[0][1] + a1[2] * f.e[0][2];
        float v[3], nv = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[k] = a1[k] - p * f.e[0][k];
            nv += v[k] * v[k];
        }
        nv = sqrtf(nv);
        f.ok = f.ok && nv > 0.f;
        const float rv = nv > 0.f ? 1.0f / nv : 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) f.e[1][k] = v[k] * rv;
        f.e[2][0] = f.e[0][1] * f.e[1][2] - f.e[0][2] * f.e[1][1];
        f.e[2][1] = f.e[0][2] * f.e[1][0] - f.e[0][0] * f.e[1][2];
        f.e[2][2] = f.e[0][0] * f.e[1][1] - f.e[0][1] * f.e[1][0];
    }
    return f;
}

// Feature j (0 <= j < INP) of a node: ring coordinates relative to the node
// in the local frame over h, then the constant inputs, then the previous
// displacement `dprev` (world) in the local frame over h; zero padding.
template <int D, class Ring>
__device__ __forceinline__ float feature(const Ring& ring, const Frame& f, const float* dprev, int j)
{
    constexpr int NR = D == 2 ? 9 : 27;
    constexpr int C = D == 2 ? 4 : 13;
    const float rh = f.ok ? 1.0f / f.h : 0.f;
    if (j < NR * D) {
        const int a = j / D, c = j % D;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) s += (ring(a, k) - ring(C, k)) * f.e[c][k];
        return s * rh;
    }
    j -= NR * D;
    constexpr int NCONST = D == 2 ? 2 : 1;  // 2D: target parameters (1, 0); 3D: 1
    if (j < NCONST) return j == 0 ? 1.0f : 0.0f;
    j -= NCONST;
    if (j < D) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) s += dprev[k] * f.e[j][k];
        return s * rh;
    }
    return 0.f;
}

// The same feature for a run-time index j, without indexing a register
// array (frame rows are selected, ring values come from the accessor), so a
// lane can compute just its own slice of the features; rh = 1/h (0 if the
// frame is degenerate).
template <int D, class Ring>
__device__ __forceinline__ float feature_rt(const Ring& ring, const Frame& f, const float* dprev, float rh, int j)
{
    constexpr int NR = D == 2 ? 9 : 27;
    constexpr int C = D == 2 ? 4 : 13;
    constexpr int NCONST = D == 2 ? 2 : 1;
    auto row = [&](int c, int k) {
        float v = f.e[0][k];
        if (c == 1) v = f.e[1][k];
        if (D == 3 && c == 2) v = f.e[2][k];
        return v;
    };
    if (j < NR * D) {
        const int a = j / D, c = j - (j / D) * D;
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) s += (ring(a, k) - ring(C, k)) * row(c, k);
        return s * rh;
    }
    j -= NR * D;
    if (j < NCONST) return j == 0 ? 1.0f : 0.0f;
    j -= NCONST;
    if (j < D) {
        float s = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) s += dprev[k] * row(j, k);
        return s * rh;
    }
    return 0.f;
}

// Local-frame output -> world displacement.
template <int D> __device__ __forceinline__ void to_world(const Frame& f, const float* pl, float* p)
{
#pragma unroll
    for (int k = 0; k < D; ++k) {
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < D; ++c) s += pl[c] * f.e[c][k];
        p[k] = s * f.h;
    }
}

// ---- launchers (mlp.hip)
struct MlpTrainArgs {
    int dim;
    long long batch;            // samples per step (multiple of 16)
    const __bf16* X;            // dataset features [n][INP]
    const float* T;             // dataset targets [n][4]
    const int* idx;             // [batch] sample ids of this step, or null: a permutation (below)
    // idx == null: sample id of batch row r is perm(perm_base + r), a keyed
    // permutation of [0, perm_n) (a balanced Feistel network on 2^perm_bits
    // >= perm_n values, cycle-walked into range), one per epoch
    long long perm_n = 0, perm_base = 0;
    int perm_bits = 0;
    unsigned long long perm_key = 0;
    float* params;              // FP32 master
    float* adam_m;
    float* adam_v;
    bf16x8* fwd;                // packed forward fragments
    void* bw;                   // packed backward fragments
    __bf16* act;                // per-batch activations (see mlp.hip)
    float* partials;            // weight-gradient partials
    float* grad;                // [n_params]
    double* loss;               // [batch/16] per-tile loss partials
};

void mlp_pack(int dim, const float* params, bf16x8* fwd, void* bw, hipStream_t s);
void mlp_forward_batch(int dim, const bf16x8* fwd, const float* params, const __bf16* X, long long n, float* Y,
                       hipStream_t s);
// One training step on the batch a.idx: a.grad receives the gradient of the
// mean squared error, a.loss[tile] the squared error of each 16-sample tile;
// lr > 0 also takes the Adam step (bias correction for step `step`, 1-based)
// and repacks the BF16 fragments.
void mlp_step(const MlpTrainArgs& a, float lr, long long step, hipStream_t s);
long long mlp_chunk();  // the batch must be a multiple of this
struct BlockDesc;
// Features (BF16 [n][INP]) of the synthesized DOFs at `slots` (block `blks`)
// from the store x (previous sweep's store xprev), with each DOF's frame
// ([n][D*D+1]: axes, then h) and position ([n][D]) kept for mlp_targets.
void mlp_features(int dim, const float* x, const float* xprev, long long plane, const int* slots, const int* blks,
                  const BlockDesc* blocks, long long n, void* Xo, float* frame, float* xt, hipStream_t s);
// Targets [n][4]: displacement since mlp_features, in that frame, over h.
void mlp_targets(int dim, const float* x, long long plane, const int* slots, long long n, const float* frame,
                 const float* xt, float* T, hipStream_t s);
long long mlp_n_params(int dim);
long long mlp_fwd_frag_count(int dim);
long long mlp_bw_bytes(int dim);
long long mlp_act_elems(int dim, long long batch);
long long mlp_partial_elems(int dim, long long batch);
int mlp_inp(int dim);

}  // namespace eggmi
