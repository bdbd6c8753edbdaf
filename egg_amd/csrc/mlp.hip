// BF16 matrix-core kernels of the learned proposal: weight packing, batched
// inference, and one training step (forward + backward, weight gradients,
// Adam). FP32 accumulation everywhere; FP32 master weights.
//
// Training step over a batch of B samples (B a multiple of 16):
//   1. fwd_bwd: one wave per 16 samples. Forward as in inference, keeping the
//      hidden pre-activations in registers; dY = 2 (y - t) / B; dH2 = W3 dY
//      (16x16x16), dZ2 = dH2 * gelu'(Z2), dH1 = W2 dZ2 (16x16x32, chained),
//      dZ1 = dH1 * gelu'(Z1). The layer inputs and the dZ's are written
//      feature-major ([feature][sample] BF16), so that the weight-gradient
//      MFMA fragments below are contiguous.
//   2. wgrad: dW = Act^T G as 16x16 output tiles, one wave per (tile, chunk of
//      256 samples), K = samples (16x16x32: each lane's A and B fragments are
//      8 consecutive samples of one feature, one 16-byte load each); bias
//      gradients as row sums of G over the chunk. Partials per chunk.
//   3. adam: one thread per parameter folds the chunk partials in chunk order
//      (deterministic), then the Adam update of the FP32 master.
//   4. pack: master -> BF16 fragments for the next step.
#include "kernels.hpp"
#include "mlp.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

namespace eggmi
{

void hip_check(hipError_t e, const char* what);

namespace
{

constexpr int kChunk = 256;  // samples per weight-gradient partial

template <int D> struct Act {
    // per-batch activation buffers (BF16, feature-major [feature][B]), offsets in elements
    using L = MlpLayout<D>;
    __host__ __device__ static long long x(long long) { return 0; }
    __host__ __device__ static long long h1(long long B) { return B * L::INP; }
    __host__ __device__ static long long dz1(long long B) { return h1(B) + B * L::H; }
    __host__ __device__ static long long h2(long long B) { return dz1(B) + B * L::H; }
    __host__ __device__ static long long dz2(long long B) { return h2(B) + B * L::H; }
    __host__ __device__ static long long dy(long long B) { return dz2(B) + B * L::H; }
    __host__ __device__ static long long total(long long B) { return dy(B) + B * 16; }
};

// Weight-gradient tile table: tiles of dW1 [INP x H], dW2 [H x H], dW3 [H x 16]
// (16x16 each), then bias column-sum tiles of dZ1 [H], dZ2 [H], dY [16].
template <int D> struct Tiles {
    using L = MlpLayout<D>;
    static constexpr int t1 = (L::INP / 16) * (L::H / 16), t2 = (L::H / 16) * (L::H / 16), t3 = L::H / 16;
    static constexpr int b1 = L::H / 16, b2 = L::H / 16, b3 = 1;
    static constexpr int n = t1 + t2 + t3 + b1 + b2 + b3;
};

__device__ __forceinline__ short bf_bits(float x)
{
    const __bf16 b = (__bf16)x;
    return __builtin_bit_cast(short, b);
}

template <int D>
__global__ __launch_bounds__(256) void pack_kernel(const float* __restrict__ P, bf16x8* __restrict__ fwd,
                                                   char* __restrict__ bw)
{
    using L = MlpLayout<D>;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long nf = (long long)L::n_fwd_frags * 64 * 8;
    const long long nb3 = (long long)L::n_bw3_frags * 64 * 4;
    const long long nb2 = (long long)L::n_bw2_frags * 64 * 8;
    if (t < nf) {
        const int f = (int)(t / 512), lane = (int)((t / 8) % 64), i = (int)(t % 8);
        const int g = lane >> 4, c = lane & 15;
        float v = 0.f;
        if (f < L::F2) {
            const int mt = f / L::KS1, ks = f % L::KS1;
            const int m = mt * 16 + c, feat = ks * 32 + 8 * g + i;
            if (feat < L::IN) v = P[L::oW1 + (long long)feat * L::H + m];
        } else if (f < L::F3) {
            const int q = f - L::F2, mt = q / L::KS2, ks = q % L::KS2;
            const int m = mt * 16 + c, feat = chain_feature(ks * 32 + 8 * g + i);
            v = P[L::oW2 + (long long)feat * L::H + m];
        } else {
            const int ks = f - L::F3;
            const int m = c, feat = chain_feature(ks * 32 + 8 * g + i);
            if (m < L::OUT) v = P[L::oW3 + (long long)feat * L::OUT + m];
        }
        reinterpret_cast<__bf16*>(fwd)[t] = (__bf16)v;
        return;
    }
    long long u = t - nf;
    if (u < nb3) {  // dH2 = W3 . dY: A[m = h][k = o]
        const int mt = (int)(u / 256), lane = (int)((u / 4) % 64), i = (int)(u % 4);
        const int h = mt * 16 + (lane & 15), o = 4 * (lane >> 4) + i;
        const float v = o < L::OUT ? P[L::oW3 + (long long)h * L::OUT + o] : 0.f;
        reinterpret_cast<__bf16*>(bw)[u] = (__bf16)v;
        return;
    }
    u -= nb3;
    if (u < nb2) {  // dH1 = W2 . dZ2: A[m = h1][slot -> h2]
        const int f = (int)(u / 512), lane = (int)((u / 8) % 64), i = (int)(u % 8);
        const int mt = f / L::KS2, ks = f % L::KS2;
        const int h1 = mt * 16 + (lane & 15), h2 = chain_feature(ks * 32 + 8 * (lane >> 4) + i);
        reinterpret_cast<__bf16*>(bw + nb3 * 2)[u] = (__bf16)P[L::oW2 + (long long)h1 * L::H + h2];
    }
}

// Lane's identity-order B fragments of rows X[r0 + l%16], or zeros past n.
template <int D>
__device__ __forceinline__ void load_x(const __bf16* X, long long row, bool valid, bf16x8* xin)
{
    using L = MlpLayout<D>;
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int ks = 0; ks < L::KS1; ++ks) {
        if (valid)
            xin[ks] = *reinterpret_cast<const bf16x8*>(X + row * L::INP + ks * 32 + 8 * g);
        else
            for (int i = 0; i < 8; ++i) xin[ks][i] = (__bf16)0.0f;
    }
}

template <int D>
__global__ __launch_bounds__(64) void forward_kernel(MlpView w, const __bf16* __restrict__ X, long long n,
                                                     float* __restrict__ Y)
{
    using L = MlpLayout<D>;
    const int lane = threadIdx.x;
    const long long row = (long long)blockIdx.x * 16 + (lane & 15);
    bf16x8 xin[L::KS1];
    load_x<D>(X, row, row < n, xin);
    const f32x4 y = mlp_forward<D, false>(w, xin);
    if (lane < 16 && row < n)
#pragma unroll
        for (int i = 0; i < L::OUT; ++i) Y[row * L::OUT + i] = y[i];
}

// D-layout tile (rows 16*mt + 4g + i, column = sample) -> feature-major
// buf[feature][B] (for each row, the 16 lanes of a group write 16 consecutive
// samples). `lane_off` = byte offset of (row 4g, this lane's sample) from
// the buffer; the row base is uniform, so the stores take a scalar base and
// a 32-bit lane offset (the activation buffer is < 2 GiB).
__device__ __forceinline__ void store_tile(__bf16* buf, unsigned lane_off, long long B, int mt, const f32x4& v)
{
    char* row = reinterpret_cast<char*>(buf + (long long)(mt * 16) * B);
    const unsigned rs = (unsigned)(B * 2);
#pragma unroll
    for (int i = 0; i < 4; ++i) *reinterpret_cast<__bf16*>(row + (lane_off + i * rs)) = (__bf16)v[i];
}

// Keyed permutation of [0, n): a 4-round balanced Feistel network on
// 2^bits >= n values (bits even; the round function is a 32-bit integer
// hash of the half and the round key), applied until the value falls in
// [0, n) (cycle walking; at most 4 n / n expected rounds of walking since
// 2^bits < 4 n).
__host__ __device__ inline unsigned mix32(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
__host__ __device__ inline long long feistel_perm(long long i, long long n, int bits, unsigned long long key)
{
    const int h = bits / 2;
    const unsigned mask = (1u << h) - 1u;
    unsigned v = (unsigned)i;
    do {
        unsigned l = v >> h, r = v & mask;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const unsigned f = mix32(r ^ (unsigned)(key >> (16 * k)) ^ (0x9e3779b9U * (unsigned)(k + 1))) & mask;
            const unsigned t = l ^ f;
            l = r;
            r = t;
        }
        v = (l << h) | r;
    } while ((long long)v >= n);
    return v;
}

// Training forward/backward: a workgroup of WAVES waves stages every weight
// fragment (forward, W3^T, W2^T) in LDS once (all loads in flight together);
// each wave then takes NT tiles of 16 samples, so each fragment read feeds
// NT independent MFMA chains. Small batches use one wave and one tile per
// workgroup (enough workgroups to spread over the chip), large batches four
// waves sharing one staged copy.
template <int D> constexpr int kFragBytes = MlpLayout<D>::n_fwd_frags * 1024 + MlpLayout<D>::n_bw3_frags * 512 +
                                            MlpLayout<D>::n_bw2_frags * 1024;

template <int D, int WAVES, int NT>
__global__ __launch_bounds__(64 * WAVES) void fwd_bwd_kernel(MlpView w, const void* __restrict__ bw,
                                                                   const __bf16* __restrict__ X,
                                                                   const float* __restrict__ T,
                                                                   const int* __restrict__ idx, long long pn,
                                                                   long long pbase, int pbits,
                                                                   unsigned long long pkey, long long B,
                                                                   __bf16* __restrict__ act,
                                                                   double* __restrict__ loss)
{
    using L = MlpLayout<D>;
    using A = Act<D>;
    // one wave reads the fragments from L2 as it goes (a staged copy would
    // serialise 24-96 KB of loads in front of its work)
    constexpr bool kStage = WAVES > 1;
    __shared__ uint4 frag[kStage ? kFragBytes<D> / 16 : 1];
    const uint4* fsrc = reinterpret_cast<const uint4*>(w.fwd);
    const uint4* bsrc = static_cast<const uint4*>(bw);
    if constexpr (kStage) {
        constexpr int nf = L::n_fwd_frags * 64, nb = (kFragBytes<D> - L::n_fwd_frags * 1024) / 16;
        for (int e = threadIdx.x; e < nf; e += 64 * WAVES) frag[e] = fsrc[e];
        for (int e = threadIdx.x; e < nb; e += 64 * WAVES) frag[nf + e] = bsrc[e];
        fsrc = frag;
        bsrc = frag + nf;
    }
    const bf16x8* wf1 = reinterpret_cast<const bf16x8*>(fsrc);
    const s16x4* bw3 = reinterpret_cast<const s16x4*>(bsrc);
    const bf16x8* bw2 = reinterpret_cast<const bf16x8*>(bsrc + L::n_bw3_frags * 32);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, rg = 4 * g;
    long long s[NT], src[NT];
    unsigned off[NT], xoff[NT];  // byte offsets of (row 4g / 8g, sample s) in a [feature][B] buffer
    bf16x8 xin[NT][L::KS1];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        s[t] = (((long long)blockIdx.x * WAVES + wave) * NT + t) * 16 + (lane & 15);  // batch row
        off[t] = (unsigned)(((long long)rg * B + s[t]) * 2);
        xoff[t] = (unsigned)(((long long)(8 * g) * B + s[t]) * 2);
        src[t] = idx ? idx[s[t]] : feistel_perm(pbase + s[t], pn, pbits, pkey);
        load_x<D>(X, src[t], true, xin[t]);
    }
    // the batch copy of the inputs (for dW1), feature-major
#pragma unroll
    for (int ks = 0; ks < L::KS1; ++ks) {
        char* row = reinterpret_cast<char*>(act + A::x(B) + (long long)(ks * 32) * B);
        const unsigned rs = (unsigned)(B * 2);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int i = 0; i < 8; ++i)
                *reinterpret_cast<__bf16*>(row + (xoff[t] + i * rs)) = xin[t][ks][i];
    }
    if constexpr (kStage) __syncthreads();
    // layer 1: Z1 = W1 X + b1, H1 = gelu(Z1) (gelu'(Z1) kept)
    f32x4 dz1[NT][L::MT1];  // gelu'(Z1)
    bf16x8 hb[NT][L::KS2];
#pragma unroll
    for (int mt = 0; mt < L::MT1; ++mt) {
        f32x4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f32x4 {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS1; ++ks) {
            const bf16x8 wf = wf1[(L::F1 + mt * L::KS1 + ks) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = mfma32(wf, xin[t][ks], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 h;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float dg;
                h[i] = gelu_and_grad(acc[t][i] + w.params[L::ob1 + mt * 16 + rg + i], dg);
                dz1[t][mt][i] = dg;
                hb[t][mt >> 1][(mt & 1) * 4 + i] = (__bf16)h[i];
            }
            store_tile(act + A::h1(B), off[t], B, mt, h);
        }
    }
    // layer 2
    f32x4 dz2[NT][L::MT2];  // gelu'(Z2)
    bf16x8 hb2[NT][L::KS3];
#pragma unroll
    for (int mt = 0; mt < L::MT2; ++mt) {
        f32x4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f32x4 {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS2; ++ks) {
            const bf16x8 wf = wf1[(L::F2 + mt * L::KS2 + ks) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = mfma32(wf, hb[t][ks], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 h;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float dg;
                h[i] = gelu_and_grad(acc[t][i] + w.params[L::ob2 + mt * 16 + rg + i], dg);
                dz2[t][mt][i] = dg;
                hb2[t][mt >> 1][(mt & 1) * 4 + i] = (__bf16)h[i];
            }
            store_tile(act + A::h2(B), off[t], B, mt, h);
        }
    }
    // layer 3, dY (rows = outputs) and the per-tile squared error
    f32x4 dy[NT];
    float se[NT];
    {
        f32x4 y[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) y[t] = f32x4 {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS3; ++ks) {
            const bf16x8 wf = wf1[(L::F3 + ks) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT; ++t) y[t] = mfma32(wf, hb2[t][ks], y[t]);
        }
        const float sc = 2.0f / (float)B;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            se[t] = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = rg + i;
                const float r = o < L::OUT ? y[t][i] + w.params[L::ob3 + o] - T[src[t] * 4 + o] : 0.f;
                se[t] += r * r;
                dy[t][i] = sc * r;
            }
            store_tile(act + A::dy(B), off[t], B, 0, dy[t]);
        }
    }
    // dH2 = W3 . dY, dZ2 = dH2 * gelu'(Z2)
    s16x4 bdy[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) bdy[t][i] = bf_bits(dy[t][i]);
    bf16x8 bdz2[NT][L::KS2];
#pragma unroll
    for (int mt = 0; mt < L::MT2; ++mt) {
        const s16x4 wf = bw3[mt * 64 + lane];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wf, bdy[t], acc, 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                acc[i] *= dz2[t][mt][i];
                bdz2[t][mt >> 1][(mt & 1) * 4 + i] = (__bf16)acc[i];
            }
            store_tile(act + A::dz2(B), off[t], B, mt, acc);
        }
    }
    // dH1 = W2 . dZ2, dZ1 = dH1 * gelu'(Z1)
#pragma unroll
    for (int mt = 0; mt < L::MT1; ++mt) {
        f32x4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f32x4 {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < L::KS2; ++ks) {
            const bf16x8 wf = bw2[(mt * L::KS2 + ks) * 64 + lane];
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[t] = mfma32(wf, bdz2[t][ks], acc[t]);
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[t][i] *= dz1[t][mt][i];
            store_tile(act + A::dz1(B), off[t], B, mt, acc[t]);
        }
    }
    // squared error of each 16-sample tile (deterministic: one lane sums in lane order)
    __shared__ float red[WAVES][NT][64];
#pragma unroll
    for (int t = 0; t < NT; ++t) red[wave][t][lane] = se[t];
    __syncthreads();
    if (lane < NT) {
        double sum = 0.0;
        for (int i = 0; i < 64; ++i) sum += red[wave][lane][i];
        loss[((long long)blockIdx.x * WAVES + wave) * NT + lane] = sum;
    }
}

// One wave per (tile, chunk): 16x16 tile of Act^T G over kChunk samples, or
// a bias column-sum tile.
template <int D>
__global__ __launch_bounds__(64) void wgrad_kernel(const __bf16* __restrict__ act, long long B, int cpw,
                                                   float* __restrict__ part)
{
    using L = MlpLayout<D>;
    using A = Act<D>;
    using TT = Tiles<D>;
    const int tile = blockIdx.x, chunk = blockIdx.y;
    const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
    const long long b0 = (long long)chunk * cpw * kChunk;
    const int ns = cpw * kChunk;  // samples of this part
    float* out = part + ((long long)chunk * TT::n + tile) * 256;
    int t = tile;
    const __bf16 *a = nullptr, *gm = nullptr;
    int m0 = 0, n0 = 0;
    if (t < TT::t1) {
        a = act + A::x(B), gm = act + A::dz1(B);
        m0 = (t / (L::H / 16)) * 16, n0 = (t % (L::H / 16)) * 16;
    } else if ((t -= TT::t1) < TT::t2) {
        a = act + A::h1(B), gm = act + A::dz2(B);
        m0 = (t / (L::H / 16)) * 16, n0 = (t % (L::H / 16)) * 16;
    } else if ((t -= TT::t2) < TT::t3) {
        a = act + A::h2(B), gm = act + A::dy(B);
        m0 = t * 16, n0 = 0;
    } else {
        // bias: row sums of dZ1 / dZ2 / dY over the chunk (feature-major:
        // each group of 16 lanes takes 16 features, the 4 groups split the
        // samples; folded in group order)
        t -= TT::t3;
        const __bf16* rows;
        if (t < TT::b1) {
            rows = act + A::dz1(B) + (long long)(t * 16) * B;
        } else if ((t -= TT::b1) < TT::b2) {
            rows = act + A::dz2(B) + (long long)(t * 16) * B;
        } else {
            rows = act + A::dy(B);
        }
        float sacc = 0.f;
        const __bf16* row = rows + (long long)c * B + b0;
        for (int r = g; r < ns; r += 4) sacc += (float)row[r];
        __shared__ float red[64];
        red[lane] = sacc;
        __syncthreads();
        if (lane < 16) out[lane] = (red[lane] + red[16 + lane]) + (red[32 + lane] + red[48 + lane]);
        return;
    }
    // A[m][k] = Act[m0 + m][sample k], B[k][n] = G[n0 + n][sample k]: lane l
    // holds 8 consecutive samples of feature l % 16 of each
    const __bf16* ar = a + (long long)(m0 + c) * B + b0 + 8 * g;
    const __bf16* gr = gm + (long long)(n0 + c) * B + b0 + 8 * g;
    // two independent accumulation chains (even / odd 32-sample steps)
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int k0 = 0; k0 < ns; k0 += 64) {
        acc0 = mfma32(*reinterpret_cast<const bf16x8*>(ar + k0), *reinterpret_cast<const bf16x8*>(gr + k0), acc0);
        acc1 = mfma32(*reinterpret_cast<const bf16x8*>(ar + k0 + 32), *reinterpret_cast<const bf16x8*>(gr + k0 + 32),
                      acc1);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(4 * g + i) * 16 + c] = acc0[i] + acc1[i];
}

// Slot of feature f in a chained (D-layout) 32-chunk: the inverse of
// chain_feature.
__device__ __forceinline__ int chain_slot(int f)
{
    const int c = f & ~31, q = f & 31;
    return q < 16 ? c + 8 * (q >> 2) + (q & 3) : c + 8 * ((q - 16) >> 2) + 4 + ((q - 16) & 3);
}

// Write parameter p's BF16 copies into the packed fragments (the entries
// pack_kernel derives from it; padding entries stay zero).
template <int D>
__device__ __forceinline__ void scatter_param(long long p, float v, __bf16* fwd, __bf16* bw3, __bf16* bw2)
{
    using L = MlpLayout<D>;
    const __bf16 b = (__bf16)v;
    auto frag = [](int f, int lane, int i) { return (long long)f * 512 + lane * 8 + i; };
    if (p < L::ob1) {  // W1[feat][m]
        const int feat = (int)(p / L::H), m = (int)(p % L::H);
        fwd[frag(L::F1 + (m / 16) * L::KS1 + feat / 32, ((feat % 32) / 8) * 16 + m % 16, feat % 8)] = b;
    } else if (p >= L::oW2 && p < L::ob2) {  // W2[feat][m]
        const int q = (int)(p - L::oW2), feat = q / L::H, m = q % L::H;
        const int sl = chain_slot(feat);
        fwd[frag(L::F2 + (m / 16) * L::KS2 + sl / 32, ((sl % 32) / 8) * 16 + m % 16, sl % 8)] = b;
        const int sm = chain_slot(m);  // W2^T: A[h1 = feat][slot(h2 = m)]
        bw2[frag((feat / 16) * L::KS2 + sm / 32, ((sm % 32) / 8) * 16 + feat % 16, sm % 8)] = b;
    } else if (p >= L::oW3 && p < L::ob3) {  // W3[feat][o]
        const int q = (int)(p - L::oW3), feat = q / L::OUT, o = q % L::OUT;
        const int sl = chain_slot(feat);
        fwd[frag(L::F3 + sl / 32, ((sl % 32) / 8) * 16 + o, sl % 8)] = b;
        bw3[(long long)(feat / 16) * 256 + ((o / 4) * 16 + feat % 16) * 4 + o % 4] = b;  // A[h][o]
    }
}

// Fold the weight-gradient partials into the gradient (fixed part order),
// take an Adam step on the FP32 master and refresh its BF16 fragment
// copies. lr = 0 only folds.
template <int D>
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ P, float* __restrict__ m,
                                                   float* __restrict__ v, float* __restrict__ grad,
                                                   const float* __restrict__ part, int n_parts, float lr,
                                                   float bc1, float bc2, __bf16* __restrict__ fwd,
                                                   __bf16* __restrict__ bw)
{
    using L = MlpLayout<D>;
    using TT = Tiles<D>;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L::n_params) return;
    int tile, pos;
    if (p < L::ob1) {  // W1[feat][h]
        const int f = (int)(p / L::H), h = (int)(p % L::H);
        tile = (f / 16) * (L::H / 16) + h / 16, pos = (f % 16) * 16 + h % 16;
    } else if (p < L::oW2) {
        const int h = (int)(p - L::ob1);
        tile = TT::t1 + TT::t2 + TT::t3 + h / 16, pos = h % 16;
    } else if (p < L::ob2) {
        const int q = (int)(p - L::oW2), f = q / L::H, h = q % L::H;
        tile = TT::t1 + (f / 16) * (L::H / 16) + h / 16, pos = (f % 16) * 16 + h % 16;
    } else if (p < L::oW3) {
        const int h = (int)(p - L::ob2);
        tile = TT::t1 + TT::t2 + TT::t3 + TT::b1 + h / 16, pos = h % 16;
    } else if (p < L::ob3) {
        const int q = (int)(p - L::oW3), f = q / L::OUT, o = q % L::OUT;
        tile = TT::t1 + TT::t2 + f / 16, pos = (f % 16) * 16 + o;
    } else {
        const int o = (int)(p - L::ob3);
        tile = TT::n - 1, pos = o;
    }
    // eight loads in flight, summed in part order
    const float* src = part + (long long)tile * 256 + pos;
    const long long stride = (long long)TT::n * 256;
    float gsum = 0.f;
    int c = 0;
    for (; c + 8 <= n_parts; c += 8) {
        float x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = src[(c + i) * stride];
#pragma unroll
        for (int i = 0; i < 8; ++i) gsum += x[i];
    }
    for (; c < n_parts; ++c) gsum += src[c * stride];
    grad[p] = gsum;
    if (lr == 0.f) return;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mm = b1 * m[p] + (1.f - b1) * gsum;
    const float vv = b2 * v[p] + (1.f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    const float np = P[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    P[p] = np;
    scatter_param<D>(p, np, fwd, bw, bw + (long long)L::n_bw3_frags * 256);
}

template <int D> constexpr long long bw_bytes()
{
    using L = MlpLayout<D>;
    return (long long)L::n_bw3_frags * 64 * 4 * 2 + (long long)L::n_bw2_frags * 64 * 8 * 2;
}

template <int D> void pack_impl(const float* P, bf16x8* fwd, void* bw, hipStream_t s)
{
    using L = MlpLayout<D>;
    const long long n = (long long)L::n_fwd_frags * 512 + (long long)L::n_bw3_frags * 256 + (long long)L::n_bw2_frags * 512;
    pack_kernel<D><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(P, fwd, static_cast<char*>(bw));
}

// Weight-gradient parts per step: each wave of wgrad_kernel sums cpw
// consecutive chunks; cpw is the largest power of two dividing the chunk
// count that still leaves 4096 waves (four per SIMD, to hide the operand
// loads), so that the fold in adam_kernel reads fewer parts.
template <int D> int wgrad_parts(long long batch)
{
    const int n_chunks = (int)(batch / kChunk);
    int cpw = 1;
    while (n_chunks % (2 * cpw) == 0 && (long long)Tiles<D>::n * (n_chunks / (2 * cpw)) >= 4096) cpw *= 2;
    return n_chunks / cpw;
}

template <int D> void step_impl(const MlpTrainArgs& a, float lr, long long step, hipStream_t s)
{
    using L = MlpLayout<D>;
    if (a.batch <= 0 || a.batch % kChunk != 0)
        throw std::invalid_argument("mlp: batch must be a positive multiple of " + std::to_string(kChunk));
    if (Act<D>::total(a.batch) * 2 >= (1LL << 31))
        throw std::invalid_argument("mlp: batch too large for one step (activations must stay below 2 GiB)");
    MlpView w;
    w.fwd = a.fwd;
    w.params = a.params;
    if (a.batch >= 16384) {
        constexpr int NT = D == 2 ? 2 : 1;
        fwd_bwd_kernel<D, 4, NT><<<(unsigned)(a.batch / (64 * NT)), 256, 0, s>>>(w, a.bw, a.X, a.T, a.idx, a.perm_n,
                                                                                  a.perm_base, a.perm_bits,
                                                                                  a.perm_key, a.batch,
                                                                                  a.act, a.loss);
    } else {
        fwd_bwd_kernel<D, 1, 1><<<(unsigned)(a.batch / 16), 64, 0, s>>>(w, a.bw, a.X, a.T, a.idx, a.perm_n,
                                                                         a.perm_base, a.perm_bits, a.perm_key,
                                                                         a.batch, a.act,
                                                                         a.loss);
    }
    const int n_parts = wgrad_parts<D>(a.batch);
    const int cpw = (int)(a.batch / kChunk) / n_parts;
    wgrad_kernel<D><<<dim3(Tiles<D>::n, n_parts), 64, 0, s>>>(a.act, a.batch, cpw, a.partials);
    float bc1 = 1.f, bc2 = 1.f;
    if (lr != 0.f) {
        bc1 = 1.f - std::pow(0.9f, (float)step);
        bc2 = 1.f - std::pow(0.999f, (float)step);
    }
    adam_kernel<D><<<(unsigned)((L::n_params + 255) / 256), 256, 0, s>>>(
        a.params, a.adam_m, a.adam_v, a.grad, a.partials, n_parts, lr, bc1, bc2,
        reinterpret_cast<__bf16*>(a.fwd), static_cast<__bf16*>(a.bw));
}

}  // namespace

int mlp_inp(int dim) { return dim == 2 ? MlpDims<2>::INP : MlpDims<3>::INP; }
long long mlp_n_params(int dim) { return dim == 2 ? MlpLayout<2>::n_params : MlpLayout<3>::n_params; }
long long mlp_fwd_frag_count(int dim) { return dim == 2 ? MlpLayout<2>::n_fwd_frags : MlpLayout<3>::n_fwd_frags; }
long long mlp_bw_bytes(int dim) { return dim == 2 ? bw_bytes<2>() : bw_bytes<3>(); }
long long mlp_act_elems(int dim, long long B) { return dim == 2 ? Act<2>::total(B) : Act<3>::total(B); }
long long mlp_partial_elems(int dim, long long B)
{
    const int nt = dim == 2 ? Tiles<2>::n : Tiles<3>::n;
    return (B / kChunk) * nt * 256;
}

void mlp_pack(int dim, const float* params, bf16x8* fwd, void* bw, hipStream_t s)
{
    if (dim == 2)
        pack_impl<2>(params, fwd, bw, s);
    else
        pack_impl<3>(params, fwd, bw, s);
}

void mlp_forward_batch(int dim, const bf16x8* fwd, const float* params, const __bf16* X, long long n, float* Y,
                       hipStream_t s)
{
    if (n <= 0) return;
    MlpView w;
    w.fwd = fwd;
    w.params = params;
    const unsigned grid = (unsigned)((n + 15) / 16);
    if (dim == 2)
        forward_kernel<2><<<grid, 64, 0, s>>>(w, X, n, Y);
    else
        forward_kernel<3><<<grid, 64, 0, s>>>(w, X, n, Y);
}

// ---- harvesting: features of synthesized interior DOFs from the store, and
// their displacement some sweeps later in the same local frame

template <int D>
__global__ __launch_bounds__(256) void features_kernel(const float* __restrict__ x, const float* __restrict__ xprev,
                                                       long long plane, const int* __restrict__ slots,
                                                       const int* __restrict__ blks,
                                                       const BlockDesc* __restrict__ blocks, long long n,
                                                       __bf16* __restrict__ Xo, float* __restrict__ frame,
                                                       float* __restrict__ xt)
{
    using L = MlpLayout<D>;
    const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const long long slot = slots[d];
    const BlockDesc bd = blocks[blks[d]];
    auto ring = [&](int r, int k) -> float {
        long long o;
        if constexpr (D == 2)
            o = (long long)(r / 3 - 1) * bd.st[0] + (r % 3 - 1);
        else
            o = (long long)(r / 9 - 1) * bd.st[0] + (long long)((r / 3) % 3 - 1) * bd.st[1] + (r % 3 - 1);
        return x[k * plane + slot + o];
    };
    const Frame f = local_frame<D>(ring);
    float dprev[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dprev[k] = x[k * plane + slot] - xprev[k * plane + slot];
#pragma unroll
    for (int j = 0; j < L::INP; ++j) Xo[d * L::INP + j] = (__bf16)feature<D>(ring, f, dprev, j);
    float* fr = frame + d * (D * D + 1);
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
        for (int k = 0; k < D; ++k) fr[c * D + k] = f.e[c][k];
    fr[D * D] = f.ok ? f.h : 0.f;
#pragma unroll
    for (int k = 0; k < D; ++k) xt[d * D + k] = x[k * plane + slot];
}

template <int D>
__global__ __launch_bounds__(256) void targets_kernel(const float* __restrict__ x, long long plane,
                                                      const int* __restrict__ slots, long long n,
                                                      const float* __restrict__ frame, const float* __restrict__ xt,
                                                      float* __restrict__ T)
{
    const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= n) return;
    const long long slot = slots[d];
    const float* fr = frame + d * (D * D + 1);
    const float h = fr[D * D];
    const float rh = h > 0.f ? 1.0f / h : 0.f;
    float dx[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dx[k] = x[k * plane + slot] - xt[d * D + k];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float v = 0.f;
        if (c < D)
#pragma unroll
            for (int k = 0; k < D; ++k) v += dx[k] * fr[c * D + k];
        T[d * 4 + c] = v * rh;
    }
}

void mlp_features(int dim, const float* x, const float* xprev, long long plane, const int* slots, const int* blks,
                  const BlockDesc* blocks, long long n, void* Xo, float* frame, float* xt, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (dim == 2)
        features_kernel<2><<<g, 256, 0, s>>>(x, xprev, plane, slots, blks, blocks, n, static_cast<__bf16*>(Xo), frame, xt);
    else
        features_kernel<3><<<g, 256, 0, s>>>(x, xprev, plane, slots, blks, blocks, n, static_cast<__bf16*>(Xo), frame, xt);
}

void mlp_targets(int dim, const float* x, long long plane, const int* slots, long long n, const float* frame,
                 const float* xt, float* T, hipStream_t s)
{
    if (n <= 0) return;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (dim == 2)
        targets_kernel<2><<<g, 256, 0, s>>>(x, plane, slots, n, frame, xt, T);
    else
        targets_kernel<3><<<g, 256, 0, s>>>(x, plane, slots, n, frame, xt, T);
}

void mlp_step(const MlpTrainArgs& a, float lr, long long step, hipStream_t s)
{
    if (a.dim == 2)
        step_impl<2>(a, lr, step, s);
    else
        step_impl<3>(a, lr, step, s);
}

long long mlp_chunk() { return kChunk; }

}  // namespace eggmi
