// HIP kernels of the structured TMOP block-Jacobi smoother (gfx950, wave64).
//
// One sweep is one launch of `sweep{2,3}d`: workgroups [0, n_tiles) each own
// a tile of interior-eligible nodes of one block and stage the tile plus a
// one-node halo of both (three) coordinate planes in LDS with coalesced
// loads; the remaining workgroups run one stored-stencil DOF per thread
// (block faces, interfaces, singular corners). Per DOF the kernel does the
// reference's metric pass, Newton solve, <= 10 halvings and omega
// post-blend (reference src/sweep.hpp:1425-1650), reading the frozen snapshot
// `xin` and writing its own slot of `xout`.
//
// The interior patch is synthesized from the tile (reference
// src/structured_patch.hpp:89-130 occurrence order: cell bits outer, corner
// bits inner, axis 0 most significant); all indices are compile-time, so the
// 16 (2D) / 64 (3D) occurrences are straight-line code over registers.
#include "kernels.hpp"
#include "tmop_math.hpp"

#include <utility>

namespace eggmi
{

template <int N, class F> __device__ __forceinline__ void static_for(F&& f)
{
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (f(std::integral_constant<int, I> {}), ...);
    }(std::make_integer_sequence<int, N> {});
}

constexpr float kInf = __builtin_huge_valf();

// ------------------------------------------------------------------ 2D tables

// Occurrence `occ` of an interior 2D patch in local 3x3 coordinates (the DOF
// sits at (1,1)): corner (cr, cc), axis-0 neighbour (ar, ac), axis-1
// neighbour (br, bc), signs and the DOF's role.
struct Occ2 {
    int cr, cc, ar, ac, br, bc, s0, s1, role;
};

__host__ __device__ constexpr Occ2 occ2(int occ)
{
    const int cell = occ >> 2, o = occ & 3;
    const int o0 = o >> 1, o1 = o & 1;
    const int cr = (cell >> 1) + o0, cc = (cell & 1) + o1;
    const int ar = cr + (o0 ? -1 : 1), ac = cc;
    const int br = cr, bc = cc + (o1 ? -1 : 1);
    const int role = (cr == 1 && cc == 1) ? 0 : (ar == 1 && ac == 1) ? 1 : (br == 1 && bc == 1) ? 2 : -1;
    return {cr, cc, ar, ac, br, bc, o0 ? -1 : 1, o1 ? -1 : 1, role};
}

__host__ __device__ constexpr int occ2_m1_index(int occ)
{
    int n = 0;
    for (int o = 0; o < occ; ++o) n += occ2(o).role < 0 ? 1 : 0;
    return n;
}

// ------------------------------------------------------------------ 3D tables

struct Occ3 {
    int c;     // corner local index a0*9 + a1*3 + a2 (DOF = 13)
    int n[3];  // axis-k neighbour local index
    int s[3];
    int role;
};

__host__ __device__ constexpr Occ3 occ3(int occ)
{
    const int cell = occ >> 3, o = occ & 7;
    int corner[3] = {0, 0, 0}, ob[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
        ob[k] = (o >> (2 - k)) & 1;
        corner[k] = ((cell >> (2 - k)) & 1) + ob[k];
    }
    Occ3 r {};
    r.c = corner[0] * 9 + corner[1] * 3 + corner[2];
    r.role = (r.c == 13) ? 0 : -1;
    for (int k = 0; k < 3; ++k) {
        int nb[3] = {corner[0], corner[1], corner[2]};
        nb[k] += ob[k] ? -1 : 1;
        r.n[k] = nb[0] * 9 + nb[1] * 3 + nb[2];
        r.s[k] = ob[k] ? -1 : 1;
        if (r.role < 0 && r.n[k] == 13) r.role = k + 1;
    }
    return r;
}

__host__ __device__ constexpr int occ3_m1_index(int occ)
{
    int n = 0;
    for (int o = 0; o < occ; ++o) n += occ3(o).role < 0 ? 1 : 0;
    return n;
}

// ------------------------------------------------------------------ common

template <int KIND> __device__ __forceinline__ bool accept(float e, float e0, float md)
{
    return isfinite(e) && (e <= e0 + Tol<float>::energy) && (KIND == kUntangle || md > 0.0f);
}

// Stored-stencil DOF (2D): one thread, occurrences read from the tables.
template <int KIND, bool TAU> __device__ void stored2d(const SweepArgs& a, int d)
{
    const StoredView& st = a.st;
    if (d >= st.n) return;
    const float* X = a.xin;
    const long long PL = a.plane;
    const int dof = st.dof[d];
    const int p0 = st.off[d], p1 = st.off[d + 1];
    const float x0 = X[dof], y0 = X[PL + dof];
    float g0 = 0, g1 = 0, h00 = 0, h01 = 0, h11 = 0, e0 = 0, md = kInf;
    for (int p = p0; p < p1; ++p) {
        const int m = st.meta[p];
        const int role = (m & 3) - 1;
        const float s0 = (m & 4) ? -1.0f : 1.0f, s1 = (m & 8) ? -1.0f : 1.0f;
        const int ic = st.gc[p], ia = st.gn[p], ib = st.gn[st.total + p];
        const float cx = X[ic], cy = X[PL + ic];
        const float A00 = s0 * (X[ia] - cx), A01 = s1 * (X[ib] - cx);
        const float A10 = s0 * (X[PL + ia] - cy), A11 = s1 * (X[PL + ib] - cy);
        const float* w = st.W ? st.W + 4ll * p : a.Wu;
        const float ta = A00 * w[0] + A01 * w[2], tb = A00 * w[1] + A01 * w[3];
        const float tc = A10 * w[0] + A11 * w[2], td = A10 * w[1] + A11 * w[3];
        const float det = A00 * A11 - A01 * A10;
        if (det < md) md = det;
        float val;
        if (role >= 0) {
            float u0, u1;
            if (role == 0) {
                u0 = -(s0 * w[0] + s1 * w[2]);
                u1 = -(s0 * w[1] + s1 * w[3]);
            } else if (role == 1) {
                u0 = s0 * w[0];
                u1 = s0 * w[1];
            } else {
                u0 = s1 * w[2];
                u1 = s1 * w[3];
            }
            val = mu2_grad_hess<KIND, float>(ta, tb, tc, td, u0, u1, a.delta, g0, g1, h00, h01, h11);
        } else {
            val = mu2_value<KIND, float>(ta, tb, tc, td, a.delta);
        }
        e0 += val;
    }
    float tau0 = 0, tau1 = 0;
    if constexpr (TAU) {
        tau0 = a.tau[dof];
        tau1 = a.tau[PL + dof];
        g0 -= tau0;
        g1 -= tau1;
        e0 -= tau0 * x0;
        e0 -= tau1 * y0;
    }
    float dx, dy;
    solve2<float>(h00, h01, h01, h11, g0, g1, dx, dy);
    float alpha = 1.0f, cx = x0, cy = y0;
    for (int it = 0; it < 10; ++it) {
        const float tx = x0 + alpha * dx, ty = y0 + alpha * dy;
        float e = 0, mt = kInf;
        for (int p = p0; p < p1; ++p) {
            const int m = st.meta[p];
            const float s0 = (m & 4) ? -1.0f : 1.0f, s1 = (m & 8) ? -1.0f : 1.0f;
            const int ic = st.gc[p], ia = st.gn[p], ib = st.gn[st.total + p];
            const float ccx = ic == dof ? tx : X[ic], ccy = ic == dof ? ty : X[PL + ic];
            const float ax = ia == dof ? tx : X[ia], ay = ia == dof ? ty : X[PL + ia];
            const float bx = ib == dof ? tx : X[ib], by = ib == dof ? ty : X[PL + ib];
            const float A00 = s0 * (ax - ccx), A01 = s1 * (bx - ccx);
            const float A10 = s0 * (ay - ccy), A11 = s1 * (by - ccy);
            const float* w = st.W ? st.W + 4ll * p : a.Wu;
            const float ta = A00 * w[0] + A01 * w[2], tb = A00 * w[1] + A01 * w[3];
            const float tc = A10 * w[0] + A11 * w[2], td = A10 * w[1] + A11 * w[3];
            e += mu2_value<KIND, float>(ta, tb, tc, td, a.delta);
            mt = fminf(mt, A00 * A11 - A01 * A10);
        }
        if constexpr (TAU) {
            e -= tau0 * tx;
            e -= tau1 * ty;
        }
        if (accept<KIND>(e, e0, mt)) {
            cx = tx;
            cy = ty;
            break;
        }
        alpha *= 0.5f;
    }
    if (a.omega != 1.0f) {
        cx = x0 + a.omega * (cx - x0);
        cy = y0 + a.omega * (cy - y0);
    }
    a.xout[dof] = cx;
    a.xout[PL + dof] = cy;
}

template <int KIND, int TX, int TY, bool TAU>
__global__ __launch_bounds__(TX* TY) void sweep2d(const SweepArgs a)
{
    if (int(blockIdx.x) >= a.n_tiles) {
        stored2d<KIND, TAU>(a, (int(blockIdx.x) - a.n_tiles) * TX * TY + int(threadIdx.x));
        return;
    }
    __shared__ float sx[TY + 2][TX + 2];
    __shared__ float sy[TY + 2][TX + 2];
    const TileDesc td = a.tiles[blockIdx.x];
    const BlockDesc bd = a.blocks[td.blk];
    const long long PL = a.plane;
    const int tid = threadIdx.x;
    for (int e = tid; e < (TY + 2) * (TX + 2); e += TX * TY) {
        const int r = e / (TX + 2), c = e - r * (TX + 2);
        const int li = min(td.i0 - 1 + r, bd.n[0]);
        const int lj = min(td.j0 - 1 + c, bd.n[1]);
        const long long node = bd.base + (long long)(li + 1) * bd.st[0] + (lj + 1);
        sx[r][c] = a.xin[node];
        sy[r][c] = a.xin[PL + node];
    }
    __syncthreads();
    const int ty = tid / TX, tx = tid - ty * TX;
    const int i = td.i0 + ty, j = td.j0 + tx;
    if (i > bd.n[0] - 2 || j > bd.n[1] - 2) return;
    const long long node = bd.base + (long long)(i + 1) * bd.st[0] + (j + 1);
    if (!a.mask[node]) return;

    float px[3][3], py[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            px[r][c] = sx[ty + r][tx + c];
            py[r][c] = sy[ty + r][tx + c];
        }
    const float delta = a.delta;
    float g0 = 0, g1 = 0, h00 = 0, h01 = 0, h11 = 0, e0 = 0, md = kInf, mdc = kInf;
    float vc[4];
    static_for<16>([&](auto I) {
        constexpr Occ2 O = occ2(decltype(I)::value);
        const float A00 = O.s0 * (px[O.ar][O.ac] - px[O.cr][O.cc]);
        const float A01 = O.s1 * (px[O.br][O.bc] - px[O.cr][O.cc]);
        const float A10 = O.s0 * (py[O.ar][O.ac] - py[O.cr][O.cc]);
        const float A11 = O.s1 * (py[O.br][O.bc] - py[O.cr][O.cc]);
        const float det = A00 * A11 - A01 * A10;
        if (det < md) md = det;
        float val;
        if constexpr (O.role >= 0) {
            constexpr float u0 = O.role == 0 ? float(-O.s0) : (O.role == 1 ? float(O.s0) : 0.0f);
            constexpr float u1 = O.role == 0 ? float(-O.s1) : (O.role == 2 ? float(O.s1) : 0.0f);
            val = mu2_grad_hess<KIND, float>(A00, A01, A10, A11, u0, u1, delta, g0, g1, h00, h01, h11);
        } else {
            val = mu2_value<KIND, float>(A00, A01, A10, A11, delta);
            vc[occ2_m1_index(decltype(I)::value)] = val;
            mdc = fminf(mdc, det);
        }
        e0 += val;
    });
    const float x0 = px[1][1], y0 = py[1][1];
    float tau0 = 0, tau1 = 0;
    if constexpr (TAU) {
        tau0 = a.tau[node];
        tau1 = a.tau[PL + node];
        g0 -= tau0;
        g1 -= tau1;
        e0 -= tau0 * x0;
        e0 -= tau1 * y0;
    }
    float dx, dy;
    solve2<float>(h00, h01, h01, h11, g0, g1, dx, dy);
    float alpha = 1.0f, cx = x0, cy = y0;
    for (int it = 0; it < 10; ++it) {
        const float tx_ = x0 + alpha * dx, ty_ = y0 + alpha * dy;
        float e = 0, mt = mdc;
        static_for<16>([&](auto I) {
            constexpr Occ2 O = occ2(decltype(I)::value);
            if constexpr (O.role < 0) {
                e += vc[occ2_m1_index(decltype(I)::value)];
            } else {
                const float cxp = (O.cr == 1 && O.cc == 1) ? tx_ : px[O.cr][O.cc];
                const float cyp = (O.cr == 1 && O.cc == 1) ? ty_ : py[O.cr][O.cc];
                const float axp = (O.ar == 1 && O.ac == 1) ? tx_ : px[O.ar][O.ac];
                const float ayp = (O.ar == 1 && O.ac == 1) ? ty_ : py[O.ar][O.ac];
                const float bxp = (O.br == 1 && O.bc == 1) ? tx_ : px[O.br][O.bc];
                const float byp = (O.br == 1 && O.bc == 1) ? ty_ : py[O.br][O.bc];
                const float A00 = O.s0 * (axp - cxp), A01 = O.s1 * (bxp - cxp);
                const float A10 = O.s0 * (ayp - cyp), A11 = O.s1 * (byp - cyp);
                e += mu2_value<KIND, float>(A00, A01, A10, A11, delta);
                mt = fminf(mt, A00 * A11 - A01 * A10);
            }
        });
        if constexpr (TAU) {
            e -= tau0 * tx_;
            e -= tau1 * ty_;
        }
        if (accept<KIND>(e, e0, mt)) {
            cx = tx_;
            cy = ty_;
            break;
        }
        alpha *= 0.5f;
    }
    if (a.omega != 1.0f) {
        cx = x0 + a.omega * (cx - x0);
        cy = y0 + a.omega * (cy - y0);
    }
    a.xout[node] = cx;
    a.xout[PL + node] = cy;
}

// ------------------------------------------------------------------ 3D

template <int KIND, bool TAU> __device__ void stored3d(const SweepArgs& a, int d)
{
    const StoredView& st = a.st;
    if (d >= st.n) return;
    const float* X = a.xin;
    const long long PL = a.plane;
    const int dof = st.dof[d];
    const int p0 = st.off[d], p1 = st.off[d + 1];
    float pos[3] = {X[dof], X[PL + dof], X[2 * PL + dof]};
    float g[3] = {0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0}, e0 = 0, md = kInf;
    auto build = [&](int p, const float* sub, float* T, float& det) {
        const int m = st.meta[p];
        const int ic = st.gc[p];
        float cn[3], nb[3][3], s[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) cn[i] = (sub && ic == dof) ? sub[i] : X[i * PL + ic];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int in = st.gn[k * st.total + p];
            s[k] = (m & (4 << k)) ? -1.0f : 1.0f;
#pragma unroll
            for (int i = 0; i < 3; ++i) nb[k][i] = (sub && in == dof) ? sub[i] : X[i * PL + in];
        }
        float A[9];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) A[i * 3 + k] = s[k] * (nb[k][i] - cn[i]);
        det = det3(A);
        const float* w = st.W ? st.W + 9ll * p : a.Wu;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int k = 0; k < 3; ++k) T[i * 3 + k] = A[i * 3] * w[k] + A[i * 3 + 1] * w[3 + k] + A[i * 3 + 2] * w[6 + k];
    };
    for (int p = p0; p < p1; ++p) {
        float T[9], det;
        build(p, nullptr, T, det);
        if (det < md) md = det;
        const int m = st.meta[p];
        const int role = (m & 3) - 1;
        float val;
        if (role >= 0) {
            const float* w = st.W ? st.W + 9ll * p : a.Wu;
            float s[3], u[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) s[k] = (m & (4 << k)) ? -1.0f : 1.0f;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (role == 0)
                    u[c] = -(s[0] * w[c] + s[1] * w[3 + c] + s[2] * w[6 + c]);
                else
                    u[c] = s[role - 1] * w[(role - 1) * 3 + c];
            }
            val = mu3_grad_hess<KIND, float>(T, u, a.delta, g, h);
        } else {
            val = mu3_value<KIND, float>(T, a.delta);
        }
        e0 += val;
    }
    float tau[3] = {0, 0, 0};
    if constexpr (TAU) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            tau[k] = a.tau[k * PL + dof];
            g[k] -= tau[k];
            e0 -= tau[k] * pos[k];
        }
    }
    const float H[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
    float dl[3];
    solve3<float>(H, g, dl);
    float alpha = 1.0f, cur[3] = {pos[0], pos[1], pos[2]};
    for (int it = 0; it < 10; ++it) {
        float tr[3] = {pos[0] + alpha * dl[0], pos[1] + alpha * dl[1], pos[2] + alpha * dl[2]};
        float e = 0, mt = kInf;
        for (int p = p0; p < p1; ++p) {
            float T[9], det;
            build(p, tr, T, det);
            e += mu3_value<KIND, float>(T, a.delta);
            mt = fminf(mt, det);
        }
        if constexpr (TAU) {
#pragma unroll
            for (int k = 0; k < 3; ++k) e -= tau[k] * tr[k];
        }
        if (accept<KIND>(e, e0, mt)) {
            cur[0] = tr[0];
            cur[1] = tr[1];
            cur[2] = tr[2];
            break;
        }
        alpha *= 0.5f;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (a.omega != 1.0f) cur[k] = pos[k] + a.omega * (cur[k] - pos[k]);
        a.xout[k * PL + dof] = cur[k];
    }
}

template <int KIND, int TX, int TY, int TZ, bool TAU>
__global__ __launch_bounds__(TX* TY* TZ) void sweep3d(const SweepArgs a)
{
    constexpr int NT = TX * TY * TZ;
    if (int(blockIdx.x) >= a.n_tiles) {
        stored3d<KIND, TAU>(a, (int(blockIdx.x) - a.n_tiles) * NT + int(threadIdx.x));
        return;
    }
    __shared__ float sh[3][TZ + 2][TY + 2][TX + 2];
    const TileDesc td = a.tiles[blockIdx.x];
    const BlockDesc bd = a.blocks[td.blk];
    const long long PL = a.plane;
    const int tid = threadIdx.x;
    for (int e = tid; e < (TZ + 2) * (TY + 2) * (TX + 2); e += NT) {
        const int r = e / ((TY + 2) * (TX + 2));
        const int rem = e - r * (TY + 2) * (TX + 2);
        const int q = rem / (TX + 2), c = rem - q * (TX + 2);
        const int l0 = min(td.i0 - 1 + r, bd.n[0]);
        const int l1 = min(td.j0 - 1 + q, bd.n[1]);
        const int l2 = min(td.k0 - 1 + c, bd.n[2]);
        const long long node = bd.base + (long long)(l0 + 1) * bd.st[0] + (long long)(l1 + 1) * bd.st[1] + (l2 + 1);
        sh[0][r][q][c] = a.xin[node];
        sh[1][r][q][c] = a.xin[PL + node];
        sh[2][r][q][c] = a.xin[2 * PL + node];
    }
    __syncthreads();
    const int tz = tid / (TX * TY);
    const int rem = tid - tz * TX * TY;
    const int ty = rem / TX, tx = rem - ty * TX;
    const int i = td.i0 + tz, j = td.j0 + ty, k = td.k0 + tx;
    if (i > bd.n[0] - 2 || j > bd.n[1] - 2 || k > bd.n[2] - 2) return;
    const long long node = bd.base + (long long)(i + 1) * bd.st[0] + (long long)(j + 1) * bd.st[1] + (k + 1);
    if (!a.mask[node]) return;

    float P[3][27];
#pragma unroll
    for (int a0 = 0; a0 < 3; ++a0)
#pragma unroll
        for (int a1 = 0; a1 < 3; ++a1)
#pragma unroll
            for (int a2 = 0; a2 < 3; ++a2)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) P[cc][a0 * 9 + a1 * 3 + a2] = sh[cc][tz + a0][ty + a1][tx + a2];

    const float delta = a.delta;
    float g[3] = {0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0}, e0 = 0, md = kInf, mdc = kInf;
    float vc[32];
    static_for<64>([&](auto I) {
        constexpr Occ3 O = occ3(decltype(I)::value);
        float A[9];
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) A[ii * 3 + kk] = float(O.s[kk]) * (P[ii][O.n[kk]] - P[ii][O.c]);
        const float det = det3(A);
        if (det < md) md = det;
        float val;
        if constexpr (O.role >= 0) {
            float u[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                u[cc] = O.role == 0 ? float(-O.s[cc]) : (cc == O.role - 1 ? float(O.s[cc]) : 0.0f);
            val = mu3_grad_hess<KIND, float>(A, u, delta, g, h);
        } else {
            val = mu3_value<KIND, float>(A, delta);
            vc[occ3_m1_index(decltype(I)::value)] = val;
            mdc = fminf(mdc, det);
        }
        e0 += val;
    });
    float pos[3] = {P[0][13], P[1][13], P[2][13]};
    float tau[3] = {0, 0, 0};
    if constexpr (TAU) {
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) {
            tau[kk] = a.tau[kk * PL + node];
            g[kk] -= tau[kk];
            e0 -= tau[kk] * pos[kk];
        }
    }
    const float H[9] = {h[0], h[1], h[2], h[1], h[3], h[4], h[2], h[4], h[5]};
    float dl[3];
    solve3<float>(H, g, dl);
    float alpha = 1.0f, cur[3] = {pos[0], pos[1], pos[2]};
    for (int it = 0; it < 10; ++it) {
        const float tr[3] = {pos[0] + alpha * dl[0], pos[1] + alpha * dl[1], pos[2] + alpha * dl[2]};
        float e = 0, mt = mdc;
        static_for<64>([&](auto I) {
            constexpr Occ3 O = occ3(decltype(I)::value);
            if constexpr (O.role < 0) {
                e += vc[occ3_m1_index(decltype(I)::value)];
            } else {
                float A[9];
#pragma unroll
                for (int ii = 0; ii < 3; ++ii) {
                    const float cv = O.c == 13 ? tr[ii] : P[ii][O.c];
#pragma unroll
                    for (int kk = 0; kk < 3; ++kk) {
                        const float nv = O.n[kk] == 13 ? tr[ii] : P[ii][O.n[kk]];
                        A[ii * 3 + kk] = float(O.s[kk]) * (nv - cv);
                    }
                }
                e += mu3_value<KIND, float>(A, delta);
                mt = fminf(mt, det3(A));
            }
        });
        if constexpr (TAU) {
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) e -= tau[kk] * tr[kk];
        }
        if (accept<KIND>(e, e0, mt)) {
            cur[0] = tr[0];
            cur[1] = tr[1];
            cur[2] = tr[2];
            break;
        }
        alpha *= 0.5f;
    }
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
        if (a.omega != 1.0f) cur[kk] = pos[kk] + a.omega * (cur[kk] - pos[kk]);
        a.xout[kk * PL + node] = cur[kk];
    }
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
    const long long stride = (long long)gridDim.x * kReduceThreads;
    for (long long cid = (long long)blockIdx.x * kReduceThreads + threadIdx.x; cid < a.ncells; cid += stride) {
        int lo = 0, hi = a.nblocks - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (a.cell_off[mid] <= cid)
                lo = mid;
            else
                hi = mid - 1;
        }
        const BlockDesc bd = a.blocks[lo];
        long long r = cid - a.cell_off[lo];
        if constexpr (DIM == 2) {
            const int w1 = bd.n[1] - 1;
            const int ci = int(r / w1), cj = int(r - (long long)ci * w1);
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
            const int ci = int(r / ((long long)w1 * w2));
            r -= (long long)ci * w1 * w2;
            const int cj = int(r / w2), ck = int(r - (long long)cj * w2);
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
        if constexpr (DIM == 2) {
            esum += mu2_d<KIND>(T[0], T[1], T[2], T[3], a.delta);
            emin = fmin(emin, A[0] * A[3] - A[1] * A[2]);
        } else {
            esum += mu3_value<KIND, double>(T, a.delta);
            emin = fmin(emin, det3(A));
        }
    }
    block_reduce_store(esum, emin, partials + 2 * blockIdx.x);
}

__global__ __launch_bounds__(kReduceThreads) void reduce_finish(const double* partials, int nparts, double* out_e,
                                                                double* out_m)
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
    }
}

// ------------------------------------------------------------------ launchers

static constexpr TileShape kTiles2[] = {{64, 4, 1}, {32, 8, 1}, {16, 16, 1}, {128, 2, 1}};
static constexpr TileShape kTiles3[] = {{16, 4, 4}, {8, 8, 4}, {32, 4, 2}};

int num_tile_variants(int dim) { return dim == 2 ? 4 : 3; }
TileShape tile_variant_shape(int dim, int v) { return dim == 2 ? kTiles2[v] : kTiles3[v]; }

template <int KIND, bool TAU> static void sweep2d_v(int v, const SweepArgs& a, hipStream_t s)
{
    const int nst_blocks = (a.st.n + 255) / 256;
    const dim3 grid(a.n_tiles + nst_blocks);
    if (grid.x == 0) return;
    switch (v) {
    case 0: sweep2d<KIND, 64, 4, TAU><<<grid, 256, 0, s>>>(a); break;
    case 1: sweep2d<KIND, 32, 8, TAU><<<grid, 256, 0, s>>>(a); break;
    case 2: sweep2d<KIND, 16, 16, TAU><<<grid, 256, 0, s>>>(a); break;
    default: sweep2d<KIND, 128, 2, TAU><<<grid, 256, 0, s>>>(a); break;
    }
}

template <int KIND, bool TAU> static void sweep3d_v(int v, const SweepArgs& a, hipStream_t s)
{
    const int nst_blocks = (a.st.n + 255) / 256;
    const dim3 grid(a.n_tiles + nst_blocks);
    if (grid.x == 0) return;
    switch (v) {
    case 0: sweep3d<KIND, 16, 4, 4, TAU><<<grid, 256, 0, s>>>(a); break;
    case 1: sweep3d<KIND, 8, 8, 4, TAU><<<grid, 256, 0, s>>>(a); break;
    default: sweep3d<KIND, 32, 4, 2, TAU><<<grid, 256, 0, s>>>(a); break;
    }
}

template <bool TAU> static void sweep_kind(int dim, int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (dim == 2) {
        if (kind == kUntangle)
            sweep2d_v<kUntangle, TAU>(v, a, s);
        else if (kind == kShapeSize)
            sweep2d_v<kShapeSize, TAU>(v, a, s);
        else
            sweep2d_v<kShape, TAU>(v, a, s);
    } else {
        if (kind == kUntangle)
            sweep3d_v<kUntangle, TAU>(v, a, s);
        else if (kind == kShapeSize)
            sweep3d_v<kShapeSize, TAU>(v, a, s);
        else
            sweep3d_v<kShape, TAU>(v, a, s);
    }
}

void launch_sweep(int dim, int kind, int v, const SweepArgs& a, hipStream_t s)
{
    if (a.tau)
        sweep_kind<true>(dim, kind, v, a, s);
    else
        sweep_kind<false>(dim, kind, v, a, s);
}

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

void launch_reduce_finish(const double* partials, int nparts, double* out_e, double* out_m, hipStream_t s)
{
    reduce_finish<<<1, kReduceThreads, 0, s>>>(partials, nparts, out_e, out_m);
}

}  // namespace eggmi
