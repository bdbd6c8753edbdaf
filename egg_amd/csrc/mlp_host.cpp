// Host side of the learned proposal (see mlp_host.hpp).
#include "mlp_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

namespace eggmi
{

namespace
{

uint16_t f2bf(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest even
    return (uint16_t)(u >> 16);
}

float bf2f(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// splitmix64 -> uniform [0, 1)
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

Mlp::Mlp(int d, uint64_t seed, long long capacity) : dim(d)
{
    if (d != 2 && d != 3) throw InvalidArgument("mlp: dim must be 2 or 3");
    if (capacity < 0) throw InvalidArgument("mlp: capacity must be >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw NoDevice("no HIP device available");
    inp = mlp_inp(d);
    in = d == 2 ? MlpDims<2>::IN : MlpDims<3>::IN;
    hidden = d == 2 ? MlpDims<2>::H : MlpDims<3>::H;
    out = d;
    n_params = mlp_n_params(d);
    EGGMI_HIP(hipStreamCreate(&stream));
    // Glorot-uniform weights, zero biases; the last layer starts small so
    // that the untrained network proposes (almost) no motion
    std::vector<float> p((size_t)n_params, 0.0f);
    Rng r {seed};
    auto fill = [&](long off, int fan_in, int fan_out, double gain) {
        const double a = gain * std::sqrt(6.0 / (fan_in + fan_out));
        for (long i = 0; i < (long)fan_in * fan_out; ++i) p[off + i] = (float)((2.0 * r.uniform() - 1.0) * a);
    };
    const long oW2 = (long)in * hidden + hidden, oW3 = oW2 + (long)hidden * hidden + hidden;
    fill(0, in, hidden, 1.0);
    fill(oW2, hidden, hidden, 1.0);
    fill(oW3, hidden, out, 0.1);
    params.upload(p);
    adam_m.alloc(n_params);
    adam_v.alloc(n_params);
    grad.alloc(n_params);
    EGGMI_HIP(hipMemset(adam_m.p, 0, n_params * sizeof(float)));
    EGGMI_HIP(hipMemset(adam_v.p, 0, n_params * sizeof(float)));
    fwd.alloc(mlp_fwd_frag_count(d) * 64);
    bw.alloc(mlp_bw_bytes(d));
    pack();
    cap = capacity;
    if (cap) {
        X.alloc((size_t)cap * inp);
        T.alloc((size_t)cap * 4);
    }
    EGGMI_HIP(hipStreamSynchronize(stream));
}

Mlp::~Mlp()
{
    if (stream) (void)hipStreamDestroy(stream);
}

void Mlp::pack() { mlp_pack(dim, params.p, fwd.p, bw.p, stream); }

MlpView Mlp::view() const
{
    MlpView v;
    v.fwd = fwd.p;
    v.params = params.p;
    return v;
}

void Mlp::get_params(float* o) const
{
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(o, params.p, n_params * sizeof(float), hipMemcpyDeviceToHost));
}

void Mlp::set_params(const float* in_)
{
    for (long long i = 0; i < n_params; ++i)
        if (!std::isfinite(in_[i])) throw InvalidArgument("mlp: non-finite parameter " + std::to_string(i));
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(params.p, in_, n_params * sizeof(float), hipMemcpyHostToDevice));
    EGGMI_HIP(hipMemset(adam_m.p, 0, n_params * sizeof(float)));
    EGGMI_HIP(hipMemset(adam_v.p, 0, n_params * sizeof(float)));
    adam_step = 0;
    pack();
    EGGMI_HIP(hipStreamSynchronize(stream));
}

void Mlp::forward(const float* Xh, long long rows, float* Yh)
{
    if (rows < 0) throw InvalidArgument("mlp: rows must be >= 0");
    if (rows == 0) return;
    std::vector<uint16_t> xb((size_t)rows * inp);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = f2bf(Xh[i]);
    DArr<uint16_t> dx;
    dx.upload(xb);
    DArr<float> dy;
    dy.alloc((size_t)rows * out);
    mlp_forward_batch(dim, fwd.p, params.p, reinterpret_cast<const __bf16*>(dx.p), rows, dy.p, stream);
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(Yh, dy.p, (size_t)rows * out * sizeof(float), hipMemcpyDeviceToHost));
}

void Mlp::data_add(const float* Xh, const float* Th, long long rows)
{
    if (rows < 0) throw InvalidArgument("mlp: rows must be >= 0");
    if (n + rows > cap)
        throw InvalidArgument("mlp: training set capacity " + std::to_string(cap) + " exceeded (" +
                              std::to_string(n + rows) + " rows)");
    std::vector<uint16_t> xb((size_t)rows * inp);
    std::vector<float> tb((size_t)rows * 4, 0.0f);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = f2bf(Xh[i]);
    for (long long r = 0; r < rows; ++r)
        for (int c = 0; c < out; ++c) tb[(size_t)r * 4 + c] = Th[(size_t)r * out + c];
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(X.p + (size_t)n * inp, xb.data(), xb.size() * 2, hipMemcpyHostToDevice));
    EGGMI_HIP(hipMemcpy(T.p + (size_t)n * 4, tb.data(), tb.size() * 4, hipMemcpyHostToDevice));
    n += rows;
}

void Mlp::data_get(float* Xh, float* Th, long long offset, long long rows) const
{
    if (offset < 0 || rows < 0 || offset + rows > n) throw InvalidArgument("mlp: data_get range outside the set");
    std::vector<uint16_t> xb((size_t)rows * inp);
    std::vector<float> tb((size_t)rows * 4);
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(xb.data(), X.p + (size_t)offset * inp, xb.size() * 2, hipMemcpyDeviceToHost));
    EGGMI_HIP(hipMemcpy(tb.data(), T.p + (size_t)offset * 4, tb.size() * 4, hipMemcpyDeviceToHost));
    if (Xh)
        for (size_t i = 0; i < xb.size(); ++i) Xh[i] = bf2f(xb[i]);
    if (Th)
        for (long long r = 0; r < rows; ++r)
            for (int c = 0; c < out; ++c) Th[(size_t)r * out + c] = tb[(size_t)r * 4 + c];
}

void Mlp::ensure_scratch(long long batch)
{
    const long long ch = mlp_chunk();
    if (batch <= 0 || batch % ch != 0)
        throw InvalidArgument("mlp: batch must be a positive multiple of " + std::to_string(ch));
    if (batch <= scratch_batch) return;
    act.alloc(mlp_act_elems(dim, batch));
    partials.alloc(mlp_partial_elems(dim, batch));
    scratch_batch = batch;
}

MlpTrainArgs Mlp::args(long long batch)
{
    MlpTrainArgs a {};
    a.dim = dim;
    a.batch = batch;
    a.X = reinterpret_cast<const __bf16*>(X.p);
    a.T = T.p;
    a.params = params.p;
    a.adam_m = adam_m.p;
    a.adam_v = adam_v.p;
    a.fwd = fwd.p;
    a.bw = bw.p;
    a.act = reinterpret_cast<__bf16*>(act.p);
    a.partials = partials.p;
    a.grad = grad.p;
    return a;
}

double Mlp::gradient(const int32_t* rows_idx, long long batch, float* grad_out)
{
    ensure_scratch(batch);
    for (long long i = 0; i < batch; ++i)
        if (rows_idx[i] < 0 || rows_idx[i] >= n) throw InvalidArgument("mlp: sample index outside the training set");
    std::vector<int> h(rows_idx, rows_idx + batch);
    idx.upload(h);
    loss.alloc(batch / 16);
    MlpTrainArgs a = args(batch);
    a.idx = idx.p;
    a.loss = loss.p;
    mlp_step(a, 0.0f, 0, stream);
    std::vector<double> l(batch / 16);
    EGGMI_HIP(hipStreamSynchronize(stream));
    EGGMI_HIP(hipMemcpy(grad_out, grad.p, n_params * sizeof(float), hipMemcpyDeviceToHost));
    EGGMI_HIP(hipMemcpy(l.data(), loss.p, l.size() * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double v : l) s += v;
    return s / (double)batch;
}

std::vector<double> Mlp::train(int epochs, long long batch, float lr, uint64_t seed)
{
    ensure_scratch(batch);
    if (!(lr > 0.0f) || !std::isfinite(lr)) throw InvalidArgument("mlp: lr must be positive");
    const long long steps = n / batch;
    if (steps == 0) throw InvalidArgument("mlp: training set smaller than one batch");
    if (n >= (1LL << 31)) throw InvalidArgument("mlp: training set too large");
    std::vector<double> hist;
    // each epoch visits the set in a fresh keyed permutation computed on the
    // device (no host shuffle or index upload)
    int bits = 2;
    while ((1LL << bits) < n) bits += 2;
    std::mt19937_64 gen(seed);
    loss.alloc(steps * (batch / 16));
    std::vector<double> l((size_t)(steps * (batch / 16)));
    for (int ep = 0; ep < epochs; ++ep) {
        const unsigned long long key = gen();
        for (long long st = 0; st < steps; ++st) {
            MlpTrainArgs a = args(batch);
            a.idx = nullptr;
            a.perm_n = n;
            a.perm_base = st * batch;
            a.perm_bits = bits;
            a.perm_key = key;
            a.loss = loss.p + st * (batch / 16);
            mlp_step(a, lr, ++adam_step, stream);
        }
        EGGMI_HIP(hipStreamSynchronize(stream));
        EGGMI_HIP(hipMemcpy(l.data(), loss.p, l.size() * sizeof(double), hipMemcpyDeviceToHost));
        double s = 0.0;
        for (double v : l) s += v;
        hist.push_back(s / (double)(steps * batch));
    }
    return hist;
}

long long harvest(Session& s, Mlp& m, int n_snap, int every, int horizon, double omega)
{
    s.need_device();
    if (m.dim != s.dim) throw InvalidArgument("harvest: network and session dimensions differ");
    if (n_snap < 0 || every < 1 || horizon < 1) throw InvalidArgument("harvest: need n_snap >= 0, every >= 1, horizon >= 1");
    Level& L = s.fine;
    L.ensure_synth_list();
    const long long nsyn = (long long)L.syn_slot.size();
    if (nsyn == 0) return 0;
    const Mlp* saved = s.proposal;
    s.proposal = nullptr;  // harvest from the plain method
    DArr<float> frame, xt;
    frame.alloc((size_t)nsyn * (s.dim * s.dim + 1));
    xt.alloc((size_t)nsyn * s.dim);
    long long added = 0;
    try {
        for (int j = 0; j < n_snap && m.n + nsyn <= m.cap; ++j) {
            s.run_jacobi(every, 0, 0.0, omega, every);
            mlp_features(s.dim, L.x[L.cur].p, L.x[1 - L.cur].p, L.plane, L.d_syn_slot.p, L.d_syn_blk.p, L.blocks.p,
                         nsyn, m.X.p + (size_t)m.n * m.inp, frame.p, xt.p, s.stream);
            s.run_jacobi(horizon, 0, 0.0, omega, horizon);
            mlp_targets(s.dim, L.x[L.cur].p, L.plane, L.d_syn_slot.p, nsyn, frame.p, xt.p, m.T.p + (size_t)m.n * 4,
                        s.stream);
            EGGMI_HIP(hipStreamSynchronize(s.stream));
            m.n += nsyn;
            added += nsyn;
        }
    } catch (...) {
        s.proposal = saved;
        throw;
    }
    s.proposal = saved;
    return added;
}

}  // namespace eggmi
