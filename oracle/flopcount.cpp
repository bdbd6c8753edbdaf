// Op-counting scalar over the sweep kernels' per-sample math (measurement
// infrastructure; pins the algorithmic flop constants the roofline uses).
//
// Instantiates egg_amd/csrc/tmop_math.hpp with `Cnt`, a double that counts
// every +, -, *, / and sqrt as one flop (an FMA is therefore two), and
// replays the per-DOF structure of the interior sweep kernels:
//   F_metric = all samples of the patch (value, plus gradient / contracted
//              Hessian for the samples the DOF belongs to) + the Newton solve
//   F_trial  = one line-search trial (value + det of the samples that move,
//              patch sum of all samples)
// Build + run:  hipcc -x c++ -std=c++20 -O1 oracle/flopcount.cpp -o oracle/_ref/flopcount
#include "../egg_amd/csrc/tmop_math.hpp"

#include <cmath>
#include <cstdio>

static long long g_flops = 0;

struct Cnt {
    double v;
    Cnt(double x = 0) : v(x) {}
};
static Cnt operator+(Cnt a, Cnt b) { ++g_flops; return {a.v + b.v}; }
static Cnt operator-(Cnt a, Cnt b) { ++g_flops; return {a.v - b.v}; }
static Cnt operator*(Cnt a, Cnt b) { ++g_flops; return {a.v * b.v}; }
static Cnt operator/(Cnt a, Cnt b) { ++g_flops; return {a.v / b.v}; }
static Cnt operator-(Cnt a) { return {-a.v}; }
static Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
static Cnt& operator-=(Cnt& a, Cnt b) { a = a - b; return a; }
static bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
static bool operator<(Cnt a, double b) { return a.v < b; }
static Cnt sqrt(Cnt a) { ++g_flops; return {std::sqrt(a.v)}; }
static bool isfinite(Cnt a) { return std::isfinite(a.v); }
namespace eggmi
{
template <> struct Tol<Cnt> {
    static constexpr double energy = 1e-6, znorm = 1e-7;
};
}  // namespace eggmi

using namespace eggmi;

template <int KIND> static void count2d()
{
    // interior 2D patch: 16 samples, 12 with the DOF (role >= 0), 4 without
    const Cnt a(1.1), b(0.1), c(-0.05), d(0.95), u0(-1), u1(-1), delta(0.01);
    Cnt g0 = 0, g1 = 0, h00 = 0, h01 = 0, h11 = 0;
    g_flops = 0;
    for (int o = 0; o < 12; ++o) {
        Cnt A00 = a - b, A01 = b - c, A10 = c - d, A11 = d - a;  // 4 differences
        (void)A00; (void)A01; (void)A10; (void)A11;
        Cnt det = a * d - b * c;
        (void)det;
        Cnt e = mu2_grad_hess<KIND, Cnt>(a, b, c, d, u0, u1, delta, g0, g1, h00, h01, h11);
        (void)e;
    }
    for (int o = 0; o < 4; ++o) {
        Cnt A00 = a - b, A01 = b - c, A10 = c - d, A11 = d - a;
        (void)A00; (void)A01; (void)A10; (void)A11;
        Cnt det = a * d - b * c;
        (void)det;
        (void)mu2_value<KIND, Cnt>(a, b, c, d, delta);
    }
    g_flops += 16;  // energy accumulation
    Cnt x0, x1;
    solve2<Cnt>(Cnt(2), Cnt(0.1), Cnt(0.1), Cnt(1), Cnt(0.3), Cnt(0.2), x0, x1);
    const long long metric = g_flops;
    g_flops = 0;
    Cnt tx = Cnt(0.5) + Cnt(1) * Cnt(0.1), ty = Cnt(0.5) + Cnt(1) * Cnt(0.1);  // trial position
    (void)tx; (void)ty;
    for (int o = 0; o < 12; ++o) {
        Cnt A00 = a - b, A01 = b - c, A10 = c - d, A11 = d - a;
        (void)A00; (void)A01; (void)A10; (void)A11;
        (void)mu2_value<KIND, Cnt>(a, b, c, d, delta);
        Cnt det = a * d - b * c;
        (void)det;
    }
    g_flops += 16;  // patch sum
    const long long trial = g_flops;
    std::printf("2d kind %d  F_metric %lld  F_trial %lld\n", KIND, metric, trial);
}

template <int KIND> static void count3d()
{
    Cnt t[9] = {1.1, 0.1, 0.0, -0.05, 0.95, 0.02, 0.03, -0.01, 1.05}, u[3] = {-1, -1, -1};
    Cnt g[3] = {0, 0, 0}, h[6] = {0, 0, 0, 0, 0, 0};
    const Cnt delta(0.01);
    g_flops = 0;
    for (int o = 0; o < 32; ++o) {
        for (int i = 0; i < 9; ++i) (void)(t[i] - t[(i + 1) % 9]);  // 9 differences
        (void)det3(t);
        (void)mu3_grad_hess<KIND, Cnt>(t, u, delta, g, h);
    }
    for (int o = 0; o < 32; ++o) {
        for (int i = 0; i < 9; ++i) (void)(t[i] - t[(i + 1) % 9]);
        (void)det3(t);
        (void)mu3_value<KIND, Cnt>(t, delta);
    }
    g_flops += 64;
    Cnt H[9] = {2, 0.1, 0, 0.1, 2, 0, 0, 0, 2}, x[3];
    solve3<Cnt>(H, g, x);
    const long long metric = g_flops;
    g_flops = 0;
    for (int i = 0; i < 3; ++i) (void)(Cnt(0.5) + Cnt(1) * Cnt(0.1));
    for (int o = 0; o < 32; ++o) {
        for (int i = 0; i < 9; ++i) (void)(t[i] - t[(i + 1) % 9]);
        (void)mu3_value<KIND, Cnt>(t, delta);
        (void)det3(t);
    }
    g_flops += 64;
    const long long trial = g_flops;
    std::printf("3d kind %d  F_metric %lld  F_trial %lld\n", KIND, metric, trial);
}

int main()
{
    count2d<kShape>();
    count2d<kUntangle>();
    count2d<kShapeSize>();
    count3d<kShape>();
    count3d<kUntangle>();
    count3d<kShapeSize>();
    return 0;
}
