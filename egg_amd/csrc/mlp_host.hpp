// Host side of the learned proposal: parameters, a device-resident training
// set, the training loop (mlp.hip kernels) and batched inference.
#pragma once

#include "mlp.hpp"
#include "session.hpp"

#include <cstdint>
#include <vector>

namespace eggmi
{

struct Mlp {
    int dim = 2;
    int inp = 32, in = 22, hidden = 64, out = 2;
    long long n_params = 0;
    hipStream_t stream = nullptr;
    DArr<float> params, adam_m, adam_v, grad;
    DArr<bf16x8> fwd;
    DArr<char> bw;
    long long adam_step = 0;
    // training set: features [cap][inp] (BF16 bits), targets [cap][4]
    DArr<uint16_t> X;
    DArr<float> T;
    long long cap = 0, n = 0;
    // per-step scratch
    DArr<uint16_t> act;
    DArr<float> partials;
    DArr<double> loss;
    DArr<int> idx;
    long long scratch_batch = 0;

    Mlp(int dim, uint64_t seed, long long capacity);
    ~Mlp();
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;

    void get_params(float* out) const;
    void set_params(const float* in);  // also resets the Adam state
    void pack();
    MlpView view() const;
    void forward(const float* Xh, long long rows, float* Yh);  // host features [rows][inp]
    void data_add(const float* Xh, const float* Th, long long rows);  // Th [rows][out]
    void data_get(float* Xh, float* Th, long long offset, long long rows) const;
    void data_clear() { n = 0; }
    // gradient of the mean squared error over dataset rows `rows_idx`
    double gradient(const int32_t* rows_idx, long long batch, float* grad_out);
    // Adam epochs over the resident set; returns the mean loss of each epoch
    std::vector<double> train(int epochs, long long batch, float lr, uint64_t seed);

  private:
    void ensure_scratch(long long batch);
    MlpTrainArgs args(long long batch);
};

// Harvest (features at sweep t, displacement over the next `horizon`
// sweeps) pairs of every synthesized interior DOF from plain block-Jacobi
// runs of a session: n_snap snapshots, `every` sweeps apart.
long long harvest(Session& s, Mlp& m, int n_snap, int every, int horizon, double omega);

}  // namespace eggmi
