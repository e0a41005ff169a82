// sem_mlp.hpp -- per-element MLP for MINRES initial guesses (mlp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace sem {

// Row-major activations, one row per element. X: [rows x in_pad],
// Y: [rows x out_pad]; T holds training targets with Y's layout.
struct Mlp {
    static constexpr int kSplits = 32;   // row-block split of the weight-gradient GEMMs
    static constexpr int kColRows = 64;  // rows per block of the bias-gradient column sums
    int in = 0, h1 = 0, h2 = 0, out = 0, in_pad = 0, out_pad = 0;
    size_t off[6] = {0, 0, 0, 0, 0, 0};  // W1, b1, W2, b2, W3, b3 in the parameter block
    size_t nparams = 0;
    float* params = nullptr;
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    long t = 0;
    int cap = 0;
    float *X = nullptr, *H1 = nullptr, *A1 = nullptr, *H2 = nullptr, *A2 = nullptr, *Y = nullptr, *T = nullptr;
    float *dY = nullptr, *dA2 = nullptr, *dA1 = nullptr, *slices = nullptr, *colpart = nullptr;
    double* red = nullptr;
    hipStream_t stream = nullptr;

    Mlp(int in, int h1, int h2, int out, unsigned long long seed, hipStream_t s);
    ~Mlp();
    Mlp(const Mlp&) = delete;
    Mlp& operator=(const Mlp&) = delete;
    void reserve(int rows);
    void set_params(const float* host);
    void get_params(float* host, bool grad);
    void forward(int rows, bool keep = false);    // X -> Y (keep: H1, A1, H2, A2 for a backward pass)
    double backward(int rows, bool host_loss);    // X, T -> grads (MSE loss); returns the loss if host_loss
    void backprop(int rows);                      // dY (and the activations of forward) -> grads
    void adam_step();
    void gemm(bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
              const float* bias, int split);
    void colsum(const float* X, int rows, int cols, float* out);
};

// X[e][g*N*N + i + j*N] = scale[g] * src[g][pos(e, i, j)], then `extra`
// constant columns extra_val[0..extra), padding columns zero
struct MlpFeatureArgs {
    const double* src[8];
    double scale[8];
    int ngroups;
    int extra;
    float extra_val[8];
    float* X;
    int ld;
    int E, N, K;
};

// out[c][pos] = base[c][pos] + scale * Y[e][c*N*N + i + j*N]
struct MlpScatterArgs {
    const float* Y;
    int ld;
    double* out;
    const double* base;  // may be null
    double scale;
    size_t len;
    int E, N, K, ncomp;
};

void launch_gather_features(const MlpFeatureArgs& a, hipStream_t s);
void launch_scatter_output(const MlpScatterArgs& a, hipStream_t s);
// T[e][c*N*N + n] = scale * (z - base)[c][pos]
void launch_gather_target(const MlpScatterArgs& a, float* T, const double* z, hipStream_t s);

}  // namespace sem
