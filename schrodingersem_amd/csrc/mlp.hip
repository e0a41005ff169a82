// mlp.hip -- per-element MLP that predicts MINRES initial guesses, trained on
// the GPU by hand-written forward, backward and Adam kernels.
//
// No counterpart in the reference, whose stage guesses are zero
// (openmp_impl/schrodinger.hpp:222,230,238,246,254,379,387,395).
//
// Network: in -> H1 -> H2 -> out, GELU after the two hidden layers,
// row-major activations [rows x features], one row per element.
// Parameters are FP32 master copies; every GEMM rounds its operands to
// bfloat16 when it stages them in LDS and accumulates in FP32 on the matrix
// cores (v_mfma_f32_32x32x8bf16_1k). GELU uses the exact erf form,
// 0.5 x (1 + erf(x / sqrt 2)) (Hendrycks & Gimpel, 2016). Adam follows
// Kingma & Ba (2015) with bias correction. Gradients of the weights are split
// over row blocks and the slices are added in a fixed order, so a training
// step is bit-reproducible.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "sem_internal.hpp"
#include "sem_mlp.hpp"

namespace sem {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x)
{
    return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// C[z] = op(A) op(B) (+ bias) over k in [z*kchunk, (z+1)*kchunk); op(A) is M x K,
// op(B) is K x N. TA: A stored K x M (A(m,k) = A[k*lda + m]); TB: B stored
// N x K (B(k,n) = B[n*ldb + k]). Split slice z is written at C + z*cstride.
// 64 x 64 output tile per 256-thread workgroup, 32 x 32 per wavefront, k
// tiles of 32. Both operands are staged in LDS as bfloat16 with k
// contiguous ([m][k] and [n][k]), so each MFMA operand is one 8-byte LDS read;
// each thread loads 8 consecutive floats of a tile (two 16-byte loads for
// full, aligned tiles), and the next tile's loads are in flight during the
// current tile's MFMAs. Epilogue EPI: 0 C = result (+ bias); 1 C = GELU(result
// + bias); 2 C = result + bias and C2 = GELU(C) (the pre-activation kept for
// the backward pass), the same values as a separate GELU pass over C.
template <bool TA, bool TB, int EPI = 0>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const float* __restrict__ A, int lda,
                                              const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                              size_t cstride, const float* __restrict__ bias, int kchunk,
                                              float* __restrict__ C2 = nullptr)
{
    constexpr int P = 40;  // LDS row pitch in bfloat16 (80 bytes: 16-byte aligned rows)
    __shared__ __align__(16) short As[64][P];
    __shared__ __align__(16) short Bt[64][P];
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const int kz0 = blockIdx.z * kchunk;
    const int kz1 = min(K, kz0 + kchunk);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w >> 1, wn = w & 1;
    // this thread's 8 values of each tile: A(ma + v*dma, ka + v*dka), B(kb + v*dkb, nb + v*dnb)
    const int ma = TA ? (t & 7) * 8 : t >> 2, ka = TA ? t >> 3 : (t & 3) * 8;
    const int nb = TB ? t >> 2 : (t & 7) * 8, kb = TB ? (t & 3) * 8 : t >> 3;
    const bool vec = (((reinterpret_cast<size_t>(A) | reinterpret_cast<size_t>(B)) & 15) == 0) && (lda & 3) == 0 &&
                     (ldb & 3) == 0 && m0 + 64 <= M && n0 + 64 <= N;
    float ra[8], rb[8];
    auto load = [&](int k0) {
        const bool full = vec && k0 + 32 <= kz1;
        const float* pa = TA ? A + (size_t)(k0 + ka) * lda + (m0 + ma) : A + (size_t)(m0 + ma) * lda + (k0 + ka);
        const float* pb = TB ? B + (size_t)(n0 + nb) * ldb + (k0 + kb) : B + (size_t)(k0 + kb) * ldb + (n0 + nb);
        if (full) {
            const float4 a0 = *reinterpret_cast<const float4*>(pa), a1 = *reinterpret_cast<const float4*>(pa + 4);
            const float4 b0 = *reinterpret_cast<const float4*>(pb), b1 = *reinterpret_cast<const float4*>(pb + 4);
            ra[0] = a0.x; ra[1] = a0.y; ra[2] = a0.z; ra[3] = a0.w; ra[4] = a1.x; ra[5] = a1.y; ra[6] = a1.z; ra[7] = a1.w;
            rb[0] = b0.x; rb[1] = b0.y; rb[2] = b0.z; rb[3] = b0.w; rb[4] = b1.x; rb[5] = b1.y; rb[6] = b1.z; rb[7] = b1.w;
        } else {
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const int m = TA ? m0 + ma + v : m0 + ma, k = TA ? k0 + ka : k0 + ka + v;
                ra[v] = (m < M && k < kz1) ? pa[v] : 0.f;
                const int n = TB ? n0 + nb : n0 + nb + v, kq = TB ? k0 + kb + v : k0 + kb;
                rb[v] = (n < N && kq < kz1) ? pb[v] : 0.f;
            }
        }
    };
    auto pack = [](const float* r, int j) {
        return (unsigned)(unsigned short)f2bf(r[j]) | ((unsigned)(unsigned short)f2bf(r[j + 1]) << 16);
    };
    auto store = [&]() {
        if (TA) {
#pragma unroll
            for (int v = 0; v < 8; ++v) As[ma + v][ka] = f2bf(ra[v]);
        } else {
            *reinterpret_cast<uint4*>(&As[ma][ka]) = make_uint4(pack(ra, 0), pack(ra, 2), pack(ra, 4), pack(ra, 6));
        }
        if (TB) {
            *reinterpret_cast<uint4*>(&Bt[nb][kb]) = make_uint4(pack(rb, 0), pack(rb, 2), pack(rb, 4), pack(rb, 6));
        } else {
#pragma unroll
            for (int v = 0; v < 8; ++v) Bt[nb + v][kb] = f2bf(rb[v]);
        }
    };
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    if (kz0 < kz1) load(kz0);
    for (int k0 = kz0; k0 < kz1; k0 += 32) {
        store();
        __syncthreads();
        if (k0 + 32 < kz1) load(k0 + 32);
#pragma unroll
        for (int ks = 0; ks < 32; ks += 8) {
            const int kq = ks + 4 * (lane >> 5);
            const bf16x4 a = *reinterpret_cast<const bf16x4*>(&As[wm * 32 + (lane & 31)][kq]);
            const bf16x4 b = *reinterpret_cast<const bf16x4*>(&Bt[wn * 32 + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    float* Cz = C + (size_t)blockIdx.z * cstride;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int col = n0 + wn * 32 + (lane & 31);
        if (row < M && col < N) {
            float v = acc[r];
            if (bias) v += bias[col];
            if constexpr (EPI == 1) {
                Cz[(size_t)row * ldc + col] = gelu(v);
            } else {
                Cz[(size_t)row * ldc + col] = v;
                if constexpr (EPI == 2) C2[(size_t)row * ldc + col] = gelu(v);
            }
        }
    }
}

// out[i] = sum over s of slices[s][i] in slice order (eight loads in flight,
// the additions in slice order)
__global__ __launch_bounds__(256) void k_sum_slices(const float* __restrict__ sl, int ns, size_t n, float* out)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float s = 0.f;
        int z = 0;
        for (; z + 8 <= ns; z += 8) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = sl[(size_t)(z + u) * n + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; z < ns; ++z) s += sl[(size_t)z * n + i];
        out[i] = s;
    }
}

// column sums of a rows x cols matrix, two passes (row blocks, then blocks in order)
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, int rows, int cols, int rb,
                                                     float* __restrict__ part)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    const int r0 = blockIdx.y * rb, r1 = min(rows, r0 + rb);
    float s = 0.f;
    int r = r0;
    for (; r + 8 <= r1; r += 8) {  // eight loads in flight, additions in row order
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = X[(size_t)(r + u) * cols + c];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; r < r1; ++r) s += X[(size_t)r * cols + c];
    part[(size_t)blockIdx.y * cols + c] = s;
}


__global__ __launch_bounds__(256) void k_gelu_bwd(const float* __restrict__ H, float* __restrict__ dA, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dA[i] *= gelu_d(H[i]);
}

// dY = 2 (Y - T) / (rows * nout) on the nout real columns (padding columns 0);
// per-block partial of sum (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T, float* dY,
                                             int rows, int ld, int nout, double* part)
{
    __shared__ double sh[4];
    double acc = 0;
    const size_t n = (size_t)rows * ld;
    const float sc = 2.f / ((float)rows * (float)nout);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int c = (int)(i % ld);
        float d = 0.f;
        if (c < nout) {
            d = Y[i] - T[i];
            acc += (double)d * (double)d;
        }
        dY[i] = sc * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// sum of n doubles in a fixed order: thread t adds p[t], p[t+256], ... in
// turn, then a pairwise tree over the 256 thread sums (one workgroup of 256)
__global__ __launch_bounds__(256) void k_sum_double(const double* p, int n, double* out)
{
    __shared__ double sh[256];
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam (Kingma & Ba 2015): m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
// w -= lr * (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps)
__global__ __launch_bounds__(256) void k_adam(float* w, const float* g, float* m, float* v, size_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// features of element rows from device-layout vectors (see sem_mlp.hpp)
__global__ __launch_bounds__(256) void k_gather_features(MlpFeatureArgs a)
{
    const int N = a.N, K = a.K, NN = N * N;
    const long total = (long)a.E * NN;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
        const int e = (int)(t / NN), n = (int)(t % NN);
        const int i = n % N, j = n / N;
        const size_t p = (size_t)(e / K) * K * NN + (size_t)i * K * N + (size_t)(e % K) * N + j;
        float* row = a.X + (size_t)e * a.ld;
        for (int g = 0; g < a.ngroups; ++g) row[g * NN + n] = (float)(a.scale[g] * a.src[g][p]);
        if (n == 0)
            for (int c = a.ngroups * NN; c < a.ld; ++c) {
                const int q = c - a.ngroups * NN;
                row[c] = q < a.extra ? a.extra_val[q] : 0.f;
            }
    }
}

// out[c][pos(e,i,j)] = base[c][pos] + scale * Y[e][c*N*N + i + j*N]
__global__ __launch_bounds__(256) void k_scatter_output(MlpScatterArgs a)
{
    const int N = a.N, K = a.K, NN = N * N;
    const long total = (long)a.E * NN;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
        const int e = (int)(t / NN), n = (int)(t % NN);
        const int i = n % N, j = n / N;
        const size_t p = (size_t)(e / K) * K * NN + (size_t)i * K * N + (size_t)(e % K) * N + j;
        for (int c = 0; c < a.ncomp; ++c) {
            const double y = (double)a.Y[(size_t)e * a.ld + c * NN + n];
            a.out[(size_t)c * a.len + p] = (a.base ? a.base[(size_t)c * a.len + p] : 0.0) + a.scale * y;
        }
    }
}

// T[e][c*N*N + n] = scale * (z - base)[c][pos]   (training targets)
__global__ __launch_bounds__(256) void k_gather_target(MlpScatterArgs a, float* T, const double* z)
{
    const int N = a.N, K = a.K, NN = N * N;
    const long total = (long)a.E * NN;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < total; t += (long)gridDim.x * 256) {
        const int e = (int)(t / NN), n = (int)(t % NN);
        const int i = n % N, j = n / N;
        const size_t p = (size_t)(e / K) * K * NN + (size_t)i * K * N + (size_t)(e % K) * N + j;
        float* row = T + (size_t)e * a.ld;
        for (int c = 0; c < a.ncomp; ++c) {
            const double d = z[(size_t)c * a.len + p] - (a.base ? a.base[(size_t)c * a.len + p] : 0.0);
            row[c * NN + n] = (float)(a.scale * d);
        }
        if (n == 0)
            for (int c = a.ncomp * NN; c < a.ld; ++c) row[c] = 0.f;
    }
}

static int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

// --------------------------------------------------------------- host --
Mlp::Mlp(int in_, int h1_, int h2_, int out_, unsigned long long seed, hipStream_t s) : stream(s)
{
    in = in_;
    h1 = h1_;
    h2 = h2_;
    out = out_;
    in_pad = (in + 63) / 64 * 64;
    out_pad = (out + 63) / 64 * 64;
    // parameter block: W1 [in_pad x h1], b1, W2 [h1 x h2], b2, W3 [h2 x out_pad], b3
    off[0] = 0;
    off[1] = off[0] + (size_t)in_pad * h1;
    off[2] = off[1] + h1;
    off[3] = off[2] + (size_t)h1 * h2;
    off[4] = off[3] + h2;
    off[5] = off[4] + (size_t)h2 * out_pad;
    nparams = off[5] + out_pad;
    SEM_HIP(hipMalloc(&params, nparams * sizeof(float)));
    SEM_HIP(hipMalloc(&grads, nparams * sizeof(float)));
    SEM_HIP(hipMalloc(&adam_m, nparams * sizeof(float)));
    SEM_HIP(hipMalloc(&adam_v, nparams * sizeof(float)));
    SEM_HIP(hipMalloc(&red, 8192 * sizeof(double)));
    // Glorot-uniform weights from a SplitMix64 stream, zero biases; the last
    // layer starts at zero so the untrained network predicts no correction
    std::vector<float> h(nparams, 0.f);
    uint64_t st = seed;
    auto uni = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    auto init = [&](size_t o, int rows, int cols, int real_rows, int real_cols) {
        const double a = std::sqrt(6.0 / (real_rows + real_cols));
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c)
                h[o + (size_t)r * cols + c] = (r < real_rows && c < real_cols) ? (float)((2 * uni() - 1) * a) : 0.f;
    };
    init(off[0], in_pad, h1, in, h1);
    init(off[2], h1, h2, h1, h2);
    set_params(h.data());
    SEM_HIP(hipMemsetAsync(adam_m, 0, nparams * sizeof(float), stream));
    SEM_HIP(hipMemsetAsync(adam_v, 0, nparams * sizeof(float), stream));
    SEM_HIP(hipStreamSynchronize(stream));
}

Mlp::~Mlp()
{
    if (stream) hipStreamSynchronize(stream);
    for (float* p : {params, grads, adam_m, adam_v, X, H1, A1, H2, A2, Y, T, dY, dA2, dA1, slices, colpart}) hipFree(p);
    hipFree(red);
}

void Mlp::set_params(const float* host)
{
    SEM_HIP(hipMemcpyAsync(params, host, nparams * sizeof(float), hipMemcpyHostToDevice, stream));
    SEM_HIP(hipStreamSynchronize(stream));
}

void Mlp::get_params(float* host, bool grad)
{
    SEM_HIP(hipMemcpyAsync(host, grad ? grads : params, nparams * sizeof(float), hipMemcpyDeviceToHost, stream));
    SEM_HIP(hipStreamSynchronize(stream));
}

void Mlp::reserve(int rows_)
{
    if (rows_ <= cap) return;
    SEM_HIP(hipStreamSynchronize(stream));
    for (float** p : {&X, &H1, &A1, &H2, &A2, &Y, &T, &dY, &dA2, &dA1, &slices, &colpart}) {
        hipFree(*p);
        *p = nullptr;
    }
    cap = rows_;
    const size_t r = (size_t)cap;
    SEM_HIP(hipMalloc(&X, r * in_pad * sizeof(float)));
    SEM_HIP(hipMalloc(&H1, r * h1 * sizeof(float)));
    SEM_HIP(hipMalloc(&A1, r * h1 * sizeof(float)));
    SEM_HIP(hipMalloc(&H2, r * h2 * sizeof(float)));
    SEM_HIP(hipMalloc(&A2, r * h2 * sizeof(float)));
    SEM_HIP(hipMalloc(&Y, r * out_pad * sizeof(float)));
    SEM_HIP(hipMalloc(&T, r * out_pad * sizeof(float)));
    SEM_HIP(hipMalloc(&dY, r * out_pad * sizeof(float)));
    SEM_HIP(hipMalloc(&dA2, r * h2 * sizeof(float)));
    SEM_HIP(hipMalloc(&dA1, r * h1 * sizeof(float)));
    const size_t big = std::max({(size_t)in_pad * h1, (size_t)h1 * h2, (size_t)h2 * out_pad});
    SEM_HIP(hipMalloc(&slices, kSplits * big * sizeof(float)));
    const size_t cb = (r + kColRows - 1) / kColRows;
    SEM_HIP(hipMalloc(&colpart, cb * std::max({h1, h2, out_pad}) * sizeof(float)));
}

// C = op(A) op(B) (+ bias); split > 1 sums row-block slices in order
void Mlp::gemm(bool ta, bool tb, int M, int N, int K, const float* A, int lda, const float* B, int ldb, float* C,
               int ldc, const float* bias, int split)
{
    dim3 g((N + 63) / 64, (M + 63) / 64, split), b(256);
    const int kchunk = split > 1 ? ((K + split * 32 - 1) / (split * 32)) * 32 : K;
    float* dst = split > 1 ? slices : C;
    const size_t cstride = (size_t)M * ldc;
    if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), g, b, 0, stream, M, N, K, A, lda, B, ldb, dst, ldc, cstride, bias, kchunk);
    else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), g, b, 0, stream, M, N, K, A, lda, B, ldb, dst, ldc, cstride, bias, kchunk);
    else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), g, b, 0, stream, M, N, K, A, lda, B, ldb, dst, ldc, cstride, bias, kchunk);
    else hipLaunchKernelGGL((k_gemm<true, true>), g, b, 0, stream, M, N, K, A, lda, B, ldb, dst, ldc, cstride, bias, kchunk);
    if (split > 1) {
        const size_t n = (size_t)M * ldc;
        hipLaunchKernelGGL(k_sum_slices, dim3(grid_for(n)), dim3(256), 0, stream, slices, split, n, C);
    }
}

void Mlp::colsum(const float* Xm, int rows, int cols, float* out)
{
    const int nb = (rows + kColRows - 1) / kColRows;
    hipLaunchKernelGGL(k_colsum_part, dim3((cols + 255) / 256, nb), dim3(256), 0, stream, Xm, rows, cols, kColRows,
                       colpart);
    hipLaunchKernelGGL(k_sum_slices, dim3(grid_for(cols)), dim3(256), 0, stream, colpart, nb, (size_t)cols, out);
}

void Mlp::forward(int rows, bool keep)
{
    reserve(rows);
    const float *W1 = params + off[0], *b1 = params + off[1], *W2 = params + off[2], *b2 = params + off[3],
                *W3 = params + off[4], *b3 = params + off[5];
    // GELU in the GEMM epilogue; the pre-activations H1, H2 are written only
    // when a backward pass follows
    auto layer = [&](int n, int k, const float* in, const float* W, const float* b, float* H, float* Aout) {
        dim3 g((n + 63) / 64, (rows + 63) / 64, 1);
        if (keep)
            hipLaunchKernelGGL((k_gemm<false, false, 2>), g, dim3(256), 0, stream, rows, n, k, in, k, W, n, H, n,
                               (size_t)0, b, k, Aout);
        else
            hipLaunchKernelGGL((k_gemm<false, false, 1>), g, dim3(256), 0, stream, rows, n, k, in, k, W, n, Aout, n,
                               (size_t)0, b, k, nullptr);
    };
    layer(h1, in_pad, X, W1, b1, H1, A1);
    layer(h2, h1, A1, W2, b2, H2, A2);
    gemm(false, false, rows, out_pad, h2, A2, h2, W3, out_pad, Y, out_pad, b3, 1);
}

// forward + mean-squared-error backward; loss (sum of squares / (rows*out)) in red[0] when host_loss
double Mlp::backward(int rows, bool host_loss)
{
    forward(rows, true);
    const int nb = std::min(grid_for((size_t)rows * out_pad), 4096);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, stream, Y, T, dY, rows, out_pad, out, red + 16);
    hipLaunchKernelGGL(k_sum_double, dim3(1), dim3(256), 0, stream, red + 16, nb, red);
    backprop(rows);
    if (!host_loss) return 0.0;
    double loss = 0;
    SEM_HIP(hipMemcpyAsync(&loss, red, sizeof(double), hipMemcpyDeviceToHost, stream));
    SEM_HIP(hipStreamSynchronize(stream));
    return loss / ((double)rows * out);
}

void Mlp::backprop(int rows)
{
    const float *W2 = params + off[2], *W3 = params + off[4];
    float *gW1 = grads + off[0], *gb1 = grads + off[1], *gW2 = grads + off[2], *gb2 = grads + off[3],
          *gW3 = grads + off[4], *gb3 = grads + off[5];
    const int split = std::min(kSplits, std::max(1, rows / 1024));
    // layer 3
    gemm(true, false, h2, out_pad, rows, A2, h2, dY, out_pad, gW3, out_pad, nullptr, split);
    colsum(dY, rows, out_pad, gb3);
    gemm(false, true, rows, h2, out_pad, dY, out_pad, W3, out_pad, dA2, h2, nullptr, 1);
    hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for((size_t)rows * h2)), dim3(256), 0, stream, H2, dA2, (size_t)rows * h2);
    // layer 2
    gemm(true, false, h1, h2, rows, A1, h1, dA2, h2, gW2, h2, nullptr, split);
    colsum(dA2, rows, h2, gb2);
    gemm(false, true, rows, h1, h2, dA2, h2, W2, h2, dA1, h1, nullptr, 1);
    hipLaunchKernelGGL(k_gelu_bwd, dim3(grid_for((size_t)rows * h1)), dim3(256), 0, stream, H1, dA1, (size_t)rows * h1);
    // layer 1
    gemm(true, false, in_pad, h1, rows, X, in_pad, dA1, h1, gW1, h1, nullptr, split);
    colsum(dA1, rows, h1, gb1);
}

void Mlp::adam_step()
{
    ++t;
    const float c1 = (float)(1.0 - std::pow((double)beta1, t)), c2 = (float)(1.0 - std::pow((double)beta2, t));
    hipLaunchKernelGGL(k_adam, dim3(grid_for(nparams)), dim3(256), 0, stream, params, grads, adam_m, adam_v, nparams, lr,
                       beta1, beta2, eps, c1, c2);
}

void launch_gather_features(const MlpFeatureArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_features, dim3(grid_for((size_t)a.E * a.N * a.N)), dim3(256), 0, s, a);
}

void launch_scatter_output(const MlpScatterArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_scatter_output, dim3(grid_for((size_t)a.E * a.N * a.N)), dim3(256), 0, s, a);
}

void launch_gather_target(const MlpScatterArgs& a, float* T, const double* z, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_target, dim3(grid_for((size_t)a.E * a.N * a.N)), dim3(256), 0, s, a, T, z);
}

}  // namespace sem
