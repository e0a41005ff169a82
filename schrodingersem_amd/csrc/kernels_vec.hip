// kernels_vec.hip -- direct-stiffness summation, vector updates and
// deterministic reductions.
//
// DSS is a gather over precomputed index lists: every interface node is one
// work item that reads all its copies, sums them in the reference's order and
// writes the result to every copy. Nothing is accumulated with atomics, so
// repeated runs are bit-identical. Edge sums are owner + other
// (global_sum.hpp:28-34); corner sums are a left fold from +0.0 in
// connected_elements order (global_sum.hpp:54-63).
//
// Reductions are two-stage: every block writes one partial (butterfly inside
// the wave, waves in index order), then a single-block kernel adds the
// partials in a fixed order. The result depends only on the launch shape.
#include <cfloat>

#include "krylov_device.hpp"
#include "sem_internal.hpp"
#include "sem_kernels.hpp"

namespace sem {

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// block sum for blockDim = 256; result valid in thread 0
__device__ __forceinline__ double block_sum256(double v)
{
    __shared__ double sh[4];
    v = wsum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) s = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    return s;
}

// owner weight of entry (i, j) of an element with ownership bits `bits`
__device__ __forceinline__ double owner_weight(int i, int j, int N, unsigned bits)
{
    const bool ri = (i == 0 || i == N - 1), rj = (j == 0 || j == N - 1);
    if (!ri && !rj) return 1.0;
    int bit;
    if (ri && rj) bit = 4 + ((i == 0) ? ((j == 0) ? 0 : 3) : ((j == 0) ? 1 : 2));
    else if (j == 0) bit = 0;
    else if (i == N - 1) bit = 1;
    else if (j == N - 1) bit = 2;
    else bit = 3;
    return ((bits >> bit) & 1u) ? 1.0 : 0.0;
}

// ------------------------------------------------------------------- DSS --
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_dss(DssArgs<T> a)
{
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const long np = a.n_pairs, nc = a.n_corners, nb = a.n_bnd;
    double part = 0;
    bool skip = false;
    if constexpr (MODE == DSS_PCG) {
        const KrylovScalars* sc = reinterpret_cast<const KrylovScalars*>(a.sc);
        skip = sc->done != 0;
    }
    if (!skip) {
        for (int c = 0; c < a.ncomp; ++c) {
            T* y = a.y + (size_t)c * a.len;
            const T post = a.post[c];
            auto finish = [&](T s, int p0) -> T {
                if constexpr (MODE == DSS_SUM_ZB) return post * s;
                if constexpr (MODE == DSS_PCG) {
                    part += (double)a.r2[(size_t)c * a.len + p0] * (double)s;
                    return s;
                }
                return s;
            };
            if constexpr (MODE == DSS_ZERO_BND) {
                if (t >= np + nc && t < np + nc + nb) y[a.bnd[t - np - nc]] = 0;
                continue;
            }
            if (t < np) {
                const int2 pr = a.pairs[t];
                if constexpr (MODE == DSS_UNMASK) {
                    y[pr.y] = y[pr.x];
                } else if constexpr (MODE == DSS_MASK) {
                    y[pr.y] = 0;
                } else {
                    T s = y[pr.x] + y[pr.y];
                    s = finish(s, pr.x);
                    y[pr.x] = s;
                    y[pr.y] = s;
                }
            } else if (t < np + nc) {
                const int k = (int)(t - np);
                const int b = a.c_off[k], e = a.c_off[k + 1];
                if constexpr (MODE == DSS_UNMASK) {
                    const T v = y[a.c_pos[b]];
                    for (int q = b + 1; q < e; ++q) y[a.c_pos[q]] = v;
                } else if constexpr (MODE == DSS_MASK) {
                    for (int q = b + 1; q < e; ++q) y[a.c_pos[q]] = 0;
                } else {
                    T s = 0;
                    for (int q = b; q < e; ++q) s += y[a.c_pos[q]];
                    s = finish(s, a.c_pos[b]);
                    for (int q = b; q < e; ++q) y[a.c_pos[q]] = s;
                }
            } else if (t < np + nc + nb) {
                if constexpr (MODE == DSS_SUM_ZB || MODE == DSS_MASK || MODE == DSS_PCG)
                    y[a.bnd[t - np - nc]] = 0;
            }
        }
    }
    if constexpr (MODE == DSS_PCG) {
        const double s = block_sum256(part);
        if (threadIdx.x == 0) a.partial[blockIdx.x] = s;
    }
}

template <typename T>
void launch_dss(int mode, const DssArgs<T>& a, hipStream_t s)
{
    const long total = (long)a.n_pairs + a.n_corners + a.n_bnd;
    const int nb = dss_blocks(a);
    if (total == 0 && mode != DSS_PCG) return;
    dim3 g(nb), b(256);
    switch (mode) {
    case DSS_SUM: hipLaunchKernelGGL((k_dss<T, DSS_SUM>), g, b, 0, s, a); break;
    case DSS_SUM_ZB: hipLaunchKernelGGL((k_dss<T, DSS_SUM_ZB>), g, b, 0, s, a); break;
    case DSS_UNMASK: hipLaunchKernelGGL((k_dss<T, DSS_UNMASK>), g, b, 0, s, a); break;
    case DSS_MASK: hipLaunchKernelGGL((k_dss<T, DSS_MASK>), g, b, 0, s, a); break;
    case DSS_PCG: hipLaunchKernelGGL((k_dss<T, DSS_PCG>), g, b, 0, s, a); break;
    case DSS_ZERO_BND: hipLaunchKernelGGL((k_dss<T, DSS_ZERO_BND>), g, b, 0, s, a); break;
    }
}
template void launch_dss<double>(int, const DssArgs<double>&, hipStream_t);
template void launch_dss<float>(int, const DssArgs<float>&, hipStream_t);

// ---------------------------------------------------- elementwise kernels --
// mode 0: out = ((x0 + c1 x1) + c2 x2) + ...
// mode 1: out = ((c0 x0 + c1 x1) + c2 x2) + ...
// mode 2: out = out + (((c0 x0 + c1 x1) + c2 x2) + ...)
template <typename T>
__global__ __launch_bounds__(256) void k_lincomb(LinComb<T> a)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < a.n; k += (size_t)gridDim.x * 256) {
        T s;
        if (a.mode == 0) s = a.x[0][k];
        else s = (T)a.c[0] * a.x[0][k];
        for (int t = 1; t < a.m; ++t) s = s + (T)a.c[t] * a.x[t][k];
        if (a.mode == 2) a.out[k] = a.out[k] + s;
        else a.out[k] = s;
    }
}

template <typename T>
void launch_lincomb(const LinComb<T>& a, hipStream_t s)
{
    if (a.n == 0) return;
    const int nb = (int)std::min<size_t>((a.n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_lincomb<T>), dim3(nb), dim3(256), 0, s, a);
}
template void launch_lincomb<double>(const LinComb<double>&, hipStream_t);
template void launch_lincomb<float>(const LinComb<float>&, hipStream_t);

// one workgroup: the records go to LDS with all 256 threads (up to 1536 of
// them), then one thread runs the sequential backward recurrence from LDS
constexpr int kCoefLds = 1536;
__global__ __launch_bounds__(256) void k_basis_coef(const double* __restrict__ rec, int k, double* __restrict__ coef)
{
#pragma clang fp contract(off)
    __shared__ double sr[5 * kCoefLds];
    const bool staged = k <= kCoefLds;
    if (staged) {
        for (int q = threadIdx.x; q < 5 * k; q += 256) sr[q] = rec[q];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double* rr = staged ? sr : rec;
    double t1 = 0.0, t2 = 0.0;  // t_{j+1}, t_{j+2}
    for (int j = k; j >= 1; --j) {
        const double* r = rr + 5 * (size_t)(j - 1);
        double v = r[4];
        if (j + 1 <= k) v -= t1 * rr[5 * (size_t)j + 1] * rr[5 * (size_t)j + 3];
        if (j + 2 <= k) v -= t2 * rr[5 * (size_t)(j + 1) + 1] * rr[5 * (size_t)(j + 1) + 2];
        t2 = t1;
        t1 = v;
        coef[j - 1] = v * r[1] * r[0];
    }
}

void launch_basis_coef(const double* rec, int k, double* coef, hipStream_t s)
{
    hipLaunchKernelGGL(k_basis_coef, dim3(1), dim3(256), 0, s, rec, k, coef);
}

// x (+)= sum_j coef[j] * basis_j: each thread accumulates its entries in FP64
// over the stored vectors in j order (eight loads in flight per round)
template <typename T>
__global__ __launch_bounds__(256) void k_basis_combine(T* __restrict__ x, int accumulate, const T* __restrict__ basis,
                                                       size_t stride, size_t n, const double* __restrict__ coef, int k)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double acc = accumulate ? (double)x[i] : 0.0;
        int j = 0;
        for (; j + 8 <= k; j += 8) {
            T v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = basis[(size_t)(j + u) * stride + i];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += coef[j + u] * (double)v[u];
        }
        for (; j < k; ++j) acc += coef[j] * (double)basis[(size_t)j * stride + i];
        x[i] = (T)acc;
    }
}

template <typename T>
void launch_basis_combine(T* x, bool accumulate, const T* basis, size_t stride, size_t n, const double* coef, int k,
                          hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_basis_combine<T>), dim3(nb), dim3(256), 0, s, x, accumulate ? 1 : 0, basis, stride, n,
                       coef, k);
}
template void launch_basis_combine<float>(float*, bool, const float*, size_t, size_t, const double*, int, hipStream_t);
template void launch_basis_combine<double>(double*, bool, const double*, size_t, size_t, const double*, int,
                                           hipStream_t);

template <typename D, typename S>
__global__ __launch_bounds__(256) void k_cast(D* out, const S* in, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) out[k] = (D)in[k];
}

template <typename D, typename S>
void launch_cast(D* out, const S* in, size_t n, hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_cast<D, S>), dim3(nb), dim3(256), 0, s, out, in, n);
}
template void launch_cast<float, double>(float*, const double*, size_t, hipStream_t);
template void launch_cast<double, float>(double*, const float*, size_t, hipStream_t);
template void launch_cast<double, double>(double*, const double*, size_t, hipStream_t);
template void launch_cast<float, float>(float*, const float*, size_t, hipStream_t);

// out = in * S, component-wise with the per-node scale S (length len) on
// each of the ncomp components; the product is formed in double.
template <typename D, typename Src>
__global__ __launch_bounds__(256) void k_scale_cast(D* out, const Src* in, const float* sc, size_t len, int ncomp)
{
    const size_t n = len * (size_t)ncomp;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
        out[k] = (D)((double)in[k] * (double)sc[k % len]);
}

template <typename D, typename Src>
void launch_scale_cast(D* out, const Src* in, const float* sc, size_t len, int ncomp, hipStream_t s)
{
    const size_t n = len * (size_t)ncomp;
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_scale_cast<D, Src>), dim3(nb), dim3(256), 0, s, out, in, sc, len, ncomp);
}
template void launch_scale_cast<float, double>(float*, const double*, const float*, size_t, int, hipStream_t);
template void launch_scale_cast<double, float>(double*, const float*, const float*, size_t, int, hipStream_t);

// Diagonal scaling of the inner Schrodinger-stage system: with the assembled
// mass diagonal m, stiffness diagonal k and potential-mass diagonal c,
// S = 1 / sqrt(m + |alpha| |k + c|) (any positive diagonal keeps S A S
// symmetric; 0 where the denominator is not positive: padding entries).
__global__ __launch_bounds__(256) void k_inner_scale(float* out, const double* m, const double* k, const double* c,
                                                     double alpha, size_t n)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        const double den = m[q] + fabs(alpha) * fabs(k[q] + c[q]);
        out[q] = den > 0.0 ? (float)(1.0 / sqrt(den)) : 0.0f;
    }
}

void launch_inner_scale(float* out, const double* m, const double* k, const double* c, double alpha, size_t n,
                        hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_inner_scale, dim3(nb), dim3(256), 0, s, out, m, k, c, alpha, n);
}

// streaming copy: 16-byte accesses (native vector type), four in flight per thread
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy16(dvec2* __restrict__ out, const dvec2* __restrict__ in, size_t n2)
{
    const size_t stride = (size_t)gridDim.x * 256;
    size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    for (; k + 3 * stride < n2; k += 4 * stride) {
        const dvec2 a = __builtin_nontemporal_load(in + k);
        const dvec2 b = __builtin_nontemporal_load(in + k + stride);
        const dvec2 c = __builtin_nontemporal_load(in + k + 2 * stride);
        const dvec2 d = __builtin_nontemporal_load(in + k + 3 * stride);
        __builtin_nontemporal_store(a, out + k);
        __builtin_nontemporal_store(b, out + k + stride);
        __builtin_nontemporal_store(c, out + k + 2 * stride);
        __builtin_nontemporal_store(d, out + k + 3 * stride);
    }
    for (; k < n2; k += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(in + k), out + k);
}

void launch_copy16(double* out, const double* in, size_t n, int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, s, (dvec2*)out, (const dvec2*)in, n / 2);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* out, T v, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) out[k] = v;
}

template <typename T>
void launch_fill(T* out, T v, size_t n, hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_fill<T>), dim3(nb), dim3(256), 0, s, out, v, n);
}
template void launch_fill<double>(double*, double, size_t, hipStream_t);
template void launch_fill<float>(float*, float, size_t, hipStream_t);

// x (+)= y elementwise, y = x * z (pointwise product) used to premultiply fields
template <typename T>
__global__ __launch_bounds__(256) void k_mul(T* out, const T* a, const T* b, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) out[k] = a[k] * b[k];
}

template <typename T>
void launch_mul(T* out, const T* a, const T* b, size_t n, hipStream_t s)
{
    if (n == 0) return;
    const int nb = (int)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((k_mul<T>), dim3(nb), dim3(256), 0, s, out, a, b, n);
}
template void launch_mul<double>(double*, const double*, const double*, size_t, hipStream_t);

// ------------------------------------------------------------ reductions --
// owner-weighted sum over ncomp components of a*b (unique-DOF dot,
// openmp_impl/dot.hpp:11-61); one block of 64 threads per chunk, lane = column.
template <typename T>
__global__ __launch_bounds__(64) void k_wdot(const T* a, const T* b, size_t len, int ncomp, int N, int K,
                                            const uint8_t* obits, double* partial)
{
    const int lane = threadIdx.x & 63, KN = K * N;  // one wave per workgroup
    double part = 0;
    if (lane < KN) {
        const int el = lane / N, j = lane % N;
        const int e = blockIdx.x * K + el;
        const unsigned bits = obits[e];
        const size_t base = (size_t)blockIdx.x * K * N * N + lane;
        for (int c = 0; c < ncomp; ++c)
            for (int i = 0; i < N; ++i) {
                const size_t k = (size_t)c * len + base + (size_t)i * KN;
                part += owner_weight(i, j, N, bits) * ((double)a[k] * (double)b[k]);
            }
    }
    part = wsum(part);
    if (lane == 0) partial[blockIdx.x] = part;
}

template <typename T>
void launch_wdot(const T* a, const T* b, size_t len, int ncomp, const Layout& L, const uint8_t* obits, double* partial,
                 hipStream_t s)
{
    hipLaunchKernelGGL((k_wdot<T>), dim3(L.nchunks), dim3(64), 0, s, a, b, len, ncomp, L.N, L.K, obits, partial);
}
template void launch_wdot<double>(const double*, const double*, size_t, int, const Layout&, const uint8_t*, double*,
                                  hipStream_t);
template void launch_wdot<float>(const float*, const float*, size_t, int, const Layout&, const uint8_t*, double*,
                                 hipStream_t);

// <a, a>, <b, b> and <a, b> in one pass, each accumulated as k_wdot does (the
// same bits as three k_wdot launches); partials at partial[q * nch + chunk]
template <typename T>
__global__ __launch_bounds__(64) void k_wdot3(const T* a, const T* b, size_t len, int ncomp, int N, int K,
                                             const uint8_t* obits, double* partial, int nch)
{
    const int lane = threadIdx.x & 63, KN = K * N;
    double paa = 0, pbb = 0, pab = 0;
    if (lane < KN) {
        const int el = lane / N, j = lane % N;
        const int e = blockIdx.x * K + el;
        const unsigned bits = obits[e];
        const size_t base = (size_t)blockIdx.x * K * N * N + lane;
        for (int c = 0; c < ncomp; ++c)
            for (int i = 0; i < N; ++i) {
                const size_t k = (size_t)c * len + base + (size_t)i * KN;
                const double x = (double)a[k], y = (double)b[k], w = owner_weight(i, j, N, bits);
                paa += w * (x * x);
                pbb += w * (y * y);
                pab += w * (x * y);
            }
    }
    paa = wsum(paa);
    pbb = wsum(pbb);
    pab = wsum(pab);
    if (lane == 0) {
        partial[blockIdx.x] = paa;
        partial[(size_t)nch + blockIdx.x] = pbb;
        partial[2 * (size_t)nch + blockIdx.x] = pab;
    }
}

template <typename T>
void launch_wdot3(const T* a, const T* b, size_t len, int ncomp, const Layout& L, const uint8_t* obits, double* partial,
                  hipStream_t s)
{
    hipLaunchKernelGGL((k_wdot3<T>), dim3(L.nchunks), dim3(64), 0, s, a, b, len, ncomp, L.N, L.K, obits, partial,
                       L.nchunks);
}
template void launch_wdot3<double>(const double*, const double*, size_t, int, const Layout&, const uint8_t*, double*,
                                   hipStream_t);

// plain Euclidean dot (generic solvers), grid-stride with fixed grid
template <typename T>
__global__ __launch_bounds__(256) void k_edot(const T* a, const T* b, size_t n, double* partial)
{
    double part = 0;
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256)
        part += (double)a[k] * (double)b[k];
    const double s = block_sum256(part);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int launch_edot(const double* a, const double* b, size_t n, double* partial, hipStream_t s)
{
    const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 1024));
    hipLaunchKernelGGL((k_edot<double>), dim3(nb), dim3(256), 0, s, a, b, n, partial);
    return nb;
}

// single-block fixed-order sum of n partials (blockDim = 256)
__device__ __forceinline__ double sum_partials(const double* p, int n)
{
    __shared__ double sh[4];
    return group_sum(p, n, sh, 4);
}

__global__ __launch_bounds__(256) void k_fin_sum(const double* partial, int n, double* dst)
{
    const double s = sum_partials(partial, n);
    if (threadIdx.x == 0) *dst = s;
}

void launch_fin_sum(const double* partial, int n, double* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(256), 0, s, partial, n, dst);
}

// per-instance sums of per-chunk partials into sc[k].red[slot]
__global__ __launch_bounds__(256) void k_fin_sum_inst(const double* partial, InstTable it, KrylovScalars* sc, int slot)
{
    const int k = blockIdx.x, p0 = it.inst_off[k];
    const double s = sum_partials(partial + p0, it.inst_off[k + 1] - p0);
    if (threadIdx.x == 0) sc[k].red[slot] = s;
}

void launch_fin_sum_inst(const double* partial, const InstTable& it, KrylovScalars* sc, int slot, hipStream_t s)
{
    hipLaunchKernelGGL(k_fin_sum_inst, dim3(it.n_inst), dim3(256), 0, s, partial, it, sc, slot);
}

// three arrays of per-chunk partials (stride nch) at once, each summed as
// k_fin_sum_inst sums it, into red[slot0 + q]
__global__ __launch_bounds__(256) void k_fin_sum3_inst(const double* partial, InstTable it, KrylovScalars* sc, int nch,
                                                       int slot0)
{
    __shared__ double sh[12];
    const int k = blockIdx.x, p0 = it.inst_off[k];
    double out[3];
    group_sum3<16>(partial + p0, nch, it.inst_off[k + 1] - p0, sh, 4, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) sc[k].red[slot0 + q] = out[q];
}

void launch_fin_sum3_inst(const double* partial, const InstTable& it, KrylovScalars* sc, int nch, int slot0,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_fin_sum3_inst, dim3(it.n_inst), dim3(256), 0, s, partial, it, sc, nch, slot0);
}

// --------------------------------------------------------- MINRES scalars --
// minres.hpp:24-44 given red[slot] = dot(r1, r1); one workgroup per instance
__global__ void k_minres_init(KrylovScalars* scs, int slot, int maxit, double tol, double atol, int bref_slot)
{
    if (threadIdx.x != 0) return;
    KrylovScalars* sc = scs + blockIdx.x;
    const double r = sc->red[slot];
    sc->b_norm = sqrt(bref_slot >= 0 ? sc->red[bref_slot] : r);
    const double beta1 = sqrt(r);
    sc->beta = beta1;
    sc->beta_old = 0;
    sc->tnorm = 0;
    sc->dbar = 0;
    sc->eps = 0;
    sc->eps_old = 0;
    sc->phibar = beta1;
    sc->cs = -1;
    sc->sn = 0;
    sc->tol = tol;
    sc->atol = atol >= 0 ? atol : tol;
    sc->alpha = 0;
    sc->ib = 1.0 / beta1;
    sc->cr1 = 0;
    sc->p3_pending = 0;
    sc->iter = 0;
    sc->maxit = maxit;
    sc->has_r1 = 0;
    sc->success = 0;
    sc->done = (maxit < 1) ? 1 : 0;
}

void launch_minres_init(KrylovScalars* sc, int n_inst, int slot, int maxit, double tol, hipStream_t s, double atol,
                        int bref_slot)
{
    hipLaunchKernelGGL(k_minres_init, dim3(n_inst), dim3(64), 0, s, sc, slot, maxit, tol, atol, bref_slot);
}

// Final pending w/x update of every instance whose last step left one
// (minres.hpp:82-86). The buffers rotate with the step index j = iter of the
// instance: r1 = R[(j+2)%3], wa = W[(j+1)%3], wb = W[(j+2)%3], wn = W[j%3].
template <typename T>
__global__ __launch_bounds__(256) void k_minres_flush(MinresFlushArgs<T> a, InstTable it, const KrylovScalars* scs)
{
    const int chunk = blockIdx.x;
    const KrylovScalars* sc = scs + chunk_instance(it, chunk).x;
    if (!sc->p3_pending) return;
    const int j = sc->iter;
    const T* r1 = a.R[(j + 2) % 3];
    const T* wa = a.W[(j + 1) % 3];
    const T* wb = a.W[(j + 2) % 3];
    T* wn = a.W[j % 3];
    const T ibp = (T)sc->p3_ibp, ig = (T)sc->p3_ig, eo = (T)sc->p3_eo, dl = (T)sc->p3_dl, phi = (T)sc->p3_phi;
    for (int c = 0; c < a.ncomp; ++c)
        for (int q = threadIdx.x; q < a.chunk_len; q += 256) {
            const size_t k = (size_t)c * a.len + (size_t)chunk * a.chunk_len + q;
            const T v = ibp * r1[k];
            const T w = ig * ((v - eo * wa[k]) - dl * wb[k]);
            wn[k] = w;
            a.x[k] += phi * w;
        }
}

__global__ void k_clear_pending(KrylovScalars* scs)
{
    if (threadIdx.x == 0) scs[blockIdx.x].p3_pending = 0;
}

template <typename T>
void launch_minres_flush(const MinresFlushArgs<T>& a, const InstTable& it, int nchunks, KrylovScalars* sc,
                         hipStream_t s)
{
    hipLaunchKernelGGL((k_minres_flush<T>), dim3(nchunks), dim3(256), 0, s, a, it, sc);
    hipLaunchKernelGGL(k_clear_pending, dim3(it.n_inst), dim3(64), 0, s, sc);
}
template void launch_minres_flush<double>(const MinresFlushArgs<double>&, const InstTable&, int, KrylovScalars*,
                                          hipStream_t);
template void launch_minres_flush<float>(const MinresFlushArgs<float>&, const InstTable&, int, KrylovScalars*,
                                         hipStream_t);

// ------------------------------------------------------------ PCG scalars --
// red[0] = dot(b,b), red[1] = dot(r,r) (pcg.hpp:62-63, :24)
__global__ void k_pcg_init(KrylovScalars* sc, int maxit, double tol)
{
    if (threadIdx.x != 0) return;
    sc->bnorm = sqrt(sc->red[0]);
    sc->rho = sc->red[1];
    sc->rho_prev = 0;
    sc->pcg_beta = 0;
    sc->pcg_alpha = 0;
    sc->tol = tol;
    sc->first = 1;
    sc->iter = 0;
    sc->maxit = maxit;
    sc->success = 0;
    sc->resid = sqrt(sc->rho);
    sc->done = (maxit < 1) ? 1 : 0;
}

void launch_pcg_init(KrylovScalars* sc, int maxit, double tol, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(64), 0, s, sc, maxit, tol);
}

// alpha = rho / dot(p, Ap) (pcg.hpp:78)
__global__ __launch_bounds__(256) void k_pcg_fin_alpha(KrylovScalars* sc, const double* partial, int n)
{
    if (sc->done) return;
    const double s = sum_partials(partial, n);
    if (threadIdx.x == 0) sc->pcg_alpha = sc->rho / s;
}

void launch_pcg_fin_alpha(KrylovScalars* sc, const double* partial, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_fin_alpha, dim3(1), dim3(256), 0, s, sc, partial, n);
}

// x += alpha p; r -= alpha Ap (pcg.hpp:79-80); partial of the next rho = dot(r,r)
template <typename T>
__global__ __launch_bounds__(64) void k_pcg_upd(T* x, T* r, const T* p, const T* ap, size_t len, int N, int K,
                                               const uint8_t* obits, const double* scp, double* partial)
{
    const KrylovScalars* sc = reinterpret_cast<const KrylovScalars*>(scp);
    const int lane = threadIdx.x & 63, KN = K * N;  // one wave per workgroup
    double part = 0;
    if (!sc->done && lane < KN) {
        const T al = (T)sc->pcg_alpha;
        const int el = lane / N, j = lane % N;
        const unsigned bits = obits[blockIdx.x * K + el];
        const size_t base = (size_t)blockIdx.x * K * N * N + lane;
        for (int i = 0; i < N; ++i) {
            const size_t k = base + (size_t)i * KN;
            x[k] += al * p[k];
            const T v = r[k] - al * ap[k];
            r[k] = v;
            part += owner_weight(i, j, N, bits) * ((double)v * (double)v);
        }
    }
    part = wsum(part);
    if (lane == 0) partial[blockIdx.x] = part;
}

void launch_pcg_upd(double* x, double* r, const double* p, const double* ap, size_t len, const Layout& L,
                    const uint8_t* obits, const double* sc, double* partial, hipStream_t s)
{
    hipLaunchKernelGGL((k_pcg_upd<double>), dim3(L.nchunks), dim3(64), 0, s, x, r, p, ap, len, L.N, L.K, obits, sc,
                       partial);
}

// stop test on the pre-update rho (pcg.hpp:82-88), then rho for the next step
__global__ __launch_bounds__(256) void k_pcg_fin_rho(KrylovScalars* sc, const double* partial, int n)
{
    if (sc->done) return;
    const double s = sum_partials(partial, n);
    if (threadIdx.x != 0) return;
    const int i = sc->iter;  // 0-based index of the step just finished
    sc->rho_prev = sc->rho;
    sc->resid = sqrt(sc->rho);
    if (sqrt(sc->rho) < sc->tol * sc->bnorm) {
        sc->success = 1;
        sc->done = 1;
        return;
    }
    sc->iter = i + 1;
    if (sc->iter >= sc->maxit) {
        sc->done = 1;
        return;
    }
    sc->rho = s;
    sc->pcg_beta = s / sc->rho_prev;
    sc->first = 0;
}

void launch_pcg_fin_rho(KrylovScalars* sc, const double* partial, int n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_fin_rho, dim3(1), dim3(256), 0, s, sc, partial, n);
}

// ------------------------------------------------- layouts and diagonals --
// Element-local arrays between the reference layout (e*N*N + i + j*N) and the
// chunked device layout (Layout::pos); one thread per entry of the E elements.
__global__ __launch_bounds__(256) void k_layout(const double* __restrict__ in, double* __restrict__ out, long E, int N,
                                                int K, int to_dev)
{
    const long total = E * N * N;
    for (long r = (long)blockIdx.x * 256 + threadIdx.x; r < total; r += (long)gridDim.x * 256) {
        const long e = r / (N * N);
        const int ij = (int)(r - e * N * N), i = ij % N, j = ij / N;
        const long d = (e / K) * K * N * N + (long)i * K * N + (e % K) * N + j;
        if (to_dev) out[d] = in[r];
        else out[r] = in[d];
    }
}

void launch_layout(const double* in, double* out, long E, int N, int K, int to_dev, hipStream_t s)
{
    const long total = E * N * N;
    const int nb = (int)std::max<long>(1, std::min<long>((total + 255) / 256, 65536));
    hipLaunchKernelGGL(k_layout, dim3(nb), dim3(256), 0, s, in, out, E, N, K, to_dev);
}

// out = 1 / |d| (0 where d is 0): a symmetric positive Jacobi preconditioner
__global__ __launch_bounds__(256) void k_abs_recip(double* out, const double* d, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (size_t)gridDim.x * 256) {
        const double a = fabs(d[k]);
        out[k] = a > 0.0 ? 1.0 / a : 0.0;
    }
}

void launch_abs_recip(double* out, const double* d, size_t n, hipStream_t s)
{
    const int nb = (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65536));
    hipLaunchKernelGGL(k_abs_recip, dim3(nb), dim3(256), 0, s, out, d, n);
}

// ---------------------------------------------------- tridiagonal operator --
__global__ __launch_bounds__(256) void k_tridiag(const double* x, double* y, long n, double lo, double di, double up)
{
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        double s = di * x[i];
        if (i > 0) s += lo * x[i - 1];
        if (i + 1 < n) s += up * x[i + 1];
        y[i] = s;
    }
}

void launch_tridiag(const double* x, double* y, long n, double lo, double di, double up, hipStream_t s)
{
    const int nb = (int)std::max<long>(1, std::min<long>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_tridiag, dim3(nb), dim3(256), 0, s, x, y, n, lo, di, up);
}

}  // namespace sem
