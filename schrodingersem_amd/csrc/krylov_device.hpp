// krylov_device.hpp -- device-side pieces shared by the Krylov kernels: the
// fixed-order sum of per-chunk partials and the MINRES recurrence. Every
// workgroup that evaluates them on the same inputs gets bit-identical scalars.
#pragma once

#include <cfloat>

#include "sem_internal.hpp"

namespace sem {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Group sums of a fixed 256-lane decomposition of p[0..n): virtual lane
// v = 64 g + l (g = 0..3) adds p[v], p[v + 256], ... in index order, and group
// g's sum is a butterfly over l. M groups (g0, g0 + gstep, ...) are handled by
// the calling wave; loads are issued 16 at a time.
template <int M>
__device__ __forceinline__ void group_sums(const double* __restrict__ p, int n, int g0, int gstep, double* sh)
{
    const int lane = threadIdx.x & 63;
    constexpr int KB = 16 / M;
    double s[M];
#pragma unroll
    for (int j = 0; j < M; ++j) s[j] = 0;
    for (int kb = 0; kb * 256 < n; kb += KB) {
        double v[M][KB];
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int k = (kb + u) * 256 + (g0 + j * gstep) * 64 + lane;
                v[j][u] = k < n ? p[k] : 0.0;
            }
#pragma unroll
        for (int j = 0; j < M; ++j)
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int k = (kb + u) * 256 + (g0 + j * gstep) * 64 + lane;
                if (k < n) s[j] += v[j][u];
            }
    }
#pragma unroll
    for (int j = 0; j < M; ++j) {
        const double t = wave_sum(s[j]);
        if (lane == 0) sh[g0 + j * gstep] = t;
    }
}

// Sum of p[0..n) by a workgroup of nw waves, ((G0 + G1) + G2) + G3 over the
// group sums above; every thread returns it, and the value does not depend
// on nw. sh: 4 doubles of LDS. Contains workgroup barriers.
__device__ __forceinline__ double group_sum(const double* __restrict__ p, int n, double* sh, int nw)
{
    const int wv = threadIdx.x >> 6;
    if (nw == 1) group_sums<4>(p, n, 0, 1, sh);
    else if (nw == 2) group_sums<2>(p, n, wv, 2, sh);
    else if (nw == 3) { if (wv == 0) group_sums<2>(p, n, 0, 3, sh); else group_sums<1>(p, n, wv, 1, sh); }
    else if (wv < 4) group_sums<1>(p, n, wv, 1, sh);
    __syncthreads();
    const double t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return t;
}

// Three sums at once: p[a*stride .. a*stride + n) for a = 0, 1, 2, each with
// the decomposition and order of group_sum (so each equals group_sum of its
// array); the loads of the three arrays are issued together, LB per array
// and thread at a time (the order of the additions does not depend on LB).
// sh: 12 doubles.
template <int M, int LB>
__device__ __forceinline__ void group_sums3(const double* __restrict__ p, int stride, int n, int g0, int gstep,
                                            double* sh)
{
    const int lane = threadIdx.x & 63;
    constexpr int KB = LB / M;
    static_assert(KB >= 1, "load budget below the virtual groups per wave");
    double s[3][M];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < M; ++j) s[a][j] = 0;
    for (int kb = 0; kb * 256 < n; kb += KB) {
        double v[3][M][KB];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    const int k = (kb + u) * 256 + (g0 + j * gstep) * 64 + lane;
                    v[a][j][u] = k < n ? p[(size_t)a * stride + k] : 0.0;
                }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    const int k = (kb + u) * 256 + (g0 + j * gstep) * 64 + lane;
                    if (k < n) s[a][j] += v[a][j][u];
                }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const double t = wave_sum(s[a][j]);
            if (lane == 0) sh[4 * a + g0 + j * gstep] = t;
        }
}

template <int LB = 16>
__device__ __forceinline__ void group_sum3(const double* __restrict__ p, int stride, int n, double* sh, int nw,
                                           double (&out)[3])
{
    const int wv = threadIdx.x >> 6;
    if (nw == 1) group_sums3<4, LB>(p, stride, n, 0, 1, sh);
    else if (nw == 2) group_sums3<2, LB>(p, stride, n, wv, 2, sh);
    else if (nw == 3) { if (wv == 0) group_sums3<2, LB>(p, stride, n, 0, 3, sh); else group_sums3<1, LB>(p, stride, n, wv, 1, sh); }
    else if (wv < 4) group_sums3<1, LB>(p, stride, n, wv, 1, sh);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = ((sh[4 * a] + sh[4 * a + 1]) + sh[4 * a + 2]) + sh[4 * a + 3];
    __syncthreads();
}

// group_sum3 in two phases for a workgroup of NW waves (NW != 3): issue()
// loads the first LB values per array and thread, finish() loads the rest,
// adds in group_sum's order (so the results equal group_sum3's bit for bit)
// and reduces. Loads issued between the two phases stay in flight while
// finish() waits for its own, which were issued first.
template <int NW, int LB>
struct GroupSum3 {
    static_assert(NW != 3, "three-wave workgroups use group_sum3");
    static constexpr int G = NW < 4 ? NW : 4;  // waves that take part
    static constexpr int M = 4 / G;            // virtual groups per wave
    static constexpr int KB = LB / M;
    static_assert(KB >= 1, "load budget below the virtual groups per wave");
    double v[3][M][KB];

    __device__ __forceinline__ void load(const double* __restrict__ p, int stride, int n, int kb)
    {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int j = 0; j < M; ++j)
#pragma unroll
                for (int u = 0; u < KB; ++u) {
                    const int k = (kb + u) * 256 + (wv + j * G) * 64 + lane;
                    v[a][j][u] = k < n ? p[(size_t)a * stride + k] : 0.0;
                }
    }

    __device__ __forceinline__ void issue(const double* __restrict__ p, int stride, int n)
    {
        if ((threadIdx.x >> 6) < G) load(p, stride, n, 0);
    }

    __device__ __forceinline__ void finish(const double* __restrict__ p, int stride, int n, double* sh,
                                           double (&out)[3])
    {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (wv < G) {
            double s[3][M];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int j = 0; j < M; ++j) s[a][j] = 0;
            for (int kb = 0; kb * 256 < n; kb += KB) {
                if (kb > 0) load(p, stride, n, kb);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int j = 0; j < M; ++j)
#pragma unroll
                        for (int u = 0; u < KB; ++u) {
                            const int k = (kb + u) * 256 + (wv + j * G) * 64 + lane;
                            if (k < n) s[a][j] += v[a][j][u];
                        }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int j = 0; j < M; ++j) {
                    const double t = wave_sum(s[a][j]);
                    if (lane == 0) sh[4 * a + wv + j * G] = t;
                }
        }
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 3; ++a) out[a] = ((sh[4 * a] + sh[4 * a + 1]) + sh[4 * a + 2]) + sh[4 * a + 3];
        __syncthreads();
    }
};

// beta = sqrt(s), s = dot(y, y) of the step just finished, and the Givens
// update of minres.hpp:62-99 (stop test minres.hpp:92 with atol = tol there).
// Contraction is off so that every workgroup, whatever code surrounds the
// call, rounds the recurrence the same way.
__device__ __forceinline__ void minres_beta_update(MinresState& sc, double s)
{
#pragma clang fp contract(off)
    const double alpha = sc.alpha;
    const double beta_prev = sc.beta;
    const double ib_prev = sc.ib;
    sc.beta_old = beta_prev;
    const double beta = sqrt(s);
    sc.beta = beta;
    sc.tnorm += alpha * alpha + beta * beta + beta_prev * beta_prev;
    const double eps_old = sc.eps;
    sc.eps_old = eps_old;
    const double cs = sc.cs, sn = sc.sn, dbar = sc.dbar;
    const double delta = cs * dbar + sn * alpha;
    const double gbar = sn * dbar - cs * alpha;
    sc.eps = sn * beta;
    sc.dbar = -cs * beta;
    const double gamma = fmax(sqrt(gbar * gbar + beta * beta), DBL_EPSILON);
    sc.cs = gbar / gamma;
    sc.sn = beta / gamma;
    const double phi = sc.cs * sc.phibar;
    sc.phibar *= sc.sn;
    // pending w/x update of this step (minres.hpp:82-86)
    sc.p3_ibp = ib_prev;
    sc.p3_ig = 1.0 / gamma;
    sc.p3_eo = eps_old;
    sc.p3_dl = delta;
    sc.p3_phi = phi;
    sc.p3_pending = 1;
    // coefficients of the next Lanczos step (minres.hpp:48-52)
    sc.ib = 1.0 / beta;
    sc.cr1 = beta / beta_prev;
    sc.has_r1 = 1;
    sc.iter += 1;
    if (sc.phibar < sc.tol * sc.b_norm + sc.atol) {
        sc.success = 1;
        sc.done = 1;
    } else if (sc.iter >= sc.maxit) {
        sc.done = 1;
    }
}

// Stored-basis MINRES: the update coefficients of iteration j = sc.iter
// (minres.hpp:82-86: w_j = (v_j - eps_j w_{j-2} - delta_j w_{j-1}) / gamma_j,
// x += phi_j w_j, v_j = y_{j-1} / beta_j), kept for forming x from the stored
// y vectors at the end
__device__ __forceinline__ void minres_record(double* rec, const MinresState& sc)
{
    double* r = rec + 5 * (size_t)(sc.iter - 1);
    r[0] = sc.p3_ibp;
    r[1] = sc.p3_ig;
    r[2] = sc.p3_eo;
    r[3] = sc.p3_dl;
    r[4] = sc.p3_phi;
}

// Recurrence at the start of fused Lanczos step i >= 2 (deferred scaling: step
// i finishes y_{i-1} = A y_{i-2} / beta_{i-1} - (beta_{i-1}/beta_{i-2}) y_{i-3}
// - (alpha_{i-1}/beta_{i-1}) y_{i-2}, then applies A to it). P are the sums of
// step i-1: P0 = ||y_{i-2}||^2 = beta_{i-1}^2, P1 = y_{i-2}.A y_{i-2},
// P2 = y_{i-2}.y_{i-3}. For i >= 3 this is the Givens update of iteration i-2
// (minres.hpp:62-99) followed by alpha_{i-1} = v.(A v - (beta/beta_old) r1)
// (minres.hpp:50-54) written with those sums.
__device__ __forceinline__ void minres_fused_prologue(MinresState& sc, int step, const double (&P)[3])
{
#pragma clang fp contract(off)
    if (step >= 3) {
        if (sc.done) {
            sc.p3_pending = 0;  // applied by the step that stopped
            return;
        }
        minres_beta_update(sc, P[0]);
        if (sc.done) return;
        sc.alpha = P[1] * (sc.ib * sc.ib) - P[2] / sc.beta_old;
    } else if (step == 2) {
        sc.alpha = P[1] * (sc.ib * sc.ib);
    }
}

}  // namespace sem
