// kernels_op.hip -- element-operator kernels for gfx950.
//
// One 64-lane wavefront (one workgroup) per chunk of K = 64/N elements; lane
// l = el*N + j holds column j of element el in registers (the device layout
// makes plane i of a chunk one contiguous K*N-value run, so every load and
// store is coalesced). The sum-factorised contractions of glaplace
// (glaplace.hpp:16-28) run in registers with D read from the kernel-argument
// segment (uniform scalar loads); the two index transposes per direction go
// through LDS. Geometric factors arrive premultiplied by w_i w_j:
//   gA = w_i w_j A, gB = w_i w_j B, gC = w_i w_j C, gJ = w_i w_j J, so
//   K u (i,l) = sum_k D(k,i) G1(k,l) + sum_k D(k,l) G2(i,k),
//   G1 = gA u_xi - gB u_eta, G2 = gC u_eta - gB u_xi,
// which is F W + W G of glaplace.hpp:22-25 with the weights moved inside.
#include <hip/hip_ext.h>

#include "krylov_device.hpp"
#include "sem_internal.hpp"

namespace sem {

template <int N>
struct Shape {
    static constexpr int K = 64 / N;
    static constexpr int KN = K * N;
    static constexpr int P = (N % 2) ? N : N + 1;  // LDS column pitch (odd)
    static constexpr int LDS = K * N * P;
};

template <int N, typename T>
struct Dm {
    T v[N * N];  // D(i,k) at i + k*N
};

__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// LDS exchange among the lanes of one wavefront. Every wave works in its own
// LDS region, so no workgroup barrier is needed: LDS instructions of a wave
// execute in order; the fences keep the compiler from moving the reads above
// the writes, and (unlike __syncthreads) they do not order global memory
// traffic, so independent global loads stay in flight across the exchange.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// thread holds column l of X: col[i] = X(i,l); returns row[k] = X(l,k)
template <int N, typename T>
__device__ __forceinline__ void xpose_c2r(T* buf, int el, int l, bool act, const T (&col)[N], T (&row)[N])
{
    constexpr int P = Shape<N>::P;
    T* b = buf + el * N * P;
    if (act) {
#pragma unroll
        for (int i = 0; i < N; ++i) b[l * P + i] = col[i];
    }
    wave_lds_sync();
#pragma unroll
    for (int k = 0; k < N; ++k) row[k] = act ? b[k * P + l] : T(0);
}

// thread holds row l of X: row[j] = X(l,j); returns col[i] = X(i,l)
template <int N, typename T>
__device__ __forceinline__ void xpose_r2c(T* buf, int el, int l, bool act, const T (&row)[N], T (&col)[N])
{
    constexpr int P = Shape<N>::P;
    T* b = buf + el * N * P;
    if (act) {
#pragma unroll
        for (int j = 0; j < N; ++j) b[j * P + l] = row[j];
    }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) col[i] = act ? b[l * P + i] : T(0);
}

// out(:,l) = (K q)(:,l), the element stiffness action of glaplace; geo(i, a,
// b, c) gives gA, gB, gC at node (i, l) where G1 and G2 are formed (FP32
// kernels compute them there from the element map, so they are not held in
// registers across the first two contractions).
template <int N, typename T, class GeoF>
__device__ __forceinline__ void stiffness(const Dm<N, T>& D, T* lds0, T* lds1, int el, int l, bool act, const T (&q)[N],
                                          const GeoF& geo, T (&out)[N])
{
    T qxi[N], tmp[N], qer[N], qe[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {  // u_xi = D u
        T s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s = fmad(D.v[i + k * N], q[k], s);
        qxi[i] = s;
    }
    xpose_c2r<N>(lds0, el, l, act, q, tmp);
#pragma unroll
    for (int j = 0; j < N; ++j) {  // u_eta = u D^T (row l)
        T s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s = fmad(D.v[j + k * N], tmp[k], s);
        qer[j] = s;
    }
    xpose_r2c<N>(lds1, el, l, act, qer, qe);
    T g1[N], g2[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        T ga, gb, gc;
        geo(i, ga, gb, gc);
        g1[i] = ga * qxi[i] - gb * qe[i];
        g2[i] = gc * qe[i] - gb * qxi[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {  // D^T G1
        T s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s = fmad(D.v[k + i * N], g1[k], s);
        out[i] = s;
    }
    xpose_c2r<N>(lds0, el, l, act, g2, tmp);
#pragma unroll
    for (int j = 0; j < N; ++j) {  // G2 D (row l)
        T s = 0;
#pragma unroll
        for (int k = 0; k < N; ++k) s = fmad(D.v[k + j * N], tmp[k], s);
        qer[j] = s;
    }
    xpose_r2c<N>(lds1, el, l, act, qer, qe);
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] += qe[i];
}

// the same with gA, gB, gC in registers
template <int N, typename T>
__device__ __forceinline__ void stiffness(const Dm<N, T>& D, T* lds0, T* lds1, int el, int l, bool act, const T (&q)[N],
                                          const T (&ga)[N], const T (&gb)[N], const T (&gc)[N], T (&out)[N])
{
    stiffness<N>(D, lds0, lds1, el, l, act, q,
                 [&](int i, T& a, T& b, T& c) {
                     a = ga[i];
                     b = gb[i];
                     c = gc[i];
                 },
                 out);
}

// ----------------------------------------------- MFMA stacked contractions --
// The D-contractions of the stiffness action as dense 16 x 16 x 16 bf16 MFMA
// tiles with FP32 accumulation (v_mfma_f32_16x16x16bf16_1k). Lane l holds
// A[l%16][4(l/16)+r] and B[4(l/16)+r][l%16] of the operands and receives
// C[4(l/16)+r][l%16], r = 0..3. A wave's chunk is stacked two ways:
//   xi direction:  C = M X with the 64 lanes (element columns) as the columns of X;
//   eta direction: C = X M with the K*N (plane, element) pairs of the chunk as rows
//                  and the element's column index as the contracted index.
// Both go through one LDS scratch of 64 columns x 17 (odd pitch) per wave.
// SPLIT = 1 rounds operands to bf16; SPLIT = 3 represents each operand as
// hi + lo bf16 and sums hi*hi + hi*lo + lo*hi (about 2^-16 relative error).
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMfPitch = 17;

__device__ __forceinline__ short f2bf(float f)  // round to nearest even (finite values)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

__device__ __forceinline__ void split_bf(const float (&v)[4], bf16x4& hi, bf16x4& lo)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const short h = f2bf(v[r]);
        hi[r] = h;
        lo[r] = f2bf(v[r] - bf2f(h));
    }
}

template <int SPLIT>
__device__ __forceinline__ f32x4 mm16(bf16x4 ahi, bf16x4 alo, bf16x4 bhi, bf16x4 blo)
{
    f32x4 c = {0.f, 0.f, 0.f, 0.f};
    if constexpr (SPLIT == 3) {  // small terms first
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, blo, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bhi, c, 0, 0, 0);
    }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ahi, bhi, c, 0, 0, 0);
}

// D in the two operand layouts: P[r] = D(l%16, 4(l/16)+r), Q[r] = D(4(l/16)+r, l%16)
// (zero outside N x N), hi and lo parts
struct MfD {
    bf16x4 p_hi, p_lo, q_hi, q_lo;
};

template <int N>
__device__ __forceinline__ MfD mf_dmat(const Dm<N, float>& D, int lane)
{
    const int m = lane % 16, kb = 4 * (lane / 16);
    float p[4], q[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = kb + r;
        const bool in = m < N && k < N;
        p[r] = in ? D.v[m + k * N] : 0.f;
        q[r] = in ? D.v[k + m * N] : 0.f;
    }
    MfD d;
    split_bf(p, d.p_hi, d.p_lo);
    split_bf(q, d.q_hi, d.q_lo);
    return d;
}

// y = M x, column per lane (xi direction); A operand (ahi, alo) holds M
template <int N, int SPLIT>
__device__ __forceinline__ void mf_xi(float* S, int lane, const float (&x)[N], bf16x4 ahi, bf16x4 alo, float (&y)[N])
{
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) S[lane * kMfPitch + i] = x[i];
    wave_lds_sync();
    const int c16 = lane % 16, kb = 4 * (lane / 16);
    f32x4 c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int col = 16 * t + c16;
        float b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) b[r] = (kb + r < N) ? S[col * kMfPitch + kb + r] : 0.f;
        bf16x4 bhi, blo;
        split_bf(b, bhi, blo);
        c[t] = mm16<SPLIT>(ahi, alo, bhi, blo);
    }
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            if (kb + r < N) S[(16 * t + c16) * kMfPitch + kb + r] = c[t][r];
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = S[lane * kMfPitch + i];
}

// y(i, el, j) = sum_k x(i, el, k) M(k, j), column per lane (eta direction);
// B operand (bhi, blo) holds M
template <int N, int SPLIT>
__device__ __forceinline__ void mf_eta(float* S, int lane, const float (&x)[N], bf16x4 bhi, bf16x4 blo, float (&y)[N])
{
    constexpr int K = 64 / N, R = K * N;
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) S[lane * kMfPitch + i] = x[i];
    wave_lds_sync();
    const int c16 = lane % 16, kb = 4 * (lane / 16);
    f32x4 c[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int rho = 16 * t + c16, i = rho / K, el = rho % K;
        float a[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = (rho < R && kb + r < N) ? S[(el * N + kb + r) * kMfPitch + i] : 0.f;
        bf16x4 ahi, alo;
        split_bf(a, ahi, alo);
        c[t] = mm16<SPLIT>(ahi, alo, bhi, blo);
    }
    wave_lds_sync();
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rho = 16 * t + kb + r;
            if (rho < R && c16 < N) S[((rho % K) * N + c16) * kMfPitch + rho / K] = c[t][r];
        }
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < N; ++i) y[i] = S[lane * kMfPitch + i];
}

// the stiffness action of `stiffness` (same inputs and outputs per lane) with
// the contractions on MFMA; q, ga, gb, gc are zero on inactive lanes
template <int N, int SPLIT>
__device__ __forceinline__ void stiffness_mf(const MfD& d, float* S, int lane, const float (&q)[N], const float (&ga)[N],
                                             const float (&gb)[N], const float (&gc)[N], float (&out)[N])
{
    float qxi[N], qe[N];
    mf_xi<N, SPLIT>(S, lane, q, d.p_hi, d.p_lo, qxi);   // D u
    mf_eta<N, SPLIT>(S, lane, q, d.p_hi, d.p_lo, qe);   // u D^T
    float g1[N], g2[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        g1[i] = ga[i] * qxi[i] - gb[i] * qe[i];
        g2[i] = gc[i] * qe[i] - gb[i] * qxi[i];
    }
    float o2[N];
    mf_xi<N, SPLIT>(S, lane, g1, d.q_hi, d.q_lo, out);  // D^T G1
    mf_eta<N, SPLIT>(S, lane, g2, d.q_hi, d.q_lo, o2);  // G2 D
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] += o2[i];
}

template <int N>
__device__ __forceinline__ bool is_interior(int i, int l)
{
    return i > 0 && i < N - 1 && l > 0 && l < N - 1;
}

// Column loaders / storers of the device layout. NT: nontemporal accesses,
// for launches whose arrays do not fit the Infinity Cache (streamed once).
template <int N, typename T, bool NT = false>
struct Col {
    size_t base;
    bool act;
    __device__ __forceinline__ void ld(const T* __restrict__ p, T (&r)[N]) const
    {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const T* q = p + base + (size_t)i * Shape<N>::KN;
            if constexpr (NT) r[i] = act ? __builtin_nontemporal_load(q) : T(0);
            else r[i] = act ? *q : T(0);
        }
    }
    __device__ __forceinline__ void st(T* __restrict__ p, const T (&r)[N]) const
    {
        if (act) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                T* q = p + base + (size_t)i * Shape<N>::KN;
                if constexpr (NT) __builtin_nontemporal_store(r[i], q);
                else *q = r[i];
            }
        }
    }
};

// FP32 geometric factors of column l of device element e, formed from the
// element's bilinear map (QuadMap): the arithmetic of quad_metrics
// (Quad.hpp:39-54, the reference's C index included) times w_i w_j, with
// explicit FMAs so that every kernel forms the same bits. A zero map (padding
// elements, ghosts) gives zero factors.
template <int N>
struct QuadColumn {
    float4 u, v;  // (ax, bx, ay, by), (cx, dx, cy, dy)
    float xxi, yxi, wl;  // X_xi, Y_xi at node l; w_l
    __device__ __forceinline__ QuadColumn(const QuadMap& q, int e, int l, bool act)
    {
        u = act ? q.map[2 * e] : make_float4(0.f, 0.f, 0.f, 0.f);
        v = act ? q.map[2 * e + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        float zl = 0.f;
        wl = 0.f;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (l == k) {
                zl = q.z[k];
                wl = q.w[k];
            }
        xxi = __builtin_fmaf(u.y, zl, u.x);
        yxi = __builtin_fmaf(u.w, zl, u.z);
    }
    // J(i, l), with X_eta(i) and Y_eta(i)
    __device__ __forceinline__ float jac(const QuadMap& q, int i, float& xe, float& ye) const
    {
        xe = __builtin_fmaf(v.y, q.z[i], v.x);
        ye = __builtin_fmaf(v.w, q.z[i], v.z);
        return __builtin_fmaf(xxi, ye, -(xe * yxi));
    }
    // gA, gB, gC at node (i, l)
    __device__ __forceinline__ void at(const QuadMap& q, int i, float& ga, float& gb, float& gc) const
    {
        float xe, ye;
        const float j = jac(q, i, xe, ye);
        const float s = (j != 0.f) ? (q.w[i] * wl) * __builtin_amdgcn_rcpf(j) : 0.f;
        const float yxi_i = __builtin_fmaf(u.w, q.z[i], u.z);
        ga = __builtin_fmaf(ye, ye, xe * xe) * s;
        gb = __builtin_fmaf(ye, yxi, xe * xxi) * s;
        gc = __builtin_fmaf(yxi_i, yxi_i, xxi * xxi) * s;
    }
    __device__ __forceinline__ void stiff(const QuadMap& q, float (&ga)[N], float (&gb)[N], float (&gc)[N]) const
    {
#pragma unroll
        for (int i = 0; i < N; ++i) at(q, i, ga[i], gb[i], gc[i]);
    }
    __device__ __forceinline__ void mass(const QuadMap& q, float (&gj)[N]) const
    {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            float xe, ye;
            gj[i] = (q.w[i] * wl) * jac(q, i, xe, ye);
        }
    }
};

// gA, gB, gC of column l of element e: formed from the map (FP32 where
// map_geometry(N)), loaded otherwise
template <int N, typename T, class C>
__device__ __forceinline__ void geo_stiff(const T* geo, size_t len, const QuadMap& qm, const C& c, int e, int l,
                                          bool act, T (&ga)[N], T (&gb)[N], T (&gc)[N])
{
    if constexpr (sizeof(T) == 4 && map_geometry(N)) {
        QuadColumn<N>(qm, e, l, act).stiff(qm, ga, gb, gc);
    } else {
        c.ld(geo, ga);
        c.ld(geo + len, gb);
        c.ld(geo + 2 * len, gc);
    }
}

// gJ of column l of element e
template <int N, typename T, class C>
__device__ __forceinline__ void geo_mass(const T* geo, size_t len, const QuadMap& qm, const C& c, int e, int l,
                                         bool act, T (&gj)[N])
{
    if constexpr (sizeof(T) == 4 && map_geometry(N)) QuadColumn<N>(qm, e, l, act).mass(qm, gj);
    else c.ld(geo + 3 * len, gj);
}

// Element operator of every sem_op_kind; ring entries are written raw (the DSS
// kernel assembles them), interior entries are final and scaled by `post`.
template <int N, typename T, int KIND, bool NT, int MF = 0>
__global__ __launch_bounds__(64) void k_element_op(OpArgs<T> a, Dm<N, T> D)
{
    using S = Shape<N>;
    static_assert(MF == 0 || sizeof(T) == 4, "MFMA contractions serve the FP32 apply");
    __shared__ T lds0[S::LDS], lds1[S::LDS];
    __shared__ float mfs[MF ? 64 * kMfPitch : 1];
    // the mask tells the compiler lane < 64: for N = 8 (K N = 64) the active-lane
    // predicate then folds away, and with it an exec-mask branch around every
    // load, store and LDS exchange; the SGPRs those masks take would otherwise
    // push D's 64 uniform entries out of scalar registers (measured: the FP64
    // Res apply at 256 x 256, degree 7, runs in 52.6 us instead of 95.6 us)
    const int lane = threadIdx.x & 63;
    const bool act = lane < S::KN;
    const int el = act ? lane / N : 0;
    const int l = act ? lane % N : 0;
    const Col<N, T, NT> c{(size_t)blockIdx.x * S::K * N * N + lane, act};
    const size_t len = a.len;
    const int e = blockIdx.x * S::K + el;
    T ga[N], gb[N], gc[N];
    geo_stiff<N>(a.geo, len, a.qm, c, e, l, act, ga, gb, gc);
    auto finish = [&](T* dst, T (&y)[N]) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (is_interior<N>(i, l)) y[i] = a.post * y[i];
        c.st(dst, y);
    };
    if constexpr (KIND == K_RES || KIND == K_F) {
        T q0[N], q1[N], cf[N], ky[N], y[N];
        c.ld(a.in, q0);
        c.ld(a.in + len, q1);
        c.ld(a.cf, cf);
        auto stiff = [&](const T (&q)[N], T (&o)[N]) {
            if constexpr (MF != 0) stiffness_mf<N, MF>(mf_dmat<N>(D, lane), mfs, lane, q, ga, gb, gc, o);
            else stiffness<N>(D, lds0, lds1, el, l, act, q, ga, gb, gc, o);
        };
        // component 0 output from K q1
        stiff(q1, ky);
        if constexpr (KIND == K_RES) {
            T gj[N];
            geo_mass<N>(a.geo, len, a.qm, c, e, l, act, gj);
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = gj[i] * q0[i] + a.alpha * (ky[i] + cf[i] * q1[i]);
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = -(ky[i] + cf[i] * q1[i]);
        }
        finish(a.out, y);
        stiff(q0, ky);
        if constexpr (KIND == K_RES) {
            T gj[N];
            geo_mass<N>(a.geo, len, a.qm, c, e, l, act, gj);
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = a.alpha * (ky[i] + cf[i] * q0[i]) - gj[i] * q1[i];
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = -(ky[i] + cf[i] * q0[i]);
        }
        finish(a.out + len, y);
    } else {
        T q[N], y[N];
        c.ld(a.in, q);
        if constexpr (KIND == K_MASS) {
            T gj[N];
            geo_mass<N>(a.geo, len, a.qm, c, e, l, act, gj);
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = gj[i] * q[i];
        } else {
            T ky[N];
            stiffness<N>(D, lds0, lds1, el, l, act, q, ga, gb, gc, ky);
            if constexpr (KIND == K_LAPLACE) {
#pragma unroll
                for (int i = 0; i < N; ++i) y[i] = ky[i];
            } else if constexpr (KIND == K_HELMHOLTZ) {
                T cf[N];
                c.ld(a.cf, cf);
#pragma unroll
                for (int i = 0; i < N; ++i) y[i] = cf[i] * q[i] + ky[i];
            } else {  // K_L
                T cf[N];
                c.ld(a.cf, cf);
#pragma unroll
                for (int i = 0; i < N; ++i) y[i] = ky[i] + cf[i] * q[i];
            }
        }
        finish(a.out, y);
    }
}


// Workgroups are handed to the 8 XCDs round robin by workgroup index, and
// each XCD has its own L2. This maps workgroup b to a work item so that every
// XCD gets one contiguous range of items: neighbouring elements (whose ring
// values the update kernel gathers) then share an L2. A bijection on [0, n).
__device__ __forceinline__ int xcd_item(int b, int n)
{
    constexpr int nx = 8;
    const int q = n / nx, r = n % nx;
    const int x = b % nx, k = b / nx;
    return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Persistent form of the same idea: XCD x owns the contiguous item range
// that xcd_item gives it, and its workgroups stride through that range.
// Grids smaller than 8 fall back to a plain grid stride. body(item) returns
// -1, or an item index to skip ahead to (the end of a stopped instance's
// items); the walk continues at the first of its own items at or after it.
template <typename F>
__device__ __forceinline__ void for_items_persistent(int n, F&& body)
{
    constexpr int nx = 8;
    const int G = gridDim.x, b = blockIdx.x;
    int start, end, step;
    if (G < nx) {
        start = b;
        end = n;
        step = G;
    } else {
        const int q = n / nx, r = n % nx;
        const int x = b % nx, k = b / nx;
        const int lo = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
        start = lo + k;
        end = lo + q + (x < r ? 1 : 0);
        step = (G - x + nx - 1) / nx;  // workgroups on XCD x
    }
    for (int it = start; it < end;) {
        const int skip = body(it);
        if (skip > it) it = start + (skip - start + step - 1) / step * step;
        else it += step;
    }
}

// owner test of entry (i, l) of an element with ownership bits (bit s-1:
// side s owned; bit 4+k: corner k owned)
template <int N>
__device__ __forceinline__ bool owns(int i, int l, unsigned bits)
{
    const bool ri = (i == 0 || i == N - 1), rj = (l == 0 || l == N - 1);
    if (!ri && !rj) return true;
    int bit;
    if (ri && rj) bit = 4 + ((i == 0) ? ((l == 0) ? 0 : 3) : ((l == 0) ? 1 : 2));
    else if (l == 0) bit = 0;
    else if (i == N - 1) bit = 1;
    else if (l == N - 1) bit = 2;
    else bit = 3;
    return (bits >> bit) & 1u;
}

// MINRES Lanczos step i, operator part (minres.hpp:48-55): v = ib*r2,
// y = A v - cr1*r1. One wavefront per output component (workgroup = NC
// waves on the same chunk, so the second wave's geometry loads hit in cache).
// The subtraction is applied on element-interior entries and on the owner
// copy of every interface node; the update kernel sums the copies, which gives
// the assembled (A v)_owner - cr1*r1 on every copy. alpha = dot(v, y) is
// therefore the sum of v*y over interior entries plus v*(copy value) over
// every interface copy, computed here as one partial per workgroup.
// The FP32 instantiation asks for 5 waves per SIMD (a 128 x 128, degree-8 mesh
// then fits in one wave round on 256 CUs); the FP64 one keeps the default,
// where forcing occupancy costs more in register spills than it gains.
template <int N, typename T, int KIND, int CPB, bool PERSIST>
__global__ __launch_bounds__(64 * CPB * ((KIND == K_RES) ? 2 : 1))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1, 8))) void k_minres_op(MinresOpArgs<T> a, Dm<N, T> D)
{
    using S = Shape<N>;
    constexpr int NC = (KIND == K_RES) ? 2 : 1;
    constexpr int NW = NC * CPB;
    __shared__ T lds[NW][2][S::LDS];
    __shared__ double wpart[NW];
    __shared__ double red[4];
    const int wv = threadIdx.x >> 6;
    const int w = wv % NC;    // output component of this wave
    const int sub = wv / NC;  // chunk of the work item
    const int lane = threadIdx.x & 63;
    const bool act = lane < S::KN;
    const int el = act ? lane / N : 0;
    const int l = act ? lane % N : 0;
    const size_t len = a.len;
    const int n_items = a.nchunks / CPB;
    int cur_inst = -1;
    MinresState sc;
    bool pend = false;
    auto body = [&](int item) {
        const int c0 = item * CPB;
        const int4 ci = chunk_instance(a.it, c0);
        const int inst = ci.x;
        if (inst != cur_inst) {
            cur_inst = inst;
            sc = *static_cast<const MinresState*>(a.st_in + inst);
            if (!a.first) {
                // the previous step's beta recurrence, evaluated by every workgroup
                if (!sc.done) {
                    const int p0 = ci.y;
                    minres_beta_update(sc, group_sum(a.prev_partial + p0, ci.z - p0, red, NW));
                } else {
                    sc.p3_pending = 0;  // applied by the kernel of the stopping step
                }
            }
            // this kernel applies the pending w/x update; the state it stores drops it
            pend = sc.p3_pending != 0;
            sc.p3_pending = 0;
            if (c0 == ci.y && threadIdx.x == 0) *static_cast<MinresState*>(a.st_out + inst) = sc;
        }
        // a stopped instance has no work left: skip the rest of its items
        // (its partials are never read again)
        if (sc.done && !pend) return ci.z / CPB;
        const int chunk = c0 + sub;
        const Col<N, T> c{(size_t)chunk * S::K * N * N + lane, act};
        double part = 0;
        T r[N];  // r1 of this wave's component
        if (!sc.done) {
            const T ib = (T)sc.ib, cr1 = (T)sc.cr1;
            const int has_r1 = sc.has_r1;
            const unsigned bits = act ? a.obits[chunk * S::K + el] : 0u;
            T ga[N], gb[N], gc[N], qs[N], y[N];
            // stiffness input: the other component for Res, the only one otherwise
            c.ld(a.r2 + (NC == 2 ? (1 - w) : 0) * len, qs);
            geo_stiff<N>(a.geo, len, a.qm, c, chunk * S::K + el, l, act, ga, gb, gc);
#pragma unroll
            for (int i = 0; i < N; ++i) qs[i] = ib * qs[i];
            stiffness<N>(D, lds[wv][0], lds[wv][1], el, l, act, qs, ga, gb, gc, y);
            T qm[N];  // v of the output component (alpha partial, mass term)
            if constexpr (KIND == K_RES) {
                T cf[N], gj[N];
                c.ld(a.r2 + w * len, qm);
                c.ld(a.cf, cf);
                geo_mass<N>(a.geo, len, a.qm, c, chunk * S::K + el, l, act, gj);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    qm[i] = ib * qm[i];
                    const T t = a.alpha * (y[i] + cf[i] * qs[i]);
                    y[i] = (w == 0) ? gj[i] * qm[i] + t : t - gj[i] * qm[i];
                }
            } else if constexpr (KIND == K_HELMHOLTZ) {
                T cf[N];
                c.ld(a.cf, cf);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    qm[i] = qs[i];
                    y[i] = cf[i] * qs[i] + y[i];
                }
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i) qm[i] = qs[i];
            }
            if (has_r1 || pend) c.ld(a.r1 + w * len, r);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                if (has_r1 && owns<N>(i, l, bits)) y[i] = y[i] - cr1 * r[i];
                part += (double)qm[i] * (double)y[i];
            }
            c.st(a.y + w * len, y);
        } else if (pend) {
            c.ld(a.r1 + w * len, r);
        }
        if (pend) {
            // pending w/x update of the previous step on this component
            // (minres.hpp:82-86): w_new = ig*((ibp*r1 - eo*wa) - dl*wb), x += phi*w_new
            const T ibp = (T)sc.p3_ibp, ig = (T)sc.p3_ig, eo = (T)sc.p3_eo, dl = (T)sc.p3_dl, phi = (T)sc.p3_phi;
            T wa[N], wb[N], x[N], wn[N];
            c.ld(a.wa + w * len, wa);
            c.ld(a.wb + w * len, wb);
            c.ld(a.x + w * len, x);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const T v = ibp * r[i];
                wn[i] = ig * ((v - eo * wa[i]) - dl * wb[i]);
                x[i] = x[i] + phi * wn[i];
            }
            c.st(a.wn + w * len, wn);
            c.st(a.x + w * len, x);
        }
        part = wave_sum(part);
        if (lane == 0) wpart[wv] = part;
        __syncthreads();
        if (threadIdx.x < CPB) {
            double t = wpart[threadIdx.x * NC];
            if constexpr (NC == 2) t += wpart[threadIdx.x * NC + 1];
            a.partial[c0 + threadIdx.x] = t;
        }
        __syncthreads();
        return -1;
    };
    if constexpr (PERSIST) {
        for_items_persistent(n_items, body);
    } else if ((int)blockIdx.x < n_items) {
        body(xcd_item(blockIdx.x, n_items));
    }
}

// Closes a batch of Lanczos steps: the beta recurrence of the last step with
// the same reduction as the operator kernel of the step that would follow,
// so the scalars do not depend on where batches end. One workgroup per
// instance reads in[k] and writes out[k] (in may equal out). Defined in the
// FP64 translation unit only.
#ifndef SEM_OP_FP32
__global__ __launch_bounds__(256) void k_minres_close(InstTable it, const KrylovScalars* in, KrylovScalars* out,
                                                      const double* partial)
{
    __shared__ double red[4];
    const int k = blockIdx.x;
    const int p0 = it.inst_off[k];
    const double s2 = group_sum(partial + p0, it.inst_off[k + 1] - p0, red, 4);
    if (threadIdx.x == 0) {
        MinresState sc = *static_cast<const MinresState*>(in + k);
        if (!sc.done) minres_beta_update(sc, s2);
        else sc.p3_pending = 0;  // the update kernel of the stopped step applied it
        *static_cast<MinresState*>(out + k) = sc;
    }
}

void launch_minres_close(int kind, const InstTable& it, const KrylovScalars* in, KrylovScalars* out,
                         const double* partial, hipStream_t s)
{
    (void)kind;
    hipLaunchKernelGGL(k_minres_close, dim3(it.n_inst), dim3(256), 0, s, it, in, out, partial);
}
#endif


// Assembled values of column j of element e (lane position base of the
// chunk's first plane) from element-local values yr, in two phases so that
// the loads are in flight while the caller does other work: issue() loads the
// element's own column and every other copy of its ring nodes, finish()
// combines them in v. Interior planes are the element's own; ring entries sum
// every copy of their node in the reference's order (edge: owner + other,
// global_sum.hpp:28-34; corner: left fold from +0.0 in connected_elements
// order, global_sum.hpp:54-63), so every copy gets the same bits; Dirichlet
// entries are zero.
template <int N, typename T>
struct ColumnGather {
    static constexpr int NE = (N - 2 > 2) ? N - 2 : 2;
    T v[N];        // own column, then the assembled column
    T o[NE];       // other copies: planes 1..N-2 (side column) or planes 0 and N-1
    T c[2][4];     // copies of the low and high corner nodes (side column)
    int d0, d1;    // side descriptors
    unsigned cm;   // copies present: bits 0-3 low corner, 4-7 high corner
    bool side;     // column on side 1 (j = 0) or side 3 (j = N-1)

    // position of entry k of side descriptor d's side (d: (element << 3) |
    // (side - 1) << 1 | reversed)
    __device__ static __forceinline__ long side_pos(int d, int k)
    {
        using S = Shape<N>;
        constexpr int K = S::K, KN = S::KN;
        const int e2 = d >> 3, s2 = ((d >> 1) & 3) + 1;
        const int k2 = (d & 1) ? N - 1 - k : k;
        int i2, j2;
        if (s2 == 1) { i2 = k2; j2 = 0; }
        else if (s2 == 2) { i2 = N - 1; j2 = k2; }
        else if (s2 == 3) { i2 = k2; j2 = N - 1; }
        else { i2 = 0; j2 = k2; }
        return (long)(e2 / K) * K * N * N + (long)i2 * KN + (long)(e2 % K) * N + j2;
    }

    __device__ __forceinline__ void issue(const T* __restrict__ yr, size_t base, int e, int j,
                                          const int* __restrict__ eside, const int4* __restrict__ ecq,
                                          const int* __restrict__ c_off, const int* __restrict__ c_pos)
    {
        constexpr int KN = Shape<N>::KN;
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = yr[base + (size_t)i * KN];
        side = (j == 0 || j == N - 1);
        cm = 0;
        if (side) {
            d0 = eside[4 * e + (j == 0 ? 0 : 2)];
            d1 = -1;
            if (d0 >= 0) {
#pragma unroll
                for (int i = 1; i < N - 1; ++i) o[i - 1] = yr[side_pos(d0, i)];
            }
            const int4 q[2] = {ecq[4 * e + (j == 0 ? 0 : 3)], ecq[4 * e + (j == 0 ? 1 : 2)]};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (q[h].x == -2) {
                    // more than 4 copies: fold them from the corner's list
                    T sum = 0;
                    for (int t = c_off[q[h].y]; t < c_off[q[h].y + 1]; ++t) sum += yr[c_pos[t]];
                    c[h][0] = sum;
                    cm |= 1u << (4 * h);
                } else {
                    const int pp[4] = {q[h].x, q[h].y, q[h].z, q[h].w};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        c[h][t] = pp[t] >= 0 ? yr[pp[t]] : T(0);
                        if (pp[t] >= 0) cm |= 1u << (4 * h + t);
                    }
                }
            }
        } else {
            d0 = eside[4 * e + 3];  // plane 0 is side 4
            d1 = eside[4 * e + 1];  // plane N-1 is side 2
            o[0] = d0 >= 0 ? yr[side_pos(d0, j)] : T(0);
            o[1] = d1 >= 0 ? yr[side_pos(d1, j)] : T(0);
        }
    }

    __device__ __forceinline__ void finish()
    {
        if (side) {
#pragma unroll
            for (int i = 1; i < N - 1; ++i) v[i] = (d0 >= 0) ? v[i] + o[i - 1] : T(0);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                T sum = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t)
                    if ((cm >> (4 * h + t)) & 1u) sum += c[h][t];
                v[h ? N - 1 : 0] = sum;
            }
        } else {
            v[0] = (d0 >= 0) ? v[0] + o[0] : T(0);
            v[N - 1] = (d1 >= 0) ? v[N - 1] + o[1] : T(0);
        }
    }
};

// Second kernel of a Lanczos step: work items of CPB chunks, one wave per
// chunk and component. It finishes component w of y: interface copies gather
// the unassembled values of all copies of their node (edge: owner + other,
// global_sum.hpp:28-34; corner: left fold from +0.0 in connected_elements
// order, global_sum.hpp:54-63), so every copy computes the same sum bit for
// bit; Dirichlet entries become zero; then y -= c2*r2 (minres.hpp:55) and the
// beta^2 partial with unique-DOF weights (minres.hpp:63).
template <int N, typename T, int CPB, bool PERSIST>
__global__ __launch_bounds__(128 * CPB) void k_minres_upd(MinresUpdArgs<T> a)
{
    using S = Shape<N>;
    constexpr int K = S::K, KN = S::KN;
    __shared__ double wpart[2 * CPB];
    __shared__ double red[4];
    const int nc = a.ncomp;
    const int wall = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int sub = wall / nc, wv = wall % nc;
    const bool act = lane < KN;
    const size_t len = a.len;
    const int n_items = a.nchunks / CPB;
    int cur_inst = -1;
    double c2d = 0;
    auto body = [&](int item) {
        const int c0 = item * CPB;
        const int4 ci = chunk_instance(a.it, c0);
        const int inst = ci.x;
        const KrylovScalars* sc = a.st + inst;
        if (sc->done) return ci.z / CPB;  // stopped: its partials are never read again
        const int chunk = c0 + sub;
        const size_t base = (size_t)chunk * K * N * N + lane;
        const int el = act ? lane / N : 0, j = act ? lane % N : 0;
        const int e = chunk * K + el;
        // loads first: they do not depend on alpha
        ColumnGather<N, T> g;
        T r[N];
        if (act) {
            g.issue(a.yraw + (size_t)wv * len, base, e, j, a.eside, a.ecq, a.c_off, a.c_pos);
#pragma unroll
            for (int i = 0; i < N; ++i) r[i] = a.r2[(size_t)wv * len + base + (size_t)i * KN];
        }
        if (inst != cur_inst) {
            cur_inst = inst;
            // alpha = dot(v, y) from the operator kernel's partials (minres.hpp:54),
            // c2 = alpha / beta (minres.hpp:55); the instance's first workgroup keeps alpha
            const int p0 = ci.y;
            const double alpha = group_sum(a.op_partial + p0, ci.z - p0, red, nc * CPB);
            c2d = alpha / sc->beta;
            if (c0 == p0 && threadIdx.x == 0) a.st[inst].alpha = alpha;
        }
        double part = 0;
        if (act) {
            const T c2 = (T)c2d;
            const unsigned bits = a.obits[e];
            g.finish();
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const T yv = g.v[i] - c2 * r[i];
                a.y[(size_t)wv * len + base + (size_t)i * KN] = yv;
                if (owns<N>(i, j, bits)) part += (double)yv * (double)yv;
            }
        }
        part = wave_sum(part);
        if (lane == 0) wpart[wall] = part;
        __syncthreads();
        if (threadIdx.x < CPB) {
            double t = wpart[threadIdx.x * nc];
            if (nc == 2) t += wpart[threadIdx.x * nc + 1];
            a.partial[c0 + threadIdx.x] = t;
        }
        __syncthreads();
        return -1;
    };
    if constexpr (PERSIST) {
        for_items_persistent(n_items, body);
    } else if ((int)blockIdx.x < n_items) {
        body(xcd_item(blockIdx.x, n_items));
    }
}

// Fused Lanczos step (one launch per step, deferred scaling; see
// MinresFusedArgs). Work items of CPB chunks, one wave per chunk and output
// component. Every workgroup first evaluates the recurrence of this step
// (minres_fused_prologue) from the previous step's per-chunk sums; the
// workgroup holding an instance's first chunk stores the state in the other
// buffer. In step mode 2 the closing kernel evaluates the recurrence after
// every step and each launch starts from that state (first = 1).
// FORM 0 serves every case; FORM 3 is compiled for step mode 2 with the stored
// basis (the default for large meshes): without the in-kernel reduction and
// the w/x update its registers fit four waves per SIMD.
template <int N, typename T, int KIND, int CPB, int FORM, int MF = 0>
__global__ __launch_bounds__(64 * CPB * ((KIND == K_RES) ? 2 : 1))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 3 : 1, 8))) void k_minres_fused(MinresFusedArgs<T> a, Dm<N, T> D)
{
    using S = Shape<N>;
    constexpr bool EXT = (FORM & 1) != 0;    // step mode 2: the recurrence is closed outside
    constexpr bool BASIS = (FORM & 2) != 0;  // stored basis: no w/x update
    constexpr int NC = (KIND == K_RES) ? 2 : 1;
    constexpr int NW = NC * CPB;
    // MF: the stiffness contractions on MFMA (1 bf16, 3 split bf16), FP32 only
    static_assert(MF == 0 || sizeof(T) == 4, "MFMA contractions serve the FP32 step");
    __shared__ T lds[MF ? 1 : NW][2][MF ? 1 : S::LDS];
    __shared__ float mfs[MF ? NW : 1][MF ? 64 * kMfPitch : 1];
    // component exchange (Res): with VALU contractions it lives in the wave's
    // own transpose buffers (2 S::LDS >= 64 N values), free until the
    // stiffness apply that follows the exchange's second barrier; the LDS a
    // workgroup holds decides how many fit on a CU. Form 3 keeps it in its own
    // buffer instead and reads both components back from there after the
    // stiffness apply, so they are not held in registers across it (its
    // occupancy is set by registers, not LDS).
    constexpr bool XALIAS = MF == 0 && FORM != 3;
    __shared__ T xch_own[(NC == 2 && !XALIAS) ? NW : 1][(NC == 2 && !XALIAS) ? 64 * N : 1];
    static_assert(!XALIAS || 2 * S::LDS >= 64 * N, "exchange does not fit the transpose buffers");
    __shared__ double wpart[3][NW];
    __shared__ double red[12];
    const int wv = threadIdx.x >> 6;
    const int w = wv % NC;    // output component of this wave
    const int sub = wv / NC;  // chunk of the work item
    const int lane = threadIdx.x & 63;
    const bool act = lane < S::KN;
    const int el = act ? lane / N : 0;
    const int l = act ? lane % N : 0;
    const size_t len = a.len;
    const int nch = a.nchunks;
    const int n_items = nch / CPB;
    int cur_inst = -1;
    MinresState sc;
    bool pend = false;
    auto body = [&](int item) {
        const int c0 = item * CPB;
        const int4 ci = chunk_instance(a.it, c0);
        const int inst = ci.x;
        const bool fresh = inst != cur_inst;
        if (fresh) {
            cur_inst = inst;
            sc = *static_cast<const MinresState*>(a.st_in + inst);
        }
        const int chunk = c0 + sub;
        const size_t base = (size_t)chunk * S::K * N * N + lane;
        const Col<N, T> c{base, act};
        // loads that do not depend on this step's recurrence scalars go out
        // before the recurrence is evaluated (an instance that has stopped
        // stays stopped; y_{i-3} is needed from step 3 on)
        const bool early = !sc.done && a.step >= 2;
        const bool early3 = early && (a.step >= 3 || sc.has_r1);
        // the recurrence's sums are loaded first, so that its wait does not
        // include the gather
        const bool rec = !EXT && fresh && !a.first && a.step >= 2 && !sc.done;
        // four loads per lane and array: the in-kernel reduction serves step
        // mode 1, meshes (instances) of under 1024 chunks, so one or two
        // rounds; more loads per lane held registers across the gather (155
        // instead of 112 VGPRs at N = 8) for no gain. The order of the
        // additions does not depend on the loads per round.
        GroupSum3<NW, 4> gs;
        if (rec) gs.issue(a.prev_sums + ci.y, nch, ci.z - ci.y);
        ColumnGather<N, T> g;
        T r2[N], r3[N];  // y_{i-2}, y_{i-3} of this component
        if (early) {
            if (act) g.issue(a.z_in + w * len, base, chunk * S::K + el, l, a.eside, a.ecq, a.c_off, a.c_pos);
            c.ld(a.y2 + w * len, r2);
            if (early3) c.ld(a.y3 + w * len, r3);
        }
        if (fresh) {
            if (!a.first && a.step >= 2) {  // never in step mode 2 (first = 1)
                double P[3] = {0.0, 0.0, 0.0};  // not read once the instance has stopped
                if (rec) gs.finish(a.prev_sums + ci.y, nch, ci.z - ci.y, red, P);
                minres_fused_prologue(sc, a.step, P);
            }
            // this kernel applies the pending w/x update; the state it stores drops it
            pend = sc.p3_pending != 0;
            if (c0 == ci.y && threadIdx.x == 0 && pend && a.rec) minres_record(a.rec, sc);
            sc.p3_pending = 0;
            if (c0 == ci.y && threadIdx.x == 0) *static_cast<MinresState*>(a.st_out + inst) = sc;
            // stored basis: no w/x update (the solution is formed at the end)
            if (BASIS || a.rec) pend = false;
        }
        // a stopped instance has no work left: skip the rest of its items
        // (its sums are never read again)
        if (sc.done && !pend) return ci.z / CPB;
        const unsigned bits = act ? a.obits[chunk * S::K + el] : 0u;
        double p0 = 0, p1 = 0, p2 = 0;
        const bool has_r1 = sc.has_r1 != 0;
        if (!sc.done) {
            T yc[N];  // y_{i-1} of this component
            if (a.step == 1) {
                c.ld(a.y_out + w * len, yc);
            } else {
                const T ib = (T)sc.ib, cr1 = (T)sc.cr1, c2 = (T)(sc.alpha * sc.ib);
                if (act) g.finish();
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    T t = act ? g.v[i] * ib : T(0);
                    if (has_r1) t = t - cr1 * r3[i];
                    yc[i] = t - c2 * r2[i];
                    if (owns<N>(i, l, bits)) p2 += (double)yc[i] * (double)r2[i];
                }
                c.st(a.y_out + w * len, yc);
            }
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (owns<N>(i, l, bits)) p0 += (double)yc[i] * (double)yc[i];
            // stiffness input: the other component for Res, the only one otherwise
            T qs[N];
            // Res with a diagonal scaling S: the operator acts on S y and its
            // output is scaled by S again (S A S)
            const bool scl = (KIND == K_RES) && a.dscale != nullptr;
            T sv[N];
            if (scl) {
#pragma unroll
                for (int i = 0; i < N; ++i) sv[i] = act ? (T)a.dscale[base + (size_t)i * S::KN] : T(0);
            }
            T* mine = nullptr;   // this wave's y (scaled), then the other wave's
            T* other = nullptr;
            if constexpr (NC == 2) {
                mine = XALIAS ? &lds[wv][0][0] : xch_own[wv];
                other = XALIAS ? &lds[sub * NC + (1 - w)][0][0] : xch_own[sub * NC + (1 - w)];
#pragma unroll
                for (int i = 0; i < N; ++i) mine[i * 64 + lane] = scl ? yc[i] * sv[i] : yc[i];
                __syncthreads();
#pragma unroll
                for (int i = 0; i < N; ++i) qs[i] = other[i * 64 + lane];
                __syncthreads();
            } else {
#pragma unroll
                for (int i = 0; i < N; ++i) qs[i] = yc[i];
            }
            T z[N];
            if constexpr (sizeof(T) == 4 && MF == 0 && map_geometry(N)) {
                const QuadColumn<N> qc(a.qm, chunk * S::K + el, l, act);
                stiffness<N>(D, lds[wv][0], lds[wv][1], el, l, act, qs,
                             [&](int i, T& ga, T& gb, T& gc) { qc.at(a.qm, i, ga, gb, gc); }, z);
            } else {
                T ga[N], gb[N], gc[N];
                geo_stiff<N>(a.geo, len, a.qm, c, chunk * S::K + el, l, act, ga, gb, gc);
                if constexpr (MF != 0) stiffness_mf<N, MF>(mf_dmat<N>(D, lane), mfs[wv], lane, qs, ga, gb, gc, z);
                else stiffness<N>(D, lds[wv][0], lds[wv][1], el, l, act, qs, ga, gb, gc, z);
            }
            if constexpr (KIND == K_RES) {
                // u = A (S y) on this component, z = S u; y.z = (S y).u
                T cf[N], gj[N];
                c.ld(a.cf, cf);
                geo_mass<N>(a.geo, len, a.qm, c, chunk * S::K + el, l, act, gj);
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    T qv, yo;
                    if constexpr (!XALIAS) {
                        qv = other[i * 64 + lane];
                        yo = mine[i * 64 + lane];
                    } else {
                        qv = qs[i];
                        yo = scl ? yc[i] * sv[i] : yc[i];
                    }
                    const T t = a.alpha * (z[i] + cf[i] * qv);
                    const T u = (w == 0) ? gj[i] * yo + t : t - gj[i] * yo;
                    p1 += (double)yo * (double)u;
                    z[i] = scl ? u * sv[i] : u;
                }
            } else {
                if constexpr (KIND == K_HELMHOLTZ) {
                    T cf[N];
                    c.ld(a.cf, cf);
#pragma unroll
                    for (int i = 0; i < N; ++i) z[i] = cf[i] * qs[i] + z[i];
                }
#pragma unroll
                for (int i = 0; i < N; ++i) p1 += (double)yc[i] * (double)z[i];
            }
            c.st(a.z_out + w * len, z);
        }
        if (pend) {
            // pending w/x update of iteration i-2 (minres.hpp:82-86) with
            // v = y_{i-3}/beta_{i-2}: w_new = ig*((v - eo*wa) - dl*wb), x += phi*w_new
            if (!early3) c.ld(a.y3 + w * len, r3);
            const T ibp = (T)sc.p3_ibp, ig = (T)sc.p3_ig, eo = (T)sc.p3_eo, dl = (T)sc.p3_dl, phi = (T)sc.p3_phi;
            T wa[N], wb[N], x[N], wn[N];
            c.ld(a.wa + w * len, wa);
            c.ld(a.wb + w * len, wb);
            c.ld(a.x + w * len, x);
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const T v = ibp * r3[i];
                wn[i] = ig * ((v - eo * wa[i]) - dl * wb[i]);
                x[i] = x[i] + phi * wn[i];
            }
            c.st(a.wn + w * len, wn);
            c.st(a.x + w * len, x);
        }
        p0 = wave_sum(p0);
        p1 = wave_sum(p1);
        p2 = wave_sum(p2);
        if (lane == 0) {
            wpart[0][wv] = p0;
            wpart[1][wv] = p1;
            wpart[2][wv] = p2;
        }
        __syncthreads();
        if (threadIdx.x < 3 * CPB) {
            const int q = threadIdx.x / CPB, ch = threadIdx.x % CPB;
            double t = wpart[q][ch * NC];
            if constexpr (NC == 2) t += wpart[q][ch * NC + 1];
            a.sums[(size_t)q * nch + c0 + ch] = t;
        }
        __syncthreads();
        return -1;
    };
    if ((int)blockIdx.x < n_items) body(xcd_item(blockIdx.x, n_items));
}

#ifndef SEM_OP_FP32
// Closes a batch of fused steps: the recurrence step `step` would evaluate at
// its start, with the same reduction; one workgroup per instance.
__global__ __launch_bounds__(768) void k_minres_fused_close(InstTable it, const KrylovScalars* in, KrylovScalars* out,
                                                            const double* sums, int nch, int step, double* rec)
{
    __shared__ double red[12];
    const int k = blockIdx.x;
    const int p0 = it.inst_off[k];
    const int n = it.inst_off[k + 1] - p0;
    // twelve waves: wave 4a + g adds virtual lanes 64 g .. 64 g + 63 of array a
    // (group_sum3's decomposition and order, so the same bits), every load of
    // an instance of up to 8192 chunks in one round
    {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int a = wv >> 2, g = wv & 3;
        const double* pa = sums + p0 + (size_t)a * nch;
        constexpr int KB = 32;
        double sacc = 0;
        for (int kb = 0; kb * 256 < n; kb += KB) {
            double v[KB];
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int q = (kb + u) * 256 + g * 64 + lane;
                v[u] = q < n ? pa[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < KB; ++u) {
                const int q = (kb + u) * 256 + g * 64 + lane;
                if (q < n) sacc += v[u];
            }
        }
        const double t = wave_sum(sacc);
        if (lane == 0) red[4 * a + g] = t;
    }
    __syncthreads();
    double P[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) P[a] = ((red[4 * a] + red[4 * a + 1]) + red[4 * a + 2]) + red[4 * a + 3];
    if (threadIdx.x == 0) {
        MinresState sc = *static_cast<const MinresState*>(in + k);
        minres_fused_prologue(sc, step, P);
        if (rec && sc.p3_pending) minres_record(rec, sc);
        *static_cast<MinresState*>(out + k) = sc;
    }
}

void launch_minres_fused_close(const InstTable& it, const KrylovScalars* in, KrylovScalars* out, const double* sums,
                               int step, hipStream_t s, double* rec)
{
    hipLaunchKernelGGL(k_minres_fused_close, dim3(it.n_inst), dim3(768), 0, s, it, in, out, sums, it.nchunks, step,
                       rec);
}
#endif

#ifndef SEM_OP_FP32
// Diagonal of the element stiffness (glaplace) at every element-local node:
// the stiffness action on each unit vector e_(a,b) of the element, read at
// (a, b). One wave per chunk, N*N actions; used once per scaled system.
template <int N>
__global__ __launch_bounds__(64) void k_stiff_diag(const double* __restrict__ geo, size_t len,
                                                   double* __restrict__ out, Dm<N, double> D)
{
    using S = Shape<N>;
    __shared__ double lds0[S::LDS], lds1[S::LDS];
    const int lane = threadIdx.x & 63;  // one wave per workgroup (see k_element_op)
    const bool act = lane < S::KN;
    const int el = act ? lane / N : 0;
    const int l = act ? lane % N : 0;
    const Col<N, double> c{(size_t)blockIdx.x * S::K * N * N + lane, act};
    double ga[N], gb[N], gc[N], dg[N];
    c.ld(geo, ga);
    c.ld(geo + len, gb);
    c.ld(geo + 2 * len, gc);
#pragma unroll
    for (int i = 0; i < N; ++i) dg[i] = 0.0;
    for (int b = 0; b < N; ++b)
        for (int a = 0; a < N; ++a) {
            double q[N], y[N];
#pragma unroll
            for (int i = 0; i < N; ++i) q[i] = (l == b && i == a) ? 1.0 : 0.0;
            stiffness<N>(D, lds0, lds1, el, l, act, q, ga, gb, gc, y);
#pragma unroll
            for (int i = 0; i < N; ++i) dg[i] = (l == b && i == a) ? y[i] : dg[i];
        }
    c.st(out, dg);
}

#endif

// PCG step (pcg.hpp:70-80, identity preconditioner): p = r (first step) or
// r + beta*p; Ap on interior entries; pAp partial over interior entries.

template <int N, typename T>
__global__ __launch_bounds__(64) void k_pcg_op(PcgOpArgs<T> a, Dm<N, T> D)
{
    using S = Shape<N>;
    __shared__ T lds0[S::LDS], lds1[S::LDS];
    const KrylovScalars* sc = reinterpret_cast<const KrylovScalars*>(a.sc);
    const int lane = threadIdx.x & 63;  // one wave per workgroup (see k_element_op)
    const bool act = lane < S::KN;
    const int el = act ? lane / N : 0;
    const int l = act ? lane % N : 0;
    const Col<N, T> c{(size_t)blockIdx.x * S::K * N * N + lane, act};
    double part = 0;
    if (!sc->done) {
        const int first = sc->first;
        const T beta = (T)sc->pcg_beta;
        T q[N], ga[N], gb[N], gc[N], y[N];
        c.ld(a.r, q);
        if (!first) {
            T pp[N];
            c.ld(a.p, pp);
#pragma unroll
            for (int i = 0; i < N; ++i) q[i] = q[i] + beta * pp[i];
        }
        c.st(a.p, q);
        c.ld(a.geo, ga);
        c.ld(a.geo + a.len, gb);
        c.ld(a.geo + 2 * a.len, gc);
        stiffness<N>(D, lds0, lds1, el, l, act, q, ga, gb, gc, y);
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (is_interior<N>(i, l)) part += (double)q[i] * (double)y[i];
        c.st(a.ap, y);
    }
    part = wave_sum(part);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = part;
}

// ---------------------------------------------------------------- launch --
template <int N, typename T>
static Dm<N, T> make_dm(const T* d)
{
    Dm<N, T> m;
    for (int k = 0; k < N * N; ++k) m.v[k] = d[k];
    return m;
}

// Arrays of one element-operator launch (inputs, geometry, coefficient, outputs)
// larger than the 256 MB Infinity Cache are streamed with nontemporal accesses;
// smaller ones stay cacheable for the kernel that reads the result next.
constexpr double kMallBytes = 256.0 * 1024 * 1024;

template <int N, typename T, bool NT>
static void element_op_NT(int kind, const OpArgs<T>& a, const Dm<N, T>& D, hipStream_t s)
{
    dim3 g(a.nchunks), b(64);
    if constexpr (sizeof(T) == 4) {
        if (a.mf != 0) {
            if (kind != K_RES) throw std::runtime_error("MFMA contractions serve the stage operator Res");
            if (a.mf == 1) hipLaunchKernelGGL((k_element_op<N, T, K_RES, NT, 1>), g, b, 0, s, a, D);
            else hipLaunchKernelGGL((k_element_op<N, T, K_RES, NT, 3>), g, b, 0, s, a, D);
            return;
        }
    } else {
        if (a.mf != 0) throw std::runtime_error("MFMA contractions serve the FP32 apply");
    }
    switch (kind) {
    case K_LAPLACE: hipLaunchKernelGGL((k_element_op<N, T, K_LAPLACE, NT>), g, b, 0, s, a, D); break;
    case K_HELMHOLTZ: hipLaunchKernelGGL((k_element_op<N, T, K_HELMHOLTZ, NT>), g, b, 0, s, a, D); break;
    case K_MASS: hipLaunchKernelGGL((k_element_op<N, T, K_MASS, NT>), g, b, 0, s, a, D); break;
    case K_RES: hipLaunchKernelGGL((k_element_op<N, T, K_RES, NT>), g, b, 0, s, a, D); break;
    case K_F: hipLaunchKernelGGL((k_element_op<N, T, K_F, NT>), g, b, 0, s, a, D); break;
    case K_L: hipLaunchKernelGGL((k_element_op<N, T, K_L, NT>), g, b, 0, s, a, D); break;
    }
}

template <int N, typename T>
static void element_op_N(int kind, const OpArgs<T>& a, const T* d, hipStream_t s)
{
    auto D = make_dm<N, T>(d);
    const double comps = (kind == K_RES || kind == K_F) ? 2 : 1;
    const double arrays = 2 * comps + 4 + ((kind == K_HELMHOLTZ || kind == K_RES || kind == K_L) ? 1 : 0);
    if (arrays * (double)a.len * sizeof(T) > kMallBytes) element_op_NT<N, T, true>(kind, a, D, s);
    else element_op_NT<N, T, false>(kind, a, D, s);
}

// Launch; with events, they take the kernel's own start and end timestamps.
#define SEM_LAUNCH_TIMED(KERNEL, G, B, S, E0, E1, ...)                                  \
    do {                                                                                \
        if (E0) hipExtLaunchKernelGGL(KERNEL, G, B, 0, S, E0, E1, 0, __VA_ARGS__);      \
        else hipLaunchKernelGGL(KERNEL, G, B, 0, S, __VA_ARGS__);                       \
    } while (0)

template <int N, typename T, int CPB, bool P>
static void minres_op_NC(int kind, const MinresOpArgs<T>& a, const Dm<N, T>& D, int grid, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1)
{
    const int items = a.nchunks / CPB;
    dim3 g(P ? std::min(grid, items) : items);
    dim3 b1(64 * CPB), b2(128 * CPB);
    switch (kind) {
    case K_LAPLACE: SEM_LAUNCH_TIMED((k_minres_op<N, T, K_LAPLACE, CPB, P>), g, b1, s, e0, e1, a, D); break;
    case K_HELMHOLTZ: SEM_LAUNCH_TIMED((k_minres_op<N, T, K_HELMHOLTZ, CPB, P>), g, b1, s, e0, e1, a, D); break;
    case K_RES: SEM_LAUNCH_TIMED((k_minres_op<N, T, K_RES, CPB, P>), g, b2, s, e0, e1, a, D); break;
    default: throw std::runtime_error("minres: unsupported operator kind");
    }
}

// grid > 0 selects the persistent form (grid-stride over work items)
template <int N, typename T>
static void minres_op_N(int kind, const MinresOpArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1)
{
    auto D = make_dm<N, T>(d);
    if (cpb == 2) {
        if (grid > 0) minres_op_NC<N, T, 2, true>(kind, a, D, grid, s, e0, e1);
        else minres_op_NC<N, T, 2, false>(kind, a, D, grid, s, e0, e1);
    } else {
        if (grid > 0) minres_op_NC<N, T, 1, true>(kind, a, D, grid, s, e0, e1);
        else minres_op_NC<N, T, 1, false>(kind, a, D, grid, s, e0, e1);
    }
}

template <int N, typename T>
static void pcg_op_N(const PcgOpArgs<T>& a, int nchunks, const T* d, hipStream_t s)
{
    auto D = make_dm<N, T>(d);
    hipLaunchKernelGGL((k_pcg_op<N, T>), dim3(nchunks), dim3(64), 0, s, a, D);
}

template <int N, typename T, int CPB, bool P>
static void minres_upd_NC(const MinresUpdArgs<T>& a, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    const int items = a.nchunks / CPB;
    dim3 g(P ? std::min(grid, items) : items), b(64 * a.ncomp * CPB);
    SEM_LAUNCH_TIMED((k_minres_upd<N, T, CPB, P>), g, b, s, e0, e1, a);
}

template <int N, typename T>
static void minres_upd_N(const MinresUpdArgs<T>& a, int cpb, int grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1)
{
    if (cpb == 2) {
        if (grid > 0) minres_upd_NC<N, T, 2, true>(a, grid, s, e0, e1);
        else minres_upd_NC<N, T, 2, false>(a, grid, s, e0, e1);
    } else {
        if (grid > 0) minres_upd_NC<N, T, 1, true>(a, grid, s, e0, e1);
        else minres_upd_NC<N, T, 1, false>(a, grid, s, e0, e1);
    }
}

// The fused step runs the FP32 inner solver of the mixed-precision stage
// solves (operator Res) only.
template <int N, typename T, int CPB, bool P>
static void minres_fused_NC(int kind, const MinresFusedArgs<T>& a, const Dm<N, T>& D, int grid, hipStream_t s,
                            hipEvent_t e0, hipEvent_t e1)
{
    (void)grid;
    // One work item per workgroup for every shape: the persistent form of
    // this kernel spilled registers and moved 1.4-1.9x its algorithmic bytes
    // (and was 3-4x slower), so a persistent request runs the one-shot grid;
    // results do not depend on the shape.
    const dim3 g(a.nchunks / CPB);
    const dim3 b2(128 * CPB);
    if (kind != K_RES) throw std::runtime_error("minres: the one-launch step serves the Schrodinger stage operator");
    // step mode 2 with the stored basis runs its own compiled form
    const bool f3 = a.ext && a.rec != nullptr;
    if (a.mf == 0 && f3) SEM_LAUNCH_TIMED((k_minres_fused<N, T, K_RES, CPB, 3>), g, b2, s, e0, e1, a, D);
    else if (a.mf == 0) SEM_LAUNCH_TIMED((k_minres_fused<N, T, K_RES, CPB, 0>), g, b2, s, e0, e1, a, D);
    else if (a.mf == 1) SEM_LAUNCH_TIMED((k_minres_fused<N, T, K_RES, CPB, 0, 1>), g, b2, s, e0, e1, a, D);
    else SEM_LAUNCH_TIMED((k_minres_fused<N, T, K_RES, CPB, 0, 3>), g, b2, s, e0, e1, a, D);
}

template <int N, typename T>
static void minres_fused_N(int kind, const MinresFusedArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                           hipEvent_t e0, hipEvent_t e1)
{
    auto D = make_dm<N, T>(d);
    if (cpb == 2) {
        if (grid > 0) minres_fused_NC<N, T, 2, true>(kind, a, D, grid, s, e0, e1);
        else minres_fused_NC<N, T, 2, false>(kind, a, D, grid, s, e0, e1);
    } else {
        if (grid > 0) minres_fused_NC<N, T, 1, true>(kind, a, D, grid, s, e0, e1);
        else minres_fused_NC<N, T, 1, false>(kind, a, D, grid, s, e0, e1);
    }
}

// The DSS of an element operator's output as one gather per element column
// (ColumnGather), writing the assembled vector out of place: one workgroup per
// chunk, one wave per component.
template <int N, typename T>
__global__ __launch_bounds__(128) void k_gather_assemble(GatherArgs<T> a)
{
    using S = Shape<N>;
    const int chunk = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane >= S::KN) return;  // no LDS and no barrier below
    const int el = lane / N, l = lane % N;
    const size_t base = (size_t)chunk * S::K * N * N + lane;
    ColumnGather<N, T> g;
    g.issue(a.in + w * a.len, base, chunk * S::K + el, l, a.eside, a.ecq, a.c_off, a.c_pos);
    g.finish();
    const bool side = (l == 0 || l == N - 1);
    T* o = a.out + w * a.len + base;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool ring = side || i == 0 || i == N - 1;
        o[(size_t)i * S::KN] = ring ? a.post * g.v[i] : g.v[i];
    }
}

template <int N, typename T>
static void gather_assemble_N(const GatherArgs<T>& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_gather_assemble<N, T>), dim3(a.nchunks), dim3(64 * a.ncomp), 0, s, a);
}

#define SEM_DISPATCH_N(N, CALL)                                   \
    switch (N) {                                                  \
    case 2: CALL(2); break;                                       \
    case 3: CALL(3); break;                                       \
    case 4: CALL(4); break;                                       \
    case 5: CALL(5); break;                                       \
    case 6: CALL(6); break;                                       \
    case 7: CALL(7); break;                                       \
    case 8: CALL(8); break;                                       \
    case 9: CALL(9); break;                                       \
    case 10: CALL(10); break;                                     \
    case 11: CALL(11); break;                                     \
    case 12: CALL(12); break;                                     \
    case 13: CALL(13); break;                                     \
    case 14: CALL(14); break;                                     \
    case 15: CALL(15); break;                                     \
    case 16: CALL(16); break;                                     \
    default: throw std::runtime_error("unsupported N");           \
    }

template <typename T>
void launch_element_op_d(int N, int kind, const OpArgs<T>& a, const T* d, hipStream_t s)
{
#define CALL_EO(n) element_op_N<n, T>(kind, a, d, s)
    SEM_DISPATCH_N(N, CALL_EO)
#undef CALL_EO
}

template <typename T>
void launch_gather_assemble(int N, const GatherArgs<T>& a, hipStream_t s)
{
    if (a.ncomp < 1 || a.ncomp > 2 || a.in == a.out) throw std::runtime_error("gather_assemble: bad arguments");
#define CALL_GA(n) gather_assemble_N<n, T>(a, s)
    SEM_DISPATCH_N(N, CALL_GA)
#undef CALL_GA
}

template <typename T>
void launch_minres_op_d(int N, int kind, const MinresOpArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1)
{
    if (a.nchunks % kChunkAlign != 0) throw std::runtime_error("internal error: chunk count not aligned");
#define CALL_MO(n) minres_op_N<n, T>(kind, a, d, cpb, grid, s, e0, e1)
    SEM_DISPATCH_N(N, CALL_MO)
#undef CALL_MO
}

template <typename T>
void launch_minres_fused(int N, int kind, const MinresFusedArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                         hipEvent_t e0, hipEvent_t e1)
{
    if (a.nchunks % kChunkAlign != 0) throw std::runtime_error("internal error: chunk count not aligned");
#define CALL_MF(n) minres_fused_N<n, T>(kind, a, d, cpb, grid, s, e0, e1)
    SEM_DISPATCH_N(N, CALL_MF)
#undef CALL_MF
}

template <typename T>
void launch_minres_upd(int N, const MinresUpdArgs<T>& a, int cpb, int grid, hipStream_t s, hipEvent_t e0,
                       hipEvent_t e1)
{
    if (a.nchunks % kChunkAlign != 0) throw std::runtime_error("internal error: chunk count not aligned");
#define CALL_MU(n) minres_upd_N<n, T>(a, cpb, grid, s, e0, e1)
    SEM_DISPATCH_N(N, CALL_MU)
#undef CALL_MU
}

template <typename T>
void launch_pcg_op_d(int N, const PcgOpArgs<T>& a, int nchunks, const T* d, hipStream_t s)
{
#define CALL_PO(n) pcg_op_N<n, T>(a, nchunks, d, s)
    SEM_DISPATCH_N(N, CALL_PO)
#undef CALL_PO
}

// The translation unit is compiled once per value type (SEM_OP_FP32 set or
// not) so the two halves of the kernel set build in parallel.
#ifndef SEM_OP_FP32
template <int N>
static void stiff_diag_N(const double* geo, size_t len, int nchunks, double* out, const double* d, hipStream_t s)
{
    hipLaunchKernelGGL(k_stiff_diag<N>, dim3(nchunks), dim3(64), 0, s, geo, len, out, make_dm<N, double>(d));
}

void launch_stiff_diag(int N, const double* geo, size_t len, int nchunks, double* out, const double* d, hipStream_t s)
{
#define CALL_SD(n) stiff_diag_N<n>(geo, len, nchunks, out, d, s)
    SEM_DISPATCH_N(N, CALL_SD)
#undef CALL_SD
}

template void launch_element_op_d<double>(int, int, const OpArgs<double>&, const double*, hipStream_t);
template void launch_gather_assemble<double>(int, const GatherArgs<double>&, hipStream_t);
template void launch_minres_op_d<double>(int, int, const MinresOpArgs<double>&, const double*, int, int, hipStream_t,
                                          hipEvent_t, hipEvent_t);
template void launch_minres_upd<double>(int, const MinresUpdArgs<double>&, int, int, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_pcg_op_d<double>(int, const PcgOpArgs<double>&, int, const double*, hipStream_t);
// the one-launch step is an FP32 kernel only (FP64 solves use the two-launch step)
template <>
void launch_minres_fused<double>(int, int, const MinresFusedArgs<double>&, const double*, int, int, hipStream_t,
                                 hipEvent_t, hipEvent_t)
{
    throw std::runtime_error("the one-launch Lanczos step runs in FP32 only");
}
#else
template void launch_element_op_d<float>(int, int, const OpArgs<float>&, const float*, hipStream_t);
template void launch_gather_assemble<float>(int, const GatherArgs<float>&, hipStream_t);
template void launch_minres_op_d<float>(int, int, const MinresOpArgs<float>&, const float*, int, int, hipStream_t,
                                          hipEvent_t, hipEvent_t);
template void launch_minres_upd<float>(int, const MinresUpdArgs<float>&, int, int, hipStream_t, hipEvent_t, hipEvent_t);
template void launch_minres_fused<float>(int, int, const MinresFusedArgs<float>&, const float*, int, int, hipStream_t,
                                           hipEvent_t, hipEvent_t);
#endif

}  // namespace sem
