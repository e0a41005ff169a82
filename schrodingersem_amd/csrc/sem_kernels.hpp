// sem_kernels.hpp -- argument structs and launchers of kernels_vec.hip.
#pragma once

#include "sem_internal.hpp"

namespace sem {

// DSS work modes (one launch covers interior-edge pairs, interior corners and
// Dirichlet entries):
//   DSS_SUM     global_sum (global_sum.hpp:65-77): sums to every copy, boundary untouched
//   DSS_SUM_ZB  global_sum then zero Dirichlet entries, sums scaled by post[c]
//   DSS_UNMASK  unmask (unmask.hpp:43-55): owner value to every copy
//   DSS_MASK    mask (mask.hpp:16-28): zero non-owner copies and Dirichlet entries
//   DSS_PCG     DSS_SUM_ZB fused with the p.Ap partial
//   DSS_ZERO_BND zero Dirichlet entries only
enum DssMode { DSS_SUM = 0, DSS_SUM_ZB = 1, DSS_UNMASK = 2, DSS_MASK = 3, DSS_PCG = 5, DSS_ZERO_BND = 6 };

template <typename T>
struct DssArgs {
    T* y;
    size_t len;
    int ncomp;
    const int2* pairs;
    int n_pairs;
    const int* c_off;
    const int* c_pos;
    int n_corners;
    const int* bnd;
    int n_bnd;
    T post[2];
    const T* r2;  // PCG: search direction p
    const double* sc;
    double* partial;
};

template <typename T>
inline int dss_blocks(const DssArgs<T>& a)
{
    const long total = (long)a.n_pairs + a.n_corners + a.n_bnd;
    return (int)std::max<long>(1, (total + 255) / 256);
}

template <typename T>
struct LinComb {
    T* out;
    const T* x[6];
    double c[6];
    int m;
    int mode;
    size_t n;
};

template <typename T>
void launch_dss(int mode, const DssArgs<T>& a, hipStream_t s);
template <typename T>
void launch_lincomb(const LinComb<T>& a, hipStream_t s);
template <typename D, typename S>
void launch_cast(D* out, const S* in, size_t n, hipStream_t s);
// out = in * S on each of ncomp components (S: one value per node, len nodes)
template <typename D, typename S>
void launch_scale_cast(D* out, const S* in, const float* sc, size_t len, int ncomp, hipStream_t s);
// S = 1 / sqrt(m + |alpha| |k + c|) (0 where that is not positive)
void launch_inner_scale(float* out, const double* m, const double* k, const double* c, double alpha, size_t n,
                        hipStream_t s);
template <typename T>
void launch_fill(T* out, T v, size_t n, hipStream_t s);
// n doubles (even), 16-byte nontemporal accesses
void launch_copy16(double* out, const double* in, size_t n, int blocks, hipStream_t s);
template <typename T>
void launch_mul(T* out, const T* a, const T* b, size_t n, hipStream_t s);
template <typename T>
void launch_wdot(const T* a, const T* b, size_t len, int ncomp, const Layout& L, const uint8_t* obits,
                 double* partial, hipStream_t s);
int launch_edot(const double* a, const double* b, size_t n, double* partial, hipStream_t s);
void launch_fin_sum(const double* partial, int n, double* dst, hipStream_t s);

void launch_fin_sum_inst(const double* partial, const InstTable& it, KrylovScalars* sc, int slot, hipStream_t s);

// MINRES set-up of every instance from sc[k].red[slot] = dot(r1, r1)
// the MINRES start of minres.hpp:24-40 from ||r1||^2 in red[slot]; atol < 0
// means atol = tol (minres.hpp:92); bref_slot >= 0 takes the stop test's
// reference norm from red[bref_slot] instead (||b||, the test with x0 = 0)
// <a,a>, <b,b>, <a,b> per chunk in one pass (partials at q * nchunks + chunk), and
// their per-instance fixed-order sums into red[slot0 .. slot0 + 2]
template <typename T>
void launch_wdot3(const T* a, const T* b, size_t len, int ncomp, const Layout& L, const uint8_t* obits, double* partial,
                  hipStream_t s);
void launch_fin_sum3_inst(const double* partial, const InstTable& it, KrylovScalars* sc, int nch, int slot0,
                          hipStream_t s);
void launch_minres_init(KrylovScalars* sc, int n_inst, int slot, int maxit, double tol, hipStream_t s,
                        double atol = -1.0, int bref_slot = -1);

template <typename T>
struct MinresFlushArgs {
    const T* R[3];
    T* W[3];
    T* x;
    size_t len;     // entries per component
    int chunk_len;  // entries per chunk and component (K*N*N)
    int ncomp;
};
template <typename T>
void launch_minres_flush(const MinresFlushArgs<T>& a, const InstTable& it, int nchunks, KrylovScalars* sc,
                         hipStream_t s);

void launch_pcg_init(KrylovScalars* sc, int maxit, double tol, hipStream_t s);
void launch_pcg_fin_alpha(KrylovScalars* sc, const double* partial, int n, hipStream_t s);
void launch_pcg_upd(double* x, double* r, const double* p, const double* ap, size_t len, const Layout& L,
                    const uint8_t* obits, const double* sc, double* partial, hipStream_t s);
void launch_pcg_fin_rho(KrylovScalars* sc, const double* partial, int n, hipStream_t s);

void launch_tridiag(const double* x, double* y, long n, double lo, double di, double up, hipStream_t s);
// element-local arrays of E elements: reference layout -> device layout (to_dev = 1) or back
void launch_layout(const double* in, double* out, long E, int N, int K, int to_dev, hipStream_t s);
// out = 1 / |d| (0 where d = 0)
void launch_abs_recip(double* out, const double* d, size_t n, hipStream_t s);

}  // namespace sem
