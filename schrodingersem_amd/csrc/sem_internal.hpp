// sem_internal.hpp -- host/device data model of libsem_hip.so.
//
// Device layout ("chunked pencil" layout). Elements are grouped in chunks of
// K = floor(64 / N) elements, one 64-lane wavefront per chunk, lane
// l = el*N + j owning column j of element el. Inside a chunk the N x N blocks
// are stored row-plane by row-plane:
//     pos(e, i, j) = (e / K) * K*N*N + i * K*N + (e % K) * N + j
// so the operator kernel's load of plane i is one contiguous K*N-value
// stream (coalesced), and every lane holds its column in registers.
// Vectors live in "consistent" form inside the solvers: every copy of a shared
// node holds the same value and Dirichlet entries are zero; the reference's
// masked form (owner copy only, openmp_impl/mask.hpp:16-28) is produced at the
// ABI boundary. Both carry the same owner values, so the Krylov recurrences see
// identical numbers up to summation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "sem.h"

namespace sem {

constexpr int kMinN = 2;
constexpr int kMaxN = 16;  // p <= 15

void set_error(const std::string& s);

struct HipError {
    hipError_t code;
    std::string where;
};

#define SEM_HIP(call)                                                     \
    do {                                                                  \
        hipError_t _e = (call);                                           \
        if (_e != hipSuccess) throw ::sem::HipError{_e, #call};           \
    } while (0)

// ------------------------------------------------------------- host mesh --
enum { INTERIOR = 0, BOUNDARY = 1 };

struct EdgeRec {        // one unique edge (Edge.hpp:6-27)
    int type = BOUNDARY;
    int el[2] = {-1, -1};
    int side[2] = {0, 0};  // side of el[1] is negative when reversed
    int start = 0, delta = 0;
    int n0 = -1, n1 = -1;  // sorted corner ids (the edge key)
};

struct HostMesh {
    int p = 0, N = 0;
    std::vector<double> x, w, D;  // GLL nodes, weights, D column-major
    int n_nodes = 0, n_el = 0;
    std::vector<double> node_xy;
    std::vector<int> quads;
    std::vector<EdgeRec> edges;
    std::vector<int> node_type;
    std::vector<int> conn_off, conn;  // per node: (element*4 + corner) in element order
    std::vector<double> J, A, B, C;   // reference element-local layout
    long dof = 0;
};

void gll_rule(int p, std::vector<double>& x, std::vector<double>& w);
std::vector<double> diff_matrix(const std::vector<double>& x);
void build_host_mesh(HostMesh& hm, int p, const double* xy, int n_nodes, const int* quads, int n_el);
void read_mesh_folder(const std::string& folder, std::vector<double>& xy, std::vector<int>& quads);

// ---------------------------------------------------------------- layout --
struct Layout {
    int N = 0, K = 0, E = 0, nchunks = 0;
    size_t len = 0;  // padded entries per component
    void init(int N_, int E_)
    {
        N = N_;
        E = E_;
        K = 64 / N;
        nchunks = (E + K - 1) / K;
        len = (size_t)nchunks * K * N * N;
    }
    size_t pos(int e, int i, int j) const
    {
        return (size_t)(e / K) * K * N * N + (size_t)i * K * N + (size_t)(e % K) * N + j;
    }
    size_t ref(int e, int i, int j) const { return (size_t)e * N * N + i + (size_t)j * N; }
};

// host-side permutations between the reference layout and the device layout
void to_device_layout(const Layout& L, const double* ref, double* dev);
void from_device_layout(const Layout& L, const double* dev, double* ref);

// ---------------------------------------------------------- device mesh --
struct DssLists {
    // interior edges: pairs (owner copy, other copy), (N-2) per interior edge
    int n_pairs = 0;
    int2* pairs = nullptr;
    // interior corners: CSR over copies in connected_elements order
    int n_corners = 0;
    int* corner_off = nullptr;
    int* corner_pos = nullptr;
    // Dirichlet entries: boundary-edge owner sides and every copy of a boundary corner
    int n_bnd = 0;
    int* bnd = nullptr;
    // non-owner copies (for converting to the reference's masked form)
    int n_nonowner = 0;
    int* nonowner = nullptr;
};

struct Workspace;

struct DeviceMesh {
    Layout L;
    hipStream_t stream = nullptr;
    // geometry, premultiplied by w_i w_j: gA, gB, gC, gJ (FP64 and FP32 copies)
    double* geo = nullptr;  // 4 * L.len
    float* geo32 = nullptr;
    uint8_t* obits = nullptr;  // per element: bit s-1 side s owned, bit 4+k corner k owned
    DssLists dss;
    std::vector<uint8_t> h_obits;
    Workspace* ws = nullptr;
};

// ------------------------------------------------------------ op params --
enum OpKind { K_LAPLACE = 0, K_HELMHOLTZ = 1, K_MASS = 2, K_RES = 3, K_F = 4, K_L = 5 };
inline int op_ncomp(int kind) { return (kind == K_RES || kind == K_F) ? 2 : 1; }

// Arguments of the element-operator kernels. `cf` is the premultiplied
// coefficient field gJ*c (Helmholtz) or gJ*V (Schrodinger), device layout.
template <typename T>
struct OpArgs {
    const T* in;   // ncomp * len
    T* out;        // ncomp * len
    const T* geo;  // gA, gB, gC, gJ (stride len)
    const T* cf;
    T alpha;
    T post;        // scale applied to element-interior outputs (ring entries stay raw)
    size_t len;
    int nchunks;
};

// Arguments of the MINRES operator kernel (one Lanczos step, minres.hpp:48-55).
template <typename T>
struct MinresOpArgs {
    const T* r2;  // R[i-1]: v_i = ib * r2
    const T* r1;  // R[i-2]
    T* y;         // element results: final inside, ring copies unassembled (owner copy minus cr1*r1)
    const T* geo;
    const T* cf;
    T alpha;
    size_t len;
    int nchunks;
    const uint8_t* obits;  // ownership bits per element
    double* partial;       // per-workgroup alpha partial sums (out)
    // Krylov scalars: step 1 reads st_in as is; later steps apply the beta
    // recurrence of the previous step to st_in with the update kernel's
    // partials (prev_partial, n_prev) and workgroup 0 stores it in st_out
    const double* st_in;
    double* st_out;
    const double* prev_partial;
    int n_prev;
    int first;
};

// Arguments of the MINRES update kernel: interface assembly by gather,
// y -= c2*r2, beta^2 partial, and the pending w/x update of the previous step
// (w_new = ig*((ibp*r1 - eo*wa) - dl*wb); x += phi*w_new, minres.hpp:82-86).
template <typename T>
struct MinresUpdArgs {
    const T* yraw;
    T* y;         // R[i]
    const T* r2;  // R[i-1]
    const T* r1;  // R[i-2]
    const T* wa;  // W[i-3]
    const T* wb;  // W[i-2]
    T* wn;        // W[i-1]
    T* x;
    size_t len;
    int ncomp, nchunks;
    const uint8_t* obits;
    const int* eside;    // per element side: (neighbour << 3) | (neighbour side - 1) << 1 | reversed, or -1
    const int* ecorner;  // per element corner: interior corner index, or -1 on the Dirichlet boundary
    const int* c_off;
    const int* c_pos;
    double* partial;           // per-workgroup beta^2 partial sums (out)
    double* st;                // scalars after the previous step (workgroup 0 stores alpha here)
    const double* op_partial;  // the operator kernel's alpha partials
    int n_op;
};

// Device-resident Krylov scalars (FP64), written by single-block finalize kernels.
// MINRES recurrence state (minres.hpp:28-40 names); the Lanczos-step kernels
// copy and store only this part of KrylovScalars.
struct MinresState {
    double beta, beta_old, tnorm, dbar, eps, eps_old, phibar, cs, sn, b_norm, tol, atol;
    double alpha, ib, cr1;
    double p3_ibp, p3_ig, p3_eo, p3_dl, p3_phi;
    int p3_pending, iter, done, success, maxit, has_r1;
};

struct KrylovScalars : MinresState {
    // PCG (pcg.hpp:62-89)
    double rho, rho_prev, pcg_beta, pcg_alpha, bnorm, resid;
    int first;
    int pad_;
    // generic reduction results
    double red[8];
};

// Arguments of the PCG-fused operator kernel (pcg.hpp:70-80).
template <typename T>
struct PcgOpArgs {
    const T* r;
    T* p;
    T* ap;
    const T* geo;
    size_t len;
    double* partial;
    const double* sc;
};

// Launchers of the element-operator kernels (kernels_op.hip); d = D matrix
// (column-major, N*N values) copied into the kernel-argument segment.
template <typename T>
void launch_element_op_d(int N, int kind, const OpArgs<T>& a, const T* d, hipStream_t s);
template <typename T>
void launch_minres_op_d(int N, int kind, const MinresOpArgs<T>& a, const T* d, hipStream_t s);
template <typename T>
void launch_pcg_op_d(int N, const PcgOpArgs<T>& a, int nchunks, const T* d, hipStream_t s);
// beta recurrence closing a batch of Lanczos steps (reads *in, writes *out)
void launch_minres_close(int kind, const KrylovScalars* in, KrylovScalars* out, const double* partial, int n,
                         hipStream_t s);
// update kernel of a Lanczos step; returns the number of partials written
template <typename T>
int launch_minres_upd_n(int N, const MinresUpdArgs<T>& a, hipStream_t s);

}  // namespace sem
