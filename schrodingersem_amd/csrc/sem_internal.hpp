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
// Chunk counts are padded to a multiple of this many chunks, so that a
// workgroup of up to kChunkAlign chunks never straddles two instances of a
// batched mesh.
constexpr int kChunkAlign = 4;

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
void quad_metrics(const double* xs, const double* ys, const std::vector<double>& z, double* J, double* A,
                  double* B, double* C);
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
        nchunks = (nchunks + kChunkAlign - 1) / kChunkAlign * kChunkAlign;
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

// ------------------------------------------------------------ op params --
enum OpKind { K_LAPLACE = 0, K_HELMHOLTZ = 1, K_MASS = 2, K_RES = 3, K_F = 4, K_L = 5 };
inline int op_ncomp(int kind) { return (kind == K_RES || kind == K_F) ? 2 : 1; }

// Geometry of the FP32 operator applies: the metric terms of Quad.hpp:39-54
// are built from four functions of the GLL node that are linear in it,
//   X_xi(t) = ax + bx z_t, Y_xi(t) = ay + by z_t,
//   X_eta(t) = cx + dx z_t, Y_eta(t) = cy + dy z_t,
// so the FP32 kernels can form gA, gB, gC, gJ from 8 values per element (32
// bytes) instead of loading 4 N^2 stored values. That trades loads for about
// 20 FP32 operations per node, which pays where the kernels are bound by
// memory and occupancy and not by their arithmetic: the degrees for which
// map_geometry(N) holds (chosen by measurement). The others, and every FP64
// apply, load the stored factors.
struct QuadMap {
    const float4* map;  // per device element: (ax, bx, ay, by), (cx, dx, cy, dy); zero for padding and ghosts
    float z[kMaxN];     // GLL nodes
    float w[kMaxN];     // GLL weights
};

// the map of corners (xs, ys) (counter-clockwise, Quad.hpp's corner order)
void quad_map(const double* xs, const double* ys, float* out8);

// FP32 kernels of N nodes per direction form their geometry from the map
// when bit N of the mask is set. Measured on MI355X, fused Lanczos step of the
// scaled stage system at about 4.2 M nodes, stored factors vs map (us):
// N = 3: 78.7 / 71.1, 4: 60.8 / 52.9, 5: 62.9 / 55.5, 6: 64.2 / 58.5,
// 7: 61.9 / 61.6, 8: 51.3 / 48.0, 9: 59.5 / 61.5, 10: 76.4 / 88.0,
// 11: 93.0 / 104.4, 12: 112.5 / 108.8, 13: 129.5 / 142.2. So the map up to
// N = 8; above it the kernels are bound by arithmetic and registers.
#ifndef SEM_MAPGEO_MASK
#define SEM_MAPGEO_MASK 0x1FCu
#endif
constexpr bool map_geometry(int N) { return ((SEM_MAPGEO_MASK >> N) & 1u) != 0; }

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
    int mf;        // FP32 Res only: stiffness contractions on MFMA (0 VALU, 1 bf16, 3 split bf16)
    QuadMap qm;    // FP32 with map_geometry(N): the geometry
};

// Device-resident Krylov scalars (FP64), one record per instance.
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

// Instances of a batched mesh: instance k owns chunks [inst_off[k], inst_off[k+1])
// (multiples of kChunkAlign); chunk_info[c] = (instance, first chunk, end chunk, 0)
// of chunk c's instance. A plain mesh is one instance and skips the lookup.
struct InstTable {
    const int4* chunk_info;
    const int* inst_off;
    int n_inst;
    int nchunks;
};

__device__ __forceinline__ int4 chunk_instance(const InstTable& it, int chunk)
{
    return (it.n_inst == 1) ? make_int4(0, 0, it.nchunks, 0) : it.chunk_info[chunk];
}

// Arguments of the MINRES operator kernel (one Lanczos step, minres.hpp:48-55).
// Work items are groups of CPB chunks (template parameter of the kernel),
// visited grid-stride, so any grid size covers the mesh.
template <typename T>
struct MinresOpArgs {
    const T* r2;  // R[i-1]: v_i = ib * r2
    const T* r1;  // R[i-2]
    T* y;         // element results: final inside, ring copies unassembled (owner copy minus cr1*r1)
    // pending w/x update of the previous step (minres.hpp:82-86):
    // w_new = ig*((ibp*r1 - eo*wa) - dl*wb), x += phi*w_new
    const T* wa;  // W[i-3]
    const T* wb;  // W[i-2]
    T* wn;        // W[i-1]
    T* x;
    const T* geo;
    const T* cf;
    T alpha;
    size_t len;
    int nchunks;
    QuadMap qm;            // FP32 with map_geometry(N): the geometry
    const uint8_t* obits;  // ownership bits per element
    double* partial;       // per-chunk alpha partial sums (out)
    InstTable it;
    // Krylov scalars: step 1 of a batch reads st_in as is; later steps apply
    // the beta recurrence of the previous step to st_in with the update
    // kernel's per-chunk partials, and the workgroup holding an instance's
    // first chunk stores the result in st_out
    const KrylovScalars* st_in;
    KrylovScalars* st_out;
    const double* prev_partial;
    int first;
};

// Arguments of the MINRES update kernel: interface assembly by gather,
// y -= c2*r2 and the beta^2 partial.
template <typename T>
struct MinresUpdArgs {
    const T* yraw;
    T* y;         // R[i]
    const T* r2;  // R[i-1]
    size_t len;
    int ncomp, nchunks;
    const uint8_t* obits;
    const int* eside;    // per element side: (neighbour << 3) | (neighbour side - 1) << 1 | reversed, or -1
    // per element corner: positions of the corner node's copies in fold order
    // (-1 padded) when there are at most 4, {-2, corner index, 0, 0} when
    // there are more (then c_off/c_pos list them), all -1 on the Dirichlet boundary
    const int4* ecq;
    const int* c_off;
    const int* c_pos;
    double* partial;           // per-chunk beta^2 partial sums (out)
    InstTable it;
    KrylovScalars* st;         // scalars after the previous step (alpha is stored here)
    const double* op_partial;  // the operator kernel's per-chunk alpha partials
};

// Arguments of the fused Lanczos-step kernel (one launch per step, deferred
// scaling). Step i: finishes y_{i-1} from z_in = A y_{i-2} (raw, element-local;
// interface copies gathered), y_{i-2} and y_{i-3}; stores it; applies A to it
// (z_out, raw); writes per-chunk sums ||y_{i-1}||^2, y_{i-1}.z_out and
// y_{i-1}.y_{i-2}; applies the pending w/x update of iteration i-2 with
// v = y_{i-3}/beta_{i-2}. Step 1 reads y_0 = r0 from y_out and stores nothing.
template <typename T>
struct MinresFusedArgs {
    const T* z_in;
    T* z_out;
    const T* y2;  // y_{i-2}
    const T* y3;  // y_{i-3}
    T* y_out;     // y_{i-1}
    const T* wa;  // w_{i-4}
    const T* wb;  // w_{i-3}
    T* wn;        // w_{i-2}
    T* x;
    const T* geo;
    const T* cf;
    T alpha;
    size_t len;
    int nchunks;
    QuadMap qm;  // FP32 with map_geometry(N): the geometry
    const uint8_t* obits;
    const int* eside;
    const int4* ecq;
    const int* c_off;
    const int* c_pos;
    double* sums;              // 3 * nchunks per-chunk sums of this step (out)
    const double* prev_sums;   // the same of the previous step
    int ext;                   // step mode 2: a closing kernel evaluates the recurrence after every step
                               // and every launch starts from its state (first = 1)
    // Schrodinger stage (Res) only: diagonal scaling S of the system, one
    // value per node; the step applies S A S instead of A (nullptr: none)
    const float* dscale;
    InstTable it;
    const KrylovScalars* st_in;
    KrylovScalars* st_out;
    int step;   // i (1-based)
    int first;  // first launch of a batch: the state already holds this step's recurrence
    int mf;     // stiffness contractions on MFMA (0 VALU, 1 bf16, 3 split bf16)
    // stored-basis mode (non-null): every y_j stays in its own slot, the step
    // skips the w/x update and records iteration j's update coefficients
    // (1/beta_j, 1/gamma_j, eps_j, delta_j, phi_j) at rec[5 (j - 1)]; the
    // solution is formed from the stored vectors at the end
    double* rec;
};

// Work-item shape of the two Lanczos-step kernels: CPB chunks per workgroup
// and the grid size (0: one workgroup per item).
struct StepShape {
    int op_cpb = 1, op_grid = 0;
    int upd_cpb = 1, upd_grid = 0;
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
// Assembled, consistent form of an element operator's output (the DSS as a
// gather): every ring entry sums its node's copies in the reference's order
// (global_sum.hpp:28-34, :54-63) and is scaled by post, interior entries are
// copied, Dirichlet entries are zero. out must not alias in.
template <typename T>
struct GatherArgs {
    const T* in;
    T* out;
    T post;
    size_t len;
    int nchunks;
    int ncomp;
    const int* eside;
    const int4* ecq;
    const int* c_off;
    const int* c_pos;
};
template <typename T>
void launch_gather_assemble(int N, const GatherArgs<T>& a, hipStream_t s);
// e0/e1 (optional): events that take the kernel's own start and end timestamps
template <typename T>
void launch_minres_op_d(int N, int kind, const MinresOpArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                        hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
template <typename T>
void launch_pcg_op_d(int N, const PcgOpArgs<T>& a, int nchunks, const T* d, hipStream_t s);
// beta recurrence closing a batch of Lanczos steps, one workgroup per instance
// (reads in[k], writes out[k])
void launch_minres_close(int kind, const InstTable& it, const KrylovScalars* in, KrylovScalars* out,
                         const double* partial, hipStream_t s);
// element stiffness diagonal at every element-local node (FP64 geometry)
void launch_stiff_diag(int N, const double* geo, size_t len, int nchunks, double* out, const double* d, hipStream_t s);
// fused Lanczos step (one launch per step)
template <typename T>
void launch_minres_fused(int N, int kind, const MinresFusedArgs<T>& a, const T* d, int cpb, int grid, hipStream_t s,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// recurrence closing a batch of fused steps: what step `step` would compute
// at its start, one workgroup per instance (reads in[k], writes out[k])
void launch_minres_fused_close(const InstTable& it, const KrylovScalars* in, KrylovScalars* out, const double* sums,
                               int step, hipStream_t s, double* rec = nullptr);
// coefficients of the stored-basis MINRES solution from the records of
// iterations 1..k (rec[5 (j-1) ..]: 1/beta_j, 1/gamma_j, eps_j, delta_j, phi_j):
// t_j = phi_j - t_{j+1} delta_{j+1} / gamma_{j+1} - t_{j+2} eps_{j+2} / gamma_{j+2},
// coef[j-1] = t_j / (beta_j gamma_j); coef holds k + 2 doubles of workspace after the k results
void launch_basis_coef(const double* rec, int k, double* coef, hipStream_t s);
// x[i] (+)= sum_j coef[j] * basis[j * stride + i], i < n, the sum in FP64 in j order
template <typename T>
void launch_basis_combine(T* x, bool accumulate, const T* basis, size_t stride, size_t n, const double* coef, int k,
                          hipStream_t s);
// update kernel of a Lanczos step
template <typename T>
void launch_minres_upd(int N, const MinresUpdArgs<T>& a, int cpb, int grid, hipStream_t s, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr);

}  // namespace sem
