// engine.hip -- device mesh, Krylov drivers, PDE drivers and the C-ABI of
// libsem_hip.so. Host code: it sequences the kernels of kernels_op.hip and
// kernels_vec.hip on the handle's stream; all vector work stays on the GPU.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "sem_internal.hpp"
#include "sem_kernels.hpp"
#include "sem_mlp.hpp"

namespace sem {

static thread_local std::string g_error;
void set_error(const std::string& s) { g_error = s; }

// --------------------------------------------------------------- buffers --
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Results {
    int success = 0;
    int n_iter = 0;
    double residual = 0;
};

// Identity of a captured batch of Lanczos steps: every pointer and value the
// captured launches bake in besides the mesh's own arrays.
struct GraphKey {
    std::array<const void*, 13> ptr{};
    int kind = 0, tsize = 0, steps = 0;
    int variant = 0;  // 0 two-launch step; fused step: 1 + 2 (step mode 2) + 4 * contraction
    double alpha = 0;
    StepShape shape;
    bool operator==(const GraphKey& o) const
    {
        return ptr == o.ptr && kind == o.kind && tsize == o.tsize && steps == o.steps && variant == o.variant &&
               alpha == o.alpha &&
               shape.op_cpb == o.shape.op_cpb && shape.op_grid == o.shape.op_grid &&
               shape.upd_cpb == o.shape.upd_cpb && shape.upd_grid == o.shape.upd_grid;
    }
};

struct KrylovOpts {
    int maxit = 0;
    double tol = 0;
    double atol = 0;      // absolute term of the stop test (tol in the reference)
    bool rel_to_b = false; // stop relative to ||b|| instead of ||b - A x0||
    int hint_slot = -1;    // iteration-count history slot (default: by precision)
    // batched meshes: instances (nonzero entries) that take no part in this solve;
    // their Krylov state starts stopped, so x keeps its value there
    const uint8_t* frozen = nullptr;
    // batched meshes: per-instance relative tolerance and iteration limit
    // (replacing tol and maxit for instance k; maxit bounds the launches)
    const double* inst_tol = nullptr;
    const int* inst_maxit = nullptr;
};

// Workspace of sem_minres_noalloc: one array per GPU, held until the process
// ends. Host code must not pass `scratch` itself to a kernel (it names the
// host-side shadow); it uses the device address noalloc_scratch() returns.
constexpr long kNoAllocMaxN = 262144;                      // largest n of a solve
constexpr size_t kNoAllocDoubles = 8 * (size_t)kNoAllocMaxN + 1024 + 8;  // 8 vectors, partials, result
static __device__ __align__(256) double scratch[kNoAllocDoubles];

// marks instance k's Krylov state stopped before the first step (not converged)
inline void freeze_instance(KrylovScalars& h, const uint8_t* frozen, int k)
{
    if (frozen && frozen[k]) {
        h.done = 1;
        h.success = 0;
    }
}

}  // namespace sem

using namespace sem;

struct sem_mesh {
    HostMesh hm;
    Layout L;
    hipStream_t stream = nullptr;
    double* geo = nullptr;
    float* geo32 = nullptr;  // FP32 factors where !map_geometry(N)
    float* qmap = nullptr;   // QuadMap::map (8 floats per device element)
    QuadMap qm{};
    uint8_t* obits = nullptr;
    std::vector<uint8_t> h_obits;
    // DSS lists (device layout positions)
    int n_pairs = 0, n_corners = 0, n_bnd = 0;
    int2* pairs = nullptr;
    int* c_off = nullptr;
    int* c_pos = nullptr;
    int* bnd = nullptr;
    int dss_nb = 1;
    int* eside = nullptr;    // per element side: neighbour descriptor or -1
    int4* ecq = nullptr;     // per element corner: copy positions in fold order (see the gather kernels)
    // D matrix
    std::vector<double> D64;
    std::vector<float> D32;
    // scalars and partials
    // instances (a batched mesh holds several independent meshes of one
    // degree; a plain mesh is one instance): element and chunk ranges
    int n_inst = 1;
    std::vector<int> inst_el_off;        // first element of each instance (+ end); empty: one instance
    std::vector<int> inst_el_count;      // real elements (the rest up to the next offset are padding)
    std::vector<int> h_inst_off;         // chunk ranges
    int* d_inst_off = nullptr;
    int4* d_chunk_info = nullptr;
    std::vector<Results> inst_results;   // per-instance results of the last Krylov solve
    std::vector<uint8_t> ghost;          // per element: padding element of a batched mesh
    double last_true_rel = 0;            // ||b - A x|| / ||b|| after the last FP64-refined solve
    // scalars: one record per instance
    KrylovScalars* sc = nullptr;
    KrylovScalars* sc_alt = nullptr;  // second state buffer of the Lanczos-step kernels
    KrylovScalars* h_sc = nullptr;    // pinned
    StepShape shape[2];               // Lanczos-step work shape (FP64, FP32)
    // MINRES Lanczos step per precision (FP64, FP32): 0 two launches per step,
    // 1 one fused launch (not with graphs); FP32 defaults to fused (faster on
    // the DIRK4S3 workload), FP64 to two launches
    int step_mode[2] = {0, 1};
    // diagonal scaling of the FP32 inner Schrodinger-stage solves (on by
    // default); the assembled stiffness and mass diagonals are kept once
    // computed, the scale itself for the last (field, generation, alpha)
    int inner_scale = 1;
    // stiffness contractions of the FP32 inner solver's operator: 0 VALU FMAs,
    // 1 bf16 MFMA tiles, 3 split-bf16 MFMA tiles (FP32 accumulation)
    int contraction = 0;
    // FP32 one-launch MINRES keeps every Lanczos vector in HBM (stored basis,
    // no w/x streams per step) within basis_budget bytes; 0: the w/x recurrence
    int basis = 1;
    size_t basis_budget = (size_t)8 << 30;
    bool is_geo = false;
    const double* is_cf = nullptr;
    long is_gen = -1;
    double is_alpha = 0;
    long field_gen = 0;  // bumped whenever a coefficient field is (re)written
    // captured Lanczos-step batches (graph path), most recent last
    bool use_graphs = false;
    int graph_steps = 12;
    struct CachedGraph {
        GraphKey key;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<CachedGraph> graphs;
    hipGraphExec_t find_graph(const GraphKey& k) const
    {
        for (const auto& g : graphs)
            if (g.key == k) return g.exec;
        return nullptr;
    }
    void add_graph(const GraphKey& k, hipGraph_t g, hipGraphExec_t e)
    {
        if (graphs.size() >= 16) {  // keep the cache small: drop the oldest
            hipGraphExecDestroy(graphs.front().exec);
            hipGraphDestroy(graphs.front().graph);
            graphs.erase(graphs.begin());
        }
        graphs.push_back(CachedGraph{k, g, e});
    }
    void clear_graphs()
    {
        for (auto& g : graphs) {
            hipGraphExecDestroy(g.exec);
            hipGraphDestroy(g.graph);
        }
        graphs.clear();
    }
    double* partial = nullptr;
    int n_partial = 0;
    std::unordered_map<std::string, DevBuf> pool;
    sem_run_stats stats{};
    int batch = 8;          // Lanczos steps launched between host checks of the stop flag
    int iter_hint[8] = {};  // iterations of the last MINRES solve per hint slot
    // live timing of the two Lanczos-step kernels: 0 off, 1 one Lanczos step
    // in sixteen (sampled_step), 2 every step
    int timing = 0;
    std::vector<hipEvent_t> ev;
    // slots: 0/1 operator kernel FP64/FP32, 2/3 update kernel FP64/FP32
    long kt_launches[4] = {0, 0, 0, 0};
    double kt_us[4] = {0, 0, 0, 0};
    double kt_bytes[4] = {0, 0, 0, 0};

    ~sem_mesh()
    {
        if (stream) hipStreamSynchronize(stream);
        clear_graphs();
        for (hipEvent_t e : ev) hipEventDestroy(e);
        for (auto& kv : pool) hipFree(kv.second.p);
        hipFree(geo);
        hipFree(geo32);
        hipFree(qmap);
        hipFree(obits);
        hipFree(pairs);
        hipFree(c_off);
        hipFree(c_pos);
        hipFree(bnd);
        hipFree(eside);
        hipFree(ecq);
        hipFree(sc);
        hipFree(sc_alt);
        hipFree(d_inst_off);
        hipFree(d_chunk_info);
        hipFree(partial);
        if (h_sc) hipHostFree(h_sc);
        if (stream) hipStreamDestroy(stream);
    }

    template <typename T>
    T* buf(const std::string& name, size_t n)
    {
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        auto it = pool.find(name);
        if (it != pool.end() && it->second.bytes >= bytes) return (T*)it->second.p;
        if (it != pool.end()) {
            SEM_HIP(hipStreamSynchronize(stream));
            clear_graphs();  // captured launches may name the old buffer
            hipFree(it->second.p);
            pool.erase(it);
        }
        void* p = nullptr;
        SEM_HIP(hipMalloc(&p, bytes));
        pool[name] = DevBuf{p, bytes};
        return (T*)p;
    }

    size_t len() const { return L.len; }
    template <typename T>
    const T* geo_t() const;
    template <typename T>
    const T* dmat() const;
    void sync() { SEM_HIP(hipStreamSynchronize(stream)); }

    // reference layout host -> device layout buffer (ncomp components)
    void upload(double* dst, const double* src, int ncomp)
    {
        std::vector<double> tmp(L.len);
        const size_t rl = (size_t)L.E * L.N * L.N;
        for (int c = 0; c < ncomp; ++c) {
            to_device_layout(L, src + c * rl, tmp.data());
            SEM_HIP(hipMemcpyAsync(dst + c * L.len, tmp.data(), L.len * sizeof(double), hipMemcpyHostToDevice, stream));
            sync();
        }
    }
    void download(double* dst, const double* src, int ncomp)
    {
        std::vector<double> tmp(L.len);
        const size_t rl = (size_t)L.E * L.N * L.N;
        for (int c = 0; c < ncomp; ++c) {
            SEM_HIP(hipMemcpyAsync(tmp.data(), src + c * L.len, L.len * sizeof(double), hipMemcpyDeviceToHost, stream));
            sync();
            from_device_layout(L, tmp.data(), dst + c * rl);
        }
    }
    void read_scalars()
    {
        SEM_HIP(hipMemcpyAsync(h_sc, sc, n_inst * sizeof(KrylovScalars), hipMemcpyDeviceToHost, stream));
        sync();
    }
    InstTable it() const { return InstTable{d_chunk_info, d_inst_off, n_inst, L.nchunks}; }
};

template <>
const double* sem_mesh::geo_t<double>() const { return geo; }
template <>
const float* sem_mesh::geo_t<float>() const { return geo32; }  // nullptr where the FP32 kernels use qm
template <>
const double* sem_mesh::dmat<double>() const { return D64.data(); }
template <>
const float* sem_mesh::dmat<float>() const { return D32.data(); }

namespace sem {

// -------------------------------------------------------- mesh creation --
static void setup_device(sem_mesh* m)
{
    HostMesh& hm = m->hm;
    const int N = hm.N;
    m->L.init(N, hm.n_el);
    const Layout& L = m->L;
    SEM_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->D64 = hm.D;
    m->D32.assign(hm.D.begin(), hm.D.end());

    // premultiplied geometry
    const size_t rl = (size_t)hm.n_el * N * N;
    std::vector<double> g(4 * L.len, 0.0), tmp(rl);
    const std::vector<double>* src[4] = {&hm.A, &hm.B, &hm.C, &hm.J};
    for (int f = 0; f < 4; ++f) {
        for (int e = 0; e < hm.n_el; ++e)
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < N; ++i) {
                    size_t k = (size_t)e * N * N + i + (size_t)j * N;
                    tmp[k] = (hm.w[i] * hm.w[j]) * (*src[f])[k];
                    if (!m->ghost.empty() && m->ghost[e]) tmp[k] = 0.0;  // ghosts carry no operator
                }
        to_device_layout(L, tmp.data(), g.data() + f * L.len);
    }
    SEM_HIP(hipMalloc(&m->geo, g.size() * sizeof(double)));
    SEM_HIP(hipMemcpy(m->geo, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice));
    if (!map_geometry(N)) {  // the FP32 kernels of this degree load stored factors
        std::vector<float> g32(g.begin(), g.end());
        SEM_HIP(hipMalloc(&m->geo32, g32.size() * sizeof(float)));
        SEM_HIP(hipMemcpy(m->geo32, g32.data(), g32.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    // the FP32 applies' geometry where map_geometry(N): each element's
    // bilinear map (zero for padding elements and ghosts, whose factors are zero)
    {
        std::vector<float> qm((size_t)L.nchunks * L.K * 8, 0.0f);
        for (int e = 0; e < hm.n_el; ++e) {
            if (!m->ghost.empty() && m->ghost[e]) continue;
            double xs[4], ys[4];
            for (int k = 0; k < 4; ++k) {
                xs[k] = hm.node_xy[2 * (size_t)hm.quads[4 * (size_t)e + k]];
                ys[k] = hm.node_xy[2 * (size_t)hm.quads[4 * (size_t)e + k] + 1];
            }
            quad_map(xs, ys, qm.data() + 8 * (size_t)e);
        }
        SEM_HIP(hipMalloc(&m->qmap, qm.size() * sizeof(float)));
        SEM_HIP(hipMemcpy(m->qmap, qm.data(), qm.size() * sizeof(float), hipMemcpyHostToDevice));
        m->qm.map = reinterpret_cast<const float4*>(m->qmap);
        for (int k = 0; k < kMaxN; ++k) {
            m->qm.z[k] = k < N ? (float)hm.x[k] : 0.0f;
            m->qm.w[k] = k < N ? (float)hm.w[k] : 0.0f;
        }
    }

    // DSS lists
    auto side_ij = [&](int s, int k, int& i, int& j) {
        const int sm = (s == 1 || s == 4) ? 0 : N - 1;
        if (s == 2 || s == 4) { i = sm; j = k; }
        else { i = k; j = sm; }
    };
    static const int ci[4] = {0, 1, 1, 0}, cj[4] = {0, 0, 1, 1};
    auto corner_pos = [&](int code) {
        int e = code / 4, k = code % 4;
        return (int)L.pos(e, ci[k] * (N - 1), cj[k] * (N - 1));
    };
    if (4 * L.len > (size_t)INT32_MAX) throw std::runtime_error("mesh too large for 32-bit DSS indices");
    std::vector<int2> pairs;
    std::vector<int> bnd, c_off{0}, c_pos;
    m->h_obits.assign((size_t)L.nchunks * L.K, 0);
    for (const EdgeRec& r : hm.edges) {
        m->h_obits[r.el[0]] |= (uint8_t)(1u << (r.side[0] - 1));
        for (int k = 1; k <= N - 2; ++k) {
            int i0, j0;
            side_ij(r.side[0], k, i0, j0);
            int p0 = (int)L.pos(r.el[0], i0, j0);
            if (r.type == BOUNDARY) {
                bnd.push_back(p0);
            } else {
                int kk = r.start + (k - 1) * r.delta, i1, j1;
                side_ij(std::abs(r.side[1]), kk, i1, j1);
                pairs.push_back(make_int2(p0, (int)L.pos(r.el[1], i1, j1)));
            }
        }
    }
    for (int c = 0; c < hm.n_nodes; ++c) {
        const int b = hm.conn_off[c], e = hm.conn_off[c + 1];
        if (e > b) {
            int first = hm.conn[b];
            m->h_obits[first / 4] |= (uint8_t)(1u << (4 + first % 4));
        }
        if (hm.node_type[c] == BOUNDARY) {
            for (int q = b; q < e; ++q) bnd.push_back(corner_pos(hm.conn[q]));
        } else {
            for (int q = b; q < e; ++q) c_pos.push_back(corner_pos(hm.conn[q]));
            c_off.push_back((int)c_pos.size());
        }
    }
    // every DSS position must address the padded device arrays
    for (const int2& pr : pairs)
        if (pr.x < 0 || pr.y < 0 || (size_t)pr.x >= L.len || (size_t)pr.y >= L.len)
            throw std::runtime_error("internal error: DSS pair position out of range");
    for (int v : bnd)
        if (v < 0 || (size_t)v >= L.len) throw std::runtime_error("internal error: Dirichlet position out of range");
    for (int v : c_pos)
        if (v < 0 || (size_t)v >= L.len) throw std::runtime_error("internal error: corner position out of range");
    // gather descriptors of the MINRES update kernel
    std::vector<int> eside((size_t)L.nchunks * L.K * 4, -1), ecorner((size_t)L.nchunks * L.K * 4, -1);
    for (const EdgeRec& r : hm.edges) {
        if (r.type == BOUNDARY) continue;
        const int s0 = r.side[0], s1 = std::abs(r.side[1]), rev = r.side[1] < 0 ? 1 : 0;
        eside[4 * (size_t)r.el[0] + s0 - 1] = (r.el[1] << 3) | ((s1 - 1) << 1) | rev;
        eside[4 * (size_t)r.el[1] + s1 - 1] = (r.el[0] << 3) | ((s0 - 1) << 1) | rev;
    }
    {
        int ci = 0;
        for (int c = 0; c < hm.n_nodes; ++c) {
            if (hm.node_type[c] == BOUNDARY) continue;
            for (int q = hm.conn_off[c]; q < hm.conn_off[c + 1]; ++q) ecorner[hm.conn[q]] = ci;
            ++ci;
        }
    }
    if ((size_t)hm.n_el >= (1u << 27)) throw std::runtime_error("mesh too large for packed side descriptors");
    {
        // the gather descriptors must name exactly the partners of the DSS pair lists
        std::unordered_map<int, int> partner;
        partner.reserve(2 * pairs.size());
        for (const int2& pr : pairs) {
            partner[pr.x] = pr.y;
            partner[pr.y] = pr.x;
        }
        size_t seen = 0;
        for (int e = 0; e < hm.n_el; ++e)
            for (int sd = 1; sd <= 4; ++sd) {
                const int d = eside[4 * (size_t)e + sd - 1];
                if (d < 0) continue;
                const int e2 = d >> 3, s2 = ((d >> 1) & 3) + 1;
                for (int k = 1; k <= N - 2; ++k) {
                    const int k2 = (d & 1) ? N - 1 - k : k;
                    int i0, j0, i1, j1;
                    side_ij(sd, k, i0, j0);
                    side_ij(s2, k2, i1, j1);
                    auto it = partner.find((int)L.pos(e, i0, j0));
                    if (it == partner.end() || it->second != (int)L.pos(e2, i1, j1))
                        throw std::runtime_error("internal error: side descriptor disagrees with the DSS pairs");
                    ++seen;
                }
            }
        if (seen != 2 * pairs.size()) throw std::runtime_error("internal error: side descriptors miss DSS pairs");
        for (int e = 0; e < hm.n_el; ++e)
            for (int k = 0; k < 4; ++k) {
                const int ci = ecorner[4 * (size_t)e + k];
                if (ci < -1 || ci >= (int)c_off.size() - 1) throw std::runtime_error("internal error: corner index");
                if (ci >= 0) {
                    bool found = false;
                    for (int q = c_off[ci]; q < c_off[ci + 1]; ++q) found |= c_pos[q] == corner_pos(4 * e + k);
                    if (!found) throw std::runtime_error("internal error: corner descriptor");
                }
            }
    }
    // per element corner: the positions of every copy of the corner node in
    // fold order when there are at most 4 (-1 padded), {-2, corner index, 0, 0}
    // when there are more, all -1 on the Dirichlet boundary
    std::vector<int4> ecq(ecorner.size(), make_int4(-1, -1, -1, -1));
    for (size_t q = 0; q < ecorner.size(); ++q) {
        const int ci = ecorner[q];
        if (ci < 0) continue;
        const int b = c_off[ci], n = c_off[ci + 1] - b;
        if (n > 4) {
            ecq[q] = make_int4(-2, ci, 0, 0);
        } else {
            int v[4] = {-1, -1, -1, -1};
            for (int t = 0; t < n; ++t) v[t] = c_pos[b + t];
            ecq[q] = make_int4(v[0], v[1], v[2], v[3]);
        }
    }
    m->n_pairs = (int)pairs.size();
    m->n_corners = (int)c_off.size() - 1;
    m->n_bnd = (int)bnd.size();
    auto up = [&](auto*& dst, const auto& v) {
        using E = typename std::decay_t<decltype(v)>::value_type;
        SEM_HIP(hipMalloc(&dst, std::max<size_t>(v.size(), 1) * sizeof(E)));
        if (!v.empty()) SEM_HIP(hipMemcpy(dst, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice));
    };
    up(m->pairs, pairs);
    up(m->c_off, c_off);
    up(m->c_pos, c_pos);
    up(m->bnd, bnd);
    up(m->obits, m->h_obits);
    up(m->eside, eside);
    up(m->ecq, ecq);
    const long total = (long)m->n_pairs + m->n_corners + m->n_bnd;
    m->dss_nb = (int)std::max<long>(1, (total + 255) / 256);

    // instance table: chunk ranges from the element ranges (each instance
    // starts on a kChunkAlign-chunk boundary)
    if (m->inst_el_off.empty()) {  // a plain mesh is one instance (a batch sets its own tables)
        m->n_inst = 1;
        m->inst_el_off = {0, hm.n_el};
        m->inst_el_count = {hm.n_el};
    }
    if ((int)m->inst_el_off.size() != m->n_inst + 1 || (int)m->inst_el_count.size() != m->n_inst ||
        m->inst_el_off.back() != hm.n_el)
        throw std::runtime_error("internal error: instance table does not cover the mesh");
    m->h_inst_off.assign(m->n_inst + 1, 0);
    std::vector<int4> chunk_info(L.nchunks);
    for (int k = 0; k < m->n_inst; ++k) {
        const int e0 = m->inst_el_off[k];
        if (e0 % (L.K * kChunkAlign) != 0) throw std::runtime_error("internal error: instance not chunk aligned");
        m->h_inst_off[k] = e0 / L.K;
    }
    m->h_inst_off[m->n_inst] = L.nchunks;
    for (int k = 0; k < m->n_inst; ++k) {
        if (m->h_inst_off[k + 1] <= m->h_inst_off[k]) throw std::runtime_error("internal error: empty instance");
        for (int c = m->h_inst_off[k]; c < m->h_inst_off[k + 1]; ++c)
            chunk_info[c] = make_int4(k, m->h_inst_off[k], m->h_inst_off[k + 1], 0);
    }
    up(m->d_inst_off, m->h_inst_off);
    up(m->d_chunk_info, chunk_info);
    m->inst_results.assign(m->n_inst, Results{});
    // default FP32 step: the recurrence's reduction in its own kernel once an
    // instance has many chunks (every workgroup re-reading all the per-chunk
    // sums then costs more than a launch: 136 -> 97 us per step at 8192 chunks,
    // 37.5 -> 32.5 us at 2344; at 128 chunks the extra launch is 7% slower)
    int max_chunks = 0;
    for (int k = 0; k < m->n_inst; ++k) max_chunks = std::max(max_chunks, m->h_inst_off[k + 1] - m->h_inst_off[k]);
    m->step_mode[1] = max_chunks >= 1024 ? 2 : 1;

    const size_t scb = (size_t)m->n_inst * sizeof(KrylovScalars);
    SEM_HIP(hipMalloc(&m->sc, scb));
    SEM_HIP(hipMemset(m->sc, 0, scb));
    SEM_HIP(hipMalloc(&m->sc_alt, scb));
    SEM_HIP(hipMemset(m->sc_alt, 0, scb));
    SEM_HIP(hipHostMalloc(&m->h_sc, scb, hipHostMallocDefault));
    m->n_partial = 3 * L.nchunks + m->dss_nb + 1024;
    SEM_HIP(hipMalloc(&m->partial, (size_t)m->n_partial * sizeof(double)));
}

template <typename T>
static DssArgs<T> dss_args(sem_mesh* m, T* y, int ncomp)
{
    DssArgs<T> d{};
    d.y = y;
    d.len = m->len();
    d.ncomp = ncomp;
    d.pairs = m->pairs;
    d.n_pairs = m->n_pairs;
    d.c_off = m->c_off;
    d.c_pos = m->c_pos;
    d.n_corners = m->n_corners;
    d.bnd = m->bnd;
    d.n_bnd = m->n_bnd;
    d.post[0] = d.post[1] = 1;
    d.sc = reinterpret_cast<const double*>(m->sc);
    d.partial = m->partial;
    return d;
}

// element op (raw ring entries) + DSS with Dirichlet zeroing: the consistent
// form of mask(global_sum(op(unmask(in)))).
template <typename T>
static void apply_op(sem_mesh* m, int kind, const T* in, T* out, const T* cf, double alpha, double post = 1.0)
{
    OpArgs<T> a{};
    a.in = in;
    a.out = out;
    a.geo = m->geo_t<T>();
    a.qm = m->qm;
    a.cf = cf;
    a.alpha = (T)alpha;
    a.post = (T)post;
    a.len = m->len();
    a.nchunks = m->L.nchunks;
    const int nc = op_ncomp(kind);
    if (post > 0) {
        // the element operator into scratch, then the DSS as one gather per
        // column into out (every entry written once, coalesced); the same bits
        // as the pair/corner DSS below for a positive post
        T* raw = m->buf<T>(sizeof(T) == 8 ? "ap_raw_d" : "ap_raw_f", (size_t)nc * m->len());
        a.out = raw;
        launch_element_op_d<T>(m->L.N, kind, a, m->dmat<T>(), m->stream);
        GatherArgs<T> g{};
        g.in = raw;
        g.out = out;
        g.post = (T)post;
        g.len = m->len();
        g.nchunks = m->L.nchunks;
        g.ncomp = nc;
        g.eside = m->eside;
        g.ecq = m->ecq;
        g.c_off = m->c_off;
        g.c_pos = m->c_pos;
        launch_gather_assemble<T>(m->L.N, g, m->stream);
        return;
    }
    launch_element_op_d<T>(m->L.N, kind, a, m->dmat<T>(), m->stream);
    DssArgs<T> d = dss_args<T>(m, out, nc);
    d.post[0] = d.post[1] = (T)post;
    launch_dss<T>(DSS_SUM_ZB, d, m->stream);
}

template <typename T>
static void dss(sem_mesh* m, int mode, T* y, int ncomp)
{
    DssArgs<T> d = dss_args<T>(m, y, ncomp);
    launch_dss<T>(mode, d, m->stream);
}

template <typename T>
static void lincomb(sem_mesh* m, T* out, int mode, std::initializer_list<std::pair<double, const T*>> terms, size_t n)
{
    LinComb<T> a{};
    a.out = out;
    a.mode = mode;
    a.n = n;
    int t = 0;
    for (auto& pr : terms) {
        a.c[t] = pr.first;
        a.x[t] = pr.second;
        ++t;
    }
    a.m = t;
    launch_lincomb<T>(a, m->stream);
}

// owner-weighted dot of every instance into sc[k].red[slot] (device)
template <typename T>
static void wdot_dev(sem_mesh* m, const T* a, const T* b, int ncomp, int slot)
{
    launch_wdot<T>(a, b, m->len(), ncomp, m->L, m->obits, m->partial, m->stream);
    launch_fin_sum_inst(m->partial, m->it(), m->sc, slot, m->stream);
}

// <a,a>, <b,b>, <a,b> of every instance into sc[k].red[slot0 .. slot0 + 2] (device), one pass
static void wdot3_dev(sem_mesh* m, const double* a, const double* b, int ncomp, int slot0)
{
    launch_wdot3<double>(a, b, m->len(), ncomp, m->L, m->obits, m->partial, m->stream);
    launch_fin_sum3_inst(m->partial, m->it(), m->sc, m->L.nchunks, slot0, m->stream);
}

template <typename T>
static double wdot_host(sem_mesh* m, const T* a, const T* b, int ncomp)
{
    wdot_dev<T>(m, a, b, ncomp, 7);
    m->read_scalars();
    return m->h_sc->red[7];
}

// the owner-weighted dot of every instance (one entry per instance)
template <typename T>
static std::vector<double> wdot_inst(sem_mesh* m, const T* a, const T* b, int ncomp)
{
    wdot_dev<T>(m, a, b, ncomp, 7);
    m->read_scalars();
    std::vector<double> v(m->n_inst);
    for (int k = 0; k < m->n_inst; ++k) v[k] = m->h_sc[k].red[7];
    return v;
}

// Krylov options that differ from launch_minres_init's defaults, per instance:
// the stop test phibar < tol * b_ref + atol (minres.hpp:92 has b_ref = ||r0||,
// atol = tol), b_ref = ||b|| (red[1]) when rel_to_b, instances frozen out of
// the solve, per-instance tolerances and iteration limits
static void patch_instances(sem_mesh* m, const KrylovOpts& o)
{
    if (!(o.rel_to_b || o.atol != o.tol || o.frozen || o.inst_tol || o.inst_maxit)) return;
    m->read_scalars();
    for (int k = 0; k < m->n_inst; ++k) {
        KrylovScalars& h = m->h_sc[k];
        if (o.rel_to_b) h.b_norm = std::sqrt(h.red[1]);
        h.atol = o.atol;
        if (o.inst_tol) h.tol = o.inst_tol[k];
        if (o.inst_maxit) {
            h.maxit = o.inst_maxit[k];
            h.done = h.maxit < 1 ? 1 : 0;
        }
        freeze_instance(h, o.frozen, k);
    }
    SEM_HIP(hipMemcpyAsync(m->sc, m->h_sc, m->n_inst * sizeof(KrylovScalars), hipMemcpyHostToDevice, m->stream));
}

// ---------------------------------------------------------- fused MINRES --
// minres.hpp:19-107 with one launch per Lanczos step (deferred scaling, see
// MinresFusedArgs): step i finishes y_{i-1} from A y_{i-2}, applies A to it and
// writes the sums its successor needs; the Givens update of iteration i-2 and
// alpha_{i-1} are evaluated at the start of step i in every workgroup. Same
// recurrence as the two-launch path; the operator is applied before the 1/beta
// scaling instead of after, so results agree to rounding. Buffers: y_k in
// Y[k % 3], A y in Z[step % 2], w_k in W[k % 3], sums in S[step % 2].
// Kernel timing mode 1 times one Lanczos step in sixteen, picked by a hash of
// the step index so that the sample does not follow the batch structure
// (the first launch after a host check starts on an idle GPU)
// Upper bound of the first batch of Lanczos steps launched before a host
// check of the stop flags (and of the timing events kept per batch).
constexpr int kMaxFirstBatch = 2048;
constexpr int kBasisMax = 4096;  // most Lanczos vectors a stored-basis MINRES solve keeps

static bool sampled_step(int i) { return (((unsigned)i * 2654435761u) >> 20 & 15u) == 0; }

// the one-launch Lanczos step runs for precision index ti (0 FP64, 1 FP32)
static bool fused_step(const sem_mesh* m, int ti) { return m->step_mode[ti] >= 1; }

template <typename T>
static Results minres_fused1(sem_mesh* m, int kind, const T* cf, double alpha, const T* b, T* x, bool x_zero,
                             const KrylovOpts& o, const float* dscale = nullptr)
{
    const int nc = op_ncomp(kind);
    const size_t n = (size_t)nc * m->len();
    const std::string tp = sizeof(T) == 8 ? "d" : "f";
    // the Y and W slots are the two-launch path's R and W buffers
    T* Y[3] = {m->buf<T>("mr_R0" + tp, n), m->buf<T>("mr_R1" + tp, n), m->buf<T>("mr_R2" + tp, n)};
    T* W[3] = {m->buf<T>("mr_W0" + tp, n), m->buf<T>("mr_W1" + tp, n), m->buf<T>("mr_W2" + tp, n)};
    T* Z[2] = {m->buf<T>("mr_Yraw" + tp, n), m->buf<T>("mr_Z1" + tp, n)};
    const int nch = m->L.nchunks;
    double* S[2] = {m->buf<double>("mr_S0", 3 * (size_t)nch), m->buf<double>("mr_S1", 3 * (size_t)nch)};
    hipStream_t s = m->stream;
    const int ni = m->n_inst;
    const bool graphs = m->use_graphs && !m->timing;
    // Stored basis (FP32, one instance, no graph replay): every Lanczos vector
    // y_j keeps its own slot of HBM, the steps skip the w/x update (five of
    // the step's vector streams per component) and record each iteration's
    // update coefficients; x = x0 + sum_j c_j y_{j-1} is formed once at the
    // end (minres.hpp:82-86 unrolled: W R = V, x = x0 + V R^-1 t). The basis
    // holds at most basis_budget bytes; a solve that reaches its capacity
    // stops there unconverged, as at maxit.
    int cap = 0;
    T* Bv = nullptr;
    double* rec = nullptr;
    if (sizeof(T) == 4 && m->basis && ni == 1 && !graphs) {
        const size_t fit = m->basis_budget / (n * sizeof(T));
        cap = (int)std::min<size_t>({(size_t)o.maxit + 2, (size_t)kBasisMax, fit});
        if (cap >= 16) {
            Bv = m->buf<T>("mr_basis", (size_t)cap * n);
            rec = m->buf<double>("mr_rec", 5 * (size_t)cap);
        } else {
            cap = 0;
        }
    }
    const int maxit = cap ? std::min(o.maxit, cap - 2) : o.maxit;
    T* y0 = cap ? Bv : Y[0];
    if (x_zero) {
        if (!cap) launch_fill<T>(x, T(0), n, s);
        launch_cast<T, T>(y0, b, n, s);
    } else {
        T* Ax = m->buf<T>("mr_Ax" + tp, n);
        apply_op<T>(m, kind, x, Ax, cf, alpha);
        lincomb<T>(m, y0, 0, {{1.0, b}, {-1.0, Ax}}, n);  // y_0 = r1 = b - A(x)
    }
    if (!cap)
        for (int q = 0; q < 3; ++q) launch_fill<T>(W[q], T(0), n, s);
    wdot_dev<T>(m, y0, y0, nc, 0);
    if (o.rel_to_b) wdot_dev<T>(m, b, b, nc, 1);
    // the stop test's atol and reference norm go in with the start; only
    // per-instance options need a host round trip
    launch_minres_init(m->sc, ni, 0, maxit, o.tol, s, o.atol, o.rel_to_b ? 1 : -1);
    if (o.frozen || o.inst_tol || o.inst_maxit) patch_instances(m, o);
    const int ti = sizeof(T) == 8 ? 0 : 1;
    const StepShape shp = m->shape[ti];
    const int slot = o.hint_slot >= 0 ? std::min(o.hint_slot, 7) : ti;
    // first batch: the previous solve's count, plus the two launches by which
    // the stop test trails (iteration k is tested at the start of step k + 2),
    // plus a margin of two: launches past an instance's stop exit after
    // reading its state, which is cheaper than another host check
    // Graph mode: after the first (direct) batch, batches of G steps replay as
    // one captured graph. The step kernel's buffers rotate with period 6 and its
    // behaviour depends on the step index only below step 4, so a batch of G
    // (a multiple of 6) steps captured from a step i = 1 (mod 6), i >= 7,
    // replays every later such batch exactly; the first batch is rounded up to
    // a multiple of 6 to start the replays there
    const int G = (m->graph_steps + 5) / 6 * 6;
    int first_batch = std::max(m->batch, std::min(m->iter_hint[slot] + 4, kMaxFirstBatch));
    if (graphs) first_batch = (first_batch + 5) / 6 * 6;
    if (m->timing) {
        while ((int)m->ev.size() < 4 * first_batch) {
            hipEvent_t e;
            SEM_HIP(hipEventCreate(&e));
            m->ev.push_back(e);
        }
        // algorithmic bytes of one launch: finishing y (z in, y_{i-2}, y_{i-3} in,
        // y out), the operator apply (its inputs and outputs) and the w/x update,
        // per component, plus the diagonal scale when there is one; the
        // geometry is 4 stored factors per node in FP64 and each element's
        // 32-byte bilinear map in FP32
        const double EN2 = (double)m->L.E * m->L.N * m->L.N;
        const double cfv = (kind == K_RES || kind == K_HELMHOLTZ) ? 1 : 0;
        const double geo = (sizeof(T) == 4 && map_geometry(m->L.N)) ? 32.0 * m->L.E
                                                                    : (kind == K_RES ? 4.0 : 3.0) * sizeof(T) * EN2;
        m->kt_bytes[ti] = (cfv + (cap ? 5.0 : 10.0) * nc) * sizeof(T) * EN2 + geo + (dscale ? sizeof(float) * EN2 : 0.0);
    }
    KrylovScalars* buf[2] = {m->sc, m->sc_alt};
    const InstTable itab = m->it();
    // step mode 2: the recurrence reduction of every step runs once per
    // instance in its own small kernel (the batch-closing kernel) after the
    // step, and every step starts from the finished state (no per-workgroup
    // reduction of all the step's per-chunk sums; same order, same bits)
    const bool ext = m->step_mode[sizeof(T) == 8 ? 0 : 1] == 2;
    auto max_iter = [&]() {
        int v = 0;
        for (int k = 0; k < ni; ++k) v = std::max(v, m->h_sc[k].iter);
        return v;
    };
    auto all_done = [&]() {
        for (int k = 0; k < ni; ++k)
            if (!m->h_sc[k].done) return false;
        return true;
    };
    auto all_success = [&]() {
        for (int k = 0; k < ni; ++k)
            if (!m->h_sc[k].success) return false;
        return true;
    };
    // step i of a batch (b_: its position in the batch, cur: the state buffer
    // it reads, updated)
    auto launch_step = [&](int i, int b_, int& cur, bool tim) {
        MinresFusedArgs<T> a{};
        a.z_in = Z[(i + 1) % 2];
        a.z_out = Z[i % 2];
        a.y2 = Y[(i + 1) % 3];
        a.y3 = Y[i % 3];
        a.y_out = Y[(i + 2) % 3];
        a.wa = W[(i + 2) % 3];
        a.wb = W[i % 3];
        a.wn = W[(i + 1) % 3];
        a.x = x;
        if (cap) {  // y_j in slot j
            a.y_out = Bv + (size_t)(i - 1) * n;
            a.y2 = Bv + (size_t)std::max(i - 2, 0) * n;
            a.y3 = Bv + (size_t)std::max(i - 3, 0) * n;
            a.rec = rec;
        }
        a.geo = m->geo_t<T>();
        a.qm = m->qm;
        a.cf = cf;
        a.alpha = (T)alpha;
        a.len = m->len();
        a.nchunks = nch;
        a.obits = m->obits;
        a.eside = m->eside;
        a.ecq = m->ecq;
        a.c_off = m->c_off;
        a.c_pos = m->c_pos;
        a.sums = S[i % 2];
        a.prev_sums = S[(i + 1) % 2];
        a.it = itab;
        if (ext) {
            a.st_in = m->sc;
            a.st_out = m->sc_alt;
        } else {
            a.st_in = buf[cur];
            cur ^= 1;
            a.st_out = buf[cur];
        }
        a.step = i;
        a.first = ext || (b_ == 0);
        a.ext = ext ? 1 : 0;
        a.mf = sizeof(T) == 4 ? m->contraction : 0;
        a.dscale = dscale;
        launch_minres_fused<T>(m->L.N, kind, a, m->dmat<T>(), shp.op_cpb, shp.op_grid, s,
                               tim ? m->ev[4 * b_] : nullptr, tim ? m->ev[4 * b_ + 1] : nullptr);
        if (ext) launch_minres_fused_close(itab, m->sc_alt, m->sc, S[i % 2], i + 1, s, rec);
    };
    int i = 1;
    bool done = o.maxit < 1;
    while (!done) {
        if (graphs && i > 1) {
            GraphKey key{};
            key.ptr = {(const void*)Y[0], (const void*)Y[1], (const void*)Y[2], (const void*)W[0], (const void*)W[1],
                       (const void*)W[2], (const void*)Z[0], (const void*)Z[1], (const void*)x, (const void*)cf,
                       (const void*)dscale, (const void*)S[0], (const void*)S[1]};
            key.kind = kind;
            key.tsize = (int)sizeof(T);
            key.alpha = alpha;
            key.steps = G;
            key.shape = shp;
            key.variant = 1 + (ext ? 2 : 0) + 4 * (sizeof(T) == 4 ? m->contraction : 0);
            hipGraphExec_t ex = m->find_graph(key);
            if (!ex) {
                // captured from step 7 (= 1 mod 6); replayed from i
                hipGraph_t gr = nullptr;
                SEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int cur = 0;
                for (int b_ = 0; b_ < G; ++b_) launch_step(7 + b_, b_, cur, false);
                if (!ext) launch_minres_fused_close(itab, buf[cur], m->sc, S[(7 + G + 1) % 2], 7 + G, s);
                SEM_HIP(hipStreamEndCapture(s, &gr));
                SEM_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
                m->add_graph(key, gr, ex);
            }
            SEM_HIP(hipGraphLaunch(ex, s));
            i += G;
            m->read_scalars();
            done = all_done() || i > o.maxit + 2;
            continue;
        }
        int launched = 0;
        const int i0 = i;
        int cur = 0;
        const int nbatch = (i == 1) ? first_batch : m->batch;
        for (int b_ = 0; b_ < nbatch && i <= maxit + 2; ++b_, ++i) {
            const bool tim = m->timing == 2 || (m->timing == 1 && sampled_step(i));
            launch_step(i, b_, cur, tim);
            ++launched;
        }
        // the batch ended after step i-1; close with what step i would compute first
        if (!ext) launch_minres_fused_close(itab, buf[cur], m->sc, S[(i + 1) % 2], i, s, rec);
        m->read_scalars();
        if (m->timing) {
            const int last = max_iter() + 2;
            for (int b_ = 0; b_ < launched && i0 + b_ <= last; ++b_) {
                if (m->timing == 1 && !sampled_step(i0 + b_)) continue;
                float ms = 0;
                SEM_HIP(hipEventElapsedTime(&ms, m->ev[4 * b_], m->ev[4 * b_ + 1]));
                m->kt_us[ti] += 1000.0 * ms;
                m->kt_launches[ti] += 1;
            }
        }
        done = all_done() || i > maxit + 2;
    }
    bool pending = false;
    for (int k = 0; k < ni; ++k) pending |= m->h_sc[k].p3_pending != 0;
    if (cap) {
        // x = x0 + sum_j c_j y_{j-1}: with the records of iterations 1..k, the
        // coefficient of w_j in x is t_j = phi_j - t_{j+1} delta_{j+1} / gamma_{j+1}
        // - t_{j+2} eps_{j+2} / gamma_{j+2}, and w_j holds y_{j-1} / (beta_j gamma_j)
        // (on the device: no host round trip)
        const int k = m->h_sc[0].iter;
        if (k > 0) {
            double* dc = m->buf<double>("mr_coef", 2 * (size_t)cap + 2);
            launch_basis_coef(rec, k, dc, s);
            launch_basis_combine<T>(x, !x_zero, Bv, n, n, dc, k, s);
        } else if (x_zero) {
            launch_fill<T>(x, T(0), n, s);
        }
    } else if (pending) {
        // iteration k = iter of each instance: v = y_{k-1}, w_{k-2}, w_{k-1} -> w_k;
        // the flush kernel's rotation (R[(k+2)%3], W[(k+1)%3], W[(k+2)%3] -> W[k%3]) matches
        MinresFlushArgs<T> fa{};
        for (int q = 0; q < 3; ++q) {
            fa.R[q] = Y[q];
            fa.W[q] = W[q];
        }
        fa.x = x;
        fa.len = m->len();
        fa.chunk_len = m->L.K * m->L.N * m->L.N;
        fa.ncomp = nc;
        launch_minres_flush<T>(fa, itab, nch, m->sc, s);
    }
    for (int k = 0; k < ni; ++k) {
        Results& r = m->inst_results[k];
        const KrylovScalars& h = m->h_sc[k];
        const int mk = o.inst_maxit ? o.inst_maxit[k] : maxit;
        r.success = h.success;
        r.n_iter = r.success ? h.iter : mk + 1;
        if (mk < 1) r.n_iter = 1;
        r.residual = h.phibar;
    }
    // the next solve's first batch follows this one's count, if it converged
    // (a solve that ran to maxit says nothing about the next one)
    if (all_success()) m->iter_hint[slot] = max_iter();
    Results all = m->inst_results[0];
    for (int k = 1; k < ni; ++k) {
        all.success = all.success && m->inst_results[k].success;
        all.n_iter = std::max(all.n_iter, m->inst_results[k].n_iter);
        all.residual = std::max(all.residual, m->inst_results[k].residual);
    }
    return all;
}

// ---------------------------------------------------------------- MINRES --
// minres.hpp:19-107 with IdentityPreconditioner, two fused launches per
// Lanczos step (the scalar recurrence runs inside them). Vectors are T (FP64 or
// FP32); every reduction and every recurrence scalar is FP64. Each instance of
// a batched mesh runs its own recurrence and stop test; steps continue until
// every instance has stopped, and the kernels skip stopped instances. Returns
// the aggregate over instances (all succeeded, longest iteration count, largest
// residual); m->inst_results holds each instance's result.
template <typename T>
static Results minres_fused(sem_mesh* m, int kind, const T* cf, double alpha, const T* b, T* x, bool x_zero,
                            const KrylovOpts& o, const float* dscale = nullptr)
{
    // a scaled system (S A S, Res only) runs on the one-launch step
    if (dscale && (kind != K_RES || !fused_step(m, sizeof(T) == 8 ? 0 : 1)))
        throw std::runtime_error("internal error: diagonal scaling needs the one-launch Res step");
    if (fused_step(m, sizeof(T) == 8 ? 0 : 1)) return minres_fused1<T>(m, kind, cf, alpha, b, x, x_zero, o, dscale);
    const int nc = op_ncomp(kind);
    const size_t n = (size_t)nc * m->len();
    const std::string tp = sizeof(T) == 8 ? "d" : "f";
    T* R[3] = {m->buf<T>("mr_R0" + tp, n), m->buf<T>("mr_R1" + tp, n), m->buf<T>("mr_R2" + tp, n)};
    T* W[3] = {m->buf<T>("mr_W0" + tp, n), m->buf<T>("mr_W1" + tp, n), m->buf<T>("mr_W2" + tp, n)};
    T* Yraw = m->buf<T>("mr_Yraw" + tp, n);
    hipStream_t s = m->stream;
    const int ni = m->n_inst;
    if (x_zero) {
        launch_fill<T>(x, T(0), n, s);
        launch_cast<T, T>(R[0], b, n, s);
    } else {
        T* Ax = m->buf<T>("mr_Ax" + tp, n);
        apply_op<T>(m, kind, x, Ax, cf, alpha);
        lincomb<T>(m, R[0], 0, {{1.0, b}, {-1.0, Ax}}, n);  // r1 = b - A(x)
    }
    launch_fill<T>(W[0], T(0), n, s);
    launch_fill<T>(W[2], T(0), n, s);
    wdot_dev<T>(m, R[0], R[0], nc, 0);
    launch_minres_init(m->sc, ni, 0, o.maxit, o.tol, s);
    if (o.rel_to_b) {  // b_norm = ||b|| (fair stop test for nonzero guesses)
        wdot_dev<T>(m, b, b, nc, 1);
    }
    patch_instances(m, o);
    const int nch = m->L.nchunks;
    const int ti = sizeof(T) == 8 ? 0 : 1;
    const StepShape shp = m->shape[ti];
    // The first batch runs most of the expected iteration count without a
    // host round trip (the previous solve of this precision predicts it);
    // steps launched after the stop flag is set exit at once.
    const int slot = o.hint_slot >= 0 ? std::min(o.hint_slot, 7) : ti;
    const int first_batch = std::max(m->batch, std::min((int)(0.85 * m->iter_hint[slot]), kMaxFirstBatch));
    if (m->timing) {
        while ((int)m->ev.size() < 4 * first_batch) {
            hipEvent_t e;
            SEM_HIP(hipEventCreate(&e));
            m->ev.push_back(e);
        }
        // algorithmic bytes of one launch: the operator apply (its inputs, geometry
        // and outputs, each read or written once, as the reference stores them),
        // r1 read for y -= cr1*r1, and the w/x update (wa, wb, x in; w, x out),
        // per component
        const double EN2 = (double)m->L.E * m->L.N * m->L.N;
        const double op_vals = (kind == K_RES) ? 5 : (kind == K_HELMHOLTZ ? 4 : 2);  // vectors and coefficient
        const double geo = (sizeof(T) == 4 && map_geometry(m->L.N)) ? 32.0 * m->L.E
                                                                    : (kind == K_RES ? 4.0 : 3.0) * sizeof(T) * EN2;
        m->kt_bytes[ti] = (op_vals + 6.0 * nc) * sizeof(T) * EN2 + geo;
        // update kernel: yraw and r2 in, y out, per component
        m->kt_bytes[2 + ti] = 3.0 * nc * sizeof(T) * EN2;
    }
    // Scalar state: the operator kernel of step i (when not first in its
    // batch) applies the beta recurrence of step i-1 in every workgroup, then
    // the pending w/x update, and stores the state in the other of the two
    // buffers; the update kernel of step i reads that buffer and stores alpha
    // there. A closing kernel ends each batch with the state back in m->sc.
    KrylovScalars* buf[2] = {m->sc, m->sc_alt};
    const InstTable itab = m->it();
    auto max_iter = [&]() {
        int v = 0;
        for (int k = 0; k < ni; ++k) v = std::max(v, m->h_sc[k].iter);
        return v;
    };
    auto all_done = [&]() {
        for (int k = 0; k < ni; ++k)
            if (!m->h_sc[k].done) return false;
        return true;
    };
    auto all_success = [&]() {
        for (int k = 0; k < ni; ++k)
            if (!m->h_sc[k].success) return false;
        return true;
    };
    // launches of Lanczos step i, the b_-th of its batch (state buffer cur)
    auto launch_step = [&](int i, int b_, int& cur, bool tim) {
        MinresOpArgs<T> a{};
        a.r2 = R[(i - 1) % 3];
        a.r1 = R[(i + 1) % 3];
        a.y = Yraw;
        a.geo = m->geo_t<T>();
        a.qm = m->qm;
        a.cf = cf;
        a.alpha = (T)alpha;
        a.len = m->len();
        a.nchunks = nch;
        a.obits = m->obits;
        a.partial = m->partial;
        a.wa = W[i % 3];
        a.wb = W[(i + 1) % 3];
        a.wn = W[(i + 2) % 3];
        a.x = x;
        a.it = itab;
        a.st_in = buf[cur];
        a.first = (b_ == 0);
        cur ^= 1;
        a.st_out = buf[cur];
        a.prev_partial = m->partial + nch;
        launch_minres_op_d<T>(m->L.N, kind, a, m->dmat<T>(), shp.op_cpb, shp.op_grid, s, tim ? m->ev[4 * b_] : nullptr,
                              tim ? m->ev[4 * b_ + 1] : nullptr);
        MinresUpdArgs<T> u{};
        u.yraw = Yraw;
        u.y = R[i % 3];
        u.r2 = R[(i - 1) % 3];
        u.len = m->len();
        u.ncomp = nc;
        u.nchunks = nch;
        u.obits = m->obits;
        u.eside = m->eside;
        u.ecq = m->ecq;
        u.c_off = m->c_off;
        u.c_pos = m->c_pos;
        u.partial = m->partial + nch;
        u.it = itab;
        u.st = buf[cur];
        u.op_partial = m->partial;
        launch_minres_upd<T>(m->L.N, u, shp.upd_cpb, shp.upd_grid, s, tim ? m->ev[4 * b_ + 2] : nullptr,
                             tim ? m->ev[4 * b_ + 3] : nullptr);
    };
    int i = 1;
    bool done = o.maxit < 1;
    if (m->use_graphs && !m->timing && !done) {
        // Graph path: one captured batch of G Lanczos steps plus the closing
        // kernel, replayed until every instance has stopped. G is a multiple
        // of 3, so every batch starts at a step i = 1 (mod 3) and sees the
        // same buffers; the first host check comes after about the expected
        // iteration count. Steps past an instance's stop exit at once, so the
        // results equal the eager path's.
        const int G = m->graph_steps;
        GraphKey key{};
        key.ptr = {(const void*)R[0], (const void*)R[1], (const void*)R[2], (const void*)W[0], (const void*)W[1],
                   (const void*)W[2], (const void*)Yraw, (const void*)x, (const void*)cf};
        key.kind = kind;
        key.tsize = (int)sizeof(T);
        key.alpha = alpha;
        key.steps = G;
        key.shape = shp;
        hipGraphExec_t ex = m->find_graph(key);
        if (!ex) {
            hipGraph_t gr = nullptr;
            SEM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int cur = 0;
            for (int b_ = 0; b_ < G; ++b_) launch_step(1 + b_, b_, cur, false);
            launch_minres_close(kind, itab, buf[cur], m->sc, m->partial + nch, s);
            SEM_HIP(hipStreamEndCapture(s, &gr));
            SEM_HIP(hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0));
            m->add_graph(key, gr, ex);
        }
        int reps = std::max(1, (first_batch + G - 1) / G);
        while (!done) {
            for (int r = 0; r < reps; ++r) SEM_HIP(hipGraphLaunch(ex, s));
            i += reps * G;
            m->read_scalars();
            done = all_done() || i > o.maxit + 1;
            reps = 1;
        }
    }
    while (!done) {
        int launched = 0;
        const int i0 = i;
        int cur = 0;
        const int nbatch = (i == 1) ? first_batch : m->batch;
        for (int b_ = 0; b_ < nbatch && i <= o.maxit + 1; ++b_, ++i) {
            const bool tim = m->timing == 2 || (m->timing == 1 && sampled_step(i));
            launch_step(i, b_, cur, tim);
            ++launched;
        }
        launch_minres_close(kind, itab, buf[cur], m->sc, m->partial + nch, s);
        m->read_scalars();
        if (m->timing) {
            // launches that ran a full Lanczos step (later ones in the batch exit early)
            const int last = max_iter();
            for (int b_ = 0; b_ < launched && i0 + b_ <= last; ++b_) {
                if (m->timing == 1 && !sampled_step(i0 + b_)) continue;
                for (int q = 0; q < 2; ++q) {
                    float ms = 0;
                    SEM_HIP(hipEventElapsedTime(&ms, m->ev[4 * b_ + 2 * q], m->ev[4 * b_ + 2 * q + 1]));
                    m->kt_us[2 * q + ti] += 1000.0 * ms;
                    m->kt_launches[2 * q + ti] += 1;
                }
            }
        }
        done = all_done() || i > o.maxit + 1;
    }
    bool pending = false;
    for (int k = 0; k < ni; ++k) pending |= m->h_sc[k].p3_pending != 0;
    if (pending) {
        MinresFlushArgs<T> fa{};
        for (int q = 0; q < 3; ++q) {
            fa.R[q] = R[q];
            fa.W[q] = W[q];
        }
        fa.x = x;
        fa.len = m->len();
        fa.chunk_len = m->L.K * m->L.N * m->L.N;
        fa.ncomp = nc;
        launch_minres_flush<T>(fa, itab, nch, m->sc, s);
    }
    for (int k = 0; k < ni; ++k) {
        Results& r = m->inst_results[k];
        const KrylovScalars& h = m->h_sc[k];
        const int mk = o.inst_maxit ? o.inst_maxit[k] : o.maxit;
        r.success = h.success;
        r.n_iter = r.success ? h.iter : mk + 1;
        if (mk < 1) r.n_iter = 1;
        r.residual = h.phibar;
    }
    // the next solve's first batch follows this one's count, if it converged
    // (a solve that ran to maxit says nothing about the next one)
    if (all_success()) m->iter_hint[slot] = max_iter();
    // aggregate: success of every instance, the longest iteration count, the largest residual
    Results all = m->inst_results[0];
    for (int k = 1; k < ni; ++k) {
        all.success = all.success && m->inst_results[k].success;
        all.n_iter = std::max(all.n_iter, m->inst_results[k].n_iter);
        all.residual = std::max(all.residual, m->inst_results[k].residual);
    }
    return all;
}

// ------------------------------------------------------------- autotuner --
// Work shapes tried for the two Lanczos-step kernels: chunks per workgroup
// (1 or 2) times one workgroup per work item or a persistent grid of 2 or 4
// workgroups per compute unit. Results do not depend on the shape (partials
// are per chunk and every reduction has a fixed order), so the choice only
// changes speed.
static std::vector<std::pair<int, int>> step_candidates()
{
    int dev = 0, ncu = 256;
    SEM_HIP(hipGetDevice(&dev));
    SEM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    return {{1, 0}, {2, 0}, {1, 2 * ncu}, {2, 2 * ncu}, {1, 4 * ncu}, {2, 4 * ncu}};
}

// Times every candidate on this mesh with `steps` Lanczos steps of a
// synthetic solve (b = 1 in consistent form, zero guess, no stop) and keeps
// the fastest shape for each kernel. us[2c], us[2c+1]: operator and update
// kernel microseconds per launch of candidate c.
template <typename T>
static void autotune_steps(sem_mesh* m, int kind, int steps, std::vector<double>& us)
{
    const int nc = op_ncomp(kind);
    const size_t n = (size_t)nc * m->len();
    const int ti = sizeof(T) == 8 ? 0 : 1;
    const std::string tp = sizeof(T) == 8 ? "d" : "f";
    T* b = m->buf<T>("tune_b" + tp, n);
    T* x = m->buf<T>("tune_x" + tp, n);
    T* cf = m->buf<T>("tune_cf" + tp, m->len());
    launch_fill<T>(b, T(1), n, m->stream);
    dss<T>(m, DSS_ZERO_BND, b, nc);
    launch_fill<T>(cf, T(0), m->len(), m->stream);
    const auto cand = step_candidates();
    us.assign(2 * cand.size(), 0.0);
    const int timing = m->timing;
    const StepShape keep = m->shape[ti];
    const int hint = m->iter_hint[7];
    long saved_l[4];
    double saved_u[4];
    for (int q = 0; q < 4; ++q) {
        saved_l[q] = m->kt_launches[q];
        saved_u[q] = m->kt_us[q];
    }
    KrylovOpts o;
    o.maxit = steps;
    o.tol = 0;
    o.atol = 0;
    o.hint_slot = 7;
    m->timing = 2;
    // the one-launch FP32 step runs one work item per workgroup for every
    // shape: its persistent candidates repeat the one-shot timing of their
    // chunks per workgroup and are not picked
    const bool fused1 = fused_step(m, ti);
    // the mixed-precision stage solves run the one-launch step on the scaled
    // system (S A S): time it with a scale, so that the step loads what it
    // loads there
    float* ds = nullptr;
    if (fused1 && kind == K_RES) {
        ds = m->buf<float>("tune_ds", m->len());
        launch_fill<float>(ds, 1.0f, m->len(), m->stream);
    }
    for (int rep = 0; rep < 2; ++rep)  // first pass warms up and allocates
        for (size_t c = 0; c < cand.size(); ++c) {
            if (fused1 && cand[c].second > 0) {
                if (rep == 1)
                    for (size_t c0 = 0; c0 < cand.size(); ++c0)
                        if (cand[c0].second == 0 && cand[c0].first == cand[c].first) {
                            us[2 * c] = us[2 * c0];
                            us[2 * c + 1] = us[2 * c0 + 1];
                        }
                continue;
            }
            m->shape[ti] = StepShape{cand[c].first, cand[c].second, cand[c].first, cand[c].second};
            for (int q = 0; q < 4; ++q) {
                m->kt_launches[q] = 0;
                m->kt_us[q] = 0;
            }
            m->iter_hint[7] = steps;
            minres_fused<T>(m, kind, cf, 0.5, b, x, true, o, ds);
            if (rep == 1)
                for (int q = 0; q < 2; ++q)
                    us[2 * c + q] = m->kt_us[2 * q + ti] / std::max<long>(1, m->kt_launches[2 * q + ti]);
        }
    StepShape best = keep;
    double b_op = 1e300, b_up = 1e300;
    // in fused mode only the operator (fused) kernel runs; the update shape stays
    const bool fused = fused_step(m, ti);
    for (size_t c = 0; c < cand.size(); ++c) {
        if (fused1 && cand[c].second > 0) continue;
        if (us[2 * c] < b_op) {
            b_op = us[2 * c];
            best.op_cpb = cand[c].first;
            best.op_grid = cand[c].second;
        }
        if (!fused && us[2 * c + 1] < b_up) {
            b_up = us[2 * c + 1];
            best.upd_cpb = cand[c].first;
            best.upd_grid = cand[c].second;
        }
    }
    m->shape[ti] = best;
    m->timing = timing;
    m->iter_hint[7] = hint;
    for (int q = 0; q < 4; ++q) {
        m->kt_launches[q] = saved_l[q];
        m->kt_us[q] = saved_u[q];
    }
}

// ------------------------------------------------------------------- PCG --
// pcg.hpp:59-97 (identity preconditioner), FP64.
static Results pcg_fused(sem_mesh* m, const double* b, double* x, bool x_zero, int maxit, double tol)
{
    const size_t n = m->len();
    hipStream_t s = m->stream;
    double* r = m->buf<double>("pcg_r", n);
    double* p = m->buf<double>("pcg_p", n);
    double* ap = m->buf<double>("pcg_ap", n);
    if (x_zero) {
        launch_fill<double>(x, 0.0, n, s);
        launch_cast<double, double>(r, b, n, s);
    } else {
        apply_op<double>(m, K_LAPLACE, x, ap, nullptr, 0);
        lincomb<double>(m, r, 0, {{1.0, b}, {-1.0, ap}}, n);
    }
    wdot_dev<double>(m, b, b, 1, 0);
    wdot_dev<double>(m, r, r, 1, 1);
    launch_pcg_init(m->sc, maxit, tol, s);
    const int nch = m->L.nchunks;
    int i = 0;
    bool done = false;
    while (!done) {
        for (int b_ = 0; b_ < m->batch && i < maxit + 1; ++b_, ++i) {
            PcgOpArgs<double> a{};
            a.r = r;
            a.p = p;
            a.ap = ap;
            a.geo = m->geo;
            a.len = n;
            a.partial = m->partial;
            a.sc = reinterpret_cast<const double*>(m->sc);
            launch_pcg_op_d<double>(m->L.N, a, nch, m->D64.data(), s);
            DssArgs<double> d = dss_args<double>(m, ap, 1);
            d.r2 = p;
            d.partial = m->partial + nch;
            launch_dss<double>(DSS_PCG, d, s);
            launch_pcg_fin_alpha(m->sc, m->partial, nch + m->dss_nb, s);
            launch_pcg_upd(x, r, p, ap, n, m->L, m->obits, reinterpret_cast<const double*>(m->sc), m->partial, s);
            launch_pcg_fin_rho(m->sc, m->partial, nch, s);
        }
        m->read_scalars();
        done = m->h_sc->done != 0 || i >= maxit + 1;
    }
    Results res;
    res.success = m->h_sc->success;
    res.n_iter = m->h_sc->iter;
    res.residual = m->h_sc->resid;
    return res;
}

// ------------------------------------------------ mixed precision solve --
// Diagonal scaling of the Schrodinger-stage system Res = [[M, aL], [aL, -M]]
// (L = K + V M): S = 1 / sqrt(m + |a| |k + c|) from the assembled diagonals of
// the mass (m), stiffness (k) and potential mass (c) matrices. The inner
// FP32 MINRES then solves (S Res S) w = S r and returns d = S w; on the
// Schrodinger workloads this takes 15-20% fewer Lanczos steps per decade of
// residual than the unscaled system. The outer refinement still stops on the
// true FP64 residual of Res.
static const float* inner_scale(sem_mesh* m, const double* cf64, double alpha)
{
    const size_t len = m->len();
    hipStream_t s = m->stream;
    float* S = m->buf<float>("is_S", len);
    if (m->is_cf == cf64 && m->is_gen == m->field_gen && m->is_alpha == alpha) return S;
    double* kd = m->buf<double>("is_kd", len);
    double* md = m->buf<double>("is_md", len);
    if (!m->is_geo) {
        launch_stiff_diag(m->L.N, m->geo, len, m->L.nchunks, kd, m->dmat<double>(), s);
        dss<double>(m, DSS_SUM, kd, 1);
        launch_cast<double, double>(md, m->geo + 3 * len, len, s);
        dss<double>(m, DSS_SUM, md, 1);
        m->is_geo = true;
    }
    double* cd = m->buf<double>("is_cd", len);
    launch_cast<double, double>(cd, cf64, len, s);
    dss<double>(m, DSS_SUM, cd, 1);
    launch_inner_scale(S, md, kd, cd, alpha, len, s);
    m->is_cf = cf64;
    m->is_gen = m->field_gen;
    m->is_alpha = alpha;
    return S;
}

// Smallest relative reduction asked of one FP32 inner solve: what FP32 (or the
// split-bf16 contractions) can deliver; bf16 contractions carry about 4e-3
// relative error, so asking more of one sweep only adds Lanczos steps.
static double inner_tol_floor(const sem_mesh* m) { return m->contraction == 1 ? 5e-3 : 2e-5; }

// FP64 iterative refinement around an FP32 MINRES: r = b - A x in FP64, the
// correction A d = r is solved in FP32 (FP64 recurrence scalars and dots),
// x += d in FP64. Stops on the true FP64 residual, ||r|| < tol*||r0|| + atol
// (or tol*||b|| + atol when rel_to_b), the reference's test applied to the
// true residual instead of MINRES's recurrence estimate.
static Results solve_mixed(sem_mesh* m, int kind, const double* cf64, const float* cf32, double alpha, const double* b,
                           double* x, bool x_zero, const KrylovOpts& o, double inner_tol, double* r_given = nullptr)
{
    const int nc = op_ncomp(kind);
    const size_t n = (size_t)nc * m->len();
    hipStream_t s = m->stream;
    // the caller's buffer holding b - A x becomes the residual buffer
    double* r = r_given ? r_given : m->buf<double>("mx_r", n);
    double* Ax = m->buf<double>("mx_Ax", n);
    float* r32 = m->buf<float>("mx_r32", n);
    float* d32 = m->buf<float>("mx_d32", n);
    if (x_zero) {
        launch_fill<double>(x, 0.0, n, s);
        launch_cast<double, double>(r, b, n, s);
    } else if (r_given) {
        // the caller already holds b - A x (the learned guess's combination) in
        // r; the stop test after every sweep still uses a freshly computed residual
    } else {
        apply_op<double>(m, kind, x, Ax, cf64, alpha);
        lincomb<double>(m, r, 0, {{1.0, b}, {-1.0, Ax}}, n);
    }
    const float* S = (kind == K_RES && m->inner_scale && fused_step(m, 1)) ? inner_scale(m, cf64, alpha) : nullptr;
    // every instance of a batched mesh refines on its own: its own norms,
    // target, sweeps and inner iteration budget; an instance that has met its
    // target (or run out) sits out the later sweeps, its x untouched
    const int ni = m->n_inst;
    // ||r||^2 and ||b||^2 of every instance in one pass and one host read
    wdot3_dev(m, r, b, nc, 4);
    m->read_scalars();
    std::vector<double> rn(ni), bnorm(ni);
    for (int k = 0; k < ni; ++k) {
        rn[k] = m->h_sc[k].red[4];
        bnorm[k] = m->h_sc[k].red[5];
    }
    std::vector<double> target(ni), itol(ni);
    std::vector<long> inner(ni, 0);
    std::vector<int> sweeps(ni, 0), imax(ni, 0);
    std::vector<uint8_t> idle(ni, 0);
    for (int k = 0; k < ni; ++k) {
        rn[k] = std::sqrt(rn[k]);
        bnorm[k] = std::sqrt(bnorm[k]);
        target[k] = o.tol * (o.rel_to_b ? bnorm[k] : rn[k]) + o.atol;
    }
    auto refresh = [&]() {
        bool any = false;
        for (int k = 0; k < ni; ++k) {
            idle[k] = (o.frozen && o.frozen[k]) || rn[k] < target[k] || inner[k] >= o.maxit || sweeps[k] >= 50;
            any |= !idle[k];
        }
        return any;
    };
    const size_t cl = (size_t)m->L.K * m->L.N * m->L.N;
    while (refresh()) {
        if (S) launch_scale_cast<float, double>(r32, r, S, m->len(), nc, s);
        else launch_cast<float, double>(r32, r, n, s);
        KrylovOpts io;
        io.maxit = 1;
        bool all_active = true;
        for (int k = 0; k < ni; ++k) {
            all_active &= !idle[k];
            imax[k] = (int)std::max<long>(1, o.maxit - inner[k]);
            // relative reduction of this sweep; never ask FP32 for more than it can give
            itol[k] = std::max(inner_tol, 0.5 * target[k] / std::max(rn[k], 1e-300));
            if (!idle[k]) io.maxit = std::max(io.maxit, imax[k]);
        }
        io.tol = itol[0];
        io.atol = 0;
        if (ni > 1) {
            io.frozen = idle.data();
            io.inst_tol = itol.data();
            io.inst_maxit = imax.data();
        }
        int sw = 0;
        for (int k = 0; k < ni; ++k)
            if (!idle[k]) sw = std::max(sw, sweeps[k]);
        io.hint_slot = 2 + std::min(sw, 5);
        minres_fused<float>(m, kind, cf32, alpha, r32, d32, true, io, S);
        if (S) launch_scale_cast<double, float>(Ax, d32, S, m->len(), nc, s);
        else launch_cast<double, float>(Ax, d32, n, s);
        if (all_active) {
            lincomb<double>(m, x, 2, {{1.0, Ax}}, n);  // x += d
        } else {
            for (int k = 0; k < ni; ++k) {  // x += d on the instances of this sweep
                if (idle[k]) continue;
                const size_t off = (size_t)m->h_inst_off[k] * cl;
                const size_t cnt = (size_t)(m->h_inst_off[k + 1] - m->h_inst_off[k]) * cl;
                for (int c = 0; c < nc; ++c)
                    lincomb<double>(m, x + c * m->len() + off, 2, {{1.0, Ax + c * m->len() + off}}, cnt);
            }
        }
        apply_op<double>(m, kind, x, Ax, cf64, alpha);
        lincomb<double>(m, r, 0, {{1.0, b}, {-1.0, Ax}}, n);
        const std::vector<double> r2 = wdot_inst<double>(m, r, r, nc);
        for (int k = 0; k < ni; ++k) {
            if (idle[k]) continue;
            inner[k] += m->inst_results[k].n_iter;
            rn[k] = std::sqrt(r2[k]);
            ++sweeps[k];
        }
        m->stats.refinements++;
    }
    // aggregate over the instances of the solve: all succeeded, the most inner
    // iterations, the largest residual
    Results res;
    res.success = 1;
    m->last_true_rel = 0;
    for (int k = 0; k < ni; ++k) {
        Results& rk = m->inst_results[k];
        rk.success = rn[k] < target[k];
        rk.n_iter = (int)inner[k];
        rk.residual = rn[k];
        if (o.frozen && o.frozen[k]) continue;
        res.success = res.success && rk.success;
        res.n_iter = std::max(res.n_iter, rk.n_iter);
        res.residual = std::max(res.residual, rk.residual);
        m->last_true_rel = std::max(m->last_true_rel, rn[k] / std::max(bnorm[k], 1e-300));
    }
    return res;
}

// ---------------------------------------------------------- PDE drivers --
struct Field {
    double* c64 = nullptr;
    float* c32 = nullptr;
};

// premultiplied coefficient field gJ * f on the device, from a host field in
// the reference layout
static Field make_field(sem_mesh* m, const double* host, const std::string& name)
{
    Field f;
    const size_t n = m->len();
    f.c64 = m->buf<double>(name + "_64", n);
    f.c32 = m->buf<float>(name + "_32", n);
    double* tmp = m->buf<double>("field_tmp", n);
    m->field_gen++;
    m->upload(tmp, host, 1);
    launch_mul<double>(f.c64, m->geo + 3 * n, tmp, n, m->stream);
    launch_cast<float, double>(f.c32, f.c64, n, m->stream);
    return f;
}

// masked output of the reference's poisson/helmholtz: owner copies carry the
// solution, non-owner copies keep the caller's input (poisson.hpp:56)
static void merge_owner(sem_mesh* m, const double* xdev, double* u_host)
{
    const Layout& L = m->L;
    const int N = L.N;
    std::vector<double> xr((size_t)L.E * N * N);
    m->download(xr.data(), xdev, 1);
    for (int e = 0; e < L.E; ++e) {
        const unsigned bits = m->h_obits[e];
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const bool ri = (i == 0 || i == N - 1), rj = (j == 0 || j == N - 1);
                bool own = true;
                if (ri && rj) own = (bits >> (4 + ((i == 0) ? ((j == 0) ? 0 : 3) : ((j == 0) ? 1 : 2)))) & 1u;
                else if (j == 0) own = bits & 1u;
                else if (i == N - 1) own = (bits >> 1) & 1u;
                else if (j == N - 1) own = (bits >> 2) & 1u;
                else if (i == 0) own = (bits >> 3) & 1u;
                const size_t k = L.ref(e, i, j);
                if (own) u_host[k] = xr[k];
            }
    }
}

// Schrodinger stepper session: the state psi, the potential and the stage
// vectors stay resident on the GPU between advance() calls.
// one-hot stage columns of the guess network's input (at most 8 stages)
constexpr int kStageCols = 8;

struct Stepper {
    sem_mesh* m = nullptr;
    int scheme = 1;  // 0 bdf2, 1 dirk4s3, 2 dirk4s5
    sem_solver_opts opts{};
    Field V;
    size_t n = 0;  // 2 * len
    std::vector<void*> owned;
    double* p = nullptr;
    double* v = nullptr;
    double* q = nullptr;
    std::vector<double*> z, zprev, zprev2;
    bool have_prev = false, have_prev2 = false;
    double t = 0;
    std::vector<double> pnorm; // dirk4s5 (schrodinger.hpp:214), per instance
    double* ps[2] = {nullptr, nullptr};  // bdf2 history (schrodinger.hpp:85)
    int nps = 0;
    double alpha_bdf = 0;
    sem_run_stats stats{};
    std::vector<int> inst_ok;  // per instance: every solve so far succeeded
    // batched meshes: an instance whose solve fails stops there, as the
    // reference's stepper returns (BDF2 keeps the unconverged iterate and
    // counts the step, schrodinger.hpp:96-114; the SDIRK schemes keep psi of
    // the step's start, :222-226), while the other instances go on
    std::vector<uint8_t> skip;       // instances left out of the solves (stopped, or failed in this step)
    std::vector<uint8_t> frozen;     // stopped in an earlier step
    std::vector<uint8_t> failed_now; // a solve of this step failed
    std::vector<double> inst_t;      // per-instance time
    std::vector<int> inst_steps;     // per-instance completed steps
    bool multi() const { return m->n_inst > 1; }
    // learned guess (guess = 2): one network for all stages, told the stage
    // by one-hot input columns
    std::unique_ptr<Mlp> mlp;
    bool guard = true;  // keep the extrapolation when its residual is smaller
    bool residual_loss = false;  // train on ||Res (g - z)|| instead of the correction's squared error
    bool feat_rext = true;       // fourth input: b - Res ext (true) or b
    double* rext = nullptr;      // b - Res ext of the current stage
    double* Vraw = nullptr;
    double* ext = nullptr;
    double* tmp = nullptr;
    int learn_steps = 50;
    int epochs = 5;
    double s_feat[5][3] = {};  // per stage: scales of z, V, b features
    double s_out[5] = {};      // per stage: scale of the correction
    bool have_scale[5] = {}, have_out[5] = {};
    // the network's Adam steps on a training sample run on their own stream,
    // overlapped with the next stage solves; ev_data: the sample's features
    // and targets are in place; ev_trained: its Adam steps are done (the
    // network's inputs and parameters may be touched again)
    hipStream_t train_stream = nullptr;
    hipEvent_t ev_data = nullptr, ev_trained = nullptr;
    bool train_pending = false, loss_pending = false;

    double* alloc()
    {
        void* d = nullptr;
        SEM_HIP(hipMalloc(&d, n * sizeof(double)));
        owned.push_back(d);
        return (double*)d;
    }
    ~Stepper()
    {
        if (train_stream) hipStreamSynchronize(train_stream);
        if (m) hipStreamSynchronize(m->stream);
        for (void* d : owned) hipFree(d);
        if (ev_data) hipEventDestroy(ev_data);
        if (ev_trained) hipEventDestroy(ev_trained);
        if (train_stream) hipStreamDestroy(train_stream);
    }
    // the main stream waits for the last training sample's Adam steps
    void wait_training()
    {
        if (train_pending) SEM_HIP(hipStreamWaitEvent(m->stream, ev_trained, 0));
    }
    // host: the training stream has drained; the last loss is read
    void finish_training()
    {
        if (!train_pending) return;
        SEM_HIP(hipStreamSynchronize(train_stream));
        train_pending = false;
        if (loss_pending) {
            double l = 0;
            SEM_HIP(hipMemcpy(&l, mlp->red, sizeof(double), hipMemcpyDeviceToHost));
            stats.mlp_last_loss = l / ((double)m->L.E * mlp->out);
            loss_pending = false;
        }
    }
};

// v = dt * f(q) (schrodinger.hpp:163-183, :221): f's DSS sums scaled by dt
static void stage_rhs(Stepper& st, const double* q, double* out)
{
    apply_op<double>(st.m, K_F, q, out, st.V.c64, 0.0, st.opts.dt);
}

static Results solve_res(Stepper& st, double alpha, const double* b, double* x, bool zero, bool rel_to_b,
                         double* r_given = nullptr)
{
    sem_mesh* m = st.m;
    KrylovOpts ko;
    ko.maxit = st.opts.max_iter;
    ko.tol = st.opts.tol;
    ko.atol = st.opts.tol;
    ko.rel_to_b = rel_to_b;
    if (st.multi()) ko.frozen = st.skip.data();
    Results r;
    const long ref0 = m->stats.refinements;
    if (st.opts.precision == SEM_PREC_MIXED) {
        r = solve_mixed(m, K_RES, st.V.c64, st.V.c32, alpha, b, x, zero, ko, inner_tol_floor(m), r_given);
        st.stats.max_true_rel_residual = std::max(st.stats.max_true_rel_residual, m->last_true_rel);
    } else
        r = minres_fused<double>(m, K_RES, st.V.c64, alpha, b, x, zero, ko);
    if (st.inst_ok.size() != (size_t)m->n_inst) st.inst_ok.assign(m->n_inst, 1);
    for (int k = 0; k < m->n_inst; ++k) {
        if (st.multi() && st.skip[k]) continue;
        if (!m->inst_results[k].success) {
            st.inst_ok[k] = 0;
            if (st.multi()) st.failed_now[k] = st.skip[k] = 1;
        }
    }
    // a batch goes on with the instances that converged
    if (st.multi()) {
        r.success = 0;
        for (int k = 0; k < m->n_inst; ++k) r.success |= !st.skip[k];
    }
    st.stats.refinements += m->stats.refinements - ref0;
    st.stats.stage_solves++;
    st.stats.total_iters += r.n_iter;
    return r;
}

// root mean square of a device vector (all entries of ncomp components, every
// instance of a batched mesh)
static double rms_dev(sem_mesh* m, const double* a, int ncomp)
{
    double s = 0;
    for (double v : wdot_inst<double>(m, a, a, ncomp)) s += v;
    return std::sqrt(s / std::max(1.0, (double)ncomp * m->hm.dof));
}

static MlpScatterArgs scatter_args(Stepper& st, int stage, double* out, const double* base, double scale)
{
    sem_mesh* m = st.m;
    MlpScatterArgs a{};
    (void)stage;
    a.Y = st.mlp->Y;
    a.ld = st.mlp->out_pad;
    a.out = out;
    a.base = base;
    a.scale = scale;
    a.len = m->len();
    a.E = m->L.E;
    a.N = m->L.N;
    a.K = m->L.K;
    a.ncomp = 2;
    return a;
}

static void mlp_features(Stepper& st, int stage, const double* b)
{
    sem_mesh* m = st.m;
    const size_t len = m->len();
    const double* zp = st.zprev[stage];
    const double* zq = st.zprev2[stage];
    MlpFeatureArgs f{};
    const double* src[7] = {zp, zp + len, zq, zq + len, st.Vraw, b, b + len};
    const double sc[7] = {st.s_feat[stage][0], st.s_feat[stage][0], st.s_feat[stage][0], st.s_feat[stage][0],
                          st.s_feat[stage][1], st.s_feat[stage][2], st.s_feat[stage][2]};
    for (int g = 0; g < 7; ++g) {
        f.src[g] = src[g];
        f.scale[g] = sc[g];
    }
    f.ngroups = 7;
    f.extra = kStageCols;
    for (int k = 0; k < kStageCols; ++k) f.extra_val[k] = (k == stage) ? 1.f : 0.f;
    f.X = st.mlp->X;
    f.ld = st.mlp->in_pad;
    f.E = m->L.E;
    f.N = m->L.N;
    f.K = m->L.K;
    launch_gather_features(f, m->stream);
}

static Results stage_solve(Stepper& st, double alpha, const double* b, double* z, int stage)
{
    sem_mesh* m = st.m;
    bool zero = true;
    // the network trains on the stage solutions of the first learn_steps steps,
    // which start from the extrapolated guess (so the training window follows
    // the fallback's trajectory); afterwards it is frozen and proposes guesses
    const bool learned = st.opts.guess == 2 && st.have_prev2;
    const bool training = learned && st.stats.steps < st.learn_steps;
    double* r_given = nullptr;  // b - Res z of the start, when already computed (solve_mixed works in it)
    if (learned) {
        // extrapolation plus the network's correction, made consistent
        lincomb<double>(m, st.ext, 1, {{2.0, st.zprev[stage]}, {-1.0, st.zprev2[stage]}}, st.n);
        // the extrapolation's residual b - Res ext (an input of the network, the
        // guard's fallback residual and the residual loss's normaliser)
        double* Az = m->buf<double>("guard_Az", st.n);
        apply_op<double>(m, K_RES, st.ext, Az, st.V.c64, alpha);
        lincomb<double>(m, st.rext, 0, {{1.0, b}, {-1.0, Az}}, st.n);
        const double* fin = st.feat_rext ? st.rext : b;  // the network's fourth input
        if (!st.have_scale[stage]) {
            st.s_feat[stage][0] = 1.0 / std::max(rms_dev(m, st.zprev[stage], 2), 1e-300);
            st.s_feat[stage][1] = 1.0 / std::max(rms_dev(m, st.Vraw, 1), 1e-300);
            st.s_feat[stage][2] = 1.0 / std::max(rms_dev(m, fin, 2), 1e-300);
            st.have_scale[stage] = true;
        }
        if (st.have_out[stage] && !training) {
            Mlp& net = *st.mlp;
            st.wait_training();
            net.reserve(m->L.E);
            mlp_features(st, stage, fin);
            net.forward(m->L.E);
            launch_scatter_output(scatter_args(st, stage, z, st.ext, st.s_out[stage]), m->stream);
            dss<double>(m, DSS_UNMASK, z, 2);
            dss<double>(m, DSS_ZERO_BND, z, 2);
            // guard: the extrapolation stays the fallback. The guess is
            // ext + c (z - ext) with the c that minimises the true FP64
            // residual ||b - Res x0|| = ||r_ext - c (r_ext - r_z)||: c = 0 is
            // the extrapolation and c = 1 the network's guess, so the start is
            // never worse than either, and both residuals are already at hand
            // (no further operator apply; the refinement starts from r0).
            if (!st.guard) {
                st.stats.learned_guess_chosen++;
            } else {
                double* rz = m->buf<double>("guard_r", st.n);
                apply_op<double>(m, K_RES, z, Az, st.V.c64, alpha);
                lincomb<double>(m, rz, 0, {{1.0, b}, {-1.0, Az}}, st.n);
                wdot3_dev(m, rz, st.rext, 2, 2);  // red[2] = ||r_z||^2, red[3] = ||r_ext||^2, red[4] = <r_z, r_ext>
                m->read_scalars();
                const size_t cl = (size_t)m->L.K * m->L.N * m->L.N;
                for (int k = 0; k < m->n_inst; ++k) {
                    if (st.multi() && st.skip[k]) continue;
                    const double gg = m->h_sc[k].red[2], ee = m->h_sc[k].red[3], eg = m->h_sc[k].red[4];
                    const double dd = ee - 2.0 * eg + gg;  // ||r_ext - r_z||^2 = ||Res (z - ext)||^2
                    const double c = dd > 0 ? (ee - eg) / dd : 0.0;
                    // residual of the combination: evaluated below; the
                    // statistic uses the smaller of its two exact end points
                    // when cancellation makes the closed form unreliable
                    double r2 = ee - c * (ee - eg);
                    if (!(r2 > 1e-12 * ee)) r2 = std::min(ee, gg);
                    if (c != 0.0) {
                        st.stats.learned_guess_chosen++;
                        if (ee > 0 && r2 > 0) st.stats.learned_residual_digits += 0.5 * std::log10(ee / r2);
                    }
                    // x0 = (1 - c) ext + c z, r0 = (1 - c) r_ext + c r_z, on instance k's chunks
                    const size_t off = st.multi() ? (size_t)m->h_inst_off[k] * cl : 0;
                    const size_t cnt = st.multi() ? (size_t)(m->h_inst_off[k + 1] - m->h_inst_off[k]) * cl : m->len();
                    for (int q = 0; q < 2; ++q) {
                        const size_t o = q * m->len() + off;
                        lincomb<double>(m, z + o, 1, {{1.0 - c, st.ext + o}, {c, z + o}}, cnt);
                        lincomb<double>(m, rz + o, 1, {{1.0 - c, st.rext + o}, {c, rz + o}}, cnt);
                    }
                }
                r_given = rz;
            }
        } else {
            launch_cast<double, double>(z, st.ext, st.n, m->stream);
        }
        zero = false;
    } else if (st.opts.guess >= 1 && st.have_prev) {  // extrapolation from the previous steps' same stage
        if (st.have_prev2)
            lincomb<double>(m, z, 1, {{2.0, st.zprev[stage]}, {-1.0, st.zprev2[stage]}}, st.n);
        else
            launch_cast<double, double>(z, st.zprev[stage], st.n, m->stream);
        zero = false;
    }
    // a batch trains its one network only on stages that every instance solved
    // (an instance that has stopped holds no solution to learn from)
    bool whole = true;
    if (st.multi())
        for (int k = 0; k < m->n_inst; ++k) whole = whole && !st.skip[k];
    Results r = solve_res(st, alpha, b, z, zero, !zero, r_given);
    if (st.multi())
        for (int k = 0; k < m->n_inst; ++k) whole = whole && !st.skip[k];
    if (training && r.success && whole) {
        // train on this stage's solution: target = (z - extrapolation) / s_out
        if (!st.have_out[stage]) {
            lincomb<double>(m, st.tmp, 0, {{1.0, z}, {-1.0, st.ext}}, st.n);
            st.s_out[stage] = std::max(rms_dev(m, st.tmp, 2), 1e-300);
            st.have_out[stage] = true;
        }
        Mlp& net = *st.mlp;
        st.wait_training();  // the previous sample's Adam steps still read X and T
        net.reserve(m->L.E);
        mlp_features(st, stage, st.feat_rext ? st.rext : b);
        if (!st.residual_loss) {
            // mean squared error of the scaled correction (z - ext) / s_out;
            // the Adam steps run on the training stream while the next stage
            // solves go on (the window's solves start from the extrapolation)
            launch_gather_target(scatter_args(st, stage, nullptr, st.ext, 1.0 / st.s_out[stage]), net.T, z,
                                 m->stream);
            SEM_HIP(hipEventRecord(st.ev_data, m->stream));
            SEM_HIP(hipStreamWaitEvent(st.train_stream, st.ev_data, 0));
            net.stream = st.train_stream;
            for (int ep = 0; ep < st.epochs; ++ep) {
                net.backward(m->L.E, false);
                net.adam_step();
                st.stats.mlp_train_steps++;
            }
            net.stream = m->stream;
            SEM_HIP(hipEventRecord(st.ev_trained, st.train_stream));
            st.train_pending = st.loss_pending = true;
        } else {
            // residual loss ||Res (g - z)||^2 / ||b - Res ext||^2 of the assembled
            // guess g (the owner copy of each node, Dirichlet entries zero): the
            // norm the guard compares and MINRES reduces. Its gradient with
            // respect to the owner copy of a node is 2 (Res Res (g - z)) there
            // (Res is symmetric), zero on the other copies and on the boundary.
            double* g = m->buf<double>("rl_g", st.n);
            double* rr = m->buf<double>("rl_r", st.n);
            double* gr = m->buf<double>("rl_gr", st.n);
            double n2 = 0;
            for (double v : wdot_inst<double>(m, st.rext, st.rext, 2)) n2 += v;
            n2 = std::max(n2, 1e-300);
            for (int ep = 0; ep < st.epochs; ++ep) {
                const bool last = ep + 1 == st.epochs;
                net.forward(m->L.E, true);  // activations for backprop
                launch_scatter_output(scatter_args(st, stage, g, st.ext, st.s_out[stage]), m->stream);
                dss<double>(m, DSS_UNMASK, g, 2);
                dss<double>(m, DSS_ZERO_BND, g, 2);
                lincomb<double>(m, g, 2, {{-1.0, z}}, st.n);  // g - z
                apply_op<double>(m, K_RES, g, rr, st.V.c64, alpha);
                apply_op<double>(m, K_RES, rr, gr, st.V.c64, alpha);
                dss<double>(m, DSS_MASK, gr, 2);
                dss<double>(m, DSS_ZERO_BND, gr, 2);
                launch_gather_target(scatter_args(st, stage, nullptr, nullptr, 2.0 * st.s_out[stage] / n2), net.dY, gr,
                                     m->stream);
                net.backprop(m->L.E);
                net.adam_step();
                st.stats.mlp_train_steps++;
                if (last) {
                    double l = 0;
                    for (double v : wdot_inst<double>(m, rr, rr, 2)) l += v;
                    st.stats.mlp_last_loss = l / n2;
                }
            }
        }
    }
    return r;
}

static void record_history(Stepper& st, int nstages)
{
    if (st.opts.guess < 1) return;
    for (int k = 0; k < nstages; ++k) {
        if (st.have_prev) std::swap(st.zprev2[k], st.zprev[k]);
        launch_cast<double, double>(st.zprev[k], st.z[k], st.n, st.m->stream);
    }
    st.have_prev2 = st.have_prev;
    st.have_prev = true;
}

static bool do_callback(Stepper& st, sem_step_cb cb, void* user)
{
    if (!cb) return false;
    sem_mesh* m = st.m;
    const size_t rl = (size_t)m->L.E * m->L.N * m->L.N;
    std::vector<double> h(2 * rl);
    m->download(h.data(), st.p, 2);
    return cb(h.data(), h.data() + rl, user) != 0;
}

// One VerboseTracker line (verbose.hpp:9-44): t fixed and the error estimate
// scientific, both with precision int(2 - log10 dt) (schrodinger.hpp:28),
// and a 30-character progress bar (clamped to the bar when t passes T).
static std::string progress_line(double t, double T, double dt, const double* err)
{
    int prec = (int)(2 - std::log10(dt));
    if (prec < 0) prec = 0;
    const int w = 30;
    int q = (int)(w * (t / T));
    if (q > w) q = w;
    if (q < 0) q = 0;
    std::string bar(w + 2, ' ');
    bar[0] = '[';
    for (int i = 0; i < q; ++i) bar[i + 1] = '-';
    bar[q + 1] = '>';
    bar[w + 1] = ']';
    char buf[256];
    if (err)
        std::snprintf(buf, sizeof buf, "%.*f :: %s :: error ~ %.*e\r", prec, t, bar.c_str(), prec, *err);
    else
        std::snprintf(buf, sizeof buf, "%.*f :: %s\r", prec, t, bar.c_str());
    return buf;
}

static void progress(double t, double T, double dt, const double* err)
{
    std::fputs(progress_line(t, T, dt, err).c_str(), stdout);
    std::fflush(stdout);
}

static Stepper* stepper_create(sem_mesh* m, int scheme, const double* re, const double* im, const double* V,
                               const sem_solver_opts* o)
{
    if (!(o->dt > 0)) throw std::invalid_argument("schrodinger: dt must be positive");
    if (o->guess < 0 || o->guess > 2) throw std::invalid_argument("schrodinger: unsupported guess mode");
    if (o->guess == 2 && scheme == 0) throw std::invalid_argument("schrodinger: the learned guess serves the SDIRK stages");
    if (o->precision < 0 || o->precision > 1) throw std::invalid_argument("schrodinger: unsupported precision");
    std::unique_ptr<Stepper> st(new Stepper);
    st->m = m;
    st->scheme = scheme;
    st->opts = *o;
    st->n = 2 * m->len();
    st->V = make_field(m, V, "V");
    // the field buffers of make_field live in the mesh pool under a fixed
    // name; copy them into the session so several sessions can coexist
    {
        double* c64 = st->alloc();
        SEM_HIP(hipMemcpyAsync(c64, st->V.c64, m->len() * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
        float* c32 = (float*)st->alloc();
        SEM_HIP(hipMemcpyAsync(c32, st->V.c32, m->len() * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
        st->V.c64 = c64;
        st->V.c32 = c32;
    }
    st->p = st->alloc();
    st->v = st->alloc();
    st->q = st->alloc();
    const int nz = scheme == 0 ? 0 : (scheme == 1 ? 3 : 5);
    for (int k = 0; k < nz; ++k) st->z.push_back(st->alloc());
    if (o->guess >= 1)
        for (int k = 0; k < nz; ++k) {
            st->zprev.push_back(st->alloc());
            st->zprev2.push_back(st->alloc());
        }
    if (o->guess == 2) {
        const int N = m->L.N, NN = N * N;
        st->mlp.reset(new Mlp(7 * NN + kStageCols, 256, 256, 2 * NN, 0x5EED0001ull, m->stream));
        st->mlp->lr = 3e-3f;
        st->mlp->reserve(m->L.E);  // allocated once: the training stream uses these buffers
        SEM_HIP(hipStreamCreateWithFlags(&st->train_stream, hipStreamNonBlocking));
        SEM_HIP(hipEventCreateWithFlags(&st->ev_data, hipEventDisableTiming));
        SEM_HIP(hipEventCreateWithFlags(&st->ev_trained, hipEventDisableTiming));
        st->learn_steps = o->learn_steps > 0 ? o->learn_steps : 50;
        st->ext = st->alloc();
        st->tmp = st->alloc();
        st->rext = st->alloc();
        st->Vraw = st->alloc();
        m->upload(st->Vraw, V, 1);
    }
    if (scheme == 0) {
        st->ps[0] = st->alloc();
        st->ps[1] = st->alloc();
        st->alpha_bdf = o->dt;
    }
    const size_t rl = (size_t)m->L.E * m->L.N * m->L.N;
    std::vector<double> h(2 * rl);
    std::copy(re, re + rl, h.begin());
    std::copy(im, im + rl, h.begin() + rl);
    m->upload(st->p, h.data(), 2);
    // psi arrives unmasked; unmask copies the owner values, the device form is consistent
    dss<double>(m, DSS_UNMASK, st->p, 2);
    if (scheme == 2) {
        st->pnorm = wdot_inst<double>(m, st->p, st->p, 2);
        for (double& v : st->pnorm) v = std::sqrt(v);
    }
    return st.release();
}

// Batched meshes: the stage solutions of the instances left out of this step
// (stopped earlier, or a stage solve failed) are zeroed before the update of
// psi, so psi keeps its value there (schrodinger.hpp:222-226 returns before
// the update). Instance k's entries are one contiguous chunk range per component.
static void zero_skipped(Stepper& st, double* const* z, int nz)
{
    if (!st.multi()) return;
    sem_mesh* m = st.m;
    const size_t cl = (size_t)m->L.K * m->L.N * m->L.N;
    for (int k = 0; k < m->n_inst; ++k) {
        if (!st.skip[k]) continue;
        const size_t off = (size_t)m->h_inst_off[k] * cl;
        const size_t cnt = (size_t)(m->h_inst_off[k + 1] - m->h_inst_off[k]) * cl;
        for (int j = 0; j < nz; ++j)
            for (int c = 0; c < 2; ++c) launch_fill<double>(z[j] + c * m->len() + off, 0.0, cnt, m->stream);
    }
}

// one DIRK4S3 step, schrodinger.hpp:376-406
static bool step_dirk4s3(Stepper& st)
{
    static const double gamma = std::cos(M_PI / 18.0) / std::sqrt(3.0) + 0.5;
    static const double delta = 1.0 / 6.0 / std::pow(2 * gamma - 1, 2);
    sem_mesh* m = st.m;
    const double alpha = gamma * st.opts.dt;
    const size_t n = st.n;
    double *p = st.p, *v = st.v, *q = st.q, **z = st.z.data();
    stage_rhs(st, p, v);
    if (!stage_solve(st, alpha, v, z[0], 0).success) return false;
    lincomb<double>(m, q, 0, {{1.0, p}, {0.5 - gamma, z[0]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[1], 1).success) return false;
    lincomb<double>(m, q, 0, {{1.0, p}, {2 * gamma, z[0]}, {1 - 4 * gamma, z[1]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[2], 2).success) return false;
    zero_skipped(st, z, 3);
    lincomb<double>(m, p, 2, {{delta, z[0]}, {1 - 2 * delta, z[1]}, {delta, z[2]}}, n);
    record_history(st, 3);
    return true;
}

// one DIRK4S5 step, schrodinger.hpp:219-268
static bool step_dirk4s5(Stepper& st)
{
    sem_mesh* m = st.m;
    const double alpha = 0.25 * st.opts.dt;
    const size_t n = st.n;
    double *p = st.p, *v = st.v, *q = st.q, **z = st.z.data();
    stage_rhs(st, p, v);
    if (!stage_solve(st, alpha, v, z[0], 0).success) return false;
    lincomb<double>(m, q, 0, {{1.0, p}, {0.5, z[0]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[1], 1).success) return false;
    lincomb<double>(m, q, 0, {{1.0, p}, {17.0 / 50.0, z[0]}, {-1.0 / 25.0, z[1]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[2], 2).success) return false;
    lincomb<double>(m, q, 0, {{1.0, p}, {371.0 / 1360.0, z[0]}, {-137.0 / 2720.0, z[1]}, {15.0 / 544.0, z[2]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[3], 3).success) return false;
    lincomb<double>(m, q, 0,
                    {{1.0, p}, {25.0 / 24.0, z[0]}, {-49.0 / 48.0, z[1]}, {125.0 / 16.0, z[2]}, {-85.0 / 12.0, z[3]}}, n);
    stage_rhs(st, q, v);
    if (!stage_solve(st, alpha, v, z[4], 4).success) return false;
    zero_skipped(st, z, 5);
    lincomb<double>(m, p, 2,
                    {{25.0 / 24.0, z[0]}, {-49.0 / 48.0, z[1]}, {125.0 / 16.0, z[2]}, {-85.0 / 12.0, z[3]}, {0.25, z[4]}},
                    n);
    lincomb<double>(m, q, 1, {{-3.0 / 16.0, z[0]}, {-27.0 / 32.0, z[1]}, {25.0 / 32.0, z[2]}, {0.25, z[4]}}, n);
    // the error estimate of every instance that completed the step; a batch reports the largest
    const std::vector<double> eq = wdot_inst<double>(m, q, q, 2);
    double est = 0;
    for (int k = 0; k < m->n_inst; ++k)
        if (!st.multi() || !st.skip[k]) est = std::max(est, std::sqrt(eq[k]) / st.pnorm[k]);
    st.stats.last_error_estimate = est;
    record_history(st, 5);
    return true;
}

// one BDF2 step, schrodinger.hpp:91-102 (implicit Euler first); the
// previous psi is the initial guess (solved in place)
static bool step_bdf2(Stepper& st)
{
    sem_mesh* m = st.m;
    const size_t len = m->len(), n = st.n;
    double* p = st.p;
    if (st.nps == 2) std::swap(st.ps[0], st.ps[1]);
    double* slot = st.ps[st.nps == 2 ? 1 : st.nps];
    // proj(p) = [gproj(re), -gproj(im)] (schrodinger.hpp:69-78)
    apply_op<double>(m, K_MASS, p, slot, nullptr, 0.0, 1.0);
    apply_op<double>(m, K_MASS, p + len, slot + len, nullptr, 0.0, -1.0);
    if (st.nps < 2) ++st.nps;
    Results r;
    if (st.nps == 1) {
        r = solve_res(st, st.alpha_bdf, st.ps[0], p, false, false);
        st.alpha_bdf = (2.0 / 3.0) * st.opts.dt;
    } else {
        lincomb<double>(m, st.v, 1, {{-1.0 / 3.0, st.ps[0]}, {4.0 / 3.0, st.ps[1]}}, n);
        r = solve_res(st, st.alpha_bdf, st.v, p, false, false);
    }
    return r.success != 0;
}

// while (t < T) loop of the steppers; max_steps > 0 also bounds the count.
static bool stepper_advance(Stepper& st, double T, long max_steps, sem_step_cb cb, void* user)
{
    bool success = true;
    long k = 0;
    const double dt = st.opts.dt;
    const int ni = st.m->n_inst;
    if (st.multi() && st.skip.size() != (size_t)ni) {
        st.skip.assign(ni, 0);
        st.frozen.assign(ni, 0);
        st.failed_now.assign(ni, 0);
        st.inst_t.assign(ni, st.t);
        st.inst_steps.assign(ni, st.stats.steps);
    }
    while (st.t < T && (max_steps <= 0 || k < max_steps)) {
        bool ok;
        if (st.multi()) {
            bool active = false;
            for (int q = 0; q < ni; ++q) active |= !st.frozen[q];
            if (!active) break;
        }
        if (st.scheme == 0) ok = step_bdf2(st);
        else if (st.scheme == 1) ok = step_dirk4s3(st);
        else ok = step_dirk4s5(st);
        if (st.multi()) {
            // per-instance bookkeeping of the reference's loop; the group's clock
            // runs while any instance is still stepping
            for (int q = 0; q < ni; ++q) {
                if (st.frozen[q]) continue;
                if (st.scheme == 0 || !st.failed_now[q]) {
                    st.inst_t[q] += dt;
                    st.inst_steps[q]++;
                }
                if (st.failed_now[q]) st.frozen[q] = 1;
                st.failed_now[q] = 0;
                st.skip[q] = st.frozen[q];
            }
            st.t += dt;
            st.stats.steps++;
            ++k;
            if (st.opts.verbose) progress(st.t, T, dt, st.scheme == 2 ? &st.stats.last_error_estimate : nullptr);
            if (do_callback(st, cb, user)) break;
            continue;
        }
        if (st.scheme == 0) {
            // bdf2 counts the step before testing success (schrodinger.hpp:107-114)
            st.t += dt;
            st.stats.steps++;
            ++k;
            if (st.opts.verbose) progress(st.t, T, dt, nullptr);
            if (!ok) { success = false; break; }
        } else {
            if (!ok) { success = false; break; }
            st.t += dt;
            st.stats.steps++;
            ++k;
            if (st.opts.verbose) progress(st.t, T, dt, st.scheme == 2 ? &st.stats.last_error_estimate : nullptr);
        }
        if (do_callback(st, cb, user)) break;
    }
    if (st.opts.verbose) std::printf("\n");
    st.m->sync();
    if (st.multi())
        for (int q = 0; q < ni; ++q) success = success && !st.frozen[q];
    st.finish_training();  // the window's last Adam steps are part of this call
    return success;
}

static void stepper_read(Stepper& st, double* re, double* im)
{
    sem_mesh* m = st.m;
    const size_t rl = (size_t)m->L.E * m->L.N * m->L.N;
    std::vector<double> h(2 * rl);
    m->download(h.data(), st.p, 2);
    if (re) std::copy(h.begin(), h.begin() + rl, re);
    if (im) std::copy(h.begin() + rl, h.end(), im);
}

// ------------------------------------------------- checkpoint and resume --
// A session's checkpoint is one host array of doubles: a header, the learned
// guess's per-stage scales, psi, the BDF2 history (the two projected states
// of schrodinger.hpp:85-102), the stage solutions of the previous two steps
// (extrapolated and learned guesses) and the guess network with its Adam
// moments. Every vector is a two-component array in the reference's
// element-local layout holding the node's value on every copy, so a
// save/restore round trip is a permutation of the device arrays, bit for bit.
constexpr double kCkptMagic = 5.0293e6;
constexpr int kCkptHeader = 16;
constexpr int kCkptScales = 5 * 6;

static size_t ckpt_size(const Stepper& st)
{
    const size_t v = 2 * (size_t)st.m->L.E * st.m->L.N * st.m->L.N;
    size_t n = kCkptHeader + kCkptScales + v;
    if (st.scheme == 0) n += 2 * v;
    n += 2 * st.zprev.size() * v;
    if (st.mlp) n += 3 * st.mlp->nparams;
    return n;
}

static void stepper_save(Stepper& st, double* out)
{
    sem_mesh* m = st.m;
    const size_t v = 2 * (size_t)m->L.E * m->L.N * m->L.N;
    double* h = out;
    std::fill(h, h + kCkptHeader, 0.0);
    h[0] = kCkptMagic;
    h[1] = 1;
    h[2] = st.scheme;
    h[3] = m->L.E;
    h[4] = m->L.N;
    h[5] = st.t;
    h[6] = st.stats.steps;
    h[7] = st.nps;
    h[8] = st.alpha_bdf;
    h[9] = st.pnorm.empty() ? 0.0 : st.pnorm[0];  // sessions on plain meshes (one instance)
    h[10] = st.have_prev;
    h[11] = st.have_prev2;
    h[12] = (double)st.zprev.size();
    h[13] = st.mlp ? 1 : 0;
    h[14] = st.mlp ? (double)st.mlp->nparams : 0;
    h[15] = st.mlp ? (double)st.mlp->t : 0;
    double* s = h + kCkptHeader;
    for (int k = 0; k < 5; ++k) {
        for (int q = 0; q < 3; ++q) s[6 * k + q] = st.s_feat[k][q];
        s[6 * k + 3] = st.s_out[k];
        s[6 * k + 4] = st.have_scale[k];
        s[6 * k + 5] = st.have_out[k];
    }
    double* d = s + kCkptScales;
    m->download(d, st.p, 2);
    d += v;
    if (st.scheme == 0) {
        for (int q = 0; q < 2; ++q, d += v) {
            if (q < st.nps) m->download(d, st.ps[q], 2);
            else std::fill(d, d + v, 0.0);
        }
    }
    for (size_t k = 0; k < st.zprev.size(); ++k) {
        m->download(d, st.zprev[k], 2);
        d += v;
        m->download(d, st.zprev2[k], 2);
        d += v;
    }
    if (st.mlp) {
        const size_t np = st.mlp->nparams;
        std::vector<float> f(np);
        const float* src[3] = {st.mlp->params, st.mlp->adam_m, st.mlp->adam_v};
        for (int q = 0; q < 3; ++q, d += np) {
            SEM_HIP(hipMemcpy(f.data(), src[q], np * sizeof(float), hipMemcpyDeviceToHost));
            std::copy(f.begin(), f.end(), d);
        }
    }
}

// Restores what this session keeps from a checkpoint of the same scheme and
// mesh size: time, step count, psi and the BDF2 history always; the stage
// history when this session extrapolates its guesses; the network when this
// session has one of the same size.
static void stepper_restore(Stepper& st, const double* in, size_t n)
{
    sem_mesh* m = st.m;
    const size_t v = 2 * (size_t)m->L.E * m->L.N * m->L.N;
    const double* h = in;
    if (n < (size_t)kCkptHeader + kCkptScales + v || h[0] != kCkptMagic || h[1] != 1)
        throw std::invalid_argument("sem_stepper_restore: not a stepper checkpoint");
    if ((int)h[2] != st.scheme) throw std::invalid_argument("sem_stepper_restore: checkpoint of another scheme");
    if ((int)h[3] != m->L.E || (int)h[4] != m->L.N)
        throw std::invalid_argument("sem_stepper_restore: checkpoint of another mesh size");
    const size_t nz = (size_t)h[12], np = (size_t)h[14];
    size_t need = kCkptHeader + kCkptScales + v + (st.scheme == 0 ? 2 * v : 0) + 2 * nz * v + (h[13] != 0 ? 3 * np : 0);
    if (n < need) throw std::invalid_argument("sem_stepper_restore: checkpoint array too short");
    st.t = h[5];
    st.stats.steps = (int)h[6];
    st.nps = (int)h[7];
    st.alpha_bdf = h[8];
    st.pnorm.assign(m->n_inst, h[9]);
    const double* s = h + kCkptHeader;
    const double* d = s + kCkptScales;
    m->upload(st.p, d, 2);
    d += v;
    if (st.scheme == 0) {
        for (int q = 0; q < 2; ++q, d += v)
            if (q < st.nps) m->upload(st.ps[q], d, 2);
    }
    const bool hist = !st.zprev.empty() && st.zprev.size() == nz;
    for (size_t k = 0; k < nz; ++k) {
        if (hist) {
            m->upload(st.zprev[k], d, 2);
            m->upload(st.zprev2[k], d + v, 2);
        }
        d += 2 * v;
    }
    st.have_prev = hist && h[10] != 0;
    st.have_prev2 = hist && h[11] != 0;
    if (st.mlp && h[13] != 0 && np == st.mlp->nparams) {
        for (int k = 0; k < 5; ++k) {
            for (int q = 0; q < 3; ++q) st.s_feat[k][q] = s[6 * k + q];
            st.s_out[k] = s[6 * k + 3];
            st.have_scale[k] = s[6 * k + 4] != 0;
            st.have_out[k] = s[6 * k + 5] != 0;
        }
        std::vector<float> f(np);
        float* dst[3] = {st.mlp->params, st.mlp->adam_m, st.mlp->adam_v};
        for (int q = 0; q < 3; ++q, d += np) {
            for (size_t i = 0; i < np; ++i) f[i] = (float)d[i];
            SEM_HIP(hipMemcpy(dst[q], f.data(), np * sizeof(float), hipMemcpyHostToDevice));
        }
        st.mlp->t = (long)h[15];
    }
    m->sync();
}

// -------------------------------------------------------------- batches --
// Independent meshes advanced together. Instances of one degree share a union
// mesh: the instance meshes side by side (element and node ids offset, so
// edge owners, orientations and corner orders are those of each instance),
// each padded with ghost elements (a far-away unit square with zero geometry,
// every side on the Dirichlet boundary) to a whole number of kChunkAlign-chunk
// work items. Every workgroup of the Lanczos-step kernels then serves one
// instance, every instance keeps its own Krylov scalars and reductions over
// its own chunks, and each instance computes bit for bit what it computes
// alone. Groups of different degree run on their own streams from their own
// host threads.
static sem_mesh* make_union_mesh(const std::vector<const sem_mesh*>& parts)
{
    const int p = parts[0]->hm.p, N = p + 1, K = 64 / N, align = K * kChunkAlign;
    double far = 0;
    for (const sem_mesh* s : parts)
        for (double v : s->hm.node_xy) far = std::max(far, std::fabs(v));
    const double X0 = 4.0 * far + 16.0;
    std::vector<double> xy;
    std::vector<int> quads, el_off{0}, el_count;
    std::vector<uint8_t> ghost;
    int n_ghost = 0;
    for (const sem_mesh* s : parts) {
        const HostMesh& h = s->hm;
        if (h.p != p) throw std::invalid_argument("internal error: mixed degrees in a union mesh");
        const int base = (int)(xy.size() / 2);
        xy.insert(xy.end(), h.node_xy.begin(), h.node_xy.end());
        for (int v : h.quads) quads.push_back(v + base);
        ghost.insert(ghost.end(), h.n_el, 0);
        const int E = h.n_el, Ep = (E + align - 1) / align * align;
        for (int g = E; g < Ep; ++g, ++n_ghost) {
            const int b = (int)(xy.size() / 2);
            const double x0 = X0 + 2.0 * n_ghost, y0 = X0;
            const double c[8] = {x0, y0, x0 + 1, y0, x0 + 1, y0 + 1, x0, y0 + 1};
            xy.insert(xy.end(), c, c + 8);
            for (int q = 0; q < 4; ++q) quads.push_back(b + q);
            ghost.push_back(1);
        }
        el_count.push_back(E);
        el_off.push_back(el_off.back() + Ep);
    }
    std::unique_ptr<sem_mesh> m(new sem_mesh);
    build_host_mesh(m->hm, p, xy.data(), (int)(xy.size() / 2), quads.data(), (int)(quads.size() / 4));
    m->n_inst = (int)parts.size();
    m->inst_el_off = el_off;
    m->inst_el_count = el_count;
    m->ghost = ghost;
    m->use_graphs = true;  // the Lanczos-step batches of a batched mesh replay as graphs
    setup_device(m.get());
    return m.release();
}

}  // namespace sem

struct sem_batch {
    struct Group {
        int p = 0;
        std::vector<int> members;  // batch instance ids, in union order
        std::unique_ptr<sem_mesh> m;
    };
    std::vector<Group> groups;
    std::vector<int> inst_group, inst_slot;
    std::vector<long> inst_ref_len;  // E * N * N of each instance
    std::vector<std::unique_ptr<sem::Stepper>> session;  // one stepper session per group
    sem_run_stats stats{};
    double seconds = 0;
    ~sem_batch()
    {
        session.clear();  // before the group meshes
    }
};

namespace sem {

// instance k's reference-layout values (ncomp components of E_k*N*N) into the
// union's reference layout (ghost entries zero), and back
static void pack_union(const sem_mesh* um, int slot, const double* src, std::vector<double>& dst, int ncomp)
{
    const size_t NN = (size_t)um->L.N * um->L.N, rl = (size_t)um->L.E * NN;
    const size_t e0 = (size_t)um->inst_el_off[slot], ek = (size_t)um->inst_el_count[slot];
    for (int c = 0; c < ncomp; ++c) std::copy(src + c * ek * NN, src + (c + 1) * ek * NN, dst.begin() + c * rl + e0 * NN);
}

static void unpack_union(const sem_mesh* um, int slot, const std::vector<double>& src, double* dst, int ncomp)
{
    const size_t NN = (size_t)um->L.N * um->L.N, rl = (size_t)um->L.E * NN;
    const size_t e0 = (size_t)um->inst_el_off[slot], ek = (size_t)um->inst_el_count[slot];
    for (int c = 0; c < ncomp; ++c)
        std::copy(src.begin() + c * rl + e0 * NN, src.begin() + c * rl + (e0 + ek) * NN, dst + c * ek * NN);
}

// a stepper session (scheme 0 schrodinger_bdf2, schrodinger.hpp:22-128;
// 1 dirk4s3, :294-423; 2 dirk4s5, :138-285) on every instance of one group
static Stepper* batch_group_start(sem_batch::Group& g, int scheme, const double* const* re, const double* const* im,
                                  const double* const* V, const sem_solver_opts* o)
{
    sem_mesh* um = g.m.get();
    const size_t rl = (size_t)um->L.E * um->L.N * um->L.N;
    std::vector<double> ur(rl, 0.0), ui(rl, 0.0), uv(rl, 0.0);
    for (size_t s = 0; s < g.members.size(); ++s) {
        const int k = g.members[s];
        pack_union(um, (int)s, re[k], ur, 1);
        pack_union(um, (int)s, im[k], ui, 1);
        pack_union(um, (int)s, V[k], uv, 1);
    }
    Stepper* st = stepper_create(um, scheme, ur.data(), ui.data(), uv.data(), o);
    st->inst_ok.assign(um->n_inst, 1);
    return st;
}

static void batch_group_read(sem_batch::Group& g, Stepper& st, double* const* re, double* const* im, int* ok)
{
    sem_mesh* um = g.m.get();
    const size_t rl = (size_t)um->L.E * um->L.N * um->L.N;
    std::vector<double> ur(rl), ui(rl);
    stepper_read(st, ur.data(), ui.data());
    for (size_t s = 0; s < g.members.size(); ++s) {
        const int k = g.members[s];
        if (re) unpack_union(um, (int)s, ur, re[k], 1);
        if (im) unpack_union(um, (int)s, ui, im[k], 1);
        if (ok) ok[k] = st.inst_ok[s];
    }
}

// ------------------------------------------------------- generic solvers --
// minres.hpp:19-107 and pcg.hpp:12-97 over device vectors of n doubles with a
// caller's operator, dot product and preconditioner (the reference's LinOp, Dot
// and Precond template arguments); host-driven, one scalar readback per dot.
struct GenericCtx {
    hipStream_t s = nullptr;
    long n = 0;
    double* part = nullptr;
    double* red = nullptr;
    double* h_red = nullptr;
    std::function<void(const double*, double*)> A;             // y = A(x)
    std::function<double(const double*, const double*)> dotf;  // empty: Euclidean dot on the device
    std::function<void(const double*, double*)> P;             // z = precond(r); empty: IdentityPreconditioner
    // MINRES with a preconditioner: beta = sqrt(dot(y, y)) as minres.hpp:63
    // computes it (false), or sqrt(dot(r2, y)), the preconditioned Lanczos
    // normalisation that matches beta1 = sqrt(dot(y, r1)) of minres.hpp:31 (true)
    bool consistent = false;
    double dot(const double* a, const double* b)
    {
        if (dotf) return dotf(a, b);
        int nb = launch_edot(a, b, (size_t)n, part, s);
        launch_fin_sum(part, nb, red, s);
        SEM_HIP(hipMemcpyAsync(h_red, red, sizeof(double), hipMemcpyDeviceToHost, s));
        SEM_HIP(hipStreamSynchronize(s));
        return *h_red;
    }
    void comb(double* out, int mode, std::initializer_list<std::pair<double, const double*>> t)
    {
        LinComb<double> a{};
        a.out = out;
        a.mode = mode;
        a.n = (size_t)n;
        int k = 0;
        for (auto& pr : t) {
            a.c[k] = pr.first;
            a.x[k] = pr.second;
            ++k;
        }
        a.m = k;
        launch_lincomb<double>(a, s);
    }
};

// minres.hpp:19-107; buf holds 8n doubles
static Results generic_minres(GenericCtx& g, double* x, const double* b, int maxit, double tol, double* buf)
{
    const long n = g.n;
    double *r1 = buf, *r2 = buf + n, *y = buf + 2 * n, *v = buf + 3 * n, *w = buf + 4 * n, *w1 = buf + 5 * n,
           *w2 = buf + 6 * n, *t = buf + 7 * n;
    g.A(x, t);
    g.comb(r1, 0, {{1.0, b}, {-1.0, t}});
    if (g.P) g.P(r1, y);
    else g.comb(y, 0, {{1.0, r1}});
    g.comb(r2, 0, {{1.0, r1}});
    const double b_norm = std::sqrt(g.dot(r1, r1));
    const double beta1 = std::sqrt(g.dot(y, r1));
    double beta = beta1, beta_old = 0, tnorm = 0, dbar = 0, eps = 0, eps_old = 0, phibar = beta1, cs = -1, sn = 0;
    (void)tnorm;
    g.comb(w, 1, {{0.0, x}});
    g.comb(w2, 0, {{1.0, w}});
    bool success = false;
    int i;
    for (i = 1; i <= maxit; ++i) {
        g.comb(v, 1, {{1.0 / beta, y}});
        g.A(v, y);
        if (i >= 2) g.comb(y, 0, {{1.0, y}, {-(beta / beta_old), r1}});
        const double alpha = g.dot(v, y);
        g.comb(y, 0, {{1.0, y}, {-(alpha / beta), r2}});
        std::swap(r1, r2);  // r1 = std::move(r2)
        g.comb(r2, 0, {{1.0, y}});
        if (g.P) g.P(r2, y);  // y = precond(r2); with the identity y already equals r2
        beta_old = beta;
        beta = std::sqrt(g.consistent ? g.dot(r2, y) : g.dot(y, y));
        tnorm += alpha * alpha + beta * beta + beta_old * beta_old;
        eps_old = eps;
        const double delta = cs * dbar + sn * alpha;
        const double gbar = sn * dbar - cs * alpha;
        eps = sn * beta;
        dbar = -cs * beta;
        const double gamma = std::max(std::sqrt(gbar * gbar + beta * beta), 2.220446049250313e-16);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar *= sn;
        std::swap(w1, w2);  // w1 <- w2
        std::swap(w2, w);   // w2 <- w
        g.comb(w, 1, {{1.0 / gamma, v}, {-eps_old / gamma, w1}, {-delta / gamma, w2}});
        g.comb(x, 2, {{phi, w}});
        if (phibar < tol * b_norm + tol) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, phibar};
}

// pcg.hpp:59-97 with the conj_grad steps of pcg.hpp:12-55; buf holds 4n doubles
static Results generic_pcg(GenericCtx& g, double* x, const double* b, int maxit, double tol, double* buf)
{
    const long n = g.n;
    double *r = buf, *p = buf + n, *ap = buf + 2 * n, *z = buf + 3 * n;
    g.A(x, ap);
    g.comb(r, 0, {{1.0, b}, {-1.0, ap}});
    const double bnorm = std::sqrt(g.dot(b, b));
    bool success = false;
    double rho = 0, rho_prev = 0;
    int i;
    for (i = 0; i < maxit; ++i) {
        if (i == 0) {
            if (g.P) {
                g.P(r, p);
                rho = g.dot(r, p);
            } else {
                g.comb(p, 0, {{1.0, r}});
                rho = g.dot(r, r);
            }
        } else if (g.P) {
            g.P(r, z);
            rho = g.dot(r, z);
            g.comb(p, 0, {{1.0, z}, {rho / rho_prev, p}});
        } else {
            rho = g.dot(r, r);
            g.comb(p, 0, {{1.0, r}, {rho / rho_prev, p}});
        }
        g.A(p, ap);
        const double alpha = rho / g.dot(p, ap);
        g.comb(x, 2, {{alpha, p}});
        g.comb(r, 2, {{-alpha, ap}});
        rho_prev = rho;
        if (std::sqrt(rho) < tol * bnorm) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, std::sqrt(rho)};
}

// Device workspace of one generic solve (freed on scope exit).
struct GenericWork {
    std::vector<void*> dev;
    double* h_red = nullptr;
    ~GenericWork()
    {
        for (void* p : dev) hipFree(p);
        if (h_red) hipHostFree(h_red);
    }
    double* alloc(size_t n)
    {
        void* p = nullptr;
        SEM_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)));
        dev.push_back(p);
        return (double*)p;
    }
    void init(GenericCtx& g)
    {
        g.part = alloc(1024);
        g.red = alloc(1);
        SEM_HIP(hipHostMalloc(&h_red, sizeof(double), hipHostMallocDefault));
        g.h_red = h_red;
    }
};

}  // namespace sem

// ================================================================ C-ABI ==
namespace {

template <typename F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const HipError& e) {
        set_error(std::string("HIP error ") + hipGetErrorString(e.code) + " in " + e.where);
        return SEM_EHIP;
    } catch (const std::bad_alloc&) {
        set_error("out of memory");
        return SEM_ENOMEM;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return SEM_EINVAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SEM_EINVAL;
    }
}

bool check_gpu()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
        set_error("no HIP device available (libsem_hip needs an MI355X / gfx950 GPU)");
        return false;
    }
    return true;
}

size_t ref_len(const sem_mesh* m) { return (size_t)m->L.E * m->L.N * m->L.N; }

void fill_results(sem_solver_results* out, const Results& r)
{
    if (!out) return;
    out->success = r.success;
    out->n_iter = r.n_iter;
    out->residual = r.residual;
}

}  // namespace

extern "C" {

const char* sem_last_error(void) { return g_error.c_str(); }
int sem_version(void) { return 100; }

int sem_progress_line(double t, double T, double dt, int with_error, double err, char* out, int out_len)
{
    if (!out || out_len < 1 || !(dt > 0) || !(T > 0)) return SEM_EINVAL;
    const std::string s = progress_line(t, T, dt, with_error ? &err : nullptr);
    std::snprintf(out, (size_t)out_len, "%s", s.c_str());
    return (int)s.size() < out_len ? SEM_OK : SEM_EINVAL;
}

int sem_device_info(char* name, int name_len, int* n_cu, long* hbm_bytes)
{
    return guarded([&] {
        if (!check_gpu()) return (int)SEM_EHIP;
        hipDeviceProp_t p;
        SEM_HIP(hipGetDeviceProperties(&p, 0));
        if (name && name_len > 0) std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
        if (n_cu) *n_cu = p.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = (long)p.totalGlobalMem;
        return (int)SEM_OK;
    });
}

int sem_mesh_create(int p, const double* xy, int n_nodes, const int* quads, int n_elements, sem_mesh** out)
{
    return guarded([&] {
        if (!out || !xy || !quads) throw std::invalid_argument("sem_mesh_create: null argument");
        *out = nullptr;
        if (!check_gpu()) return (int)SEM_EHIP;
        std::unique_ptr<sem_mesh> m(new sem_mesh);
        build_host_mesh(m->hm, p, xy, n_nodes, quads, n_elements);
        setup_device(m.get());
        *out = m.release();
        return (int)SEM_OK;
    });
}

int sem_mesh_load(int p, const char* folder, sem_mesh** out)
{
    std::vector<double> xy;
    std::vector<int> q;
    try {
        if (!folder) throw std::invalid_argument("null folder");
        read_mesh_folder(folder, xy, q);
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return SEM_EINVAL;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SEM_EIO;
    }
    return sem_mesh_create(p, xy.data(), (int)xy.size() / 2, q.data(), (int)q.size() / 4, out);
}

void sem_mesh_destroy(sem_mesh* m) { delete m; }

int sem_mesh_info(const sem_mesh* m, long* out)
{
    if (!m || !out) return SEM_EINVAL;
    out[0] = m->hm.N;
    out[1] = m->hm.n_el;
    out[2] = m->hm.n_nodes;
    out[3] = (long)m->hm.edges.size();
    out[4] = m->hm.dof;
    out[5] = m->L.K;
    return SEM_OK;
}

int sem_mesh_quadrature(const sem_mesh* m, double* x, double* w, double* D)
{
    if (!m) return SEM_EINVAL;
    if (x) std::copy(m->hm.x.begin(), m->hm.x.end(), x);
    if (w) std::copy(m->hm.w.begin(), m->hm.w.end(), w);
    if (D) std::copy(m->hm.D.begin(), m->hm.D.end(), D);
    return SEM_OK;
}

int sem_mesh_coords(const sem_mesh* m, double* X, double* Y)
{
    if (!m || !X || !Y) return SEM_EINVAL;
    const HostMesh& hm = m->hm;
    const int N = hm.N;
    for (int e = 0; e < hm.n_el; ++e) {
        double xs[4], ys[4];
        for (int k = 0; k < 4; ++k) {
            xs[k] = hm.node_xy[2 * hm.quads[4 * e + k]];
            ys[k] = hm.node_xy[2 * hm.quads[4 * e + k] + 1];
        }
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                const double xi = hm.x[i], eta = hm.x[j];
                const size_t k = (size_t)e * N * N + i + (size_t)j * N;
                X[k] = 0.25 * (xs[0] * (1 - xi) * (1 - eta) + xs[1] * (1 + xi) * (1 - eta) + xs[2] * (1 + xi) * (1 + eta) +
                               xs[3] * (1 - xi) * (1 + eta));
                Y[k] = 0.25 * (ys[0] * (1 - xi) * (1 - eta) + ys[1] * (1 + xi) * (1 - eta) + ys[2] * (1 + xi) * (1 + eta) +
                               ys[3] * (1 - xi) * (1 + eta));
            }
    }
    return SEM_OK;
}

int sem_mesh_metrics(const sem_mesh* m, double* J, double* A, double* B, double* C)
{
    if (!m) return SEM_EINVAL;
    if (J) std::copy(m->hm.J.begin(), m->hm.J.end(), J);
    if (A) std::copy(m->hm.A.begin(), m->hm.A.end(), A);
    if (B) std::copy(m->hm.B.begin(), m->hm.B.end(), B);
    if (C) std::copy(m->hm.C.begin(), m->hm.C.end(), C);
    return SEM_OK;
}

int sem_element_apply(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out)
{
    return guarded([&] {
        if (!m || !in || !out || kind < 0 || kind > 5) throw std::invalid_argument("sem_element_apply: bad argument");
        if ((kind == K_HELMHOLTZ || kind == K_RES || kind == K_F || kind == K_L) && !coef)
            throw std::invalid_argument("sem_element_apply: this operator needs a coefficient field");
        const int nc = op_ncomp(kind);
        const size_t n = (size_t)nc * m->len();
        double* din = m->buf<double>("api_in", n);
        double* dout = m->buf<double>("api_out", n);
        m->upload(din, in, nc);
        Field f;
        if (coef) f = make_field(m, coef, "api_cf");
        OpArgs<double> a{};
        a.in = din;
        a.out = dout;
        a.geo = m->geo;
        a.cf = f.c64;
        a.alpha = alpha;
        a.post = 1.0;
        a.len = m->len();
        a.nchunks = m->L.nchunks;
        launch_element_op_d<double>(m->L.N, kind, a, m->D64.data(), m->stream);
        m->download(out, dout, nc);
        return (int)SEM_OK;
    });
}

int sem_element_apply_fp32(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out)
{
    return guarded([&] {
        if (!m || !in || !out || kind < 0 || kind > 5) throw std::invalid_argument("sem_element_apply_fp32: bad argument");
        if ((kind == K_HELMHOLTZ || kind == K_RES || kind == K_F || kind == K_L) && !coef)
            throw std::invalid_argument("sem_element_apply_fp32: this operator needs a coefficient field");
        const int nc = op_ncomp(kind);
        const size_t n = (size_t)nc * m->len();
        double* din = m->buf<double>("api_in", n);
        double* dout = m->buf<double>("api_out", n);
        float* fin = m->buf<float>("api_in32", n);
        float* fout = m->buf<float>("api_out32", n);
        m->upload(din, in, nc);
        launch_cast<float, double>(fin, din, n, m->stream);
        Field f;
        if (coef) f = make_field(m, coef, "api_cf");
        OpArgs<float> a{};
        a.in = fin;
        a.out = fout;
        a.geo = m->geo32;
        a.qm = m->qm;
        a.cf = f.c32;
        a.alpha = (float)alpha;
        a.post = 1.0f;
        a.len = m->len();
        a.nchunks = m->L.nchunks;
        a.mf = kind == K_RES ? m->contraction : 0;
        launch_element_op_d<float>(m->L.N, kind, a, m->D32.data(), m->stream);
        launch_cast<double, float>(dout, fout, n, m->stream);
        m->download(out, dout, nc);
        return (int)SEM_OK;
    });
}

int sem_set_contraction(sem_mesh* m, int mode)
{
    if (!m || (mode != 0 && mode != 1 && mode != 3)) {
        set_error("sem_set_contraction: mode is 0 (VALU), 1 (bf16 MFMA) or 3 (split-bf16 MFMA)");
        return SEM_EINVAL;
    }
    m->contraction = mode;
    return SEM_OK;
}

int sem_set_stored_basis(sem_mesh* m, int on, long budget_bytes)
{
    if (!m || budget_bytes < 0) {
        set_error("sem_set_stored_basis: bad argument");
        return SEM_EINVAL;
    }
    m->basis = on != 0;
    if (budget_bytes > 0) m->basis_budget = (size_t)budget_bytes;
    return SEM_OK;
}

int sem_get_contraction(const sem_mesh* m, int* mode)
{
    if (!m || !mode) return SEM_EINVAL;
    *mode = m->contraction;
    return SEM_OK;
}

int sem_glaplace_apply(sem_mesh* m, const double* u, double* y)
{
    return sem_element_apply(m, SEM_OP_LAPLACE, nullptr, 0.0, u, y);
}

int sem_galerkin_apply(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out)
{
    return guarded([&] {
        if (!m || !in || !out || kind < 0 || kind > 5) throw std::invalid_argument("sem_galerkin_apply: bad argument");
        if ((kind == K_HELMHOLTZ || kind == K_RES || kind == K_F || kind == K_L) && !coef)
            throw std::invalid_argument("sem_galerkin_apply: this operator needs a coefficient field");
        const int nc = op_ncomp(kind);
        const size_t n = (size_t)nc * m->len();
        double* din = m->buf<double>("api_in", n);
        double* dout = m->buf<double>("api_out", n);
        m->upload(din, in, nc);
        Field f;
        if (coef) f = make_field(m, coef, "api_cf");
        dss<double>(m, DSS_UNMASK, din, nc);
        apply_op<double>(m, kind, din, dout, f.c64, alpha);
        dss<double>(m, DSS_MASK, dout, nc);
        m->download(out, dout, nc);
        return (int)SEM_OK;
    });
}

int sem_gproj(sem_mesh* m, const double* u, double* out)
{
    return sem_galerkin_apply(m, SEM_OP_MASS, nullptr, 0.0, u, out);
}

static int dss_host(sem_mesh* m, double* a, int mode)
{
    return guarded([&] {
        if (!m || !a) throw std::invalid_argument("null argument");
        double* d = m->buf<double>("api_in", m->len());
        m->upload(d, a, 1);
        dss<double>(m, mode, d, 1);
        m->download(a, d, 1);
        return (int)SEM_OK;
    });
}

int sem_unmask(sem_mesh* m, double* a) { return dss_host(m, a, DSS_UNMASK); }
int sem_global_sum(sem_mesh* m, double* a) { return dss_host(m, a, DSS_SUM); }
int sem_mask(sem_mesh* m, double* a) { return dss_host(m, a, DSS_MASK); }

int sem_dot(sem_mesh* m, const double* a, const double* b, double* out)
{
    return guarded([&] {
        if (!m || !a || !b || !out) throw std::invalid_argument("null argument");
        double* da = m->buf<double>("api_in", m->len());
        double* db = m->buf<double>("api_out", m->len());
        m->upload(da, a, 1);
        m->upload(db, b, 1);
        *out = wdot_host<double>(m, da, db, 1);
        return (int)SEM_OK;
    });
}

int sem_poisson(sem_mesh* m, double* u, const double* f, int max_iter, double tol, sem_solver_results* out)
{
    return guarded([&] {
        if (!m || !u || !f) throw std::invalid_argument("sem_poisson: null argument");
        const size_t n = m->len();
        double* b = m->buf<double>("ps_b", n);
        double* x = m->buf<double>("ps_x", n);
        m->upload(b, f, 1);
        dss<double>(m, DSS_UNMASK, b, 1);
        double* bb = m->buf<double>("ps_bb", n);
        apply_op<double>(m, K_MASS, b, bb, nullptr, 0.0);  // gproj(f)
        m->upload(x, u, 1);
        dss<double>(m, DSS_UNMASK, x, 1);
        Results r = pcg_fused(m, bb, x, false, max_iter, tol);
        merge_owner(m, x, u);
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

int sem_helmholtz(sem_mesh* m, double* u, const double* c, const double* f, int max_iter, double tol,
                  sem_solver_results* out)
{
    return guarded([&] {
        if (!m || !u || !c || !f) throw std::invalid_argument("sem_helmholtz: null argument");
        const size_t n = m->len();
        double* b = m->buf<double>("hh_b", n);
        double* x = m->buf<double>("hh_x", n);
        m->upload(b, f, 1);
        dss<double>(m, DSS_UNMASK, b, 1);
        double* bb = m->buf<double>("hh_bb", n);
        apply_op<double>(m, K_MASS, b, bb, nullptr, 0.0);
        Field cf = make_field(m, c, "hh_c");
        m->upload(x, u, 1);
        dss<double>(m, DSS_UNMASK, x, 1);
        KrylovOpts o;
        o.maxit = max_iter;
        o.tol = tol;
        o.atol = tol;
        Results r = minres_fused<double>(m, K_HELMHOLTZ, cf.c64, 0.0, bb, x, false, o);
        merge_owner(m, x, u);
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

static int run_stepper(int scheme, sem_mesh* m, double* re, double* im, const double* V, double T,
                       const sem_solver_opts* opts, sem_step_cb cb, void* user, int* success)
{
    return guarded([&] {
        if (!m || !re || !im || !V || !opts) throw std::invalid_argument("schrodinger: null argument");
        std::unique_ptr<Stepper> st(stepper_create(m, scheme, re, im, V, opts));
        const bool ok = stepper_advance(*st, T, 0, cb, user);
        stepper_read(*st, re, im);
        m->stats = st->stats;
        if (success) *success = ok ? 1 : 0;
        return (int)SEM_OK;
    });
}

int sem_schrodinger_bdf2(sem_mesh* m, double* re, double* im, const double* V, double T, const sem_solver_opts* opts,
                         sem_step_cb cb, void* user, int* success)
{
    return run_stepper(0, m, re, im, V, T, opts, cb, user, success);
}

int sem_schrodinger_dirk4s3(sem_mesh* m, double* re, double* im, const double* V, double T,
                            const sem_solver_opts* opts, sem_step_cb cb, void* user, int* success)
{
    return run_stepper(1, m, re, im, V, T, opts, cb, user, success);
}

int sem_schrodinger_dirk4s5(sem_mesh* m, double* re, double* im, const double* V, double T,
                            const sem_solver_opts* opts, sem_step_cb cb, void* user, int* success)
{
    return run_stepper(2, m, re, im, V, T, opts, cb, user, success);
}

int sem_last_run_stats(const sem_mesh* m, sem_run_stats* out)
{
    if (!m || !out) return SEM_EINVAL;
    *out = m->stats;
    return SEM_OK;
}

struct sem_stepper {
    std::unique_ptr<Stepper> st;
};

int sem_stepper_create(sem_mesh* m, int scheme, const double* re, const double* im, const double* V,
                       const sem_solver_opts* opts, sem_stepper** out)
{
    return guarded([&] {
        if (!m || !re || !im || !V || !opts || !out || scheme < 0 || scheme > 2)
            throw std::invalid_argument("sem_stepper_create: bad argument");
        *out = nullptr;
        std::unique_ptr<sem_stepper> h(new sem_stepper);
        h->st.reset(stepper_create(m, scheme, re, im, V, opts));
        *out = h.release();
        return (int)SEM_OK;
    });
}

int sem_stepper_advance(sem_stepper* h, double T, long max_steps, sem_step_cb cb, void* user, int* success)
{
    return guarded([&] {
        if (!h) throw std::invalid_argument("sem_stepper_advance: null handle");
        const bool ok = stepper_advance(*h->st, T, max_steps, cb, user);
        if (success) *success = ok ? 1 : 0;
        return (int)SEM_OK;
    });
}

int sem_stepper_read(sem_stepper* h, double* re, double* im, double* t)
{
    return guarded([&] {
        if (!h) throw std::invalid_argument("sem_stepper_read: null handle");
        stepper_read(*h->st, re, im);
        if (t) *t = h->st->t;
        return (int)SEM_OK;
    });
}

int sem_stepper_set_learning(sem_stepper* h, int epochs, float lr)
{
    if (!h || epochs < 1 || !(lr > 0)) return SEM_EINVAL;
    if (!h->st->mlp) {
        set_error("sem_stepper_set_learning: the stepper was not created with the learned guess");
        return SEM_EINVAL;
    }
    h->st->epochs = epochs;
    h->st->mlp->lr = lr;
    return SEM_OK;
}

int sem_stepper_set_training(sem_stepper* h, int loss, int inputs)
{
    if (!h || loss < 0 || loss > 1 || inputs < 0 || inputs > 1) return SEM_EINVAL;
    if (!h->st->mlp) {
        set_error("sem_stepper_set_training: the stepper was not created with the learned guess");
        return SEM_EINVAL;
    }
    for (bool v : h->st->have_scale)
        if (v) {
            set_error("sem_stepper_set_training: the network has started training");
            return SEM_EINVAL;
        }
    h->st->residual_loss = loss == 1;
    h->st->feat_rext = inputs == 1;
    return SEM_OK;
}

int sem_stepper_set_guard(sem_stepper* h, int on)
{
    if (!h) return SEM_EINVAL;
    h->st->guard = on != 0;
    return SEM_OK;
}

int sem_stepper_stats(const sem_stepper* h, sem_run_stats* out)
{
    if (!h || !out) return SEM_EINVAL;
    *out = h->st->stats;
    return SEM_OK;
}

int sem_stepper_checkpoint_size(const sem_stepper* h, long* n)
{
    if (!h || !n) return SEM_EINVAL;
    *n = (long)ckpt_size(*h->st);
    return SEM_OK;
}

int sem_stepper_save(sem_stepper* h, double* buf, long n)
{
    return guarded([&] {
        if (!h || !buf) throw std::invalid_argument("sem_stepper_save: null argument");
        if (n < (long)ckpt_size(*h->st)) throw std::invalid_argument("sem_stepper_save: buffer too small");
        stepper_save(*h->st, buf);
        return (int)SEM_OK;
    });
}

int sem_stepper_restore(sem_stepper* h, const double* buf, long n)
{
    return guarded([&] {
        if (!h || !buf || n < 1) throw std::invalid_argument("sem_stepper_restore: bad argument");
        stepper_restore(*h->st, buf, (size_t)n);
        return (int)SEM_OK;
    });
}

void sem_stepper_destroy(sem_stepper* h) { delete h; }

int sem_batch_create(int n, sem_mesh* const* meshes, sem_batch** out)
{
    return guarded([&] {
        if (!out || !meshes || n < 1) throw std::invalid_argument("sem_batch_create: bad arguments");
        *out = nullptr;
        std::unique_ptr<sem_batch> b(new sem_batch);
        b->inst_group.assign(n, -1);
        b->inst_slot.assign(n, -1);
        b->inst_ref_len.assign(n, 0);
        std::map<int, std::vector<int>> by_p;
        for (int k = 0; k < n; ++k) {
            if (!meshes[k]) throw std::invalid_argument("sem_batch_create: null mesh");
            if (meshes[k]->n_inst != 1) throw std::invalid_argument("sem_batch_create: instances must be plain meshes");
            by_p[meshes[k]->hm.p].push_back(k);
            b->inst_ref_len[k] = (long)meshes[k]->L.E * meshes[k]->L.N * meshes[k]->L.N;
        }
        for (auto& kv : by_p) {
            sem_batch::Group g;
            g.p = kv.first;
            g.members = kv.second;
            std::vector<const sem_mesh*> parts;
            for (int k : g.members) parts.push_back(meshes[k]);
            g.m.reset(make_union_mesh(parts));
            for (size_t s = 0; s < g.members.size(); ++s) {
                b->inst_group[g.members[s]] = (int)b->groups.size();
                b->inst_slot[g.members[s]] = (int)s;
            }
            b->groups.push_back(std::move(g));
        }
        *out = b.release();
        return (int)SEM_OK;
    });
}

void sem_batch_destroy(sem_batch* b) { delete b; }

int sem_batch_info(const sem_batch* b, int* n_groups, int* group_p, int* group_size, long* group_elements)
{
    if (!b) return SEM_EINVAL;
    if (n_groups) *n_groups = (int)b->groups.size();
    for (size_t g = 0; g < b->groups.size(); ++g) {
        if (group_p) group_p[g] = b->groups[g].p;
        if (group_size) group_size[g] = (int)b->groups[g].members.size();
        if (group_elements) group_elements[g] = b->groups[g].m->L.E;
    }
    return SEM_OK;
}

int sem_batch_autotune(sem_batch* b, int precision, int steps, int* shapes)
{
    if (!b || steps < 1 || (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED)) return SEM_EINVAL;
    return guarded([&] {
        for (size_t g = 0; g < b->groups.size(); ++g) {
            std::vector<double> t;
            sem_mesh* m = b->groups[g].m.get();
            if (precision == SEM_PREC_FP64) autotune_steps<double>(m, K_RES, steps, t);
            else autotune_steps<float>(m, K_RES, steps, t);
            if (shapes) sem_get_step_shape(m, precision, shapes + 4 * g);
        }
        return (int)SEM_OK;
    });
}

int sem_batch_set_graph_mode(sem_batch* b, int on, int steps)
{
    if (!b) return SEM_EINVAL;
    for (auto& g : b->groups) {
        const int rc = sem_set_graph_mode(g.m.get(), on, steps);
        if (rc != SEM_OK) return rc;
    }
    return SEM_OK;
}

int sem_batch_set_step_mode(sem_batch* b, int precision, int mode)
{
    if (!b) return SEM_EINVAL;
    for (auto& g : b->groups) {
        const int rc = sem_set_step_mode(g.m.get(), precision, mode);
        if (rc != SEM_OK) return rc;
    }
    return SEM_OK;
}

int sem_batch_set_step_shape(sem_batch* b, int group, int precision, const int* shape)
{
    if (!b || group < 0 || group >= (int)b->groups.size()) return SEM_EINVAL;
    return sem_set_step_shape(b->groups[group].m.get(), precision, shape);
}

int sem_batch_start(sem_batch* b, int scheme, const double* const* re, const double* const* im,
                    const double* const* V, const sem_solver_opts* opts)
{
    return guarded([&] {
        if (!b || !re || !im || !V || !opts) throw std::invalid_argument("sem_batch_start: null argument");
        if (scheme < 0 || scheme > 2) throw std::invalid_argument("sem_batch_start: unknown scheme");
        for (size_t k = 0; k < b->inst_group.size(); ++k)
            if (!re[k] || !im[k] || !V[k]) throw std::invalid_argument("sem_batch_start: null instance array");
        b->session.clear();
        for (auto& g : b->groups) b->session.emplace_back(batch_group_start(g, scheme, re, im, V, opts));
        b->stats = sem_run_stats{};
        b->seconds = 0;
        return (int)SEM_OK;
    });
}

int sem_batch_start_bdf2(sem_batch* b, const double* const* re, const double* const* im, const double* const* V,
                         const sem_solver_opts* opts)
{
    return sem_batch_start(b, 0, re, im, V, opts);
}

int sem_batch_advance(sem_batch* b, double T, long max_steps, int* success)
{
    return guarded([&] {
        if (!b || b->session.size() != b->groups.size()) throw std::invalid_argument("sem_batch_advance: no session");
        const size_t ng = b->groups.size();
        std::vector<std::exception_ptr> err(ng);
        std::vector<int> okg(ng, 1);
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> th;
        for (size_t g = 0; g < ng; ++g)
            th.emplace_back([&, g] {
                try {
                    okg[g] = stepper_advance(*b->session[g], T, max_steps, nullptr, nullptr) ? 1 : 0;
                } catch (...) {
                    err[g] = std::current_exception();
                }
            });
        for (auto& t : th) t.join();
        b->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto& e : err)
            if (e) std::rethrow_exception(e);
        sem_run_stats tot{};
        for (const auto& st : b->session) {
            const sem_run_stats& g = st->stats;
            tot.steps = std::max(tot.steps, g.steps);
            tot.stage_solves += g.stage_solves;
            tot.total_iters += g.total_iters;
            tot.refinements += g.refinements;
            tot.last_error_estimate = std::max(tot.last_error_estimate, g.last_error_estimate);
            tot.max_true_rel_residual = std::max(tot.max_true_rel_residual, g.max_true_rel_residual);
            tot.mlp_train_steps += g.mlp_train_steps;
            tot.mlp_last_loss = std::max(tot.mlp_last_loss, g.mlp_last_loss);
            tot.learned_guess_chosen += g.learned_guess_chosen;
            tot.learned_residual_digits += g.learned_residual_digits;
        }
        b->stats = tot;
        if (success) {
            *success = 1;
            for (int v : okg) *success = *success && v;
        }
        return (int)SEM_OK;
    });
}

int sem_batch_read(sem_batch* b, double* const* re, double* const* im, int* ok, double* t)
{
    return guarded([&] {
        if (!b || b->session.size() != b->groups.size()) throw std::invalid_argument("sem_batch_read: no session");
        for (size_t g = 0; g < b->groups.size(); ++g) batch_group_read(b->groups[g], *b->session[g], re, im, ok);
        if (t) *t = b->session.empty() ? 0.0 : b->session[0]->t;
        return (int)SEM_OK;
    });
}

int sem_batch_times(const sem_batch* b, double* t, int* steps)
{
    if (!b || b->session.size() != b->groups.size()) return SEM_EINVAL;
    for (size_t g = 0; g < b->groups.size(); ++g) {
        const Stepper& st = *b->session[g];
        const auto& mem = b->groups[g].members;
        for (size_t q = 0; q < mem.size(); ++q) {
            const bool per = st.m->n_inst > 1 && st.inst_t.size() == mem.size();
            if (t) t[mem[q]] = per ? st.inst_t[q] : st.t;
            if (steps) steps[mem[q]] = per ? st.inst_steps[q] : st.stats.steps;
        }
    }
    return SEM_OK;
}

static int batch_run(sem_batch* b, int scheme, double* const* re, double* const* im, const double* const* V, double T,
                     const sem_solver_opts* opts, int* ok)
{
    int rc = sem_batch_start(b, scheme, re, im, V, opts);
    if (rc == SEM_OK) rc = sem_batch_advance(b, T, 0, nullptr);
    if (rc == SEM_OK) rc = sem_batch_read(b, re, im, ok, nullptr);
    if (b) b->session.clear();
    return rc;
}

int sem_batch_schrodinger_bdf2(sem_batch* b, double* const* re, double* const* im, const double* const* V, double T,
                               const sem_solver_opts* opts, int* ok)
{
    return batch_run(b, 0, re, im, V, T, opts, ok);
}

int sem_batch_schrodinger_dirk4s3(sem_batch* b, double* const* re, double* const* im, const double* const* V,
                                  double T, const sem_solver_opts* opts, int* ok)
{
    return batch_run(b, 1, re, im, V, T, opts, ok);
}

int sem_batch_schrodinger_dirk4s5(sem_batch* b, double* const* re, double* const* im, const double* const* V,
                                  double T, const sem_solver_opts* opts, int* ok)
{
    return batch_run(b, 2, re, im, V, T, opts, ok);
}

int sem_batch_last_run_stats(const sem_batch* b, sem_run_stats* out, double* seconds)
{
    if (!b || !out) return SEM_EINVAL;
    *out = b->stats;
    if (seconds) *seconds = b->seconds;
    return SEM_OK;
}

int sem_stage_solve(sem_mesh* m, const double* V, double alpha, const double* b, double* z, int max_iter, double tol,
                    int precision, sem_solver_results* out)
{
    return guarded([&] {
        if (!m || !V || !b || !z) throw std::invalid_argument("sem_stage_solve: null argument");
        const size_t n = 2 * m->len();
        Field f = make_field(m, V, "ss_V");
        double* db = m->buf<double>("ss_b", n);
        double* dz = m->buf<double>("ss_z", n);
        m->upload(db, b, 2);
        // b arrives masked (as galerkin_op returns it); the solver works on its
        // consistent form: owner values on every copy, Dirichlet entries zero
        dss<double>(m, DSS_UNMASK, db, 2);
        dss<double>(m, DSS_ZERO_BND, db, 2);
        m->upload(dz, z, 2);
        dss<double>(m, DSS_UNMASK, dz, 2);
        KrylovOpts o;
        o.maxit = max_iter;
        o.tol = tol;
        o.atol = tol;
        Results r;
        if (precision == SEM_PREC_MIXED) r = solve_mixed(m, K_RES, f.c64, f.c32, alpha, db, dz, false, o, inner_tol_floor(m));
        else r = minres_fused<double>(m, K_RES, f.c64, alpha, db, dz, false, o);
        // unmasked (consistent) solution
        m->download(z, dz, 2);
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

struct StreamGuard {
    hipStream_t s = nullptr;
    StreamGuard() { SEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
    ~StreamGuard()
    {
        if (s) {
            hipStreamSynchronize(s);
            hipStreamDestroy(s);
        }
    }
};

static int tridiag_solve(bool use_minres, long n, double lo, double di, double up, const double* b, double* x,
                         int maxit, double tol, sem_solver_results* out)
{
    return guarded([&] {
        if (n < 1 || !b || !x) throw std::invalid_argument("tridiag solver: bad argument");
        if (!check_gpu()) return (int)SEM_EHIP;
        StreamGuard sg;
        GenericWork wk;
        GenericCtx g{};
        g.s = sg.s;
        g.n = n;
        wk.init(g);
        g.A = [&](const double* in, double* o) { launch_tridiag(in, o, n, lo, di, up, g.s); };
        double* db = wk.alloc(n);
        double* dx = wk.alloc(n);
        double* buf = wk.alloc(8 * (size_t)n);
        SEM_HIP(hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice));
        SEM_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
        Results r = use_minres ? generic_minres(g, dx, db, maxit, tol, buf) : generic_pcg(g, dx, db, maxit, tol, buf);
        SEM_HIP(hipStreamSynchronize(g.s));
        SEM_HIP(hipMemcpy(x, dx, n * sizeof(double), hipMemcpyDeviceToHost));
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

int sem_minres_tridiag(long n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                       sem_solver_results* out)
{
    return tridiag_solve(true, n, lo, di, up, b, x, maxit, tol, out);
}

int sem_pcg_tridiag(long n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                    sem_solver_results* out)
{
    return tridiag_solve(false, n, lo, di, up, b, x, maxit, tol, out);
}

// minres / pcg over the caller's operator, dot and preconditioner
// (minres.hpp:19-20, pcg.hpp:59-60, solver_base.hpp:14-22)
static int krylov_solve(bool use_minres, const sem_krylov* k, const double* b, double* x, int maxit, double tol,
                        sem_solver_results* out)
{
    return guarded([&] {
        if (!k || !k->A || k->n < 1 || !b || !x) throw std::invalid_argument("sem_minres / sem_pcg: bad argument");
        if (!check_gpu()) return (int)SEM_EHIP;
        const long n = k->n;
        StreamGuard sg;
        GenericWork wk;
        GenericCtx g{};
        g.s = sg.s;
        g.n = n;
        g.consistent = k->consistent != 0;
        wk.init(g);
        void* vs = (void*)g.s;
        auto fail = [](const char* what) { throw std::invalid_argument(std::string(what) + " callback failed"); };
        std::vector<double> ha, hb;
        if (k->host) {
            // host vectors: the library stages device vectors through host buffers
            ha.resize(n);
            hb.resize(n);
            g.A = [&](const double* in, double* o) {
                SEM_HIP(hipMemcpyAsync(ha.data(), in, n * sizeof(double), hipMemcpyDeviceToHost, g.s));
                SEM_HIP(hipStreamSynchronize(g.s));
                if (k->A(k->A_user, ha.data(), hb.data(), vs) != 0) fail("operator");
                SEM_HIP(hipMemcpyAsync(o, hb.data(), n * sizeof(double), hipMemcpyHostToDevice, g.s));
                SEM_HIP(hipStreamSynchronize(g.s));
            };
            if (k->dot)
                g.dotf = [&](const double* a, const double* c) {
                    SEM_HIP(hipMemcpyAsync(ha.data(), a, n * sizeof(double), hipMemcpyDeviceToHost, g.s));
                    SEM_HIP(hipMemcpyAsync(hb.data(), c, n * sizeof(double), hipMemcpyDeviceToHost, g.s));
                    SEM_HIP(hipStreamSynchronize(g.s));
                    double r = 0;
                    if (k->dot(k->dot_user, ha.data(), hb.data(), &r, vs) != 0) fail("dot");
                    return r;
                };
            if (k->precond)
                g.P = [&](const double* in, double* o) {
                    SEM_HIP(hipMemcpyAsync(ha.data(), in, n * sizeof(double), hipMemcpyDeviceToHost, g.s));
                    SEM_HIP(hipStreamSynchronize(g.s));
                    if (k->precond(k->precond_user, ha.data(), hb.data(), vs) != 0) fail("preconditioner");
                    SEM_HIP(hipMemcpyAsync(o, hb.data(), n * sizeof(double), hipMemcpyHostToDevice, g.s));
                    SEM_HIP(hipStreamSynchronize(g.s));
                };
        } else {
            // device vectors: the callbacks enqueue on the library's stream
            g.A = [&](const double* in, double* o) {
                if (k->A(k->A_user, in, o, vs) != 0) fail("operator");
            };
            if (k->dot)
                g.dotf = [&](const double* a, const double* c) {
                    double r = 0;
                    if (k->dot(k->dot_user, a, c, &r, vs) != 0) fail("dot");
                    return r;
                };
            if (k->precond)
                g.P = [&](const double* in, double* o) {
                    if (k->precond(k->precond_user, in, o, vs) != 0) fail("preconditioner");
                };
        }
        double* buf = wk.alloc((use_minres ? 8 : 4) * (size_t)n);
        double *db, *dx;
        if (k->host) {
            db = wk.alloc(n);
            dx = wk.alloc(n);
            SEM_HIP(hipMemcpy(db, b, n * sizeof(double), hipMemcpyHostToDevice));
            SEM_HIP(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
        } else {
            db = const_cast<double*>(b);
            dx = x;
        }
        Results r = use_minres ? generic_minres(g, dx, db, maxit, tol, buf) : generic_pcg(g, dx, db, maxit, tol, buf);
        SEM_HIP(hipStreamSynchronize(g.s));
        if (k->host) SEM_HIP(hipMemcpy(x, dx, n * sizeof(double), hipMemcpyDeviceToHost));
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
static double* noalloc_scratch()
{
    static std::mutex mu;
    static std::map<int, double*> addr;
    int dev = 0;
    SEM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* p = nullptr;
    SEM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError{hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)"};
    addr[dev] = (double*)p;
    return (double*)p;
}

int sem_minres_noalloc(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out)
{
    return guarded([&] {
        if (!k || !k->A || k->n < 1 || !b || !x) throw std::invalid_argument("sem_minres_noalloc: bad argument");
        if (k->host) throw std::invalid_argument("sem_minres_noalloc: device vectors only (host = 0)");
        if (k->n > kNoAllocMaxN)
            throw std::invalid_argument("sem_minres_noalloc: n = " + std::to_string(k->n) + " exceeds the workspace (n <= " +
                                        std::to_string(kNoAllocMaxN) + ")");
        if (!check_gpu()) return (int)SEM_EHIP;
        const long n = k->n;
        double* ws = noalloc_scratch();
        double h_result = 0;  // pageable host word for the dot results
        GenericCtx g{};
        g.s = hipStreamPerThread;
        g.n = n;
        g.consistent = k->consistent != 0;
        g.part = ws + 8 * (size_t)kNoAllocMaxN;
        g.red = g.part + 1024;
        g.h_red = &h_result;
        void* vs = (void*)g.s;
        auto fail = [](const char* what) { throw std::invalid_argument(std::string(what) + " callback failed"); };
        g.A = [&](const double* in, double* o) {
            if (k->A(k->A_user, in, o, vs) != 0) fail("operator");
        };
        if (k->dot)
            g.dotf = [&](const double* a, const double* c) {
                double r = 0;
                if (k->dot(k->dot_user, a, c, &r, vs) != 0) fail("dot");
                return r;
            };
        if (k->precond)
            g.P = [&](const double* in, double* o) {
                if (k->precond(k->precond_user, in, o, vs) != 0) fail("preconditioner");
            };
        // the eight vectors of generic_minres are written before they are read
        Results r = generic_minres(g, x, b, maxit, tol, ws);
        SEM_HIP(hipStreamSynchronize(g.s));
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

int sem_minres(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out)
{
    return krylov_solve(true, k, b, x, maxit, tol, out);
}

int sem_pcg(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out)
{
    return krylov_solve(false, k, b, x, maxit, tol, out);
}

// Poisson (K) and Helmholtz (K + M_c) solves with the generic recurrences on the
// mesh's operator, unique-DOF dot and an optional Jacobi preconditioner
// 1/|diag| of the assembled operator (zero on Dirichlet entries). MINRES uses
// the consistent preconditioned normalisation (a symmetric positive definite
// preconditioner, so the indefinite Helmholtz operator is fine).
static Results mesh_generic_solve(sem_mesh* m, bool helm, const double* cf64, const double* bdev, double* xdev,
                                  int maxit, double tol, int precond)
{
    const size_t len = m->len();
    GenericCtx g{};
    g.s = m->stream;
    g.n = (long)len;
    GenericWork wk;
    wk.init(g);
    const int kind = helm ? K_HELMHOLTZ : K_LAPLACE;
    g.A = [&](const double* in, double* o) { apply_op<double>(m, kind, in, o, cf64, 0.0); };
    g.dotf = [&](const double* a, const double* c) { return wdot_host<double>(m, a, c, 1); };
    double* dinv = nullptr;
    if (precond == SEM_PRECOND_JACOBI) {
        dinv = wk.alloc(len);
        double* kd = wk.alloc(len);
        launch_stiff_diag(m->L.N, m->geo, len, m->L.nchunks, kd, m->dmat<double>(), g.s);
        if (helm) lincomb<double>(m, kd, 2, {{1.0, cf64}}, len);  // + gJ c, the diagonal of M_c
        dss<double>(m, DSS_SUM, kd, 1);
        launch_abs_recip(dinv, kd, len, g.s);
        dss<double>(m, DSS_ZERO_BND, dinv, 1);
        g.P = [&](const double* in, double* o) { launch_mul<double>(o, dinv, in, len, g.s); };
        g.consistent = true;
    }
    double* buf = wk.alloc((helm ? 8 : 4) * len);
    Results r = helm ? generic_minres(g, xdev, bdev, maxit, tol, buf) : generic_pcg(g, xdev, bdev, maxit, tol, buf);
    m->sync();
    return r;
}

int sem_poisson_precond(sem_mesh* m, double* u, const double* f, int max_iter, double tol, int precond,
                        sem_solver_results* out)
{
    return guarded([&] {
        if (!m || !u || !f || (precond != SEM_PRECOND_IDENTITY && precond != SEM_PRECOND_JACOBI))
            throw std::invalid_argument("sem_poisson_precond: bad argument");
        const size_t n = m->len();
        double* b = m->buf<double>("ps_b", n);
        double* x = m->buf<double>("ps_x", n);
        double* bb = m->buf<double>("ps_bb", n);
        m->upload(b, f, 1);
        dss<double>(m, DSS_UNMASK, b, 1);
        apply_op<double>(m, K_MASS, b, bb, nullptr, 0.0);  // gproj(f)
        m->upload(x, u, 1);
        dss<double>(m, DSS_UNMASK, x, 1);
        Results r = mesh_generic_solve(m, false, nullptr, bb, x, max_iter, tol, precond);
        merge_owner(m, x, u);
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

int sem_helmholtz_precond(sem_mesh* m, double* u, const double* c, const double* f, int max_iter, double tol,
                          int precond, sem_solver_results* out)
{
    return guarded([&] {
        if (!m || !u || !c || !f || (precond != SEM_PRECOND_IDENTITY && precond != SEM_PRECOND_JACOBI))
            throw std::invalid_argument("sem_helmholtz_precond: bad argument");
        const size_t n = m->len();
        double* b = m->buf<double>("hh_b", n);
        double* x = m->buf<double>("hh_x", n);
        double* bb = m->buf<double>("hh_bb", n);
        m->upload(b, f, 1);
        dss<double>(m, DSS_UNMASK, b, 1);
        apply_op<double>(m, K_MASS, b, bb, nullptr, 0.0);
        Field cf = make_field(m, c, "hh_c");
        m->upload(x, u, 1);
        dss<double>(m, DSS_UNMASK, x, 1);
        Results r = mesh_generic_solve(m, true, cf.c64, bb, x, max_iter, tol, precond);
        merge_owner(m, x, u);
        fill_results(out, r);
        return (int)SEM_OK;
    });
}

// galerkin_op (galerkin.hpp:24-35) with the caller's element operator: unmask ->
// op(u) element by element -> global_sum -> mask, on device arrays in the
// reference's element-local layout
int sem_galerkin_apply_op(sem_mesh* m, sem_element_op_fn op, void* user, const double* in, double* out)
{
    return guarded([&] {
        if (!m || !op || !in || !out) throw std::invalid_argument("sem_galerkin_apply_op: null argument");
        const size_t len = m->len(), rl = ref_len(m);
        double* d = m->buf<double>("gop_d", len);
        double* u = m->buf<double>("gop_u", rl);
        double* y = m->buf<double>("gop_y", rl);
        hipStream_t s = m->stream;
        launch_fill<double>(d, 0.0, len, s);
        launch_layout(in, d, m->L.E, m->L.N, m->L.K, 1, s);   // reference -> device layout
        dss<double>(m, DSS_UNMASK, d, 1);
        launch_layout(d, u, m->L.E, m->L.N, m->L.K, 0, s);    // device -> reference layout
        if (op(user, m->L.E, m->L.N, u, y, (void*)s) != 0) throw std::invalid_argument("element operator callback failed");
        launch_fill<double>(d, 0.0, len, s);
        launch_layout(y, d, m->L.E, m->L.N, m->L.K, 1, s);
        dss<double>(m, DSS_SUM, d, 1);
        dss<double>(m, DSS_MASK, d, 1);
        launch_layout(d, out, m->L.E, m->L.N, m->L.K, 0, s);
        m->sync();
        return (int)SEM_OK;
    });
}

int sem_stream_synchronize(void* stream)
{
    return guarded([&] {
        SEM_HIP(hipStreamSynchronize((hipStream_t)stream));
        return (int)SEM_OK;
    });
}

int sem_device_alloc(long n, double** ptr)
{
    return guarded([&] {
        if (!ptr || n < 0) throw std::invalid_argument("sem_device_alloc: bad argument");
        if (!check_gpu()) return (int)SEM_EHIP;
        SEM_HIP(hipMalloc((void**)ptr, std::max<long>(n, 1) * sizeof(double)));
        return (int)SEM_OK;
    });
}

int sem_device_free(double* ptr)
{
    return guarded([&] {
        SEM_HIP(hipFree(ptr));
        return (int)SEM_OK;
    });
}

int sem_copy(double* dst, const double* src, long n, int direction)
{
    return guarded([&] {
        if (!dst || !src || n < 0 || direction < 0 || direction > 2) throw std::invalid_argument("sem_copy: bad argument");
        const hipMemcpyKind k = direction == 0 ? hipMemcpyHostToDevice
                                : direction == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        SEM_HIP(hipMemcpy(dst, src, n * sizeof(double), k));
        return (int)SEM_OK;
    });
}

int sem_time_copy(long bytes, int reps, int variant, double* us_per_launch)
{
    return guarded([&] {
        if (bytes < 16 || reps < 1) throw std::invalid_argument("sem_time_copy: bad argument");
        if (!check_gpu()) return (int)SEM_EHIP;
        const size_t n = (size_t)bytes / sizeof(double) / 2 * 2;
        double *a = nullptr, *b = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        int rc = SEM_OK;
        int ncu = 256;
        {
            hipDeviceProp_t p;
            if (hipGetDeviceProperties(&p, 0) == hipSuccess) ncu = p.multiProcessorCount;
        }
        auto launch = [&]() {
            if (variant == 0) launch_cast<double, double>(b, a, n, s);
            else launch_copy16(b, a, n, ncu * 8 * variant, s);
        };
        try {
            SEM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            SEM_HIP(hipMalloc(&a, n * sizeof(double)));
            SEM_HIP(hipMalloc(&b, n * sizeof(double)));
            SEM_HIP(hipEventCreate(&e0));
            SEM_HIP(hipEventCreate(&e1));
            launch_fill<double>(a, 1.0, n, s);
            launch();
            SEM_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) launch();
            SEM_HIP(hipEventRecord(e1, s));
            SEM_HIP(hipEventSynchronize(e1));
            float ms = 0;
            SEM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (us_per_launch) *us_per_launch = 1000.0 * ms / reps;
        } catch (const HipError&) {
            rc = SEM_EHIP;
        }
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        hipFree(a);
        hipFree(b);
        if (s) hipStreamDestroy(s);
        if (rc != SEM_OK) throw HipError{hipErrorUnknown, "sem_time_copy"};
        return rc;
    });
}

int sem_set_kernel_timing(sem_mesh* m, int on)
{
    if (!m || on < 0 || on > 2) return SEM_EINVAL;
    m->timing = on;
    for (int k = 0; k < 4; ++k) {
        m->kt_launches[k] = 0;
        m->kt_us[k] = 0;
    }
    return SEM_OK;
}

int sem_kernel_timing(const sem_mesh* m, int which, long* launches, double* total_us, double* bytes_per_launch)
{
    if (!m || which < 0 || which > 3) return SEM_EINVAL;
    if (launches) *launches = m->kt_launches[which];
    if (total_us) *total_us = m->kt_us[which];
    if (bytes_per_launch) *bytes_per_launch = m->kt_bytes[which];
    return SEM_OK;
}

int sem_set_step_shape(sem_mesh* m, int precision, const int* shape)
{
    if (!m || !shape || (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED)) return SEM_EINVAL;
    for (int q = 0; q < 4; q += 2)
        if ((shape[q] != 1 && shape[q] != 2) || shape[q + 1] < 0) return SEM_EINVAL;
    m->shape[precision == SEM_PREC_FP64 ? 0 : 1] = StepShape{shape[0], shape[1], shape[2], shape[3]};
    return SEM_OK;
}

int sem_set_step_mode(sem_mesh* m, int precision, int mode)
{
    if (!m || mode < 0 || mode > 2 || (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED))
        return SEM_EINVAL;
    if (precision == SEM_PREC_FP64 && mode != 0) {
        // the one-launch step's FP64 instantiation is switched off: one of its
        // work shapes faulted the GPU on a 256 x 256, degree-7 mesh with no
        // cause found; FP64 solves use the two-launch step
        set_error("sem_set_step_mode: the one-launch Lanczos step runs in FP32 only (FP64 uses the two-launch step)");
        return SEM_EUNSUP;
    }
    m->step_mode[precision == SEM_PREC_FP64 ? 0 : 1] = mode;
    return SEM_OK;
}

int sem_get_step_mode(const sem_mesh* m, int precision, int* mode)
{
    if (!m || !mode || (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED)) return SEM_EINVAL;
    *mode = m->step_mode[precision == SEM_PREC_FP64 ? 0 : 1];
    return SEM_OK;
}

int sem_set_inner_scaling(sem_mesh* m, int on)
{
    if (!m || (on != 0 && on != 1)) return SEM_EINVAL;
    m->inner_scale = on;
    return SEM_OK;
}

int sem_get_inner_scaling(const sem_mesh* m, int* on)
{
    if (!m || !on) return SEM_EINVAL;
    *on = m->inner_scale;
    return SEM_OK;
}

int sem_set_graph_mode(sem_mesh* m, int on, int steps)
{
    if (!m || (on && (steps < 3 || steps % 3 != 0 || steps > 3000))) return SEM_EINVAL;
    return guarded([&] {
        m->sync();
        m->clear_graphs();
        m->use_graphs = on != 0;
        if (on) m->graph_steps = steps;
        return (int)SEM_OK;
    });
}

int sem_get_step_shape(const sem_mesh* m, int precision, int* shape)
{
    if (!m || !shape || (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED)) return SEM_EINVAL;
    const StepShape& s = m->shape[precision == SEM_PREC_FP64 ? 0 : 1];
    shape[0] = s.op_cpb;
    shape[1] = s.op_grid;
    shape[2] = s.upd_cpb;
    shape[3] = s.upd_grid;
    return SEM_OK;
}

int sem_autotune(sem_mesh* m, int kind, int precision, int steps, int* n_candidates, int* cand, double* us)
{
    if (!m || steps < 1 || (kind != K_LAPLACE && kind != K_HELMHOLTZ && kind != K_RES)) return SEM_EINVAL;
    if (precision != SEM_PREC_FP64 && precision != SEM_PREC_MIXED) return SEM_EINVAL;
    return guarded([&] {
        std::vector<double> t;
        if (precision == SEM_PREC_FP64) autotune_steps<double>(m, kind, steps, t);
        else autotune_steps<float>(m, kind, steps, t);
        const auto c = step_candidates();
        if (n_candidates) *n_candidates = (int)c.size();
        for (size_t k = 0; k < c.size(); ++k) {
            if (cand) {
                cand[2 * k] = c[k].first;
                cand[2 * k + 1] = c[k].second;
            }
            if (us) {
                us[2 * k] = t[2 * k];
                us[2 * k + 1] = t[2 * k + 1];
            }
        }
        return SEM_OK;
    });
}

int sem_time_element_op(sem_mesh* m, int kind, int precision, int reps, double* us_per_launch, double* bytes_per_launch)
{
    return guarded([&] {
        if (!m || reps < 1 || kind < 0 || kind > 5) throw std::invalid_argument("sem_time_element_op: bad argument");
        const int nc = op_ncomp(kind);
        const size_t n = (size_t)nc * m->len();
        const int N = m->L.N;
        const double EN2 = (double)m->L.E * N * N;
        // algorithmic bytes per launch: the reference data model (8 B FP64 or
        // 4 B FP32 values); where the FP32 kernels form the geometric factors
        // from each element's 32-byte bilinear map, without the factors
        const double vb = precision ? 4.0 : 8.0;
        double fields, geo;
        switch (kind) {
        case K_LAPLACE: fields = 5, geo = 3; break;    // u, A, B, C in; y out
        case K_HELMHOLTZ: fields = 7, geo = 4; break;  // u, c, J, A, B, C in; y out
        case K_MASS: fields = 3, geo = 1; break;       // u, J in; y out
        case K_RES: fields = 9, geo = 4; break;        // u, v, V, J, A, B, C in; 2 out
        case K_F: fields = 9, geo = 4; break;          // u, v, V, J, A, B, C in; 2 out
        default: fields = 7, geo = 4; break;           // u, V, J, A, B, C in; y out
        }
        if (bytes_per_launch)
            *bytes_per_launch = (precision && map_geometry(N)) ? (fields - geo) * vb * EN2 + 32.0 * m->L.E
                                                               : fields * vb * EN2;
        hipEvent_t e0, e1;
        SEM_HIP(hipEventCreate(&e0));
        SEM_HIP(hipEventCreate(&e1));
        float ms = 0;
        if (precision) {
            float* in = m->buf<float>("t_in32", n);
            float* out = m->buf<float>("t_out32", n);
            float* cf = m->buf<float>("t_cf32", m->len());
            launch_fill<float>(in, 0.5f, n, m->stream);
            launch_fill<float>(cf, 0.25f, m->len(), m->stream);
            OpArgs<float> a{in, out, m->geo32, cf, 0.1f, 1.0f, m->len(), m->L.nchunks, kind == K_RES ? m->contraction : 0, m->qm};
            launch_element_op_d<float>(N, kind, a, m->D32.data(), m->stream);
            SEM_HIP(hipEventRecord(e0, m->stream));
            for (int r = 0; r < reps; ++r) launch_element_op_d<float>(N, kind, a, m->D32.data(), m->stream);
            SEM_HIP(hipEventRecord(e1, m->stream));
        } else {
            double* in = m->buf<double>("t_in", n);
            double* out = m->buf<double>("t_out", n);
            double* cf = m->buf<double>("t_cf", m->len());
            launch_fill<double>(in, 0.5, n, m->stream);
            launch_fill<double>(cf, 0.25, m->len(), m->stream);
            OpArgs<double> a{in, out, m->geo, cf, 0.1, 1.0, m->len(), m->L.nchunks};
            launch_element_op_d<double>(N, kind, a, m->D64.data(), m->stream);
            SEM_HIP(hipEventRecord(e0, m->stream));
            for (int r = 0; r < reps; ++r) launch_element_op_d<double>(N, kind, a, m->D64.data(), m->stream);
            SEM_HIP(hipEventRecord(e1, m->stream));
        }
        SEM_HIP(hipEventSynchronize(e1));
        SEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        if (us_per_launch) *us_per_launch = 1000.0 * ms / reps;
        return (int)SEM_OK;
    });
}


struct sem_mlp {
    hipStream_t stream = nullptr;
    std::unique_ptr<Mlp> net;
    ~sem_mlp()
    {
        net.reset();
        if (stream) hipStreamDestroy(stream);
    }
};

int sem_mlp_create(int in, int h1, int h2, int out, unsigned long long seed, sem_mlp** out_handle)
{
    return guarded([&] {
        if (!out_handle || in < 1 || h1 < 1 || h2 < 1 || out < 1) throw std::invalid_argument("sem_mlp_create: bad argument");
        *out_handle = nullptr;
        if (!check_gpu()) return (int)SEM_EHIP;
        std::unique_ptr<sem_mlp> h(new sem_mlp);
        SEM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        h->net.reset(new Mlp(in, h1, h2, out, seed, h->stream));
        *out_handle = h.release();
        return (int)SEM_OK;
    });
}

void sem_mlp_destroy(sem_mlp* m) { delete m; }

int sem_mlp_info(const sem_mlp* m, long* out)
{
    if (!m || !out) return SEM_EINVAL;
    out[0] = m->net->in_pad;
    out[1] = m->net->out_pad;
    out[2] = (long)m->net->nparams;
    for (int k = 0; k < 6; ++k) out[3 + k] = (long)m->net->off[k];
    return SEM_OK;
}

int sem_mlp_set_params(sem_mlp* m, const float* params)
{
    return guarded([&] {
        if (!m || !params) throw std::invalid_argument("null argument");
        m->net->set_params(params);
        return (int)SEM_OK;
    });
}

int sem_mlp_get_params(sem_mlp* m, float* params, int gradients)
{
    return guarded([&] {
        if (!m || !params) throw std::invalid_argument("null argument");
        m->net->get_params(params, gradients != 0);
        return (int)SEM_OK;
    });
}

int sem_mlp_forward(sem_mlp* m, int rows, const float* X, float* Y)
{
    return guarded([&] {
        if (!m || !X || !Y || rows < 1) throw std::invalid_argument("sem_mlp_forward: bad argument");
        Mlp& n = *m->net;
        n.reserve(rows);
        SEM_HIP(hipMemcpyAsync(n.X, X, (size_t)rows * n.in_pad * sizeof(float), hipMemcpyHostToDevice, m->stream));
        n.forward(rows);
        SEM_HIP(hipMemcpyAsync(Y, n.Y, (size_t)rows * n.out_pad * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        SEM_HIP(hipStreamSynchronize(m->stream));
        return (int)SEM_OK;
    });
}

int sem_mlp_gradients(sem_mlp* m, int rows, const float* X, const float* T, double* loss)
{
    return guarded([&] {
        if (!m || !X || !T || rows < 1) throw std::invalid_argument("sem_mlp_gradients: bad argument");
        Mlp& n = *m->net;
        n.reserve(rows);
        SEM_HIP(hipMemcpyAsync(n.X, X, (size_t)rows * n.in_pad * sizeof(float), hipMemcpyHostToDevice, m->stream));
        SEM_HIP(hipMemcpyAsync(n.T, T, (size_t)rows * n.out_pad * sizeof(float), hipMemcpyHostToDevice, m->stream));
        const double l = n.backward(rows, true);
        if (loss) *loss = l;
        return (int)SEM_OK;
    });
}

int sem_mlp_adam(sem_mlp* m, float lr)
{
    return guarded([&] {
        if (!m) throw std::invalid_argument("null argument");
        m->net->lr = lr;
        m->net->adam_step();
        SEM_HIP(hipStreamSynchronize(m->stream));
        return (int)SEM_OK;
    });
}

}  // extern "C"
