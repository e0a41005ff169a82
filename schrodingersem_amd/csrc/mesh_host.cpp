// mesh_host.cpp -- host-side mesh setup of libsem_hip.so: quadrature,
// derivative matrix, element metrics, edge/corner topology and the seeded
// structured-mesh generator. One-time O(E*N^2) work; everything on the
// solver path runs on the GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <stdexcept>
#include <unordered_map>

#include "sem_internal.hpp"

namespace sem {

// Legendre P_n(x) by the Bonnet recurrence.
static double legendre(int n, double x)
{
    if (n == 0) return 1.0;
    double a = 1.0, b = x;
    for (int k = 1; k < n; ++k) {
        double c = ((2 * k + 1) * x * b - k * a) / (k + 1);
        a = b;
        b = c;
    }
    return b;
}

// Gauss-Lobatto-Legendre rule of degree p (gauss_lobatto.hpp:43-126).
// p <= 8 uses the reference's literal table (the same digits, so the nodes are
// bit-identical to the reference's); p > 8 finds the interior nodes, roots of
// q = P_{p+1} - P_{p-1} (Kopriva alg. 25), by Newton from Chebyshev-Lobatto
// starting points, q' = (2p+1) P_p. Weights 2 / (p(p+1) P_p(x)^2).
void gll_rule(int p, std::vector<double>& x, std::vector<double>& w)
{
    if (p < 1) throw std::invalid_argument("gauss_lobatto error: require n >= 1, but n =" + std::to_string(p) + ".");
    static const std::map<int, std::vector<double>> table = {
        {1, {-1, 1}},
        {2, {-1, 0, 1}},
        {3, {-1, -0.447213595499958, 0.447213595499958, 1}},
        {4, {-1, -0.654653670707977, 0, 0.654653670707977, 1}},
        {5, {-1, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1}},
        {6, {-1, -0.830223896278567, -0.468848793470714, 0.0, 0.468848793470714, 0.830223896278567, 1}},
        {7, {-1, -0.871740148509607, -0.591700181433142, -0.209299217902479, 0.2092992179024789, 0.591700181433142,
             0.871740148509607, 1}},
        {8, {-1, -0.899757995411460, -0.677186279510738, -0.363117463826178, 0, 0.363117463826178, 0.677186279510738,
             0.899757995411460, 1}}};
    auto it = table.find(p);
    if (it != table.end()) {
        x = it->second;
    } else {
        x.assign(p + 1, 0.0);
        x[0] = -1;
        x[p] = 1;
        for (int k = 1; k <= p / 2; ++k) {
            double z = -std::cos(M_PI * k / p);
            for (int iter = 0; iter < 100; ++iter) {
                double q = legendre(p + 1, z) - legendre(p - 1, z);
                double dq = (2 * p + 1) * legendre(p, z);
                double d = q / dq;
                z -= d;
                if (std::fabs(d) < 1e-17) break;
            }
            x[k] = z;
            x[p - k] = -z;
        }
        if (p % 2 == 0) x[p / 2] = 0;
    }
    w.assign(p + 1, 0.0);
    for (int i = 0; i <= p / 2; ++i) {
        double P = legendre(p, x[i]);
        w[i] = 2.0 / (p * (p + 1) * (P * P));
        w[p - i] = w[i];
    }
}

// Lagrange derivative matrix on nodes x (derivative_matrix.hpp:12-38,
// Kopriva alg. 37): barycentric weights, off-diagonal (b_j/b_i)/(x_i - x_j),
// diagonal = minus the row sum accumulated in column order.
std::vector<double> diff_matrix(const std::vector<double>& x)
{
    const size_t n = x.size();
    std::vector<double> b(n, 1.0), D(n * n, 0.0);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j)
            if (i != j) b[i] *= x[i] - x[j];
        b[i] = 1 / b[i];
    }
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            if (j != i) {
                D[i + j * n] = (b[j] / b[i]) / (x[i] - x[j]);
                D[i + i * n] -= D[i + j * n];
            }
    return D;
}

// Bilinear-quad metrics (Quad.hpp:27-55). X_xi, Y_xi vary with eta (index j),
// X_eta, Y_eta with xi (index i); the C entry keeps the reference's Y_xi[i]
// (Quad.hpp:53), which equals Y_xi[j] on every parallelogram.
void quad_metrics(const double* xs, const double* ys, const std::vector<double>& z, double* J, double* A,
                         double* B, double* C)
{
    const size_t n = z.size();
    std::vector<double> Xxi(n), Yxi(n), Xeta(n), Yeta(n);
    for (size_t i = 0; i < n; ++i) {
        Xxi[i] = 0.25 * ((1 - z[i]) * (xs[1] - xs[0]) + (1 + z[i]) * (xs[2] - xs[3]));
        Yxi[i] = 0.25 * ((1 - z[i]) * (ys[1] - ys[0]) + (1 + z[i]) * (ys[2] - ys[3]));
        Xeta[i] = 0.25 * ((1 - z[i]) * (xs[3] - xs[0]) + (1 + z[i]) * (xs[2] - xs[1]));
        Yeta[i] = 0.25 * ((1 - z[i]) * (ys[3] - ys[0]) + (1 + z[i]) * (ys[2] - ys[1]));
    }
    for (size_t j = 0; j < n; ++j)
        for (size_t i = 0; i < n; ++i) {
            size_t k = i + j * n;
            double jac = Xxi[j] * Yeta[i] - Xeta[i] * Yxi[j];
            J[k] = jac;
            A[k] = (Yeta[i] * Yeta[i] + Xeta[i] * Xeta[i]) / jac;
            B[k] = (Yeta[i] * Yxi[j] + Xeta[i] * Xxi[j]) / jac;
            C[k] = (Yxi[i] * Yxi[i] + Xxi[j] * Xxi[j]) / jac;
        }
}

// The four metric functions of quad_metrics are linear in the node z:
// 0.25 ((1 - z) a + (1 + z) b) = 0.25 (a + b) + 0.25 (b - a) z.
void quad_map(const double* xs, const double* ys, float* out8)
{
    auto lin = [&](double a, double b, float* o) {
        o[0] = (float)(0.25 * (a + b));
        o[1] = (float)(0.25 * (b - a));
    };
    lin(xs[1] - xs[0], xs[2] - xs[3], out8 + 0);  // X_xi
    lin(ys[1] - ys[0], ys[2] - ys[3], out8 + 2);  // Y_xi
    lin(xs[3] - xs[0], xs[2] - xs[1], out8 + 4);  // X_eta
    lin(ys[3] - ys[0], ys[2] - ys[1], out8 + 6);  // Y_eta
}

// Mesh topology with load_mesh's semantics (load_mesh.hpp:76-209): the owner
// of an edge or corner is the first element, in element order, to touch it;
// the second element's side is negative when it runs against the owner's.
void build_host_mesh(HostMesh& hm, int p, const double* xy, int n_nodes, const int* quads, int n_el)
{
    if (p < 1 || p + 1 > kMaxN) throw std::invalid_argument("polynomial degree p=" + std::to_string(p) + " outside 1.." +
                                                             std::to_string(kMaxN - 1));
    if (n_nodes < 3 || n_el < 1) throw std::invalid_argument("mesh needs at least 3 nodes and 1 element");
    hm.p = p;
    gll_rule(p, hm.x, hm.w);
    hm.N = p + 1;
    hm.D = diff_matrix(hm.x);
    const int N = hm.N;
    hm.n_nodes = n_nodes;
    hm.n_el = n_el;
    hm.node_xy.assign(xy, xy + 2 * (size_t)n_nodes);
    hm.quads.assign(quads, quads + 4 * (size_t)n_el);
    for (int k = 0; k < 4 * n_el; ++k)
        if (quads[k] < 0 || quads[k] >= n_nodes) throw std::invalid_argument("element corner index out of range");

    // corner connections in element order
    std::vector<int> cnt(n_nodes, 0);
    for (int k = 0; k < 4 * n_el; ++k) cnt[quads[k]]++;
    hm.conn_off.assign(n_nodes + 1, 0);
    for (int c = 0; c < n_nodes; ++c) hm.conn_off[c + 1] = hm.conn_off[c] + cnt[c];
    hm.conn.assign(4 * (size_t)n_el, 0);
    std::vector<int> fill(hm.conn_off.begin(), hm.conn_off.end() - 1);
    for (int e = 0; e < n_el; ++e)
        for (int k = 0; k < 4; ++k) hm.conn[fill[quads[4 * e + k]]++] = 4 * e + k;

    // sides s = 1..4 run corner (0->1), (1->2), (3->2), (0->3)
    static const int sa[4] = {0, 1, 3, 0}, sb[4] = {1, 2, 2, 3};
    std::unordered_map<uint64_t, int> key2edge;
    key2edge.reserve(4 * (size_t)n_el);
    hm.edges.clear();
    for (int e = 0; e < n_el; ++e)
        for (int s = 0; s < 4; ++s) {
            int a = quads[4 * e + sa[s]], b = quads[4 * e + sb[s]];
            int lo = std::min(a, b), hi = std::max(a, b);
            uint64_t key = ((uint64_t)(uint32_t)lo << 32) | (uint32_t)hi;
            auto f = key2edge.find(key);
            if (f == key2edge.end()) {
                EdgeRec r;
                r.el[0] = e;
                r.side[0] = s + 1;
                r.n0 = lo;
                r.n1 = hi;
                key2edge.emplace(key, (int)hm.edges.size());
                hm.edges.push_back(r);
            } else {
                EdgeRec& r = hm.edges[f->second];
                if (r.type == INTERIOR) throw std::invalid_argument("non-manifold mesh: an edge is shared by more than two elements");
                int owner_start = quads[4 * r.el[0] + sa[r.side[0] - 1]];
                r.type = INTERIOR;
                r.el[1] = e;
                r.side[1] = (a == owner_start) ? (s + 1) : -(s + 1);
            }
        }
    hm.node_type.assign(n_nodes, INTERIOR);
    for (EdgeRec& r : hm.edges) {
        if (r.type == BOUNDARY) {
            hm.node_type[r.n0] = BOUNDARY;
            hm.node_type[r.n1] = BOUNDARY;
        } else if (r.side[1] > 0) {
            r.start = 1;
            r.delta = 1;
        } else {
            r.start = N - 2;
            r.delta = -1;
        }
    }
    // a node used by no element has no copies; the reference's dot() would
    // read connected_elements[0] of it, so such meshes are rejected.
    for (int c = 0; c < n_nodes; ++c)
        if (cnt[c] == 0) throw std::invalid_argument("node " + std::to_string(c) + " belongs to no element");
    hm.dof = (long)n_nodes + (long)hm.edges.size() * (N - 2) + (long)n_el * (N - 2) * (N - 2);

    // metrics
    const size_t n2 = (size_t)N * N;
    hm.J.assign(n2 * n_el, 0);
    hm.A.assign(n2 * n_el, 0);
    hm.B.assign(n2 * n_el, 0);
    hm.C.assign(n2 * n_el, 0);
    for (int e = 0; e < n_el; ++e) {
        double xs[4], ys[4];
        for (int k = 0; k < 4; ++k) {
            xs[k] = xy[2 * quads[4 * e + k]];
            ys[k] = xy[2 * quads[4 * e + k] + 1];
        }
        quad_metrics(xs, ys, hm.x, &hm.J[e * n2], &hm.A[e * n2], &hm.B[e * n2], &hm.C[e * n2]);
    }
}

// info.txt ("n_nodes n_elements"), coordinates.txt (x y per line),
// elements.txt (4 corner ids per line, counter-clockwise); load_mesh.hpp:90-144.
void read_mesh_folder(const std::string& folder, std::vector<double>& xy, std::vector<int>& quads)
{
    std::ifstream info(folder + "/info.txt");
    if (!info.is_open()) throw std::runtime_error("could not open info file.");
    int nn = 0, ne = 0;
    info >> nn >> ne;
    if (!info || nn <= 0 || ne <= 0) throw std::runtime_error("malformed info file.");
    std::ifstream coo(folder + "/coordinates.txt");
    if (!coo.is_open()) throw std::runtime_error("could not open coordinate file.");
    xy.assign(2 * (size_t)nn, 0);
    for (int k = 0; k < 2 * nn; ++k) coo >> xy[k];
    if (!coo) throw std::runtime_error("coordinate file too short.");
    std::ifstream el(folder + "/elements.txt");
    if (!el.is_open()) throw std::runtime_error("could not open element file.");
    quads.assign(4 * (size_t)ne, 0);
    for (int k = 0; k < 4 * ne; ++k) el >> quads[k];
    if (!el) throw std::runtime_error("element file too short.");
}

void to_device_layout(const Layout& L, const double* ref, double* dev)
{
    std::fill(dev, dev + L.len, 0.0);
    const int N = L.N;
    for (int e = 0; e < L.E; ++e)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) dev[L.pos(e, i, j)] = ref[L.ref(e, i, j)];
}

void from_device_layout(const Layout& L, const double* dev, double* ref)
{
    const int N = L.N;
    for (int e = 0; e < L.E; ++e)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) ref[L.ref(e, i, j)] = dev[L.pos(e, i, j)];
}

}  // namespace sem

// ----------------------------------------------------- generator (C-ABI) --
namespace {
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
}  // namespace

extern "C" int sem_mesh_structured(int n, double x0, double x1, double y0, double y1, double jitter,
                                   unsigned long long seed, double* xy, int* quads)
{
    if (n < 1 || !xy || !quads || !(x1 > x0) || !(y1 > y0) || jitter < 0 || jitter >= 0.5) {
        sem::set_error("sem_mesh_structured: invalid arguments");
        return SEM_EINVAL;
    }
    const double hx = (x1 - x0) / n, hy = (y1 - y0) / n;
    SplitMix64 rng{seed};
    for (int iy = 0; iy <= n; ++iy)
        for (int ix = 0; ix <= n; ++ix) {
            size_t k = (size_t)iy * (n + 1) + ix;
            double x = (ix == n) ? x1 : x0 + ix * hx;
            double y = (iy == n) ? y1 : y0 + iy * hy;
            if (ix > 0 && ix < n && iy > 0 && iy < n && jitter > 0) {
                x += (2 * rng.uniform() - 1) * jitter * hx;
                y += (2 * rng.uniform() - 1) * jitter * hy;
            }
            xy[2 * k] = x;
            xy[2 * k + 1] = y;
        }
    for (int ey = 0; ey < n; ++ey)
        for (int ex = 0; ex < n; ++ex) {
            size_t e = (size_t)ey * n + ex;
            int a = ey * (n + 1) + ex;
            quads[4 * e + 0] = a;
            quads[4 * e + 1] = a + 1;
            quads[4 * e + 2] = a + n + 2;
            quads[4 * e + 3] = a + n + 1;
        }
    return SEM_OK;
}

extern "C" int sem_mesh_write(const char* folder, const double* xy, int n_nodes, const int* quads, int n_elements)
{
    if (!folder || !xy || !quads || n_nodes <= 0 || n_elements <= 0) {
        sem::set_error("sem_mesh_write: invalid arguments");
        return SEM_EINVAL;
    }
    std::string f(folder);
    FILE* fi = std::fopen((f + "/info.txt").c_str(), "w");
    FILE* fc = std::fopen((f + "/coordinates.txt").c_str(), "w");
    FILE* fe = std::fopen((f + "/elements.txt").c_str(), "w");
    if (!fi || !fc || !fe) {
        if (fi) std::fclose(fi);
        if (fc) std::fclose(fc);
        if (fe) std::fclose(fe);
        sem::set_error("sem_mesh_write: cannot create files in " + f);
        return SEM_EIO;
    }
    std::fprintf(fi, "%d %d", n_nodes, n_elements);
    for (int k = 0; k < n_nodes; ++k) std::fprintf(fc, "%.18e %.18e\n", xy[2 * k], xy[2 * k + 1]);
    for (int e = 0; e < n_elements; ++e)
        std::fprintf(fe, "%d %d %d %d\n", quads[4 * e], quads[4 * e + 1], quads[4 * e + 2], quads[4 * e + 3]);
    std::fclose(fi);
    std::fclose(fc);
    std::fclose(fe);
    return SEM_OK;
}

// ------------------------------------------------ host-only basis entry points --
// gauss_lobatto<double>(n) (gauss_lobatto.hpp:43-126): n + 1 nodes and weights
extern "C" int sem_gauss_lobatto(int n, double* x, double* w)
{
    if (!x || !w || n < 1 || n > sem::kMaxN - 1) {
        sem::set_error("sem_gauss_lobatto: need 1 <= n <= " + std::to_string(sem::kMaxN - 1) + " and output arrays");
        return SEM_EINVAL;
    }
    try {
        std::vector<double> xx, ww;
        sem::gll_rule(n, xx, ww);
        std::copy(xx.begin(), xx.end(), x);
        std::copy(ww.begin(), ww.end(), w);
    } catch (const std::exception& e) {
        sem::set_error(e.what());
        return SEM_EINVAL;
    }
    return SEM_OK;
}

// derivative_matrix(x) (derivative_matrix.hpp:12-38): D (n x n, column-major)
extern "C" int sem_derivative_matrix(int n, const double* x, double* D)
{
    if (!x || !D || n < 1) {
        sem::set_error("sem_derivative_matrix: bad argument");
        return SEM_EINVAL;
    }
    const std::vector<double> d = sem::diff_matrix(std::vector<double>(x, x + n));
    std::copy(d.begin(), d.end(), D);
    return SEM_OK;
}

// Quad::compute_metrics(z) (Quad.hpp:27-55) for corners (xs, ys): J, A, B, C
// (n x n, column-major, the reference's C index included)
extern "C" int sem_quad_metrics(const double* xs, const double* ys, int n, const double* z, double* J, double* A,
                                double* B, double* C)
{
    if (!xs || !ys || !z || !J || !A || !B || !C || n < 1) {
        sem::set_error("sem_quad_metrics: bad argument");
        return SEM_EINVAL;
    }
    sem::quad_metrics(xs, ys, std::vector<double>(z, z + n), J, A, B, C);
    return SEM_OK;
}
