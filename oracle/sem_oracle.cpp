// sem_oracle.cpp -- FP64 CPU restatement of SchrodingerSEM's OpenMP hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP product is
// compared against. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it (liboracle_sem.so, through ctypes). The product
// library never links it and never falls back to it.
//
// Every function restates the reference algorithm from its source, read as
// text; each cites the reference file:line it follows. The reference itself
// cannot be compiled here (it needs Armadillo, Boost.Math and LAPACK), so
// three third-party routines are restated from their published definitions:
//   * LAPACK dsteqr (eigenvalues of a symmetric tridiagonal matrix): replaced
//     by Sturm-sequence bisection, which converges to the same eigenvalues;
//     the reference polishes them with two Newton steps either way.
//   * Boost.Math legendre_p: the three-term Bonnet recurrence.
//   * Boost.Math legendre_p_prime: P'_n = sum_{k=n-1,n-3,...} (2k+1) P_k.
//   * Armadillo/BLAS matrix products: plain loops, summing k ascending.
// These agree with the libraries to round-off; parity targets are 1e-12
// (operators) and 1e-9 (solutions), far above that.
//
// Layout: element-local arrays are flat, element e at offset e*N*N, entry
// (i,j) at i + j*N (column-major N x N, i along xi, j along eta), exactly the
// reference's std::vector<arma::Mat<real>> (types.hpp:9-10). Two-component
// (real, imaginary) vectors are [re | im], each E*N*N long.
//
// Pinning: tests/test_oracle_*.py check every function here against closed
// forms, the reference's own known-answer tests and brute-force assemblies.

#include <algorithm>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::vector<double> vec;

static thread_local std::string g_err;

// ---------------------------------------------------------------- a1: GLL --

// Boost legendre_p: Bonnet recurrence (k+1)P_{k+1} = (2k+1) x P_k - k P_{k-1}.
static double legendre_p(int n, double x)
{
    if (n == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2 * k + 1) * x * p1 - k * p0) / (k + 1);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

// Boost legendre_p_prime: P'_n = sum over k = n-1, n-3, ... >= 0 of (2k+1) P_k.
static double legendre_p_prime(int n, double x)
{
    double s = 0.0;
    for (int k = n - 1; k >= 0; k -= 2) s += (2 * k + 1) * legendre_p(k, x);
    return s;
}

// Eigenvalues (ascending) of the symmetric tridiagonal matrix with zero
// diagonal and off-diagonal e[0..m-2]; stands in for LAPACK dsteqr with
// COMPZ='N' (gauss_lobatto.hpp:22,82). Sturm-count bisection on [-1, 1].
static std::vector<double> tridiag_eigs_zero_diag(const std::vector<double>& e, int m)
{
    auto count_below = [&](double x) {
        int c = 0;
        double q = -x;
        if (q < 0) ++c;
        for (int i = 1; i < m; ++i) {
            if (q == 0.0) q = 1e-300;
            q = -x - e[i - 1] * e[i - 1] / q;
            if (q < 0) ++c;
        }
        return c;
    };
    std::vector<double> ev(m);
    for (int k = 0; k < m; ++k) {
        double lo = -1.0, hi = 1.0;  // eigenvalues are roots of P'_n, inside (-1,1)
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (lo + hi);
            if (mid == lo || mid == hi) break;
            if (count_below(mid) > k) hi = mid; else lo = mid;
        }
        ev[k] = 0.5 * (lo + hi);
    }
    return ev;
}

// gauss_lobatto.hpp:43-126. n = polynomial degree p; returns n+1 nodes/weights.
static void gauss_lobatto(int n, vec& x, vec& w)
{
    if (n < 1) throw std::invalid_argument("gauss_lobatto error: require n >= 1, but n =" + std::to_string(n) + ".");
    // literal table, gauss_lobatto.hpp:49-58 (kept digit for digit: the
    // reference uses these truncated literals for p <= 8).
    static const std::map<int, std::vector<double>> lgl = {
        {1, {-1, 1}},
        {2, {-1, 0, 1}},
        {3, {-1, -0.447213595499958, 0.447213595499958, 1}},
        {4, {-1, -0.654653670707977, 0, 0.654653670707977, 1}},
        {5, {-1, -0.765055323929465, -0.285231516480645, 0.285231516480645, 0.765055323929465, 1}},
        {6, {-1, -0.830223896278567, -0.468848793470714, 0.0, 0.468848793470714, 0.830223896278567, 1}},
        {7, {-1, -0.871740148509607, -0.591700181433142, -0.209299217902479, 0.2092992179024789, 0.591700181433142, 0.871740148509607, 1}},
        {8, {-1, -0.899757995411460, -0.677186279510738, -0.363117463826178, 0, 0.363117463826178, 0.677186279510738, 0.899757995411460, 1}}};
    auto it = lgl.find(n);
    if (it != lgl.end()) {
        x = it->second;
    } else {
        // Golub-Welsch, gauss_lobatto.hpp:65-113
        int N = n - 1;
        std::vector<double> E(std::max(n - 2, 0));
        for (int i = 0; i < n - 2; ++i) {
            double ii = i + 1;
            E[i] = std::sqrt(ii * (ii + 2) / ((2 * ii + 3) * (2 * ii + 1)));
        }
        std::vector<double> D = tridiag_eigs_zero_diag(E, N);
        std::vector<double> xs;
        for (int i = 0; i < N / 2; ++i) xs.push_back(D[i]);
        for (double& z : xs)
            for (int i = 0; i < 2; ++i) {  // gauss_lobatto.hpp:95-101
                double q = legendre_p(n + 1, z) - legendre_p(n - 1, z);
                double dq = legendre_p_prime(n + 1, z) - legendre_p_prime(n - 1, z);
                z -= q / dq;
            }
        if (n % 2 == 0) xs.push_back(0);
        for (int i = N / 2 - 1; i >= 0; --i) xs.push_back(-xs[i]);
        xs.insert(xs.begin(), -1);
        xs.push_back(1);
        x = xs;
    }
    // weights, gauss_lobatto.hpp:116-123
    w.assign(n + 1, 0.0);
    w[0] = 2.0 / (n * (n + 1));
    w[n] = w[0];
    for (int i = 0; i < n / 2 + 1; ++i) {
        double P = legendre_p(n, x[i]);
        w[i] = 2.0 / (n * (n + 1) * (P * P));
        w[n - i] = w[i];
    }
}

// ------------------------------------------------------ a2: derivative D --
// derivative_matrix.hpp:12-38. Column-major n x n: D(i,j) at i + j*n.
static vec derivative_matrix(const vec& x)
{
    size_t n = x.size();
    vec bw(n, 1.0);
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = 0; j < n; ++j)
            if (i != j) bw[i] *= x[i] - x[j];
        bw[i] = 1 / bw[i];
    }
    vec D(n * n, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < n; ++j)
            if (j != i) {
                D[i + j * n] = (bw[j] / bw[i]) / (x[i] - x[j]);
                D[i + i * n] -= D[i + j * n];
            }
    return D;
}

// --------------------------------------------------- a3/a4: element Quad --
struct Quad {
    double xs[4], ys[4];
    int corner_ids[4];
    vec J, A, B, C;

    // Quad.hpp:27-55, including the C metric's Yxi[i] index (Quad.hpp:53).
    void compute_metrics(const vec& z)
    {
        size_t n = z.size();
        J.assign(n * n, 0);
        A.assign(n * n, 0);
        B.assign(n * n, 0);
        C.assign(n * n, 0);
        vec Xxi(n), Yxi(n), Xeta(n), Yeta(n);
        for (size_t i = 0; i < n; ++i) {
            Xxi[i] = 0.25 * ((1 - z[i]) * (xs[1] - xs[0]) + (1 + z[i]) * (xs[2] - xs[3]));
            Yxi[i] = 0.25 * ((1 - z[i]) * (ys[1] - ys[0]) + (1 + z[i]) * (ys[2] - ys[3]));
            Xeta[i] = 0.25 * ((1 - z[i]) * (xs[3] - xs[0]) + (1 + z[i]) * (xs[2] - xs[1]));
            Yeta[i] = 0.25 * ((1 - z[i]) * (ys[3] - ys[0]) + (1 + z[i]) * (ys[2] - ys[1]));
        }
        for (size_t i = 0; i < n; ++i)
            for (size_t j = 0; j < n; ++j) {
                size_t k = i + j * n;
                J[k] = Xxi[j] * Yeta[i] - Xeta[i] * Yxi[j];
                A[k] = (Yeta[i] * Yeta[i] + Xeta[i] * Xeta[i]) / J[k];
                B[k] = (Yeta[i] * Yxi[j] + Xeta[i] * Xxi[j]) / J[k];
                C[k] = (Yxi[i] * Yxi[i] + Xxi[j] * Xxi[j]) / J[k];
            }
    }

    // Quad.hpp:57-62
    void from_local_coo(double xi, double eta, double& x, double& y) const
    {
        x = 0.25 * (xs[0] * (1 - xi) * (1 - eta) + xs[1] * (1 + xi) * (1 - eta) + xs[2] * (1 + xi) * (1 + eta) + xs[3] * (1 - xi) * (1 + eta));
        y = 0.25 * (ys[0] * (1 - xi) * (1 - eta) + ys[1] * (1 + xi) * (1 - eta) + ys[2] * (1 + xi) * (1 + eta) + ys[3] * (1 - xi) * (1 + eta));
    }
};

// -------------------------------------------------- a15/a16: topology -----
enum { INTERIOR = 0, BOUNDARY = 1 };

struct Edge {  // Edge.hpp:6-27
    int edge_type = BOUNDARY;
    int elements[2] = {-1, -1};
    int element_sides[2] = {0, 0};
    int startIter = 0, deltaIter = 0;
};

struct CornerInfo { int element_id, i, j; };  // CornerNode.hpp:8-13

struct Corner {  // CornerNode.hpp:15-29
    int corner_type = INTERIOR;
    double x = 0, y = 0;
    std::vector<CornerInfo> connected_elements;
};

struct Mesh {  // openmp_impl/Mesh.hpp:16-88
    int N = 0;
    vec qx, qw;  // quadrature
    vec D;
    std::vector<Edge> edges;
    std::vector<Corner> nodes;
    std::vector<Quad> elements;

    int smap(int side) const { const int s[] = {0, N - 1, N - 1, 0}; return s[side - 1]; }
    long dof() const { return (long)nodes.size() + (long)edges.size() * (N - 2) + (long)elements.size() * (N - 2) * (N - 2); }
    size_t n2() const { return (size_t)N * N; }
    size_t len() const { return elements.size() * n2(); }
};

// load_mesh.hpp:76-209, from in-memory arrays (nodes xy[2k], xy[2k+1];
// quads 4 corner ids per element, counter-clockwise).
static Mesh build_mesh(int p, const double* xy, int n_nodes, const int* quads, int n_el)
{
    Mesh mesh;
    gauss_lobatto(p, mesh.qx, mesh.qw);
    mesh.N = (int)mesh.qx.size();
    mesh.D = derivative_matrix(mesh.qx);
    const int N = mesh.N;
    const int cmap1[] = {0, N - 1, N - 1, 0};
    const int cmap2[] = {0, 0, N - 1, N - 1};
    const int emap1[] = {0, 1, 3, 0};
    const int emap2[] = {1, 2, 2, 3};

    mesh.nodes.resize(n_nodes);
    for (int k = 0; k < n_nodes; ++k) {
        mesh.nodes[k].x = xy[2 * k];
        mesh.nodes[k].y = xy[2 * k + 1];
    }
    mesh.elements.resize(n_el);
    for (int el = 0; el < n_el; ++el) {  // load_mesh.hpp:123-143
        Quad& q = mesh.elements[el];
        for (int i = 0; i < 4; ++i) {
            int c = quads[4 * el + i];
            if (c < 0 || c >= n_nodes) throw std::runtime_error("element corner index out of range");
            mesh.nodes[c].connected_elements.push_back({el, cmap1[i], cmap2[i]});
            q.corner_ids[i] = c;
            q.xs[i] = mesh.nodes[c].x;
            q.ys[i] = mesh.nodes[c].y;
        }
    }
    // unique edges; owner = first element to visit (load_mesh.hpp:147-176).
    // Edge order: creation order (the reference's unordered_map order only
    // changes the summation order of dot()'s edge term).
    std::map<std::pair<int, int>, int> key2edge;
    std::vector<std::pair<int, int>> keys;
    for (int e = 0; e < n_el; ++e) {
        const Quad& elem = mesh.elements[e];
        for (int i = 0; i < 4; ++i) {
            int start = elem.corner_ids[emap1[i]];
            int end = elem.corner_ids[emap2[i]];
            std::pair<int, int> key(std::min(start, end), std::max(start, end));
            auto f = key2edge.find(key);
            if (f != key2edge.end()) {
                Edge& ed = mesh.edges[f->second];
                int e0 = ed.elements[0];
                int s0 = ed.element_sides[0];
                int n1 = mesh.elements[e0].corner_ids[emap1[s0 - 1]];
                ed.elements[1] = e;
                ed.edge_type = INTERIOR;
                ed.element_sides[1] = (start == n1) ? (i + 1) : -(i + 1);
            } else {
                Edge ed;
                ed.edge_type = BOUNDARY;
                ed.elements[0] = e;
                ed.elements[1] = -1;
                ed.element_sides[0] = i + 1;
                ed.element_sides[1] = 0;
                key2edge[key] = (int)mesh.edges.size();
                keys.push_back(key);
                mesh.edges.push_back(ed);
            }
        }
    }
    for (size_t k = 0; k < mesh.edges.size(); ++k) {  // load_mesh.hpp:182-206
        Edge& ed = mesh.edges[k];
        if (ed.edge_type == BOUNDARY) {
            mesh.nodes[keys[k].first].corner_type = BOUNDARY;
            mesh.nodes[keys[k].second].corner_type = BOUNDARY;
        } else if (ed.element_sides[1] > 0) {
            ed.startIter = 1;
            ed.deltaIter = 1;
        } else {
            ed.startIter = N - 2;
            ed.deltaIter = -1;
        }
    }
    return mesh;
}

// openmp_impl/Mesh.hpp:70-75
static void compute_metrics(Mesh& mesh)
{
#pragma omp parallel for schedule(static) if (mesh.elements.size() > 63)
    for (long e = 0; e < (long)mesh.elements.size(); ++e) mesh.elements[e].compute_metrics(mesh.qx);
}

// ---------------------------------------------------------- a5: glaplace --
// glaplace.hpp:16-28: u_xi = D u; u_eta = u D^T; F = D^T (W (A%u_xi - B%u_eta));
// G = ((C%u_eta - B%u_xi) W) D; out = F W + W G.
static void glaplace(const double* u, const Quad& el, const vec& D, const vec& w, int n, double* out)
{
    std::vector<double> uxi(n * n), ueta(n * n), P(n * n), Q(n * n), F(n * n), G(n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0, t = 0;
            for (int k = 0; k < n; ++k) {
                s += D[i + k * n] * u[k + j * n];
                t += u[i + k * n] * D[j + k * n];
            }
            uxi[i + j * n] = s;
            ueta[i + j * n] = t;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            int k = i + j * n;
            P[k] = w[i] * (el.A[k] * uxi[k] - el.B[k] * ueta[k]);
            Q[k] = (el.C[k] * ueta[k] - el.B[k] * uxi[k]) * w[j];
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0, t = 0;
            for (int k = 0; k < n; ++k) {
                s += D[k + i * n] * P[k + j * n];  // (D^T P)(i,j)
                t += Q[i + k * n] * D[k + j * n];  // (Q D)(i,j)
            }
            F[i + j * n] = s;
            G[i + j * n] = t;
        }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) out[i + j * n] = F[i + j * n] * w[j] + w[i] * G[i + j * n];
}

// (diagmat(w) * u * diagmat(w)) % J, openmp_impl/galerkin.hpp:51
static inline double wuwJ(const vec& w, int i, int j, double u, double J) { return ((w[i] * u) * w[j]) * J; }

// ------------------------------------------------ a12-a14: DSS and masks --
// get_edge_values.hpp:7-22 / set_edge_values.hpp:9-31 index helpers
static inline size_t side_index(const Mesh& m, int e, int s, int k)
{
    // entry k (1..N-2) along side s of element e
    if (s == 2 || s == 4) return (size_t)e * m.n2() + m.smap(s) + (size_t)k * m.N;
    return (size_t)e * m.n2() + k + (size_t)m.smap(s) * m.N;
}

// openmp_impl/unmask.hpp:12-55
static void unmask(double* a, const Mesh& m)
{
    const int N = m.N;
#pragma omp parallel for schedule(static) if (m.edges.size() > 63)
    for (long t = 0; t < (long)m.edges.size(); ++t) {
        const Edge& ed = m.edges[t];
        if (ed.edge_type == BOUNDARY) continue;
        int e0 = ed.elements[0], s0 = ed.element_sides[0];
        int e1 = ed.elements[1], s1 = std::abs(ed.element_sides[1]);
        int kk = ed.startIter;
        for (int k = 1; k < N - 1; ++k) {
            a[side_index(m, e1, s1, kk)] = a[side_index(m, e0, s0, k)];
            kk += ed.deltaIter;
        }
    }
#pragma omp parallel for schedule(static) if (m.nodes.size() > 63)
    for (long t = 0; t < (long)m.nodes.size(); ++t) {
        const Corner& c = m.nodes[t];
        if (c.corner_type == BOUNDARY || c.connected_elements.empty()) continue;
        const CornerInfo& f = c.connected_elements[0];
        double cval = a[(size_t)f.element_id * m.n2() + f.i + f.j * N];
        for (size_t q = 1; q < c.connected_elements.size(); ++q) {
            const CornerInfo& g = c.connected_elements[q];
            a[(size_t)g.element_id * m.n2() + g.i + g.j * N] = cval;
        }
    }
}

// openmp_impl/global_sum.hpp:13-77
static void global_sum(double* a, const Mesh& m)
{
    const int N = m.N;
#pragma omp parallel for schedule(static) if (m.edges.size() > 63)
    for (long t = 0; t < (long)m.edges.size(); ++t) {
        const Edge& ed = m.edges[t];
        if (ed.edge_type == BOUNDARY) continue;
        int e0 = ed.elements[0], s0 = std::abs(ed.element_sides[0]);
        int e1 = ed.elements[1], s1 = std::abs(ed.element_sides[1]);
        int n = ed.startIter;
        for (int j = 1; j < N - 1; ++j) {
            size_t i0 = side_index(m, e0, s0, j), i1 = side_index(m, e1, s1, n);
            double s = a[i0] + a[i1];
            a[i0] = s;
            a[i1] = s;
            n += ed.deltaIter;
        }
    }
#pragma omp parallel for schedule(static) if (m.nodes.size() > 63)
    for (long t = 0; t < (long)m.nodes.size(); ++t) {
        const Corner& c = m.nodes[t];
        if (c.corner_type == BOUNDARY) continue;
        double csum = 0;
        for (const CornerInfo& g : c.connected_elements) csum += a[(size_t)g.element_id * m.n2() + g.i + g.j * N];
        for (const CornerInfo& g : c.connected_elements) a[(size_t)g.element_id * m.n2() + g.i + g.j * N] = csum;
    }
}

// openmp_impl/mask.hpp:16-28, mask_side.hpp:12-52, mask_node.hpp:8-22
static void mask(double* a, const Mesh& m)
{
    const int N = m.N;
#pragma omp parallel for schedule(static) if (m.edges.size() > 63)
    for (long t = 0; t < (long)m.edges.size(); ++t) {
        const Edge& ed = m.edges[t];
        int e, s;
        if (ed.edge_type == BOUNDARY) { e = ed.elements[0]; s = ed.element_sides[0]; }
        else { e = ed.elements[1]; s = std::abs(ed.element_sides[1]); }
        for (int k = 1; k < N - 1; ++k) a[side_index(m, e, s, k)] = 0;
    }
#pragma omp parallel for schedule(static) if (m.edges.size() > 63)
    for (long t = 0; t < (long)m.nodes.size(); ++t) {
        const Corner& c = m.nodes[t];
        size_t q0 = (c.corner_type != BOUNDARY) ? 1 : 0;
        for (size_t q = q0; q < c.connected_elements.size(); ++q) {
            const CornerInfo& g = c.connected_elements[q];
            a[(size_t)g.element_id * m.n2() + g.i + g.j * N] = 0.0;
        }
    }
}

// a17: openmp_impl/dot.hpp:11-61 -- unique-DOF Euclidean product; interiors,
// then edges (owner side, boundary edges included), then corners (first
// connected element, boundary corners included).
static double dot(const Mesh& m, const double* a, const double* b)
{
    const int n = m.N;
    const long E = (long)m.elements.size();
    double p_int = 0;
#pragma omp parallel for reduction(+ : p_int) schedule(static) if (E > 63)
    for (long e = 0; e < E; ++e) {
        double s = 0;
        const double* x = a + e * m.n2();
        const double* y = b + e * m.n2();
        for (int j = 1; j < n - 1; ++j)
            for (int i = 1; i < n - 1; ++i) s += x[i + j * n] * y[i + j * n];
        p_int += s;
    }
    double p_edges = 0;
#pragma omp parallel for reduction(+ : p_edges) schedule(static) if (m.edges.size() > 63)
    for (long t = 0; t < (long)m.edges.size(); ++t) {
        const Edge& ed = m.edges[t];
        int e = ed.elements[0], s = std::abs(ed.element_sides[0]);
        double pe = 0;
        for (int k = 1; k < n - 1; ++k) {
            size_t q = side_index(m, e, s, k);
            pe += a[q] * b[q];
        }
        p_edges += pe;
    }
    double p_corners = 0;
#pragma omp parallel for reduction(+ : p_corners) schedule(static) if (m.nodes.size() > 63)
    for (long t = 0; t < (long)m.nodes.size(); ++t) {
        const CornerInfo& g = m.nodes[t].connected_elements[0];
        size_t q = (size_t)g.element_id * m.n2() + g.i + g.j * n;
        p_corners += a[q] * b[q];
    }
    return p_int + p_edges + p_corners;
}

// --------------------------------------------- a6-a11: element operators --
enum OpKind { OP_LAPLACE = 0, OP_HELMHOLTZ = 1, OP_MASS = 2, OP_SCHRO_RES = 3, OP_SCHRO_F = 4, OP_SCHRO_L = 5 };

static inline int op_comps(int kind) { return (kind == OP_SCHRO_RES || kind == OP_SCHRO_F) ? 2 : 1; }

// Element-local operator (no DSS) on element e. in/out point at the element's
// N*N block of each component (component stride cs).
static void element_op(const Mesh& m, int kind, const double* coef, double alpha, long e, const double* in, double* out, size_t cs)
{
    const int n = m.N;
    const size_t n2 = m.n2();
    const Quad& el = m.elements[e];
    const vec& w = m.qw;
    std::vector<double> t0(n2), t1(n2);
    switch (kind) {
    case OP_LAPLACE:  // openmp_impl/poisson.hpp:37-42
        glaplace(in, el, m.D, w, n, out);
        break;
    case OP_HELMHOLTZ:  // openmp_impl/helmholtz.hpp:38-46: cuv + glaplace(u)
        glaplace(in, el, m.D, w, n, t0.data());
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int k = i + j * n;
                out[k] = wuwJ(w, i, j, in[k] * coef[e * n2 + k], el.J[k]) + t0[k];
            }
        break;
    case OP_MASS:  // openmp_impl/galerkin.hpp:49-51
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int k = i + j * n;
                out[k] = wuwJ(w, i, j, in[k], el.J[k]);
            }
        break;
    case OP_SCHRO_L:  // schrodinger.hpp:37-43: nlap + proj_Vu
        glaplace(in, el, m.D, w, n, t0.data());
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int k = i + j * n;
                out[k] = t0[k] + wuwJ(w, i, j, in[k] * coef[e * n2 + k], el.J[k]);
            }
        break;
    case OP_SCHRO_RES: {  // schrodinger.hpp:185-207 (also :45-67, :344-366)
        const double* u = in;
        const double* v = in + cs;
        std::vector<double> Lu(n2), Lv(n2);
        element_op(m, OP_SCHRO_L, coef, 0, e, u, Lu.data(), 0);
        element_op(m, OP_SCHRO_L, coef, 0, e, v, Lv.data(), 0);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int k = i + j * n;
                double pu = wuwJ(w, i, j, u[k], el.J[k]);
                double pv = wuwJ(w, i, j, v[k], el.J[k]);
                out[k] = pu + alpha * Lv[k];
                out[k + cs] = alpha * Lu[k] - pv;
            }
        break;
    }
    case OP_SCHRO_F: {  // schrodinger.hpp:163-183: [-L v, -L u]
        const double* u = in;
        const double* v = in + cs;
        std::vector<double> Lu(n2), Lv(n2);
        element_op(m, OP_SCHRO_L, coef, 0, e, u, Lu.data(), 0);
        element_op(m, OP_SCHRO_L, coef, 0, e, v, Lv.data(), 0);
        for (size_t k = 0; k < n2; ++k) {
            out[k] = -Lv[k];
            out[k + cs] = -Lu[k];
        }
        break;
    }
    }
}

static void element_op_all(const Mesh& m, int kind, const double* coef, double alpha, const double* in, double* out)
{
    const long E = (long)m.elements.size();
    const size_t cs = m.len();
#pragma omp parallel for schedule(static) if (E > 63)
    for (long e = 0; e < E; ++e) element_op(m, kind, coef, alpha, e, in + e * m.n2(), out + e * m.n2(), cs);
}

// galerkin_op (openmp_impl/galerkin.hpp:24-35) and the Res / f lambdas:
// copy -> unmask (each comp) -> element op -> global_sum -> mask.
static vec galerkin(const Mesh& m, int kind, const double* coef, double alpha, const double* in)
{
    const int nc = op_comps(kind);
    const size_t L = m.len();
    vec y(in, in + nc * L);
    for (int c = 0; c < nc; ++c) unmask(y.data() + c * L, m);
    vec out(nc * L);
    element_op_all(m, kind, coef, alpha, y.data(), out.data());
    for (int c = 0; c < nc; ++c) global_sum(out.data() + c * L, m);
    for (int c = 0; c < nc; ++c) mask(out.data() + c * L, m);
    return out;
}

// ------------------------------------------------ a19-a21: Krylov solvers --
struct Results { int success; int n_iter; double residual; };

// minres.hpp:19-107 over flat vectors; A and dot are callables.
template <class LinOp, class Dot>
static Results minres(vec& x, LinOp A, const vec& b, Dot dotp, int maxit, double tol)
{
    const size_t n = x.size();
    bool success = false;
    vec Ax = A(x);
    vec r1(n);
    for (size_t k = 0; k < n; ++k) r1[k] = b[k] - Ax[k];
    vec y = r1;  // IdentityPreconditioner (solver_base.hpp:14-22)
    vec r2 = r1;
    double b_norm = std::sqrt(dotp(r1, r1));
    const double machine_epsilon = std::numeric_limits<double>::epsilon();
    double beta1 = std::sqrt(dotp(y, r1));
    double beta = beta1, beta_old = 0, tnorm = 0, dbar = 0, eps = 0, eps_old = 0, phibar = beta1, cs = -1, sn = 0;
    vec w(n);
    for (size_t k = 0; k < n; ++k) w[k] = 0.0 * x[k];
    vec w2 = w, w1;
    vec v(n);
    int i;
    for (i = 1; i <= maxit; ++i) {
        double ib = 1.0 / beta;
        for (size_t k = 0; k < n; ++k) v[k] = ib * y[k];
        y = A(v);
        if (i >= 2) {
            double c = beta / beta_old;
            for (size_t k = 0; k < n; ++k) y[k] -= c * r1[k];
        }
        double alpha = dotp(v, y);
        double c2 = alpha / beta;
        for (size_t k = 0; k < n; ++k) y[k] -= c2 * r2[k];
        r1.swap(r2);
        r2 = y;
        beta_old = beta;
        beta = std::sqrt(dotp(y, y));
        tnorm += alpha * alpha + beta * beta + beta_old * beta_old;
        eps_old = eps;
        double delta = cs * dbar + sn * alpha;
        double gbar = sn * dbar - cs * alpha;
        eps = sn * beta;
        dbar = -cs * beta;
        double gamma = std::max(std::sqrt(gbar * gbar + beta * beta), machine_epsilon);
        cs = gbar / gamma;
        sn = beta / gamma;
        double phi = cs * phibar;
        phibar *= sn;
        w1.swap(w2);
        w2.swap(w);
        double ig = 1.0 / gamma;
        if (w.size() != n) w.resize(n);
        for (size_t k = 0; k < n; ++k) w[k] = ig * ((v[k] - eps_old * w1[k]) - delta * w2[k]);
        for (size_t k = 0; k < n; ++k) x[k] += phi * w[k];
        if (phibar < tol * b_norm + tol) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, phibar};
}

// pcg.hpp:59-97 with the IdentityPreconditioner conj_grad overloads (:20-25, :41-50)
template <class LinOp, class Dot>
static Results pcg(vec& x, LinOp A, const vec& b, Dot dotp, int max_iter, double tol)
{
    const size_t n = x.size();
    vec Ax = A(x);
    vec r(n);
    for (size_t k = 0; k < n; ++k) r[k] = b[k] - Ax[k];
    double bnorm = std::sqrt(dotp(b, b));
    bool success = false;
    vec p;
    double rho_prev = 0, rho = 0;
    int i;
    for (i = 0; i < max_iter; ++i) {
        if (i == 0) {
            p = r;
            rho = dotp(r, r);
        } else {
            rho = dotp(r, r);
            double beta = rho / rho_prev;
            for (size_t k = 0; k < n; ++k) p[k] = r[k] + beta * p[k];
        }
        vec Ap = A(p);
        double alpha = rho / dotp(p, Ap);
        for (size_t k = 0; k < n; ++k) x[k] += alpha * p[k];
        for (size_t k = 0; k < n; ++k) r[k] -= alpha * Ap[k];
        rho_prev = rho;
        if (std::sqrt(rho) < tol * bnorm) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, std::sqrt(rho)};
}

// minres.hpp:19-107 and pcg.hpp:12-97 with a preconditioner z = P(r) (the
// Precond argument; P empty means IdentityPreconditioner). MINRES follows
// minres.hpp:24 (y = precond(r1)), :31 (beta1 = sqrt(dot(y, r1))) and :60-63
// (y = precond(r2); beta = sqrt(dot(y, y))); consistent = true takes
// beta = sqrt(dot(r2, y)) instead.
typedef std::function<vec(const vec&)> Precond;

template <class LinOp, class Dot>
static Results minres_pc(vec& x, LinOp A, const vec& b, Dot dotp, const Precond& P, bool consistent, int maxit,
                         double tol)
{
    const size_t n = x.size();
    bool success = false;
    vec Ax = A(x);
    vec r1(n);
    for (size_t k = 0; k < n; ++k) r1[k] = b[k] - Ax[k];
    vec y = P ? P(r1) : r1;
    vec r2 = r1;
    double b_norm = std::sqrt(dotp(r1, r1));
    const double machine_epsilon = std::numeric_limits<double>::epsilon();
    double beta1 = std::sqrt(dotp(y, r1));
    double beta = beta1, beta_old = 0, tnorm = 0, dbar = 0, eps = 0, eps_old = 0, phibar = beta1, cs = -1, sn = 0;
    vec w(n);
    for (size_t k = 0; k < n; ++k) w[k] = 0.0 * x[k];
    vec w2 = w, w1;
    vec v(n);
    int i;
    for (i = 1; i <= maxit; ++i) {
        double ib = 1.0 / beta;
        for (size_t k = 0; k < n; ++k) v[k] = ib * y[k];
        y = A(v);
        if (i >= 2) {
            double c = beta / beta_old;
            for (size_t k = 0; k < n; ++k) y[k] -= c * r1[k];
        }
        double alpha = dotp(v, y);
        double c2 = alpha / beta;
        for (size_t k = 0; k < n; ++k) y[k] -= c2 * r2[k];
        r1.swap(r2);
        r2 = y;
        if (P) y = P(r2);
        beta_old = beta;
        beta = std::sqrt(consistent ? dotp(r2, y) : dotp(y, y));
        tnorm += alpha * alpha + beta * beta + beta_old * beta_old;
        eps_old = eps;
        double delta = cs * dbar + sn * alpha;
        double gbar = sn * dbar - cs * alpha;
        eps = sn * beta;
        dbar = -cs * beta;
        double gamma = std::max(std::sqrt(gbar * gbar + beta * beta), machine_epsilon);
        cs = gbar / gamma;
        sn = beta / gamma;
        double phi = cs * phibar;
        phibar *= sn;
        w1.swap(w2);
        w2.swap(w);
        double ig = 1.0 / gamma;
        if (w.size() != n) w.resize(n);
        for (size_t k = 0; k < n; ++k) w[k] = ig * ((v[k] - eps_old * w1[k]) - delta * w2[k]);
        for (size_t k = 0; k < n; ++k) x[k] += phi * w[k];
        if (phibar < tol * b_norm + tol) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, phibar};
}

// pcg.hpp:59-97 with the preconditioned conj_grad overloads (:12-17, :27-39)
template <class LinOp, class Dot>
static Results pcg_pc(vec& x, LinOp A, const vec& b, Dot dotp, const Precond& P, int max_iter, double tol)
{
    const size_t n = x.size();
    vec Ax = A(x);
    vec r(n);
    for (size_t k = 0; k < n; ++k) r[k] = b[k] - Ax[k];
    double bnorm = std::sqrt(dotp(b, b));
    bool success = false;
    vec p;
    double rho_prev = 0, rho = 0;
    int i;
    for (i = 0; i < max_iter; ++i) {
        if (i == 0) {
            p = P ? P(r) : r;
            rho = dotp(r, p);
        } else {
            vec z = P ? P(r) : r;
            rho = dotp(r, z);
            double beta = rho / rho_prev;
            for (size_t k = 0; k < n; ++k) p[k] = z[k] + beta * p[k];
        }
        vec Ap = A(p);
        double alpha = rho / dotp(p, Ap);
        for (size_t k = 0; k < n; ++k) x[k] += alpha * p[k];
        for (size_t k = 0; k < n; ++k) r[k] -= alpha * Ap[k];
        rho_prev = rho;
        if (std::sqrt(rho) < tol * bnorm) {
            success = true;
            break;
        }
    }
    return {success ? 1 : 0, i, std::sqrt(rho)};
}

// ------------------------------------------------------ a22/a23 drivers ---
// openmp_impl/poisson.hpp:17-59
static Results poisson(const Mesh& m, double* u, const double* f, int max_iter, double tol)
{
    const size_t L = m.len();
    vec b = galerkin(m, OP_MASS, nullptr, 0, f);  // gproj(f)
    vec x(u, u + L);
    auto A = [&](const vec& v) { return galerkin(m, OP_LAPLACE, nullptr, 0, v.data()); };
    auto d = [&](const vec& p, const vec& q) { return dot(m, p.data(), q.data()); };
    Results r = pcg(x, A, b, d, max_iter, tol);
    std::copy(x.begin(), x.end(), u);
    return r;
}

// openmp_impl/helmholtz.hpp:18-64
static Results helmholtz(const Mesh& m, double* u, const double* c, const double* f, int max_iter, double tol)
{
    const size_t L = m.len();
    vec b = galerkin(m, OP_MASS, nullptr, 0, f);
    vec x(u, u + L);
    auto A = [&](const vec& v) { return galerkin(m, OP_HELMHOLTZ, c, 0, v.data()); };
    auto d = [&](const vec& p, const vec& q) { return dot(m, p.data(), q.data()); };
    Results r = minres(x, A, b, d, max_iter, tol);
    std::copy(x.begin(), x.end(), u);
    return r;
}

// ----------------------------------------------------- a24-a26 steppers ---
typedef int (*StepCallback)(const double* re, const double* im, void* user);

struct StepStats {
    int steps = 0;
    long total_iters = 0;
    int stage_solves = 0;
    double last_error_estimate = 0;
};

static double dot2(const Mesh& m, const vec& a, const vec& b)  // schrodinger.hpp:209-212
{
    const size_t L = m.len();
    return dot(m, a.data(), b.data()) + dot(m, a.data() + L, b.data() + L);
}

static void unmask2(const Mesh& m, vec& p)
{
    unmask(p.data(), m);
    unmask(p.data() + m.len(), m);
}

// schrodinger.hpp:22-128
static int schrodinger_bdf2(const Mesh& m, double* re, double* im, const double* V, double T, int max_iter, double tol, double dt,
                            StepCallback cb, void* user, StepStats& st, int* stage_iters, int max_stage_records)
{
    const size_t L = m.len();
    double alpha = dt;
    vec p(2 * L);
    std::copy(re, re + L, p.begin());
    std::copy(im, im + L, p.begin() + L);
    auto Res = [&](const vec& q) { return galerkin(m, OP_SCHRO_RES, V, alpha, q.data()); };
    auto proj = [&](const vec& q) {  // schrodinger.hpp:69-78
        vec out(2 * L);
        vec a = galerkin(m, OP_MASS, nullptr, 0, q.data());
        vec b = galerkin(m, OP_MASS, nullptr, 0, q.data() + L);
        for (size_t k = 0; k < L; ++k) {
            out[k] = a[k];
            out[k + L] = -1.0 * b[k];
        }
        return out;
    };
    auto d = [&](const vec& a, const vec& b) { return dot2(m, a, b); };
    std::vector<vec> ps;
    bool success = true;
    double t = 0;
    while (t < T) {
        if (ps.size() == 2) ps.erase(ps.begin());
        ps.push_back(proj(p));
        Results r;
        if (ps.size() == 1) {
            r = minres(p, Res, ps[0], d, max_iter, tol);
            alpha = (2.0 / 3.0) * dt;
        } else {
            vec b(2 * L);
            for (size_t k = 0; k < 2 * L; ++k) b[k] = (-1.0 / 3.0) * ps[0][k] + (4.0 / 3.0) * ps[1][k];
            r = minres(p, Res, b, d, max_iter, tol);
        }
        if (stage_iters && st.stage_solves < max_stage_records) stage_iters[st.stage_solves] = r.n_iter;
        st.stage_solves++;
        st.total_iters += r.n_iter;
        unmask2(m, p);
        t += dt;
        st.steps++;
        success = r.success;
        if (!success) break;
        if (cb && cb(p.data(), p.data() + L, user)) break;
    }
    std::copy(p.begin(), p.begin() + L, re);
    std::copy(p.begin() + L, p.end(), im);
    return success ? 1 : 0;
}

// z = dt * f(q)
static vec stage_rhs(const Mesh& m, const double* V, double dt, const vec& q)
{
    vec v = galerkin(m, OP_SCHRO_F, V, 0, q.data());
    for (double& a : v) a = dt * a;
    return v;
}

// schrodinger.hpp:138-285
static int schrodinger_dirk4s5(const Mesh& m, double* re, double* im, const double* V, double T, int max_iter, double tol, double dt,
                               StepCallback cb, void* user, StepStats& st, int* stage_iters, int max_stage_records)
{
    const size_t L = m.len();
    const size_t n = 2 * L;
    const double alpha = 0.25 * dt;
    vec p(n);
    std::copy(re, re + L, p.begin());
    std::copy(im, im + L, p.begin() + L);
    vec zeros(n);
    for (size_t k = 0; k < n; ++k) zeros[k] = 0.0 * p[k];
    auto Res = [&](const vec& q) { return galerkin(m, OP_SCHRO_RES, V, alpha, q.data()); };
    auto d = [&](const vec& a, const vec& b) { return dot2(m, a, b); };
    double pnorm = std::sqrt(dot2(m, p, p));
    double t = 0;
    bool success = false;
    auto solve = [&](vec& z, const vec& v) {
        z = zeros;
        Results r = minres(z, Res, v, d, max_iter, tol);
        if (stage_iters && st.stage_solves < max_stage_records) stage_iters[st.stage_solves] = r.n_iter;
        st.stage_solves++;
        st.total_iters += r.n_iter;
        return r.success != 0;
    };
    vec z1, z2, z3, z4, z5, v(n);
    while (t < T) {
        v = stage_rhs(m, V, dt, p);
        if (!(success = solve(z1, v))) break;
        for (size_t k = 0; k < n; ++k) v[k] = p[k] + 0.5 * z1[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z2, v))) break;
        for (size_t k = 0; k < n; ++k) v[k] = (p[k] + (17.0 / 50.0) * z1[k]) + (-1.0 / 25.0) * z2[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z3, v))) break;
        for (size_t k = 0; k < n; ++k) v[k] = ((p[k] + (371.0 / 1360.0) * z1[k]) + (-137.0 / 2720.0) * z2[k]) + (15.0 / 544.0) * z3[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z4, v))) break;
        for (size_t k = 0; k < n; ++k)
            v[k] = (((p[k] + (25.0 / 24.0) * z1[k]) + (-49.0 / 48.0) * z2[k]) + (125.0 / 16.0) * z3[k]) + (-85.0 / 12.0) * z4[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z5, v))) break;
        for (size_t k = 0; k < n; ++k)
            p[k] += ((((25.0 / 24.0) * z1[k] + (-49.0 / 48.0) * z2[k]) + (125.0 / 16.0) * z3[k]) + (-85.0 / 12.0) * z4[k]) + 0.25 * z5[k];
        vec err(n);
        for (size_t k = 0; k < n; ++k)
            err[k] = (((-3.0 / 16.0) * z1[k] + (-27.0 / 32.0) * z2[k]) + (25.0 / 32.0) * z3[k]) + 0.25 * z5[k];
        st.last_error_estimate = std::sqrt(dot2(m, err, err)) / pnorm;
        unmask2(m, p);
        t += dt;
        st.steps++;
        if (cb && cb(p.data(), p.data() + L, user)) break;
    }
    std::copy(p.begin(), p.begin() + L, re);
    std::copy(p.begin() + L, p.end(), im);
    return success ? 1 : 0;
}

// schrodinger.hpp:294-423
static int schrodinger_dirk4s3(const Mesh& m, double* re, double* im, const double* V, double T, int max_iter, double tol, double dt,
                               StepCallback cb, void* user, StepStats& st, int* stage_iters, int max_stage_records)
{
    const double gamma = std::cos(M_PI / 18.0) / std::sqrt(3.0) + 0.5;
    const double delta = 1.0 / 6.0 / std::pow(2 * gamma - 1, 2);
    const size_t L = m.len();
    const size_t n = 2 * L;
    const double alpha = gamma * dt;
    vec p(n);
    std::copy(re, re + L, p.begin());
    std::copy(im, im + L, p.begin() + L);
    vec zeros(n);
    for (size_t k = 0; k < n; ++k) zeros[k] = 0 * p[k];
    auto Res = [&](const vec& q) { return galerkin(m, OP_SCHRO_RES, V, alpha, q.data()); };
    auto d = [&](const vec& a, const vec& b) { return dot2(m, a, b); };
    bool success = true;
    double t = 0;
    auto solve = [&](vec& z, const vec& v) {
        z = zeros;
        Results r = minres(z, Res, v, d, max_iter, tol);
        if (stage_iters && st.stage_solves < max_stage_records) stage_iters[st.stage_solves] = r.n_iter;
        st.stage_solves++;
        st.total_iters += r.n_iter;
        return r.success != 0;
    };
    vec z1, z2, z3, v(n);
    while (t < T) {
        v = stage_rhs(m, V, dt, p);
        if (!(success = solve(z1, v))) break;
        for (size_t k = 0; k < n; ++k) v[k] = p[k] + (0.5 - gamma) * z1[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z2, v))) break;
        for (size_t k = 0; k < n; ++k) v[k] = (p[k] + (2 * gamma) * z1[k]) + (1 - 4 * gamma) * z2[k];
        v = stage_rhs(m, V, dt, v);
        if (!(success = solve(z3, v))) break;
        for (size_t k = 0; k < n; ++k) p[k] += ((delta * z1[k] + (1 - 2 * delta) * z2[k]) + delta * z3[k]);
        unmask2(m, p);
        t += dt;
        st.steps++;
        if (cb && cb(p.data(), p.data() + L, user)) break;
    }
    std::copy(p.begin(), p.begin() + L, re);
    std::copy(p.begin() + L, p.end(), im);
    return success ? 1 : 0;
}

}  // namespace orc

// ============================================================ C entry ====
using namespace orc;

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_dynamic(0);
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads()
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_gll(int p, double* x, double* w)
{
    try {
        vec xx, ww;
        gauss_lobatto(p, xx, ww);
        std::copy(xx.begin(), xx.end(), x);
        std::copy(ww.begin(), ww.end(), w);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void orc_dmat(int n, const double* x, double* D)
{
    vec xx(x, x + n);
    vec d = derivative_matrix(xx);
    std::copy(d.begin(), d.end(), D);
}

// One element's metrics: corners xs[4], ys[4], GLL nodes z[n] -> J,A,B,C (n*n each).
void orc_quad_metrics(const double* xs, const double* ys, int n, const double* z, double* J, double* A, double* B, double* C)
{
    Quad q;
    for (int k = 0; k < 4; ++k) { q.xs[k] = xs[k]; q.ys[k] = ys[k]; }
    q.compute_metrics(vec(z, z + n));
    std::copy(q.J.begin(), q.J.end(), J);
    std::copy(q.A.begin(), q.A.end(), A);
    std::copy(q.B.begin(), q.B.end(), B);
    std::copy(q.C.begin(), q.C.end(), C);
}

// One element's glaplace from corners (glaplace.hpp:16-28).
void orc_glaplace_quad(const double* xs, const double* ys, int p, const double* u, double* out)
{
    vec x, w;
    gauss_lobatto(p, x, w);
    vec D = derivative_matrix(x);
    Quad q;
    for (int k = 0; k < 4; ++k) { q.xs[k] = xs[k]; q.ys[k] = ys[k]; }
    q.compute_metrics(x);
    glaplace(u, q, D, w, p + 1, out);
}

void* orc_mesh_create(int p, const double* xy, int n_nodes, const int* quads, int n_el)
{
    try {
        Mesh* m = new Mesh(build_mesh(p, xy, n_nodes, quads, n_el));
        compute_metrics(*m);
        return m;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// load_mesh.hpp:65-144: folder with info.txt, coordinates.txt, elements.txt
void* orc_mesh_load(int p, const char* folder)
{
    try {
        std::string f(folder);
        std::ifstream info(f + "/info.txt");
        if (!info.is_open()) throw std::runtime_error("could not open info file.");
        int nn, ne;
        info >> nn >> ne;
        std::ifstream coo(f + "/coordinates.txt");
        if (!coo.is_open()) throw std::runtime_error("could not open coordinate file.");
        std::vector<double> xy(2 * nn);
        for (int k = 0; k < nn; ++k) coo >> xy[2 * k] >> xy[2 * k + 1];
        std::ifstream el(f + "/elements.txt");
        if (!el.is_open()) throw std::runtime_error("could not open element file.");
        std::vector<int> q(4 * ne);
        for (int k = 0; k < 4 * ne; ++k) el >> q[k];
        return orc_mesh_create(p, xy.data(), nn, q.data(), ne);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_mesh_free(void* h) { delete (Mesh*)h; }

// out[0..6] = N, n_elements, n_nodes, n_edges, dof, n_interior_edges, n_boundary_nodes
void orc_mesh_info(void* h, long* out)
{
    Mesh& m = *(Mesh*)h;
    long ie = 0, bn = 0;
    for (auto& e : m.edges) ie += (e.edge_type == INTERIOR);
    for (auto& c : m.nodes) bn += (c.corner_type == BOUNDARY);
    out[0] = m.N;
    out[1] = (long)m.elements.size();
    out[2] = (long)m.nodes.size();
    out[3] = (long)m.edges.size();
    out[4] = m.dof();
    out[5] = ie;
    out[6] = bn;
}

void orc_mesh_gll(void* h, double* x, double* w)
{
    Mesh& m = *(Mesh*)h;
    std::copy(m.qx.begin(), m.qx.end(), x);
    std::copy(m.qw.begin(), m.qw.end(), w);
}

void orc_mesh_D(void* h, double* D)
{
    Mesh& m = *(Mesh*)h;
    std::copy(m.D.begin(), m.D.end(), D);
}

void orc_mesh_metrics(void* h, double* J, double* A, double* B, double* C)
{
    Mesh& m = *(Mesh*)h;
    size_t n2 = m.n2();
    for (size_t e = 0; e < m.elements.size(); ++e) {
        const Quad& q = m.elements[e];
        std::copy(q.J.begin(), q.J.end(), J + e * n2);
        std::copy(q.A.begin(), q.A.end(), A + e * n2);
        std::copy(q.B.begin(), q.B.end(), B + e * n2);
        std::copy(q.C.begin(), q.C.end(), C + e * n2);
    }
}

// physical coordinates of every element-local GLL node via from_local_coo
void orc_mesh_coords(void* h, double* X, double* Y)
{
    Mesh& m = *(Mesh*)h;
    const int N = m.N;
    for (size_t e = 0; e < m.elements.size(); ++e)
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i)
                m.elements[e].from_local_coo(m.qx[i], m.qx[j], X[e * m.n2() + i + j * N], Y[e * m.n2() + i + j * N]);
}

// per edge 7 ints: type, el0, el1, side0, side1, startIter, deltaIter
void orc_mesh_edges(void* h, int* out)
{
    Mesh& m = *(Mesh*)h;
    for (size_t k = 0; k < m.edges.size(); ++k) {
        const Edge& e = m.edges[k];
        int* o = out + 7 * k;
        o[0] = e.edge_type; o[1] = e.elements[0]; o[2] = e.elements[1];
        o[3] = e.element_sides[0]; o[4] = e.element_sides[1]; o[5] = e.startIter; o[6] = e.deltaIter;
    }
}

// corner types (n_nodes ints)
void orc_mesh_corner_types(void* h, int* out)
{
    Mesh& m = *(Mesh*)h;
    for (size_t k = 0; k < m.nodes.size(); ++k) out[k] = m.nodes[k].corner_type;
}

void orc_unmask(void* h, double* a) { unmask(a, *(Mesh*)h); }
void orc_global_sum(void* h, double* a) { global_sum(a, *(Mesh*)h); }
void orc_mask(void* h, double* a) { mask(a, *(Mesh*)h); }
double orc_dot(void* h, const double* a, const double* b) { return dot(*(Mesh*)h, a, b); }

// element-local operator over all elements, no DSS (kind: OpKind)
void orc_element_op(void* h, int kind, const double* coef, double alpha, const double* in, double* out)
{
    element_op_all(*(Mesh*)h, kind, coef, alpha, in, out);
}

// galerkin_op: copy, unmask, element op, global_sum, mask
void orc_galerkin(void* h, int kind, const double* coef, double alpha, const double* in, double* out)
{
    Mesh& m = *(Mesh*)h;
    vec y = galerkin(m, kind, coef, alpha, in);
    std::copy(y.begin(), y.end(), out);
}

// Solve Res(alpha, V) z = b with minres (one Schrodinger stage solve), z in/out.
int orc_stage_solve(void* h, const double* V, double alpha, const double* b, double* z, int max_iter, double tol, int* n_iter, double* resid)
{
    Mesh& m = *(Mesh*)h;
    const size_t n = 2 * m.len();
    vec x(z, z + n), bb(b, b + n);
    auto Res = [&](const vec& q) { return galerkin(m, OP_SCHRO_RES, V, alpha, q.data()); };
    auto d = [&](const vec& a, const vec& c) { return dot2(m, a, c); };
    Results r = minres(x, Res, bb, d, max_iter, tol);
    std::copy(x.begin(), x.end(), z);
    if (n_iter) *n_iter = r.n_iter;
    if (resid) *resid = r.residual;
    return r.success;
}

// generic solvers on a tridiagonal operator y_i = lo*x_{i-1} + di*x_i + up*x_{i+1}
// (arma::conv(u, {lo,di,up}, "same") in tests/pcg.cpp:13-16, tests/minres.cpp:14-17),
// Euclidean dot.
static vec tridiag_apply(const vec& x, double lo, double di, double up)
{
    size_t n = x.size();
    vec y(n);
    for (size_t i = 0; i < n; ++i) {
        double s = di * x[i];
        if (i > 0) s += lo * x[i - 1];
        if (i + 1 < n) s += up * x[i + 1];
        y[i] = s;
    }
    return y;
}

static double edot(const vec& a, const vec& b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

int orc_minres_tridiag(int n, double lo, double di, double up, const double* b, double* x, int maxit, double tol, int* n_iter, double* resid)
{
    vec xx(x, x + n), bb(b, b + n);
    Results r = minres(xx, [&](const vec& v) { return tridiag_apply(v, lo, di, up); }, bb, edot, maxit, tol);
    std::copy(xx.begin(), xx.end(), x);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

int orc_pcg_tridiag(int n, double lo, double di, double up, const double* b, double* x, int maxit, double tol, int* n_iter, double* resid)
{
    vec xx(x, x + n), bb(b, b + n);
    Results r = pcg(xx, [&](const vec& v) { return tridiag_apply(v, lo, di, up); }, bb, edot, maxit, tol);
    std::copy(xx.begin(), xx.end(), x);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

// the same solvers with a diagonal preconditioner z = pdiag * r (pdiag NULL: none)
int orc_minres_tridiag_pc(int n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                          const double* pdiag, int consistent, int* n_iter, double* resid)
{
    vec xx(x, x + n), bb(b, b + n);
    Precond P;
    vec pd;
    if (pdiag) {
        pd.assign(pdiag, pdiag + n);
        P = [&pd](const vec& r) {
            vec z(r.size());
            for (size_t k = 0; k < r.size(); ++k) z[k] = pd[k] * r[k];
            return z;
        };
    }
    Results r = minres_pc(xx, [&](const vec& v) { return tridiag_apply(v, lo, di, up); }, bb, edot, P, consistent != 0,
                          maxit, tol);
    std::copy(xx.begin(), xx.end(), x);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

int orc_pcg_tridiag_pc(int n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                       const double* pdiag, int* n_iter, double* resid)
{
    vec xx(x, x + n), bb(b, b + n);
    Precond P;
    vec pd;
    if (pdiag) {
        pd.assign(pdiag, pdiag + n);
        P = [&pd](const vec& r) {
            vec z(r.size());
            for (size_t k = 0; k < r.size(); ++k) z[k] = pd[k] * r[k];
            return z;
        };
    }
    Results r = pcg_pc(xx, [&](const vec& v) { return tridiag_apply(v, lo, di, up); }, bb, edot, P, maxit, tol);
    std::copy(xx.begin(), xx.end(), x);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

int orc_poisson(void* h, double* u, const double* f, int max_iter, double tol, int* n_iter, double* resid)
{
    Results r = poisson(*(Mesh*)h, u, f, max_iter, tol);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

int orc_helmholtz(void* h, double* u, const double* c, const double* f, int max_iter, double tol, int* n_iter, double* resid)
{
    Results r = helmholtz(*(Mesh*)h, u, c, f, max_iter, tol);
    *n_iter = r.n_iter;
    *resid = r.residual;
    return r.success;
}

// scheme: 0 = bdf2, 1 = dirk4s3, 2 = dirk4s5. stats_out[0..3] = steps,
// stage solves, total MINRES iterations, last DIRK4S5 error estimate.
// stage_iters (optional, max_records long) receives each stage solve's n_iter.
int orc_schrodinger(void* h, int scheme, double* re, double* im, const double* V, double T, int max_iter, double tol, double dt,
                    StepCallback cb, void* user, double* stats_out, int* stage_iters, int max_records)
{
    Mesh& m = *(Mesh*)h;
    StepStats st;
    int ok = 0;
    if (scheme == 0) ok = schrodinger_bdf2(m, re, im, V, T, max_iter, tol, dt, cb, user, st, stage_iters, max_records);
    else if (scheme == 1) ok = schrodinger_dirk4s3(m, re, im, V, T, max_iter, tol, dt, cb, user, st, stage_iters, max_records);
    else ok = schrodinger_dirk4s5(m, re, im, V, T, max_iter, tol, dt, cb, user, st, stage_iters, max_records);
    if (stats_out) {
        stats_out[0] = st.steps;
        stats_out[1] = st.stage_solves;
        stats_out[2] = (double)st.total_iters;
        stats_out[3] = st.last_error_estimate;
    }
    return ok;
}

}  // extern "C"
