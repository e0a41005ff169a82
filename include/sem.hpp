// sem.hpp -- header-only C++ mirror of the reference's schro_omp API
// (include/openmp_impl/*.hpp) over the C-ABI of libsem_hip.so.
//
// Code written against the reference switches by replacing
//   #include "openmp_impl/schrodinger.hpp"  /  using namespace schro_omp;
// with
//   #include "sem.hpp"                       /  using namespace schro_hip;
// The containers stay the reference's: std::vector<Mat> with one N x N
// column-major matrix per element, where Mat is any type with memptr() and
// n_elem (arma::Mat<double> included) or std::vector<double> of N*N values.
// Only real = double is provided (the GPU path keeps FP64 at the interface).
#pragma once

#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sem.h"

namespace schro_hip {

template <typename real>
struct solver_results {  // solver_base.hpp:6-12
    bool success;
    int n_iter;
    real residual;
};

template <typename real>
struct solver_opts {  // ode_base.hpp:6-13
    int max_iter;
    real tol;
    real dt;
    bool verbose;
};

class NullCallback {  // ode_base.hpp:15-23
public:
    template <typename... Args>
    constexpr bool operator()(const Args&...)
    {
        return false;
    }
};

inline void check(int rc)
{
    if (rc != SEM_OK) throw std::runtime_error(std::string("libsem_hip: ") + sem_last_error());
}

namespace detail {
template <class M>
inline const double* data(const M& m)
{
    if constexpr (std::is_same_v<M, std::vector<double>>) return m.data();
    else return m.memptr();
}
template <class M>
inline double* data(M& m)
{
    if constexpr (std::is_same_v<M, std::vector<double>>) return m.data();
    else return m.memptr();
}
template <class M>
inline size_t size(const M& m)
{
    if constexpr (std::is_same_v<M, std::vector<double>>) return m.size();
    else return (size_t)m.n_elem;
}
template <class M>
std::vector<double> pack(const std::vector<M>& v, size_t n2)
{
    std::vector<double> out(v.size() * n2);
    for (size_t e = 0; e < v.size(); ++e) {
        if (detail::size(v[e]) != n2) throw std::invalid_argument("element block has the wrong size");
        const double* s = detail::data(v[e]);
        std::copy(s, s + n2, out.begin() + e * n2);
    }
    return out;
}
template <class M>
void unpack(const double* src, std::vector<M>& v, size_t n2)
{
    for (size_t e = 0; e < v.size(); ++e) std::copy(src + e * n2, src + (e + 1) * n2, detail::data(v[e]));
}
}  // namespace detail

// quad_rule / gauss_lobatto<real>(n): n + 1 GLL nodes and weights (gauss_lobatto.hpp:27-126)
template <typename real>
struct quad_rule {
    std::vector<real> x;
    std::vector<real> w;
};

template <typename real = double>
quad_rule<real> gauss_lobatto(int n)
{
    static_assert(std::is_same_v<real, double>, "the GPU library provides real = double");
    quad_rule<real> q;
    q.x.resize(n + 1);
    q.w.resize(n + 1);
    check(sem_gauss_lobatto(n, q.x.data(), q.w.data()));
    return q;
}

// derivative_matrix(x) (derivative_matrix.hpp:12-38): n x n, column-major
inline std::vector<double> derivative_matrix(const std::vector<double>& x)
{
    std::vector<double> D(x.size() * x.size());
    check(sem_derivative_matrix((int)x.size(), x.data(), D.data()));
    return D;
}

// Quad<real> (Quad.hpp:10-62): corners, metrics (column-major n x n), bilinear map
template <typename real = double>
struct Quad {
    real xs[4];
    real ys[4];
    int corner_ids[4];
    std::vector<real> J, A, B, C;
    unsigned long n = 0;
    void compute_metrics(const std::vector<real>& z)
    {
        n = z.size();
        J.assign(n * n, 0);
        A.assign(n * n, 0);
        B.assign(n * n, 0);
        C.assign(n * n, 0);
        check(sem_quad_metrics(xs, ys, (int)n, z.data(), J.data(), A.data(), B.data(), C.data()));
    }
    std::pair<double, double> from_local_coo(double xi, double eta) const
    {
        const double x = 0.25 * (xs[0] * (1 - xi) * (1 - eta) + xs[1] * (1 + xi) * (1 - eta) +
                                 xs[2] * (1 + xi) * (1 + eta) + xs[3] * (1 - xi) * (1 + eta));
        const double y = 0.25 * (ys[0] * (1 - xi) * (1 - eta) + ys[1] * (1 + xi) * (1 - eta) +
                                 ys[2] * (1 + xi) * (1 + eta) + ys[3] * (1 - xi) * (1 + eta));
        return {x, y};
    }
};

// schro_omp::Mesh (openmp_impl/Mesh.hpp:16-88): a handle to the GPU-resident mesh.
class Mesh {
public:
    typedef double real_t;
    explicit Mesh(sem_mesh* h) : h_(h)
    {
        long info[6];
        check(sem_mesh_info(h_, info));
        N = (int)info[0];
        n_elements = (int)info[1];
        dof_ = info[4];
    }
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;
    Mesh(Mesh&& o) noexcept : h_(o.h_), N(o.N), n_elements(o.n_elements), dof_(o.dof_) { o.h_ = nullptr; }
    ~Mesh()
    {
        if (h_) sem_mesh_destroy(h_);
    }
    long dof() const { return dof_; }
    size_t n2() const { return (size_t)N * N; }
    sem_mesh* handle() const { return h_; }
    // physical coordinates of element e's node (i, j), Quad::from_local_coo at GLL nodes
    std::pair<double, double> node(int e, int i, int j) const
    {
        if (X_.empty()) {
            X_.resize(n2() * n_elements);
            Y_.resize(n2() * n_elements);
            check(sem_mesh_coords(h_, X_.data(), Y_.data()));
        }
        const size_t k = (size_t)e * n2() + i + (size_t)j * N;
        return {X_[k], Y_[k]};
    }

private:
    sem_mesh* h_;

public:
    int N;
    int n_elements;

private:
    long dof_;
    mutable std::vector<double> X_, Y_;
};

// load_mesh<Mesh>(p, folder) + Mesh::compute_metrics() (load_mesh.hpp:76-209)
inline Mesh load_mesh(int p, const std::string& folder)
{
    sem_mesh* h = nullptr;
    check(sem_mesh_load(p, folder.c_str(), &h));
    return Mesh(h);
}

// galerkin_op with one of the library's element operators (galerkin.hpp:24-35)
template <class M>
std::vector<M> galerkin_op(sem_op_kind kind, const std::vector<M>& u, const Mesh& mesh,
                           const std::vector<M>* coef = nullptr, double alpha = 0.0)
{
    std::vector<double> in = detail::pack(u, mesh.n2()), out(in.size());
    std::vector<double> c;
    if (coef) c = detail::pack(*coef, mesh.n2());
    check(sem_galerkin_apply(mesh.handle(), kind, coef ? c.data() : nullptr, alpha, in.data(), out.data()));
    std::vector<M> y = u;
    detail::unpack(out.data(), y, mesh.n2());
    return y;
}

template <class M>
std::vector<M> gproj(const std::vector<M>& u, const Mesh& mesh)  // galerkin.hpp:42-57
{
    return galerkin_op(SEM_OP_MASS, u, mesh);
}

// glaplace(u, el, D, rule) on one element (glaplace.hpp:16-28): the GPU apply on
// a one-element mesh with el's corners and rule's degree (D is implied by the rule)
template <class M>
M glaplace(const M& u, const Quad<double>& el, const M& /*D*/, const quad_rule<double>& rule)
{
    const int p = (int)rule.x.size() - 1;
    const double xy[8] = {el.xs[0], el.ys[0], el.xs[1], el.ys[1], el.xs[2], el.ys[2], el.xs[3], el.ys[3]};
    const int q[4] = {0, 1, 2, 3};
    sem_mesh* h = nullptr;
    check(sem_mesh_create(p, xy, 4, q, 1, &h));
    Mesh m(h);
    M y = u;
    check(sem_glaplace_apply(h, detail::data(u), detail::data(y)));
    return y;
}

// the DSS passes (openmp_impl/unmask.hpp:43-55, global_sum.hpp:65-77, mask.hpp:16-28), in place
template <class M>
void unmask(std::vector<M>& u, const Mesh& mesh)
{
    std::vector<double> a = detail::pack(u, mesh.n2());
    check(sem_unmask(mesh.handle(), a.data()));
    detail::unpack(a.data(), u, mesh.n2());
}
template <class M>
void global_sum(std::vector<M>& u, const Mesh& mesh)
{
    std::vector<double> a = detail::pack(u, mesh.n2());
    check(sem_global_sum(mesh.handle(), a.data()));
    detail::unpack(a.data(), u, mesh.n2());
}
template <class M>
void mask(std::vector<M>& u, const Mesh& mesh)
{
    std::vector<double> a = detail::pack(u, mesh.n2());
    check(sem_mask(mesh.handle(), a.data()));
    detail::unpack(a.data(), u, mesh.n2());
}

// galerkin_op(op, u, mesh) with the caller's host Op, op(u, mesh) in place per
// element (galerkin.hpp:24-35): unmask -> op -> global_sum -> mask
template <class Op, class M>
std::enable_if_t<!std::is_same_v<Op, sem_op_kind>, std::vector<M>> galerkin_op(Op op, std::vector<M> u,
                                                                                const Mesh& mesh)
{
    unmask(u, mesh);
    op(u, mesh);
    global_sum(u, mesh);
    mask(u, mesh);
    return u;
}

template <class M>
double dot(const Mesh& mesh, const std::vector<M>& a, const std::vector<M>& b)  // dot.hpp:11-61
{
    std::vector<double> x = detail::pack(a, mesh.n2()), y = detail::pack(b, mesh.n2());
    double out = 0;
    check(sem_dot(mesh.handle(), x.data(), y.data(), &out));
    return out;
}

template <class M>
solver_results<double> poisson(std::vector<M>& u, const std::vector<M>& f, const Mesh& mesh, int max_iter,
                               double tol)  // poisson.hpp:17-59
{
    std::vector<double> x = detail::pack(u, mesh.n2()), b = detail::pack(f, mesh.n2());
    sem_solver_results r{};
    check(sem_poisson(mesh.handle(), x.data(), b.data(), max_iter, tol, &r));
    detail::unpack(x.data(), u, mesh.n2());
    return {r.success != 0, r.n_iter, r.residual};
}

template <class M>
solver_results<double> helmholtz(std::vector<M>& u, const std::vector<M>& c, const std::vector<M>& f,
                                 const Mesh& mesh, int max_iter, double tol)  // helmholtz.hpp:18-64
{
    std::vector<double> x = detail::pack(u, mesh.n2()), cc = detail::pack(c, mesh.n2()), b = detail::pack(f, mesh.n2());
    sem_solver_results r{};
    check(sem_helmholtz(mesh.handle(), x.data(), cc.data(), b.data(), max_iter, tol, &r));
    detail::unpack(x.data(), u, mesh.n2());
    return {r.success != 0, r.n_iter, r.residual};
}

namespace detail {
template <class M, class Callback>
struct CbCtx {
    Callback* cb;
    size_t n2, nel;
    std::vector<M>* re;
    std::vector<M>* im;
};

template <class M, class Callback>
int trampoline(const double* re, const double* im, void* user)
{
    auto* c = static_cast<CbCtx<M, Callback>*>(user);
    unpack(re, *c->re, c->n2);
    unpack(im, *c->im, c->n2);
    return (*c->cb)(*c->re, *c->im) ? 1 : 0;
}

template <class M, class Callback>
bool stepper(int scheme, std::vector<M>& psi_real, std::vector<M>& psi_imag, const std::vector<M>& potential,
             const Mesh& mesh, double T, const solver_opts<double>& opts, Callback& callback)
{
    std::vector<double> re = pack(psi_real, mesh.n2()), im = pack(psi_imag, mesh.n2()),
                        V = pack(potential, mesh.n2());
    sem_solver_opts o{opts.max_iter, opts.tol, opts.dt, opts.verbose ? 1 : 0, SEM_PREC_FP64, 0, 0};
    int ok = 0;
    CbCtx<M, Callback> ctx{&callback, mesh.n2(), psi_real.size(), &psi_real, &psi_imag};
    sem_step_cb cb = std::is_same_v<Callback, NullCallback> ? nullptr : &trampoline<M, Callback>;
    int rc;
    if (scheme == 0) rc = sem_schrodinger_bdf2(mesh.handle(), re.data(), im.data(), V.data(), T, &o, cb, &ctx, &ok);
    else if (scheme == 1) rc = sem_schrodinger_dirk4s3(mesh.handle(), re.data(), im.data(), V.data(), T, &o, cb, &ctx, &ok);
    else rc = sem_schrodinger_dirk4s5(mesh.handle(), re.data(), im.data(), V.data(), T, &o, cb, &ctx, &ok);
    check(rc);
    unpack(re.data(), psi_real, mesh.n2());
    unpack(im.data(), psi_imag, mesh.n2());
    return ok != 0;
}
}  // namespace detail

// schrodinger.hpp:22-128
template <class M, class Callback = NullCallback>
bool schrodinger_bdf2(std::vector<M>& psi_real, std::vector<M>& psi_imag, const std::vector<M>& potential,
                      const Mesh& mesh, double T, const solver_opts<double>& opts, Callback callback = NullCallback{})
{
    return detail::stepper(0, psi_real, psi_imag, potential, mesh, T, opts, callback);
}

// schrodinger.hpp:294-423
template <class M, class Callback = NullCallback>
bool schrodinger_dirk4s3(std::vector<M>& psi_real, std::vector<M>& psi_imag, const std::vector<M>& potential,
                         const Mesh& mesh, double T, const solver_opts<double>& opts, Callback callback = NullCallback{})
{
    return detail::stepper(1, psi_real, psi_imag, potential, mesh, T, opts, callback);
}

// schrodinger.hpp:138-285
template <class M, class Callback = NullCallback>
bool schrodinger_dirk4s5(std::vector<M>& psi_real, std::vector<M>& psi_imag, const std::vector<M>& potential,
                         const Mesh& mesh, double T, const solver_opts<double>& opts, Callback callback = NullCallback{})
{
    return detail::stepper(2, psi_real, psi_imag, potential, mesh, T, opts, callback);
}

// ------------------------------------------------------ generic solvers --
// IdentityPreconditioner (solver_base.hpp:14-22)
class IdentityPreconditioner {
public:
    template <typename vec>
    vec operator()(const vec& x)
    {
        return x;
    }
};

namespace detail {
template <class vec>
struct SolveCtx {
    vec work;
    size_t n;
    void* A;
    void* D;
    void* P;
};

template <class vec>
void set_from(vec& v, const double* src, size_t n)
{
    std::copy(src, src + n, detail::data(v));
}

template <class vec, class LinOp>
int linop_tr(void* user, const double* x, double* y, void*)
{
    auto* c = static_cast<SolveCtx<vec>*>(user);
    set_from(c->work, x, c->n);
    vec r = (*static_cast<LinOp*>(c->A))(c->work);
    std::copy(detail::data(r), detail::data(r) + c->n, y);
    return 0;
}

template <class vec, class Dot>
int dot_tr(void* user, const double* a, const double* b, double* out, void*)
{
    auto* c = static_cast<SolveCtx<vec>*>(user);
    vec va = c->work, vb = c->work;
    set_from(va, a, c->n);
    set_from(vb, b, c->n);
    *out = (double)(*static_cast<Dot*>(c->D))(va, vb);
    return 0;
}

template <class vec, class Precond>
int precond_tr(void* user, const double* r, double* z, void*)
{
    auto* c = static_cast<SolveCtx<vec>*>(user);
    set_from(c->work, r, c->n);
    vec out = (*static_cast<Precond*>(c->P))(c->work);
    std::copy(detail::data(out), detail::data(out) + c->n, z);
    return 0;
}

template <class vec, class LinOp, class Dot, class Precond>
solver_results<double> solve(bool use_minres, vec& x, LinOp& A, const vec& b, Dot& dot, Precond& precond, int maxit,
                             double tol)
{
    SolveCtx<vec> ctx{x, detail::size(x), &A, &dot, &precond};
    sem_krylov k{};
    k.n = (long)ctx.n;
    k.host = 1;
    k.A = &linop_tr<vec, LinOp>;
    k.A_user = &ctx;
    k.dot = &dot_tr<vec, Dot>;
    k.dot_user = &ctx;
    if constexpr (!std::is_same_v<Precond, IdentityPreconditioner>) {
        k.precond = &precond_tr<vec, Precond>;
        k.precond_user = &ctx;
    }
    sem_solver_results r{};
    check(use_minres ? sem_minres(&k, detail::data(b), detail::data(x), maxit, tol, &r) : sem_pcg(&k, detail::data(b), detail::data(x), maxit, tol, &r));
    return {r.success != 0, r.n_iter, r.residual};
}
}  // namespace detail

// minres(x, A, b, dot, precond, maxit, tol) (minres.hpp:19-107) and pcg
// (pcg.hpp:59-97): vec is a contiguous vector of doubles (std::vector<double>,
// or a type with memptr() and n_elem such as arma::vec); the iterations run on
// the GPU, A, dot and precond are the caller's host callables.
template <typename real, typename vec, typename LinOp, typename Dot, typename Precond>
solver_results<real> minres(vec& x, LinOp A, const vec& b, Dot dot, Precond precond, int maxit, real tol)
{
    return detail::solve(true, x, A, b, dot, precond, maxit, (double)tol);
}

template <typename real, typename vec, typename LinOp, typename Dot, typename Precond>
solver_results<real> pcg(vec& x, LinOp A, const vec& b, Dot dot, Precond precond, int max_iter, real tol)
{
    return detail::solve(false, x, A, b, dot, precond, max_iter, (double)tol);
}

// ------------------------------------------------------ vector wrappers --
// solution_wrapper / v_wrapper (openmp_impl/solution_wrapper.hpp:20-201,
// v_wrapper.hpp:8-80): per-element vectors with +, -, scalar * and += / -=,
// evaluated at once (the GPU solvers keep their own vectors on the device)
template <typename real, class M>
class solution_wrapper {
public:
    std::vector<M> values;
    solution_wrapper() = default;
    explicit solution_wrapper(std::vector<M> v) : values(std::move(v)) {}
    M& operator[](size_t e) { return values[e]; }
    const M& operator[](size_t e) const { return values[e]; }
    size_t size() const { return values.size(); }
    template <class F>
    solution_wrapper map2(const solution_wrapper& o, F f) const
    {
        solution_wrapper r = *this;
        for (size_t e = 0; e < values.size(); ++e) {
            double* d = detail::data(r.values[e]);
            const double* a = detail::data(o.values[e]);
            for (size_t k = 0; k < detail::size(r.values[e]); ++k) d[k] = f(d[k], a[k]);
        }
        return r;
    }
    solution_wrapper operator+(const solution_wrapper& o) const { return map2(o, [](real a, real b) { return a + b; }); }
    solution_wrapper operator-(const solution_wrapper& o) const { return map2(o, [](real a, real b) { return a - b; }); }
    solution_wrapper& operator+=(const solution_wrapper& o) { return *this = *this + o; }
    solution_wrapper& operator-=(const solution_wrapper& o) { return *this = *this - o; }
    friend solution_wrapper operator*(real c, const solution_wrapper& x)
    {
        return x.map2(x, [c](real a, real) { return c * a; });
    }
};

template <typename real, class M>
class v_wrapper {
public:
    std::vector<std::vector<M>> values;
    std::vector<M>& operator[](size_t c) { return values[c]; }
    const std::vector<M>& operator[](size_t c) const { return values[c]; }
    size_t size() const { return values.size(); }
    template <class F>
    v_wrapper map2(const v_wrapper& o, F f) const
    {
        v_wrapper r = *this;
        for (size_t c = 0; c < values.size(); ++c)
            r.values[c] = solution_wrapper<real, M>(values[c]).map2(solution_wrapper<real, M>(o.values[c]), f).values;
        return r;
    }
    v_wrapper operator+(const v_wrapper& o) const { return map2(o, [](real a, real b) { return a + b; }); }
    v_wrapper operator-(const v_wrapper& o) const { return map2(o, [](real a, real b) { return a - b; }); }
    v_wrapper& operator+=(const v_wrapper& o) { return *this = *this + o; }
    v_wrapper& operator-=(const v_wrapper& o) { return *this = *this - o; }
    friend v_wrapper operator*(real c, const v_wrapper& x)
    {
        return x.map2(x, [c](real a, real) { return c * a; });
    }
};

}  // namespace schro_hip
