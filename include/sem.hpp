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
        if (size(v[e]) != n2) throw std::invalid_argument("element block has the wrong size");
        const double* s = data(v[e]);
        std::copy(s, s + n2, out.begin() + e * n2);
    }
    return out;
}
template <class M>
void unpack(const double* src, std::vector<M>& v, size_t n2)
{
    for (size_t e = 0; e < v.size(); ++e) std::copy(src + e * n2, src + (e + 1) * n2, data(v[e]));
}
}  // namespace detail

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

}  // namespace schro_hip
