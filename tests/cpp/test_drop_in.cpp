// The reference's own test sequence (tests/tests.cpp:27-40, the 14 tests of
// tests/*.cpp) written against include/sem.hpp, with `schro_omp` swapped for
// `schro_hip`. Mat mimics the part of arma::Mat the tests use (memptr, n_elem,
// (i, j) and linear [] indexing); Vec plays arma::vec.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sem.hpp"

using namespace schro_hip;

struct Mat {
    std::vector<double> v;
    size_t n_elem;
    int n;
    Mat(int n_ = 0, double fill = 0.0) : v((size_t)n_ * n_, fill), n_elem((size_t)n_ * n_), n(n_) {}
    double* memptr() { return v.data(); }
    const double* memptr() const { return v.data(); }
    double& operator()(int i, int j) { return v[i + (size_t)j * n]; }
    double& operator[](size_t k) { return v[k]; }
    double operator[](size_t k) const { return v[k]; }
};
typedef std::vector<double> Vec;

static std::vector<Mat> sample(const Mesh& mesh, double (*f)(double, double))
{
    std::vector<Mat> F(mesh.n_elements, Mat(mesh.N));
    for (int e = 0; e < mesh.n_elements; ++e)
        for (int i = 0; i < mesh.N; ++i)
            for (int j = 0; j < mesh.N; ++j) {
                auto [x, y] = mesh.node(e, i, j);
                F[e](i, j) = f(x, y);
            }
    return F;
}

static const char* g_folder = "tests/golden/small_mesh";

// tests/gauss_lobatto.cpp: integral of exp(sin(pi x)) = 2 I0(1) (Bessel value
// 2 I0(1) = 2.5321317555040164, in place of boost::math::cyl_bessel_i)
static bool test_gauss_lobatto()
{
    const double exact = 2.5321317555040164;
    auto f = [](double x) { return std::exp(std::sin(M_PI * x)); };
    quad_rule<double> a = gauss_lobatto<double>(9);
    double I = 0;
    for (size_t i = 0; i < a.x.size(); ++i) I += f(a.x[i]) * a.w[i];
    const double e1 = std::abs(I - exact);
    quad_rule<double> b = gauss_lobatto<double>(10);
    I = 0;
    for (size_t i = 0; i < b.x.size(); ++i) I += f(b.x[i]) * b.w[i];
    const double e2 = std::abs(I - exact);
    bool passed = true;
    if (e1 > 1e-4) {
        std::printf("gauss_lobatto() failed accuracy test.\n");
        passed = false;
    }
    if (e2 > e1) {
        std::printf("gauss_lobatto() failed convergence test.\n");
        passed = false;
    }
    return passed;
}

// tests/derivative_matrix.cpp: D p = p' for a cubic at n = 8
static bool test_derivative_matrix()
{
    auto [x, w] = gauss_lobatto<double>(8);
    std::vector<double> D = derivative_matrix(x);
    const size_t n = x.size();
    double e = 0;
    for (size_t i = 0; i < n; ++i) {
        double s = 0;
        for (size_t k = 0; k < n; ++k) s += D[i + k * n] * (-1 + 2 * x[k] + 5 * x[k] * x[k] - 3 * x[k] * x[k] * x[k]);
        e = std::max(e, std::abs(s - (2 + 10 * x[i] - 9 * x[i] * x[i])));
    }
    if (e > 1e-10) {
        std::printf("derivative_matrix() failed accuracy test\n");
        return false;
    }
    return true;
}

// tests/quad.cpp: metrics of the reference square on three points
static bool test_Quad()
{
    std::vector<double> x = {-1, 0, 1};
    Quad<double> element;
    element.xs[0] = -1; element.xs[1] = 1; element.xs[2] = 1; element.xs[3] = -1;
    element.ys[0] = -1; element.ys[1] = -1; element.ys[2] = 1; element.ys[3] = 1;
    element.compute_metrics(x);
    return true;
}

// tests/glaplace.cpp: the stiffness action kills constants
static bool test_glaplace()
{
    quad_rule<double> rule = gauss_lobatto<double>(5);
    Quad<double> element;
    element.xs[0] = -1; element.xs[1] = 1; element.xs[2] = 1; element.xs[3] = -1;
    element.ys[0] = -1; element.ys[1] = -1; element.ys[2] = 1; element.ys[3] = 1;
    element.compute_metrics(rule.x);
    const int n = (int)rule.x.size();
    Mat D(n);
    Mat u(n, 1.0);
    Mat lap_u = glaplace(u, element, D, rule);
    bool success = true;
    for (double v : lap_u.v) success = success && (std::abs(v) < 1e-10);
    return success;
}

// tests/load_mesh.cpp
static bool test_load_mesh()
{
    Mesh mesh = load_mesh(3, g_folder);
    return mesh.n_elements == 45;
}

// tests/solution_wrapper.cpp
static bool test_solution_wrapper()
{
    const int n = 3;
    solution_wrapper<double, Mat> x(std::vector<Mat>(n, Mat(2, 1.0)));
    solution_wrapper<double, Mat> y = 0.5 * x;
    x += 2.0 * (3.0 * x - y);
    bool success = true;
    if (x[0][0] != 6.0) {
        std::printf("solution_wrapper failed template expression computations\n");
        success = false;
    }
    if (x.size() != (size_t)n) {
        std::printf("solution_wrapper failed to preserve structure of original vector\n");
        success = false;
    }
    return success;
}

// tests/pcg.cpp: finite-difference Poisson -u'' = x^2, u(-1) = u(1) = 0
static Vec conv_same(const Vec& u, double lo, double di, double up)
{
    Vec y(u.size());
    for (size_t i = 0; i < u.size(); ++i) {
        double s = di * u[i];
        if (i > 0) s += lo * u[i - 1];
        if (i + 1 < u.size()) s += up * u[i + 1];
        y[i] = s;
    }
    return y;
}
static double vdot(const Vec& a, const Vec& b)
{
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
static double vnorm(const Vec& a) { return std::sqrt(vdot(a, a)); }

static bool test_pcg()
{
    const int n = 20;
    Vec x(n), f(n);
    for (int i = 0; i < n; ++i) {
        x[i] = -1 + 2.0 * i / (n - 1);
        f[i] = x[i] * x[i];
    }
    const double h = x[1] - x[0];
    auto L = [h](const Vec& u) { return conv_same(u, -1 / (h * h), 2 / (h * h), -1 / (h * h)); };
    Vec u(n, 0.0);
    const double tol = 1e-5;
    auto rslts = pcg(u, L, f, vdot, IdentityPreconditioner{}, n, tol);
    (void)rslts;
    Vec r = L(u);
    for (int i = 0; i < n; ++i) r[i] -= f[i];
    if (vnorm(r) > tol * vnorm(f)) {
        std::printf("pcg returned bad residual\n");
        return false;
    }
    return true;
}

// tests/minres.cpp: indefinite finite-difference Helmholtz -u'' - 2u = x^2
static bool test_minres()
{
    const int n = 20;
    const double k = 2;
    Vec x(n), f(n);
    for (int i = 0; i < n; ++i) {
        x[i] = -1 + 2.0 * i / (n - 1);
        f[i] = x[i] * x[i];
    }
    const double h = x[1] - x[0];
    auto L = [h, k](const Vec& u) { return conv_same(u, -1 / (h * h), 2 / (h * h) - k, -1 / (h * h)); };
    Vec u(n, 0.0);
    const double tol = 1e-5;
    auto rslts = minres(u, L, f, vdot, IdentityPreconditioner{}, n, tol);
    (void)rslts;
    Vec r = L(u);
    for (int i = 0; i < n; ++i) r[i] -= f[i];
    if (vnorm(r) > tol * vnorm(f)) {
        std::printf("pcg returned bad residual\n");
        return false;
    }
    return true;
}

static double xx_minus_yy(double x, double y) { return x * x - y * y; }
static double zero(double, double) { return 0.0; }
static double u0(double x, double y) { return std::sin(M_PI * x) * std::sin(M_PI * y); }
static double c_indef(double, double) { return -100.0 - std::sqrt(2.0); }

// tests/poisson.cpp
static bool test_poisson()
{
    Mesh mesh = load_mesh(6, g_folder);
    auto F = sample(mesh, xx_minus_yy);
    std::vector<Mat> u(mesh.n_elements, Mat(mesh.N));
    auto r = poisson(u, F, mesh, (int)mesh.dof(), 1e-10);
    std::printf("poisson returned after %d iteration with residual %g\n", r.n_iter, r.residual);
    return r.success;
}

// tests/helmholtz.cpp
static bool test_helmholtz()
{
    Mesh mesh = load_mesh(6, g_folder);
    auto F = sample(mesh, xx_minus_yy);
    auto c = sample(mesh, c_indef);
    std::vector<Mat> u(mesh.n_elements, Mat(mesh.N));
    auto r = helmholtz(u, c, F, mesh, (int)mesh.dof() * 10, 1e-10);
    std::printf("helmholtz returned after %d iteration with residual %g\n", r.n_iter, r.residual);
    return r.success;
}

// tests/v_wrapper.cpp
static bool test_v_wrapper()
{
    const int n = 3;
    std::vector<Mat> u(n, Mat(2, 1.0));
    std::vector<Mat> v(n, Mat(2, 2.0));
    v_wrapper<double, Mat> U;
    U.values.push_back(std::move(u));
    U.values.push_back(std::move(v));
    v_wrapper<double, Mat> W = 0.5 * U;
    U += 2.0 * (3.0 * U - W);
    bool success = true;
    if ((U[0][0][0] != 6.0) || (U[1][0][0] != 12.0)) {
        std::printf("v_wrapper failed template expression computations\n");
        success = false;
    }
    if ((U.size() != 2) || (U[0].size() != (size_t)n) || (U[1].size() != (size_t)n)) {
        std::printf("v_wrapper failed to preserve structure of original vectors\n");
        success = false;
    }
    return success;
}

// tests/schrodinger.cpp (the three steppers; the callback counts the steps)
static bool test_schrodinger(int scheme)
{
    Mesh mesh = load_mesh(3, g_folder);
    auto V = sample(mesh, zero);
    auto re = sample(mesh, u0);
    auto im = sample(mesh, zero);
    double T = 0.01;
    solver_opts<double> opts;
    opts.dt = T / 10;
    opts.max_iter = 20 * (int)mesh.dof();
    opts.tol = 1e-8;
    opts.verbose = true;
    int steps = 0;
    auto cb = [&](const std::vector<Mat>&, const std::vector<Mat>&) {
        ++steps;
        return false;
    };
    const char* name[3] = {"BDF2", "dirk4s3", "dirk4s5"};
    std::printf("solving Schrodinger equation with %s...\n", name[scheme]);
    bool ok;
    if (scheme == 0) ok = schrodinger_bdf2(re, im, V, mesh, T, opts, cb);
    else if (scheme == 1) ok = schrodinger_dirk4s3(re, im, V, mesh, T, opts, cb);
    else ok = schrodinger_dirk4s5(re, im, V, mesh, T, opts, cb);
    if (!ok) std::printf("schrodinger %s failed test\n", name[scheme]);
    return ok && steps == 10;
}

static bool test_schrodinger_bdf2() { return test_schrodinger(0); }
static bool test_schrodinger_dirk4s3() { return test_schrodinger(1); }
static bool test_schrodinger_dirk4s5() { return test_schrodinger(2); }

int main(int argc, char** argv)
{
    if (argc > 1) g_folder = argv[1];
    bool success = true;
    int count = 0;
    auto run = [&](const char* name, bool (*t)()) {
        const bool ok = t();
        ++count;
        std::printf("%s: %s\n", name, ok ? "passed" : "FAILED");
        success = ok && success;
    };
    run("test_gauss_lobatto", test_gauss_lobatto);
    run("test_derivative_matrix", test_derivative_matrix);
    run("test_Quad", test_Quad);
    run("test_glaplace", test_glaplace);
    run("test_load_mesh", test_load_mesh);
    run("test_solution_wrapper", test_solution_wrapper);
    run("test_pcg", test_pcg);
    run("test_minres", test_minres);
    run("test_poisson", test_poisson);
    run("test_helmholtz", test_helmholtz);
    run("test_v_wrapper", test_v_wrapper);
    run("test_schrodinger_bdf2", test_schrodinger_bdf2);
    run("test_schrodinger_dirk4s3", test_schrodinger_dirk4s3);
    run("test_schrodinger_dirk4s5", test_schrodinger_dirk4s5);
    std::printf("%d of 14 reference tests run\n", count);
    if (success) std::printf("all tests passed! :)\n");
    return success ? 0 : 1;  // unlike tests/tests.cpp:45, failures set the exit status
}
