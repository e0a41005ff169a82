// C++ host code written the way the reference's tests are (tests/poisson.cpp,
// tests/helmholtz.cpp, tests/schrodinger.cpp), with `schro_omp` swapped for
// `schro_hip`. Mat mimics the part of arma::Mat the wrapper uses.
#include <cmath>
#include <cstdio>
#include <vector>

#include "sem.hpp"

using namespace schro_hip;

struct Mat {
    std::vector<double> v;
    size_t n_elem;
    int n;
    Mat(int n_ = 0) : v((size_t)n_ * n_, 0.0), n_elem((size_t)n_ * n_), n(n_) {}
    double* memptr() { return v.data(); }
    const double* memptr() const { return v.data(); }
    double& operator()(int i, int j) { return v[i + (size_t)j * n]; }
};

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

static double xx_minus_yy(double x, double y) { return x * x - y * y; }
static double zero(double, double) { return 0.0; }
static double u0(double x, double y) { return std::sin(M_PI * x) * std::sin(M_PI * y); }
static double c_indef(double, double) { return -100.0 - std::sqrt(2.0); }

static bool test_poisson(const char* folder)
{
    Mesh mesh = load_mesh(6, folder);
    auto F = sample(mesh, xx_minus_yy);
    std::vector<Mat> u(mesh.n_elements, Mat(mesh.N));
    auto r = poisson(u, F, mesh, (int)mesh.dof(), 1e-10);
    std::printf("poisson returned after %d iteration with residual %g\n", r.n_iter, r.residual);
    return r.success;
}

static bool test_helmholtz(const char* folder)
{
    Mesh mesh = load_mesh(6, folder);
    auto F = sample(mesh, xx_minus_yy);
    auto c = sample(mesh, c_indef);
    std::vector<Mat> u(mesh.n_elements, Mat(mesh.N));
    auto r = helmholtz(u, c, F, mesh, (int)mesh.dof() * 10, 1e-10);
    std::printf("helmholtz returned after %d iteration with residual %g\n", r.n_iter, r.residual);
    return r.success;
}

static bool test_schrodinger(const char* folder, int scheme)
{
    Mesh mesh = load_mesh(3, folder);
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
    auto cb = [&](const std::vector<Mat>&, const std::vector<Mat>&) { ++steps; return false; };
    bool ok;
    if (scheme == 0) ok = schrodinger_bdf2(re, im, V, mesh, T, opts, cb);
    else if (scheme == 1) ok = schrodinger_dirk4s3(re, im, V, mesh, T, opts, cb);
    else ok = schrodinger_dirk4s5(re, im, V, mesh, T, opts);
    if (scheme != 2 && steps != 10) return false;
    return ok;
}

int main(int argc, char** argv)
{
    const char* folder = argc > 1 ? argv[1] : "tests/golden/small_mesh";
    bool success = true;
    success = test_poisson(folder) && success;
    success = test_helmholtz(folder) && success;
    success = test_schrodinger(folder, 0) && success;
    success = test_schrodinger(folder, 1) && success;
    success = test_schrodinger(folder, 2) && success;
    if (success) std::printf("all tests passed! :)\n");
    return success ? 0 : 1;  // unlike tests/tests.cpp:45, failures set the exit status
}
