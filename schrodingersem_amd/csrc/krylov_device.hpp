// krylov_device.hpp -- the MINRES recurrence on device scalars, shared by the
// Lanczos-step kernels (every workgroup evaluates it on the same inputs, so
// every workgroup holds bit-identical scalars).
#pragma once

#include <cfloat>

#include "sem_internal.hpp"

namespace sem {

// beta = sqrt(s), s = dot(y, y) of the step just finished, and the Givens
// update of minres.hpp:62-99 (stop test minres.hpp:92 with atol = tol there).
// Contraction is off so that every workgroup, whatever code surrounds the
// call, rounds the recurrence the same way.
__device__ __forceinline__ void minres_beta_update(MinresState& sc, double s)
{
#pragma clang fp contract(off)
    const double alpha = sc.alpha;
    const double beta_prev = sc.beta;
    const double ib_prev = sc.ib;
    sc.beta_old = beta_prev;
    const double beta = sqrt(s);
    sc.beta = beta;
    sc.tnorm += alpha * alpha + beta * beta + beta_prev * beta_prev;
    const double eps_old = sc.eps;
    sc.eps_old = eps_old;
    const double cs = sc.cs, sn = sc.sn, dbar = sc.dbar;
    const double delta = cs * dbar + sn * alpha;
    const double gbar = sn * dbar - cs * alpha;
    sc.eps = sn * beta;
    sc.dbar = -cs * beta;
    const double gamma = fmax(sqrt(gbar * gbar + beta * beta), DBL_EPSILON);
    sc.cs = gbar / gamma;
    sc.sn = beta / gamma;
    const double phi = sc.cs * sc.phibar;
    sc.phibar *= sc.sn;
    // pending w/x update of this step (minres.hpp:82-86)
    sc.p3_ibp = ib_prev;
    sc.p3_ig = 1.0 / gamma;
    sc.p3_eo = eps_old;
    sc.p3_dl = delta;
    sc.p3_phi = phi;
    sc.p3_pending = 1;
    // coefficients of the next Lanczos step (minres.hpp:48-52)
    sc.ib = 1.0 / beta;
    sc.cr1 = beta / beta_prev;
    sc.has_r1 = 1;
    sc.iter += 1;
    if (sc.phibar < sc.tol * sc.b_norm + sc.atol) {
        sc.success = 1;
        sc.done = 1;
    } else if (sc.iter >= sc.maxit) {
        sc.done = 1;
    }
}

}  // namespace sem
