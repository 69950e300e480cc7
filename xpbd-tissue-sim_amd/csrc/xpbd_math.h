// Per-element XPBD maths shared by the HIP kernels.
//
// The device keeps vertex *displacements* u = x - X from the rest positions X
// rather than absolute positions.  The deformation gradient is then
//     F = X_def Q = (Dm + DU) Q = I + DU Q,     DU = [u0-u3, u1-u3, u2-u3]
// (since Q = Dm^-1), which is exact in real arithmetic and keeps FP32 rounding
// relative to the displacement, not to the absolute coordinate.  At rest F is
// exactly I in any precision.
//
// Within a time step the displacement is further split into the value at the
// start of the step (base, fixed during the solve) and the motion of the
// current step (delta, small, updated by every projection).  Corrections that
// are far below one FP32 ulp of the displacement still accumulate exactly in
// delta; base absorbs delta once per step.
//
// Reference maths (2nd-order combined Stable Neo-Hookean projection):
//   F                 include/solver/constraint/ElementConstraint.hpp:78-119
//   C_D, grad C_D     include/solver/constraint/DeviatoricConstraint.hpp:112-152
//   C_H, grad C_H     include/solver/constraint/HydrostaticConstraint.hpp:123-221
//   2x2 block solve   src/solver/xpbd_projector/CombinedNeohookeanConstraintProjector.cpp:98-179
// Layout: Q[3*c + r] = Q(r, c) (column-major), u/dx vertex-major xyz.
#pragma once

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define XPBD_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define XPBD_HD inline
#endif

namespace xpbd {

// FP32 device code uses the hardware sqrt, reciprocal and log2 (at most 1 ulp
// from the correctly rounded result) instead of the multi-instruction IEEE
// sequences; FP64 and host code stay correctly rounded.
template <typename R> XPBD_HD R xsqrt(R v);
template <typename R> XPBD_HD R xrcp(R v);
template <typename R> XPBD_HD R xlog(R v);
template <> XPBD_HD double xsqrt<double>(double v) { return sqrt(v); }
template <> XPBD_HD double xrcp<double>(double v) { return 1.0 / v; }
template <> XPBD_HD double xlog<double>(double v) { return log(v); }
#if defined(__HIP_DEVICE_COMPILE__)
template <> XPBD_HD float xsqrt<float>(float v) { return __builtin_amdgcn_sqrtf(v); }
template <> XPBD_HD float xrcp<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> XPBD_HD float xlog<float>(float v) { return __builtin_amdgcn_logf(v) * 0.693147180559945309f; }
#else
template <> XPBD_HD float xsqrt<float>(float v) { return sqrtf(v); }
template <> XPBD_HD float xrcp<float>(float v) { return 1.0f / v; }
template <> XPBD_HD float xlog<float>(float v) { return logf(v); }
#endif

// F = I + DU Q from the four vertex displacements.
template <typename R>
XPBD_HD void deformation_gradient(const R* u0, const R* u1, const R* u2, const R* u3,
                                  const R* Q, R* F)
{
    R D[9];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        D[k] = u0[k] - u3[k];
        D[3 + k] = u1[k] - u3[k];
        D[6 + k] = u2[k] - u3[k];
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++)
            F[3 * c + r] = (r == c ? R(1) : R(0)) + D[r] * Q[3 * c] + D[3 + r] * Q[3 * c + 1]
                         + D[6 + r] * Q[3 * c + 2];
}

template <typename R>
XPBD_HD R det3cm(const R* F)
{
    return F[0] * F[4] * F[8] - F[0] * F[7] * F[5] - F[3] * F[1] * F[8]
         + F[3] * F[7] * F[2] + F[6] * F[1] * F[5] - F[6] * F[4] * F[2];
}

// C_D = |F|_F and C_H (log J - gamma for J >= 1, cubic Taylor series below).
template <typename R>
XPBD_HD void element_constraints(const R* F, R gamma, R* CD, R* CH, R* J_out)
{
    R s = R(0);
#pragma unroll
    for (int i = 0; i < 9; i++) s += F[i] * F[i];
    *CD = xsqrt<R>(s);
    const R J = det3cm<R>(F);
    if (J >= R(1)) {
        *CH = -gamma + xlog<R>(J);
    } else {
        const R a = J - R(1);
        *CH = -gamma + a - a * a * R(0.5) + a * a * a * (R(1) / R(3));
    }
    if (J_out) *J_out = J;
}

// F = I + DU Q where each vertex position is base + delta and the differences
// are formed per part: DU = (b_i - b_3) + (d_i - d_3).
template <typename R>
XPBD_HD void deformation_gradient_split(const R* b, const R* d, const R* Q, R* F)
{
    R D[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) D[3 * i + k] = (b[3 * i + k] - b[9 + k]) + (d[3 * i + k] - d[9 + k]);
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++)
            F[3 * c + r] = (r == c ? R(1) : R(0)) + D[r] * Q[3 * c] + D[3 + r] * Q[3 * c + 1]
                         + D[6 + r] * Q[3 * c + 2];
}

// Gradients of C_D and C_H at the four vertices (vertex-major xyz) from F and
// the C_D and J that element_constraints returned for it:
//   grad_i C_D = (F Q^T)_i / C_D,   grad_i C_H = dC_H/dJ (cof(F) Q^T)_i   (i < 3),
// and vertex 3 the negative sum of the other three.
template <typename R>
XPBD_HD void constraint_gradients(const R* F, const R* Q, R CD, R J, R* gD, R* gH)
{
    const R inv_C = xrcp<R>(CD);
    const R a = J - R(1);
    const R fac = (J >= R(1)) ? xrcp<R>(J) : R(1) - a + a * a;

    // P = F Q^T and Rm = Fcross Q^T ; column i is the (unscaled) gradient at vertex i
    R Fc[9];
    Fc[0] = F[4] * F[8] - F[5] * F[7];
    Fc[1] = F[5] * F[6] - F[3] * F[8];
    Fc[2] = F[3] * F[7] - F[4] * F[6];
    Fc[3] = F[7] * F[2] - F[8] * F[1];
    Fc[4] = F[8] * F[0] - F[6] * F[2];
    Fc[5] = F[6] * F[1] - F[7] * F[0];
    Fc[6] = F[1] * F[5] - F[2] * F[4];
    Fc[7] = F[2] * F[3] - F[0] * F[5];
    Fc[8] = F[0] * F[4] - F[1] * F[3];

#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            gD[3 * i + r] = inv_C * (F[r] * Q[i] + F[3 + r] * Q[3 + i] + F[6 + r] * Q[6 + i]);
            gH[3 * i + r] = fac * (Fc[r] * Q[i] + Fc[3 + r] * Q[3 + i] + Fc[6 + r] * Q[6 + i]);
        }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        gD[9 + r] = -gD[r] - gD[3 + r] - gD[6 + r];
        gH[9 + r] = -gH[r] - gH[3 + r] - gH[6 + r];
    }
}

// One combined deviatoric + hydrostatic projection.
//   u[12]   vertex displacements (vertex-major)
//   w[4]    inverse masses
//   Q[9]    rest-shape inverse, column-major
//   atD/atH compliance divided by dt^2 (alpha tilde)
//   lam[2]  multipliers, updated in place
//   dx[12]  position corrections for the four vertices
template <typename R>
XPBD_HD void project_tet_F(const R* F, const R* w, const R* Q, R atD, R atH, R gamma,
                           R* lam, R* dx)
{
    R CD, CH, J;
    element_constraints<R>(F, gamma, &CD, &CH, &J);
    R gD[12], gH[12];
    constraint_gradients<R>(F, Q, CD, J, gD, gH);

    R s00 = R(0), s01 = R(0), s11 = R(0);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const R dd = gD[3 * i] * gD[3 * i] + gD[3 * i + 1] * gD[3 * i + 1] + gD[3 * i + 2] * gD[3 * i + 2];
        const R dh = gD[3 * i] * gH[3 * i] + gD[3 * i + 1] * gH[3 * i + 1] + gD[3 * i + 2] * gH[3 * i + 2];
        const R hh = gH[3 * i] * gH[3 * i] + gH[3 * i + 1] * gH[3 * i + 1] + gH[3 * i + 2] * gH[3 * i + 2];
        s00 += w[i] * dd;
        s01 += w[i] * dh;
        s11 += w[i] * hh;
    }
    const R A00 = atD + s00, A11 = atH + s11, A01 = s01;
    const R b0 = -CD - atD * lam[0];
    const R b1 = -CH - atH * lam[1];
    // A is a Gram matrix plus the positive compliances, so det > 0 in exact
    // arithmetic; on sliver elements (compliance ~1e-6 of the Gram terms, whose
    // determinant cancels at rest) FP32 rounding can make it <= 0, and the
    // projection is then skipped instead of dividing by it
    const R det = A00 * A11 - A01 * A01;
    const R inv_det = det > R(0) ? xrcp<R>(det) : R(0);
    const R d0 = (b0 * A11 - b1 * A01) * inv_det;
    const R d1 = (b1 * A00 - b0 * A01) * inv_det;
    lam[0] += d0;
    lam[1] += d1;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int k = 0; k < 3; k++) dx[3 * i + k] = w[i] * (gD[3 * i + k] * d0 + gH[3 * i + k] * d1);
}

template <typename R>
XPBD_HD void project_tet(const R* u, const R* w, const R* Q, R atD, R atH, R gamma,
                         R* lam, R* dx)
{
    R F[9];
    deformation_gradient<R>(u, u + 3, u + 6, u + 9, Q, F);
    project_tet_F<R>(F, w, Q, atD, atH, gamma, lam, dx);
}

// Projection with vertex positions given as base + delta (see deformation_gradient_split).
template <typename R>
XPBD_HD void project_tet_split(const R* b, const R* d, const R* w, const R* Q, R atD, R atH,
                               R gamma, R* lam, R* dx)
{
    R F[9];
    deformation_gradient_split<R>(b, d, Q, F);
    project_tet_F<R>(F, w, Q, atD, atH, gamma, lam, dx);
}

}  // namespace xpbd
