/*
 * xpbd_oracle.c -- CPU FP64 oracle for the XPBD Neo-Hookean solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline legs may
 * load liboracle.so.  The shipped solver (the HIP library) never links it.
 * Its colour-parallel mode (OpenMP) serves only the all-core CPU baseline.
 *
 * It is a clean-room restatement, in plain C and double precision, of the
 * reference CPU solver (MAVYWAVY902/xpbd-tissue-sim), which cannot be built
 * here (Eigen, yaml-cpp, gmsh, Embree, ... are absent).  Each function names
 * the reference file:line it follows.  Loop orders that decide floating-point
 * rounding (Gauss-Seidel order, Jacobi buffer order, accumulation order of the
 * 2x2 system) follow the reference so that the oracle reproduces it in FP64.
 *
 * Pinning: every function here is checked by tests/test_oracle.py against
 * closed forms, invariants, finite differences, an independent float64
 * autograd restatement, or brute force (see the test docstrings).  Functions
 * whose bit-level rounding the reference decides through Eigen (the 3x3
 * inverse in the rest-state precompute) are pinned by Q*Dm = I instead.
 *
 * Layout conventions (match the reference's Eigen storage):
 *   positions  x[3*i + k]            (Eigen 3xN column-major = AoS xyz)
 *   tets       tets[4*e + j]
 *   Q          Q[9*e + 3*c + r] = Q(r, c)   (column-major 3x3, like Mat3r)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

enum { ORC_GS = 0, ORC_JACOBI = 1 };
enum { ORC_SDF_NONE = 0, ORC_SDF_SPHERE = 1, ORC_SDF_BOX = 2 };

/* ------------------------------------------------------------------------ */
/* a1  material: include/simobject/ElasticMaterial.hpp:22-33                 */
/* ------------------------------------------------------------------------ */
ORC_API void orc_lame(double E, double nu, double* mu, double* lam)
{
    *mu = E / (2 * (1 + nu));
    *lam = (E * nu) / ((1 + nu) * (1 - 2 * nu));
}

/* ------------------------------------------------------------------------ */
/* a2  rest state: include/solver/constraint/ElementConstraint.hpp:36-50     */
/*     Dm = [X0-X3, X1-X3, X2-X3] (columns), Q = Dm^-1, V = |det Dm| / 6.    */
/*     The inverse is the closed-form adjugate / determinant.                */
/* ------------------------------------------------------------------------ */
static double det3(const double* M) /* column-major */
{
    return M[0] * (M[4] * M[8] - M[7] * M[5])
         - M[3] * (M[1] * M[8] - M[7] * M[2])
         + M[6] * (M[1] * M[5] - M[4] * M[2]);
}

ORC_API int orc_rest_tet(const double* X0, const double* X1, const double* X2,
                         const double* X3, double* Q, double* vol)
{
    double D[9];
    for (int k = 0; k < 3; k++) {
        D[0 + k] = X0[k] - X3[k];
        D[3 + k] = X1[k] - X3[k];
        D[6 + k] = X2[k] - X3[k];
    }
    const double det = det3(D);
    /* inverse(r, c) = cofactor(c, r) / det ; D(r, c) = D[3c + r] */
#define DM(r, c) D[3 * (c) + (r)]
    double inv[9];
    inv[3 * 0 + 0] = (DM(1, 1) * DM(2, 2) - DM(1, 2) * DM(2, 1)) / det;
    inv[3 * 1 + 0] = (DM(0, 2) * DM(2, 1) - DM(0, 1) * DM(2, 2)) / det;
    inv[3 * 2 + 0] = (DM(0, 1) * DM(1, 2) - DM(0, 2) * DM(1, 1)) / det;
    inv[3 * 0 + 1] = (DM(1, 2) * DM(2, 0) - DM(1, 0) * DM(2, 2)) / det;
    inv[3 * 1 + 1] = (DM(0, 0) * DM(2, 2) - DM(0, 2) * DM(2, 0)) / det;
    inv[3 * 2 + 1] = (DM(0, 2) * DM(1, 0) - DM(0, 0) * DM(1, 2)) / det;
    inv[3 * 0 + 2] = (DM(1, 0) * DM(2, 1) - DM(1, 1) * DM(2, 0)) / det;
    inv[3 * 1 + 2] = (DM(0, 1) * DM(2, 0) - DM(0, 0) * DM(2, 1)) / det;
    inv[3 * 2 + 2] = (DM(0, 0) * DM(1, 1) - DM(0, 1) * DM(1, 0)) / det;
#undef DM
    memcpy(Q, inv, sizeof(inv));
    *vol = fabs(det / 6);
    return det == 0.0 ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* a5  F = X Q : include/solver/constraint/ElementConstraint.hpp:78-119       */
/* ------------------------------------------------------------------------ */
static void compute_F(const double* p0, const double* p1, const double* p2,
                      const double* p3, const double* Q, double* F)
{
    double X[9];
    for (int k = 0; k < 3; k++) {
        X[0 + k] = p0[k] - p3[k];
        X[3 + k] = p1[k] - p3[k];
        X[6 + k] = p2[k] - p3[k];
    }
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++)
            F[3 * c + r] = X[0 + r] * Q[3 * c + 0] + X[3 + r] * Q[3 * c + 1]
                         + X[6 + r] * Q[3 * c + 2];
}

/* a3  deviatoric: include/solver/constraint/DeviatoricConstraint.hpp:112-152 */
static double dev_eval(const double* F)
{
    double s = 0;
    for (int i = 0; i < 9; i++) s += F[i] * F[i];
    return sqrt(s);
}

static void dev_grad(const double* F, const double* Q, double C, double* g)
{
    const double inv_C = 1.0 / C;
    /* column i of F Q^T (vertex i), i = 0..2 : sum_c F(r,c) Q(i,c) */
    for (int i = 0; i < 3; i++)
        for (int r = 0; r < 3; r++)
            g[3 * i + r] = inv_C * (F[r] * Q[i] + F[3 + r] * Q[3 + i] + F[6 + r] * Q[6 + i]);
    for (int r = 0; r < 3; r++) g[9 + r] = -g[r] - g[3 + r] - g[6 + r];
}

/* a4  hydrostatic: include/solver/constraint/HydrostaticConstraint.hpp:123-229 */
static double det_F(const double* F)
{
    return F[0] * F[4] * F[8] - F[0] * F[7] * F[5] - F[3] * F[1] * F[8]
         + F[3] * F[7] * F[2] + F[6] * F[1] * F[5] - F[6] * F[4] * F[2];
}

static double hyd_eval(const double* F, double gamma)
{
    const double J = det_F(F);
    if (J >= 1) return -gamma + log(J);
    const double a = J - 1;
    return -gamma + a - a * a / 2.0 + a * a * a / 3.0;
}

static void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void hyd_grad(const double* F, const double* Q, double* g)
{
    double Fc[9];
    cross3(F + 3, F + 6, Fc);     /* f1 x f2 */
    cross3(F + 6, F, Fc + 3);     /* f2 x f0 */
    cross3(F, F + 3, Fc + 6);     /* f0 x f1 */
    const double J = det_F(F);
    double fac;
    if (J >= 1.0) fac = 1 / J;
    else fac = 1 - (J - 1) + (J - 1) * (J - 1);
    for (int i = 0; i < 3; i++)
        for (int r = 0; r < 3; r++)
            g[3 * i + r] = fac * (Fc[r] * Q[i] + Fc[3 + r] * Q[3 + i] + Fc[6 + r] * Q[6 + i]);
    for (int r = 0; r < 3; r++) g[9 + r] = -g[r] - g[3 + r] - g[6 + r];
}

/* Evaluate both element constraints and their 12-vector gradients.
 * C[0] = C_D, C[1] = C_H; grad[0..11] = dC_D, grad[12..23] = dC_H. */
ORC_API void orc_eval_tet(const double* x12, const double* Q, double gamma,
                          double* C, double* grad24)
{
    double F[9];
    compute_F(x12, x12 + 3, x12 + 6, x12 + 9, Q, F);
    C[0] = dev_eval(F);
    dev_grad(F, Q, C[0], grad24);
    C[1] = hyd_eval(F, gamma);
    hyd_grad(F, Q, grad24 + 12);
}

ORC_API void orc_deformation_gradient(const double* x12, const double* Q, double* F)
{
    compute_F(x12, x12 + 3, x12 + 6, x12 + 9, Q, F);
}

/* ------------------------------------------------------------------------ */
/* a7  combined 2x2 projection (2nd order):                                  */
/*     src/solver/xpbd_projector/CombinedNeohookeanConstraintProjector.cpp:98-179 */
/*     Inputs are the four vertex positions (p[k] points at 3 doubles), the   */
/*     four inverse masses, Q, compliances alpha_D/alpha_H (not yet divided   */
/*     by dt^2), gamma, dt and the multipliers lam[2] (updated in place).     */
/*     dx[12] receives the coordinate updates in vertex-major order.          */
/* ------------------------------------------------------------------------ */
/* Relaxation factor on the multiplier step (1 = the reference projector).
 * Scaling dlambda scales the position update identically, so the momentum
 * relation between x and lambda is kept. */
static double g_relax = 1.0;
ORC_API void orc_set_relaxation(double omega) { g_relax = omega; }

/* ts is the time scale of the compliance: alpha~ = alpha / ts with ts = dt^2
 * (second order, CombinedNeohookeanConstraintProjector.cpp:117) or ts = dt
 * (first order, CombinedNeohookeanConstraintProjector.cpp:35). */
static void project_combined(const double* const p[4], const double* w,
                             const double* Q, double aD, double aH, double gamma,
                             double ts, double* lam, double* dx)
{
    double F[9], C[2], g[24];
    compute_F(p[0], p[1], p[2], p[3], Q, F);
    C[0] = dev_eval(F);
    dev_grad(F, Q, C[0], g);
    C[1] = hyd_eval(F, gamma);
    hyd_grad(F, Q, g + 12);

    const double at[2] = { aD / ts, aH / ts };
    /* A(ci, cj) accumulated from alpha_tilde on the diagonal, vertex by vertex */
    double A00 = at[0], A01 = 0, A11 = at[1];
    for (int i = 0; i < 4; i++)
        A00 += w[i] * (g[3 * i] * g[3 * i] + g[3 * i + 1] * g[3 * i + 1] + g[3 * i + 2] * g[3 * i + 2]);
    for (int i = 0; i < 4; i++)
        A01 += w[i] * (g[3 * i] * g[12 + 3 * i] + g[3 * i + 1] * g[12 + 3 * i + 1] + g[3 * i + 2] * g[12 + 3 * i + 2]);
    for (int i = 0; i < 4; i++)
        A11 += w[i] * (g[12 + 3 * i] * g[12 + 3 * i] + g[12 + 3 * i + 1] * g[12 + 3 * i + 1] + g[12 + 3 * i + 2] * g[12 + 3 * i + 2]);

    const double b0 = -C[0] - at[0] * lam[0];
    const double b1 = -C[1] - at[1] * lam[1];
    const double det = A00 * A11 - A01 * A01;
    const double d0 = g_relax * (b0 * A11 - b1 * A01) / det;
    const double d1 = g_relax * (b1 * A00 - b0 * A01) / det;
    lam[0] += d0;
    lam[1] += d1;
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++)
            dx[3 * i + k] = w[i] * (g[3 * i + k] * d0 + g[12 + 3 * i + k] * d1);
}

ORC_API void orc_project_tet(const double* x12, const double* w4, const double* Q,
                             double aD, double aH, double gamma, double dt,
                             double* lam2, double* dx12)
{
    const double* p[4] = { x12, x12 + 3, x12 + 6, x12 + 9 };
    project_combined(p, w4, Q, aD, aH, gamma, dt * dt, lam2, dx12);
}

/* ------------------------------------------------------------------------ */
/* a18 static SDF collision, single-constraint projector with alpha = 0:     */
/*     include/solver/xpbd_projector/ConstraintProjector.hpp:111-172 and     */
/*     src/solver/constraint/StaticDeformableCollisionConstraint.cpp:20-55   */
/*     Returns 0 and leaves dx untouched when the inequality is inactive.    */
/* ------------------------------------------------------------------------ */
static int project_contact(const double* const p[3], const double* w,
                           const double* cp, const double* cn, const double* bary,
                           double* lam, double* dx9)
{
    double a[3];
    for (int k = 0; k < 3; k++)
        a[k] = bary[0] * p[0][k] + bary[1] * p[1][k] + bary[2] * p[2][k];
    const double C = cn[0] * (a[0] - cp[0]) + cn[1] * (a[1] - cp[1]) + cn[2] * (a[2] - cp[2]);
    if (C > 0) return 0;
    double g[9];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) g[3 * i + k] = bary[i] * cn[k];
    double L = 0; /* alpha_tilde = 0 */
    for (int i = 0; i < 3; i++)
        L += w[i] * (g[3 * i] * g[3 * i] + g[3 * i + 1] * g[3 * i + 1] + g[3 * i + 2] * g[3 * i + 2]);
    const double dl = (-C - 0.0 * (*lam)) / L;
    *lam += dl;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) dx9[3 * i + k] = w[i] * g[3 * i + k] * dl;
    return 1;
}

/* One contact projection on three given positions (x9), for unit checks.
 * Returns 1 when the constraint was active and dx9 / *lam were updated. */
ORC_API int orc_project_contact(const double* x9, const double* w3, const double* cp,
                                const double* cn, const double* bary, double* lam, double* dx9)
{
    const double* p[3] = { x9, x9 + 3, x9 + 6 };
    return project_contact(p, w3, cp, cn, bary, lam, dx9);
}

/* ------------------------------------------------------------------------ */
/* a19 SDFs: src/geometry/SphereSDF.cpp:16-26, src/geometry/BoxSDF.cpp:19-84 */
/*     Quaternions are (x, y, z, w) as in src/utils/GeometryUtils.cpp:92-102  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int kind;
    double pos[3];
    double quat[4];
    double prm[3]; /* sphere: radius in prm[0]; box: full size */
} orc_sdf;

static void quat_rotate(const double* v, const double* q, double* out)
{
    /* 2 (u.v) u + (2 w^2 - 1) v + 2 w (u x v),  GeometryUtils.cpp:92-97 */
    const double* u = q;
    const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
    double c[3];
    cross3(u, v, c);
    const double s = 2 * q[3] * q[3] - 1;
    for (int k = 0; k < 3; k++) out[k] = 2 * uv * u[k] + s * v[k] + 2 * q[3] * c[k];
}

static void quat_inverse(const double* q, double* qi)
{
    const double n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    qi[0] = -q[0] / n2; qi[1] = -q[1] / n2; qi[2] = -q[2] / n2; qi[3] = q[3] / n2;
}

static double sdf_eval(const orc_sdf* s, const double* x)
{
    if (s->kind == ORC_SDF_SPHERE) {
        const double d0 = x[0] - s->pos[0], d1 = x[1] - s->pos[1], d2 = x[2] - s->pos[2];
        return sqrt(d0 * d0 + d1 * d1 + d2 * d2) - s->prm[0];
    }
    double rel[3] = { x[0] - s->pos[0], x[1] - s->pos[1], x[2] - s->pos[2] }, qi[4], xb[3];
    quat_inverse(s->quat, qi);
    quat_rotate(rel, qi, xb);
    double q[3], qc[3];
    for (int k = 0; k < 3; k++) {
        q[k] = fabs(xb[k]) - 0.5 * s->prm[k];
        qc[k] = q[k] > 0.0 ? q[k] : 0.0;
    }
    double mq = q[0];
    if (q[1] > mq) mq = q[1];
    if (q[2] > mq) mq = q[2];
    const double outside = sqrt(qc[0] * qc[0] + qc[1] * qc[1] + qc[2] * qc[2]);
    const double inside = mq < 0.0 ? mq : 0.0;
    return outside + inside;
}

static void sdf_grad(const orc_sdf* s, const double* x, double* g)
{
    if (s->kind == ORC_SDF_SPHERE) {
        double d[3] = { x[0] - s->pos[0], x[1] - s->pos[1], x[2] - s->pos[2] };
        const double n = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; k++) g[k] = n > 0 ? d[k] / n : d[k];
        return;
    }
    double rel[3] = { x[0] - s->pos[0], x[1] - s->pos[1], x[2] - s->pos[2] }, qi[4], xb[3];
    quat_inverse(s->quat, qi);
    quat_rotate(rel, qi, xb);
    double q[3], qc[3], sg[3];
    for (int k = 0; k < 3; k++) {
        q[k] = fabs(xb[k]) - s->prm[k] / 2.0;
        qc[k] = q[k] > 0.0 ? q[k] : 0.0;
        sg[k] = xb[k] > 0 ? 1.0 : -1.0;
    }
    double mq = q[0];
    if (q[1] > mq) mq = q[1];
    if (q[2] > mq) mq = q[2];
    const double min_dist = mq < 0.0 ? mq : 0.0;
    double go[3] = { sg[0] * qc[0], sg[1] * qc[1], sg[2] * qc[2] };
    const double n = sqrt(go[0] * go[0] + go[1] * go[1] + go[2] * go[2]);
    if (n > 0) for (int k = 0; k < 3; k++) go[k] /= n;  /* zero vector stays zero */
    double gb[3];
    for (int k = 0; k < 3; k++) gb[k] = go[k] + sg[k] * (q[k] == min_dist ? 1.0 : 0.0);
    quat_rotate(gb, s->quat, g);
}

ORC_API void orc_sdf_eval(int kind, const double* pos, const double* quat, const double* prm,
                          const double* x, double* d, double* grad)
{
    orc_sdf s;
    s.kind = kind;
    memcpy(s.pos, pos, sizeof(s.pos));
    if (quat) memcpy(s.quat, quat, sizeof(s.quat));
    else { s.quat[0] = s.quat[1] = s.quat[2] = 0; s.quat[3] = 1; }
    memcpy(s.prm, prm, sizeof(s.prm));
    *d = sdf_eval(&s, x);
    if (grad) sdf_grad(&s, x, grad);
}

/* Frank-Wolfe closest point on a triangle: src/collision/CollisionScene.cpp:253-283 */
static void frank_wolfe(const orc_sdf* s, const double* p1, const double* p2,
                        const double* p3, double* x)
{
    const double d1 = sdf_eval(s, p1), d2 = sdf_eval(s, p2), d3 = sdf_eval(s, p3);
    const double* start;
    if (d1 <= d2 && d1 <= d3) start = p1;
    else if (d2 <= d1 && d2 <= d3) start = p2;
    else start = p3;
    memcpy(x, start, 3 * sizeof(double));
    for (int i = 0; i < 32; i++) {
        const double alpha = 2.0 / (i + 3);
        double g[3];
        sdf_grad(s, x, g);
        const double s1 = p1[0] * g[0] + p1[1] * g[1] + p1[2] * g[2];
        const double s2 = p2[0] * g[0] + p2[1] * g[1] + p2[2] * g[2];
        const double s3 = p3[0] * g[0] + p3[1] * g[1] + p3[2] * g[2];
        const double* sv;
        if (s1 < s2 && s1 < s3) sv = p1;
        else if (s2 < s1 && s2 < s3) sv = p2;
        else sv = p3;
        for (int k = 0; k < 3; k++) x[k] = x[k] + alpha * (sv[k] - x[k]);
    }
}

/* src/utils/GeometryUtils.cpp:6-24 */
static void barycentric(const double* p, const double* a, const double* b,
                        const double* c, double* uvw)
{
    double v0[3], v1[3], v2[3];
    for (int k = 0; k < 3; k++) { v0[k] = b[k] - a[k]; v1[k] = c[k] - a[k]; v2[k] = p[k] - a[k]; }
#define DOT(u, v) (u[0] * v[0] + u[1] * v[1] + u[2] * v[2])
    const double d00 = DOT(v0, v0), d01 = DOT(v0, v1), d11 = DOT(v1, v1);
    const double d20 = DOT(v2, v0), d21 = DOT(v2, v1);
#undef DOT
    const double denom = d00 * d11 - d01 * d01;
    const double v = (d11 * d20 - d01 * d21) / denom;
    const double w = (d00 * d21 - d01 * d20) / denom;
    uvw[0] = 1 - v - w;
    uvw[1] = v;
    uvw[2] = w;
}

ORC_API void orc_frank_wolfe(int kind, const double* pos, const double* quat, const double* prm,
                             const double* p1, const double* p2, const double* p3, double* x)
{
    orc_sdf s;
    s.kind = kind;
    memcpy(s.pos, pos, sizeof(s.pos));
    if (quat) memcpy(s.quat, quat, sizeof(s.quat));
    else { s.quat[0] = s.quat[1] = s.quat[2] = 0; s.quat[3] = 1; }
    memcpy(s.prm, prm, sizeof(s.prm));
    frank_wolfe(&s, p1, p2, p3, x);
}

ORC_API void orc_barycentric(const double* p, const double* a, const double* b,
                             const double* c, double* uvw)
{
    barycentric(p, a, b, c, uvw);
}

/* ------------------------------------------------------------------------ */
/* Whole-object simulation state (XPBDMeshObject_ + XPBDSolver + Simulation) */
/* ------------------------------------------------------------------------ */
typedef struct orc_sim {
    int nv, nt, nf, mode, iters;
    double dt, g;
    double *x, *xp, *v, *m, *w;
    unsigned char* pin;
    int* tets;
    double *Q, *vol, *aD, *aH, *gam, *lam;
    int* order;            /* GS visiting order of tets (identity = reference order) */
    int* faces;            /* surface triangles, 3 per face */
    orc_sdf sdf[1];
    int nsdf;
    /* contacts: static collision constraints in insertion order */
    int nc, ccap;
    int* c_face;
    double *c_p, *c_n, *c_bary, *c_lam;
    int* face_rank;        /* optional visiting rank of faces for contact projection */
    int* corder;           /* contact visiting order (indices into the contact list) */
    double last_res;       /* ||r||_2 at the end of the last solve (before pin reset) */
    int want_pres;         /* record the primary residual at the end of every solve */
    double* xt;            /* the step's inertial prediction (when want_pres) */
    double* pres;          /* primary residual at the end of the last solve (3*nv, when want_pres) */
    double last_pres;      /* its norm */
    double* jbuf;          /* Jacobi coordinate-update buffer (values) */
    int* jidx;             /* Jacobi coordinate-update buffer (coordinate index) */
    double ts;             /* compliance time scale: dt^2 (second order) or dt (first order) */
    int first_order;       /* 1: first-order XPBD (damping replaces mass) */
    double* dz1;           /* first order: per-vertex inertial move -g m dt / B */
    double *E, *nu;        /* element materials (kept for the first-order damping) */
    int na;                /* attachments: vertex, target point, multiplier */
    int* att_v;
    double *att_p, *att_lam;
    int local_iters;       /* num-local-collision-iters */
    int order_set;         /* s->order was set by orc_set_order (device order) */
    int *v2e_off, *v2e;    /* elements attached to each vertex, ascending (TetMesh.cpp:88-96) */
    int last_local_n;      /* element projections per local pass in the last step */
    int ncolor;            /* > 0: s->order is colour-major over units of consecutive positions: */
    int nunit;             /*   unit u at positions [unit_off[u], unit_off[u+1]), colour c holds */
    int *unit_off;         /*   units [color_off[c], color_off[c+1]), which share no vertex */
    int* color_off;
    int nthreads;          /* OpenMP threads of the colour-parallel CPU baseline (<= 1: sequential) */
} orc_sim;

/* XPBDMeshObject.cpp:102-160 (setup) + 277-331 (masses) + 352-417 (constraints) */
ORC_API orc_sim* orc_create(int nv, int nt, const double* x0, const double* v0,
                            const int* tets, const double* E, const double* nu,
                            const double* rho, const unsigned char* pinned,
                            double dt, double g, int iters, int mode)
{
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->nv = nv; s->nt = nt; s->dt = dt; s->g = g; s->iters = iters; s->mode = mode;
    s->x = (double*)malloc(sizeof(double) * 3 * nv);
    s->xp = (double*)malloc(sizeof(double) * 3 * nv);
    s->v = (double*)calloc(3 * (size_t)nv, sizeof(double));
    s->m = (double*)calloc(nv, sizeof(double));
    s->w = (double*)malloc(sizeof(double) * nv);
    s->pin = (unsigned char*)calloc(nv, 1);
    s->tets = (int*)malloc(sizeof(int) * 4 * (size_t)nt);
    s->Q = (double*)malloc(sizeof(double) * 9 * (size_t)nt);
    s->vol = (double*)malloc(sizeof(double) * nt);
    s->aD = (double*)malloc(sizeof(double) * nt);
    s->aH = (double*)malloc(sizeof(double) * nt);
    s->gam = (double*)malloc(sizeof(double) * nt);
    s->lam = (double*)calloc(2 * (size_t)nt, sizeof(double));
    s->order = (int*)malloc(sizeof(int) * nt);
    s->jbuf = (double*)malloc(sizeof(double) * 12 * (size_t)nt);
    s->jidx = (int*)malloc(sizeof(int) * 12 * (size_t)nt);
    memcpy(s->x, x0, sizeof(double) * 3 * nv);
    memcpy(s->xp, x0, sizeof(double) * 3 * nv);
    if (v0) memcpy(s->v, v0, sizeof(double) * 3 * nv);
    if (pinned) memcpy(s->pin, pinned, nv);
    memcpy(s->tets, tets, sizeof(int) * 4 * (size_t)nt);
    int bad = 0;
    for (int e = 0; e < nt; e++) {
        const int* t = tets + 4 * e;
        if (orc_rest_tet(x0 + 3 * t[0], x0 + 3 * t[1], x0 + 3 * t[2], x0 + 3 * t[3],
                         s->Q + 9 * e, s->vol + e)) bad = 1;
        s->order[e] = e;
    }
    /* lumped masses, element order: XPBDMeshObject.cpp:288-313 */
    for (int e = 0; e < nt; e++) {
        const double em = s->vol[e] * rho[e];
        for (int j = 0; j < 4; j++) s->m[tets[4 * e + j]] += em / 4.0;
    }
    for (int i = 0; i < nv; i++) s->w[i] = 1.0 / s->m[i];  /* PositionReference.hpp:30-33 */
    /* DeviatoricConstraint.cpp:13, HydrostaticConstraint.cpp:15-16 */
    for (int e = 0; e < nt; e++) {
        double mu, lam;
        orc_lame(E[e], nu[e], &mu, &lam);
        s->aD[e] = 1 / (mu * s->vol[e]);
        s->aH[e] = 1 / (lam * s->vol[e]);
        s->gam[e] = mu / lam;
    }
    s->ts = dt * dt;
    s->E = (double*)malloc(sizeof(double) * nt);
    s->nu = (double*)malloc(sizeof(double) * nt);
    memcpy(s->E, E, sizeof(double) * nt);
    memcpy(s->nu, nu, sizeof(double) * nt);
    (void)bad;
    return s;
}

/* First-order XPBD (FirstOrderXPBDMeshObject):
 *   per-vertex damping B_i = V_i b E_i / (1 + nu_i) with adjust, else V_i b,
 *   V_i = sum of element volume / 4, E_i and nu_i = means over the elements
 *   attached to the vertex (XPBDMeshObject.cpp:277-348);
 *   B_i is the constraint inertia (XPBDMeshObjectBase.hpp:125-136), so w_i = 1/B_i;
 *   alpha~ = alpha / dt (CombinedNeohookeanConstraintProjector.cpp:16-96);
 *   inertial move z -= g m_i dt / B_i, no velocity term (XPBDMeshObject.cpp:450-457). */
ORC_API void orc_set_first_order(orc_sim* s, double damping_multiplier, int adjust_to_material)
{
    const int nv = s->nv, nt = s->nt;
    double* V = (double*)calloc(nv, sizeof(double));
    double* Ev = (double*)calloc(nv, sizeof(double));
    double* Nv = (double*)calloc(nv, sizeof(double));
    int* cnt = (int*)calloc(nv, sizeof(int));
    for (int e = 0; e < nt; e++)
        for (int j = 0; j < 4; j++) cnt[s->tets[4 * e + j]]++;
    for (int e = 0; e < nt; e++)
        for (int j = 0; j < 4; j++) {
            const int v = s->tets[4 * e + j];
            V[v] += s->vol[e] / 4.0;
            Ev[v] += s->E[e] / cnt[v];
            Nv[v] += s->nu[e] / cnt[v];
        }
    free(s->dz1);
    s->dz1 = (double*)malloc(sizeof(double) * nv);
    for (int i = 0; i < nv; i++) {
        const double B = adjust_to_material ? V[i] * damping_multiplier * Ev[i] / (1 + Nv[i])
                                            : V[i] * damping_multiplier;
        s->w[i] = 1.0 / B;
        s->dz1[i] = -s->g * s->m[i] * s->dt / B;
    }
    s->ts = s->dt;
    s->first_order = 1;
    free(V); free(Ev); free(Nv); free(cnt);
}

ORC_API void orc_get_inverse_inertia(const orc_sim* s, double* out) { memcpy(out, s->w, sizeof(double) * s->nv); }

/* the inertial prediction of the step, kept for the primary residual */
static void keep_prediction(orc_sim* s)
{
    if (s->want_pres) memcpy(s->xt, s->x, sizeof(double) * 3 * (size_t)s->nv);
}

/* inertial move of one step, both orders (XPBDMeshObject.cpp:430, 446-470) */
static void move_inertially(orc_sim* s)
{
    const int nv = s->nv;
    const double dt = s->dt;
    memcpy(s->xp, s->x, sizeof(double) * 3 * nv);
    if (s->first_order) {
        for (int i = 0; i < nv; i++) s->x[3 * i + 2] += s->dz1[i];
    } else {
        for (int i = 0; i < nv; i++)
            for (int k = 0; k < 3; k++) s->x[3 * i + k] += dt * s->v[3 * i + k];
        const double dz = -s->g * dt * dt;
        for (int i = 0; i < nv; i++) s->x[3 * i + 2] += dz;
    }
    keep_prediction(s);
}

ORC_API void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->x); free(s->xp); free(s->v); free(s->m); free(s->w); free(s->pin);
    free(s->tets); free(s->Q); free(s->vol); free(s->aD); free(s->aH); free(s->gam);
    free(s->lam); free(s->order); free(s->faces); free(s->jbuf); free(s->jidx);
    free(s->c_face); free(s->c_p); free(s->c_n); free(s->c_bary); free(s->c_lam);
    free(s->face_rank); free(s->corder); free(s->dz1); free(s->E); free(s->nu); free(s->color_off); free(s->unit_off);
    free(s->att_v); free(s->att_p); free(s->att_lam);
    free(s->v2e_off); free(s->v2e);
    free(s->xt); free(s->pres);
    free(s);
}

ORC_API void orc_set_order(orc_sim* s, const int* order)
{
    memcpy(s->order, order, sizeof(int) * s->nt);
    s->order_set = 1;
    s->ncolor = 0;
}

/* Colour-parallel CPU baseline (an order of this build, not the reference's).
 * Units: with cluster = 0 every element is a unit; otherwise a unit is a run
 * of consecutive elements (element order) spanning at most 8 vertices and 8
 * elements -- one hex cell of a mesh generated cell by cell.  Units are given
 * colours greedily (first fit, unit order) so that units sharing a vertex
 * differ; the Gauss-Seidel order becomes colour-major (unit order within a
 * colour, element order within a unit).  With orc_set_threads > 1 the units of
 * one colour are then projected concurrently, each unit by one thread in
 * order.  Returns the number of colours, or -1 (order unchanged) if more than
 * 64 would be needed. */
ORC_API int orc_color_elements(orc_sim* s, int cluster)
{
    const int nt = s->nt;
    int* uoff = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    int nu = 0;
    for (int e = 0; e < nt;) {
        uoff[nu++] = e;
        int vs[8], nvs = 0, k = 0;
        for (; e < nt && k < 8; e++, k++) {
            const int* t = s->tets + 4 * e;
            int add[4], na = 0;
            for (int i = 0; i < 4; i++) {
                int seen = 0;
                for (int j = 0; j < nvs; j++) seen |= vs[j] == t[i];
                for (int j = 0; j < na; j++) seen |= add[j] == t[i];
                if (!seen) add[na++] = t[i];
            }
            if (k > 0 && (!cluster || nvs + na > 8)) break;
            for (int j = 0; j < na && nvs < 8; j++) vs[nvs++] = add[j];
        }
    }
    uoff[nu] = nt;
    unsigned long long* used = (unsigned long long*)calloc((size_t)s->nv, sizeof(unsigned long long));
    int* col = (int*)malloc(sizeof(int) * (size_t)(nu > 0 ? nu : 1));
    int ncol = 0;
    for (int u = 0; u < nu; u++) {
        unsigned long long m = 0ull;
        for (int e = uoff[u]; e < uoff[u + 1]; e++)
            for (int i = 0; i < 4; i++) m |= used[s->tets[4 * e + i]];
        if (m == ~0ull) {
            free(uoff);
            free(used);
            free(col);
            return -1;
        }
        const int c = __builtin_ctzll(~m);
        col[u] = c;
        for (int e = uoff[u]; e < uoff[u + 1]; e++)
            for (int i = 0; i < 4; i++) used[s->tets[4 * e + i]] |= 1ull << c;
        if (c + 1 > ncol) ncol = c + 1;
    }
    /* colour-major unit order */
    free(s->color_off);
    free(s->unit_off);
    s->color_off = (int*)calloc((size_t)ncol + 1, sizeof(int));
    for (int u = 0; u < nu; u++) s->color_off[col[u] + 1]++;
    for (int c = 0; c < ncol; c++) s->color_off[c + 1] += s->color_off[c];
    int* fill = (int*)malloc(sizeof(int) * ((size_t)ncol + 1));
    memcpy(fill, s->color_off, sizeof(int) * ((size_t)ncol + 1));
    int* uorder = (int*)malloc(sizeof(int) * (size_t)(nu > 0 ? nu : 1));
    for (int u = 0; u < nu; u++) uorder[fill[col[u]]++] = u;
    s->unit_off = (int*)malloc(sizeof(int) * ((size_t)nu + 1));
    int pos = 0;
    for (int i = 0; i < nu; i++) {
        const int u = uorder[i];
        s->unit_off[i] = pos;
        for (int e = uoff[u]; e < uoff[u + 1]; e++) s->order[pos++] = e;
    }
    s->unit_off[nu] = pos;
    s->nunit = nu;
    s->ncolor = ncol;
    s->order_set = 1;
    free(uorder);
    free(fill);
    free(uoff);
    free(used);
    free(col);
    return ncol;
}

/* Union-find root of vertex v with path halving (parent[v] < 0: v is a root). */
static int uf_root(int* parent, int v)
{
    while (parent[v] >= 0) {
        if (parent[parent[v]] >= 0) parent[v] = parent[parent[v]];
        v = parent[v];
    }
    return v;
}

/* Gauss-Seidel order given as levels: order[] visits the elements sequentially
 * and level_off[0..nlevels] cuts it into ranges (the colours of a
 * colour-ordered solver).  Within a level, elements are grouped into units =
 * connected components of the shared-vertex graph restricted to that level;
 * units keep their elements in order and share no vertex with each other, so
 * projecting the units of a level concurrently (orc_set_threads > 1), each in
 * order, is bit for bit the sequential pass over order[].  s->order becomes the
 * level-major, unit-major arrangement of order[], which visits every pair of
 * elements sharing a vertex in their relative order in order[].  Returns the
 * number of units, or -1 on bad offsets. */
ORC_API int orc_set_order_levels(orc_sim* s, const int* order, int nlevels, const int* level_off)
{
    const int nt = s->nt, nv = s->nv;
    if (nlevels < 1 || level_off[0] != 0 || level_off[nlevels] != nt) return -1;
    for (int l = 0; l < nlevels; l++)
        if (level_off[l + 1] < level_off[l]) return -1;
    int* parent = (int*)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
    for (int v = 0; v < nv; v++) parent[v] = -1;
    int* comp = (int*)malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));   /* unit of position i (level-local) */
    int* root_unit = (int*)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
    for (int v = 0; v < nv; v++) root_unit[v] = -1;
    int* unit_off = (int*)malloc(sizeof(int) * ((size_t)nt + 1));
    int* color_off = (int*)malloc(sizeof(int) * ((size_t)nlevels + 1));
    int* neword = (int*)malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
    int nunit = 0, pos = 0;
    for (int l = 0; l < nlevels; l++) {
        const int b = level_off[l], en = level_off[l + 1];
        color_off[l] = nunit;
        for (int i = b; i < en; i++) {
            const int* t = s->tets + 4 * order[i];
            int r = uf_root(parent, t[0]);
            for (int j = 1; j < 4; j++) {
                const int q = uf_root(parent, t[j]);
                if (q != r) parent[q] = r;
            }
        }
        /* units numbered by first appearance; count their sizes */
        const int u0 = nunit;
        for (int i = b; i < en; i++) {
            const int r = uf_root(parent, s->tets[4 * order[i]]);
            if (root_unit[r] < 0) root_unit[r] = nunit++;
            comp[i] = root_unit[r] - u0;
        }
        const int nu = nunit - u0;
        int* cnt = (int*)calloc((size_t)nu + 1, sizeof(int));
        for (int i = b; i < en; i++) cnt[comp[i] + 1]++;
        for (int u = 0; u < nu; u++) cnt[u + 1] += cnt[u];
        for (int u = 0; u < nu; u++) unit_off[u0 + u] = pos + cnt[u];
        for (int i = b; i < en; i++) neword[pos + cnt[comp[i]]++] = order[i];
        pos = en;
        free(cnt);
        /* reset the touched vertices for the next level */
        for (int i = b; i < en; i++) {
            const int* t = s->tets + 4 * order[i];
            for (int j = 0; j < 4; j++) {
                root_unit[t[j]] = -1;
                parent[t[j]] = -1;
            }
        }
    }
    color_off[nlevels] = nunit;
    unit_off[nunit] = nt;
    memcpy(s->order, neword, sizeof(int) * (size_t)nt);
    free(s->color_off);
    free(s->unit_off);
    s->color_off = color_off;
    s->unit_off = unit_off;
    s->nunit = nunit;
    s->ncolor = nlevels;
    s->order_set = 1;
    free(parent);
    free(comp);
    free(root_unit);
    free(neword);
    return nunit;
}

ORC_API void orc_set_threads(orc_sim* s, int n) { s->nthreads = n; }

/* After orc_color_elements: the order (nt), unit offsets (nunit + 1, positions)
 * and colour offsets (ncolor + 1, units).  Returns the number of units. */
ORC_API int orc_get_colors(const orc_sim* s, int* order, int* unit_off, int* color_off)
{
    if (s->ncolor <= 0) return 0;
    if (order) memcpy(order, s->order, sizeof(int) * (size_t)s->nt);
    if (unit_off) memcpy(unit_off, s->unit_off, sizeof(int) * ((size_t)s->nunit + 1));
    if (color_off) memcpy(color_off, s->color_off, sizeof(int) * ((size_t)s->ncolor + 1));
    return s->nunit;
}

ORC_API void orc_set_mode(orc_sim* s, int mode, int iters)
{
    s->mode = mode;
    s->iters = iters;
}

ORC_API void orc_set_faces(orc_sim* s, int nf, const int* faces)
{
    free(s->faces);
    s->nf = nf;
    s->faces = (int*)malloc(sizeof(int) * 3 * (size_t)(nf > 0 ? nf : 1));
    memcpy(s->faces, faces, sizeof(int) * 3 * (size_t)nf);
}

ORC_API void orc_set_sdf(orc_sim* s, int kind, const double* pos, const double* quat,
                         const double* prm)
{
    s->nsdf = kind == ORC_SDF_NONE ? 0 : 1;
    s->sdf[0].kind = kind;
    memcpy(s->sdf[0].pos, pos, 3 * sizeof(double));
    if (quat) memcpy(s->sdf[0].quat, quat, 4 * sizeof(double));
    else { s->sdf[0].quat[0] = s->sdf[0].quat[1] = s->sdf[0].quat[2] = 0; s->sdf[0].quat[3] = 1; }
    memcpy(s->sdf[0].prm, prm, 3 * sizeof(double));
}

ORC_API void orc_set_state(orc_sim* s, const double* x, const double* v)
{
    if (x) memcpy(s->x, x, sizeof(double) * 3 * s->nv);
    if (v) memcpy(s->v, v, sizeof(double) * 3 * s->nv);
}

static void add_contact(orc_sim* s, int face, const double* p, const double* n, const double* bary)
{
    if (s->nc == s->ccap) {
        s->ccap = s->ccap ? 2 * s->ccap : 64;
        s->c_face = (int*)realloc(s->c_face, sizeof(int) * s->ccap);
        s->c_p = (double*)realloc(s->c_p, sizeof(double) * 3 * s->ccap);
        s->c_n = (double*)realloc(s->c_n, sizeof(double) * 3 * s->ccap);
        s->c_bary = (double*)realloc(s->c_bary, sizeof(double) * 3 * s->ccap);
        s->c_lam = (double*)realloc(s->c_lam, sizeof(double) * s->ccap);
    }
    const int c = s->nc++;
    s->c_face[c] = face;
    memcpy(s->c_p + 3 * c, p, 3 * sizeof(double));
    memcpy(s->c_n + 3 * c, n, 3 * sizeof(double));
    memcpy(s->c_bary + 3 * c, bary, 3 * sizeof(double));
    s->c_lam[c] = 0;
}

static void order_contacts(orc_sim* s);

/* Face-vs-SDF detection, face order: src/collision/CollisionScene.cpp:171-214.
 * Replaces the current contact set (XPBDMeshObject.cpp:233-251 clears first). */
ORC_API int orc_detect(orc_sim* s)
{
    s->nc = 0;
    if (!s->nsdf) return 0;
    const orc_sdf* sd = &s->sdf[0];
    for (int f = 0; f < s->nf; f++) {
        const double* p1 = s->x + 3 * s->faces[3 * f];
        const double* p2 = s->x + 3 * s->faces[3 * f + 1];
        const double* p3 = s->x + 3 * s->faces[3 * f + 2];
        double e12 = 0, e13 = 0, e23 = 0, cen[3];
        for (int k = 0; k < 3; k++) {
            e12 += (p2[k] - p1[k]) * (p2[k] - p1[k]);
            e13 += (p3[k] - p1[k]) * (p3[k] - p1[k]);
            e23 += (p3[k] - p2[k]) * (p3[k] - p2[k]);
            cen[k] = (p1[k] + p2[k] + p3[k]) / 3;
        }
        double me = e12;
        if (e13 > me) me = e13;
        if (e23 > me) me = e23;
        const double cd = sdf_eval(sd, cen);
        if (cd * cd > me) continue;
        double x[3];
        frank_wolfe(sd, p1, p2, p3, x);
        const double dist = sdf_eval(sd, x);
        if (dist <= 1e-4) {
            double uvw[3], gr[3], sx[3];
            barycentric(x, p1, p2, p3, uvw);
            sdf_grad(sd, x, gr);
            for (int k = 0; k < 3; k++) sx[k] = x[k] - gr[k] * dist;
            add_contact(s, f, sx, gr, uvw);
        }
    }
    order_contacts(s);
    return s->nc;
}

/* Contacts are projected in insertion (face) order, as the reference does
 * (VariadicVectorContainer insertion order).  A face rank, when set, permutes
 * that order: contacts are visited by increasing rank of their face, which is
 * how the colour-ordered device solver sequences them. */
/* sort keys of the qsort comparators, one per thread: oracle instances may be
 * stepped concurrently from several threads */
static __thread const int* g_sort_rank;
static __thread const int* g_sort_face;
static int cmp_contact(const void* a, const void* b)
{
    const int ra = g_sort_rank[g_sort_face[*(const int*)a]];
    const int rb = g_sort_rank[g_sort_face[*(const int*)b]];
    return (ra > rb) - (ra < rb);
}

static void order_contacts(orc_sim* s)
{
    free(s->corder);
    s->corder = (int*)malloc(sizeof(int) * (size_t)(s->nc > 0 ? s->nc : 1));
    for (int c = 0; c < s->nc; c++) s->corder[c] = c;
    if (s->face_rank && s->nc > 1) {
        g_sort_rank = s->face_rank;
        g_sort_face = s->c_face;
        qsort(s->corder, s->nc, sizeof(int), cmp_contact);
    }
}

ORC_API void orc_set_contact_order(orc_sim* s, const int* face_rank)
{
    free(s->face_rank);
    s->face_rank = NULL;
    if (face_rank) {
        s->face_rank = (int*)malloc(sizeof(int) * (size_t)(s->nf > 0 ? s->nf : 1));
        memcpy(s->face_rank, face_rank, sizeof(int) * s->nf);
    }
    order_contacts(s);
}

/* Replace the contact set with n given contacts (face index, surface point p,
 * normal nrm, barycentrics bary; 3 doubles each per contact), multipliers 0:
 * the same clear-and-add as orc_detect (XPBDMeshObject.cpp:180-251), with the
 * contacts supplied instead of detected.  Used to project a contact set that
 * another detector (the device) produced from the same state, so that
 * projection is compared independently of detection.  Visiting order: face
 * rank if set (orc_set_contact_order), else the given order. */
ORC_API void orc_set_contacts(orc_sim* s, int n, const int* face, const double* p, const double* nrm,
                              const double* bary)
{
    s->nc = 0;
    for (int c = 0; c < n; c++) add_contact(s, face[c], p + 3 * c, nrm + 3 * c, bary + 3 * c);
    order_contacts(s);
}

ORC_API int orc_num_contacts(const orc_sim* s) { return s->nc; }
ORC_API double orc_last_step_residual(const orc_sim* s) { return s->last_res; }

ORC_API void orc_get_contacts(const orc_sim* s, int* face, double* p, double* n, double* bary)
{
    if (face) memcpy(face, s->c_face, sizeof(int) * s->nc);
    if (p) memcpy(p, s->c_p, sizeof(double) * 3 * s->nc);
    if (n) memcpy(n, s->c_n, sizeof(double) * 3 * s->nc);
    if (bary) memcpy(bary, s->c_bary, sizeof(double) * 3 * s->nc);
}

/* One solver iteration.
 * GS:     include/solver/xpbd_solver/XPBDGaussSeidelSolver.hpp:29-80
 *         (all combined projectors in s->order, then contacts in insertion order,
 *          each projector's 12 (or 9) updates applied immediately in coordinate order)
 * Jacobi: include/solver/xpbd_solver/XPBDJacobiSolver.hpp:79-114
 *         (all projectors read the same positions, updates buffered in projector
 *          order and summed into x afterwards, no relaxation) */
/* Attachment of vertex att_v[a] to the point att_p[a] = attached point + offset
 * (AttachmentConstraint.hpp:33-88 with ConstraintProjector.hpp:111-172,
 * alpha = 0): C = |x - p|, gradient (x - p) / |x - p|.  Below |x - p| = 1e-12
 * the reference substitutes a 1e-12 gradient (AttachmentConstraint.hpp:76-81),
 * which with C = |x - p| moves the vertex by ~3e11 |x - p| along -(1, 1, 1);
 * its GPU twin (GPUAttachmentConstraint.cuh:74-86) divides by zero there.  The
 * constraint is met at that distance, so this makes no correction. */
static void project_attachment(orc_sim* s, int a, double* dx3)
{
    const double* x = s->x + 3 * s->att_v[a];
    const double* p = s->att_p + 3 * a;
    const double d[3] = { x[0] - p[0], x[1] - p[1], x[2] - p[2] };
    const double dist = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (dist < 1e-12) {
        dx3[0] = dx3[1] = dx3[2] = 0.0;
        return;
    }
    double g[3];
    for (int k = 0; k < 3; k++) g[k] = d[k] / dist;
    const double w = s->w[s->att_v[a]];
    const double lhs = w * (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double dlam = (-dist - 0.0 * s->att_lam[a]) / lhs;
    s->att_lam[a] += dlam;
    for (int k = 0; k < 3; k++) dx3[k] = w * g[k] * dlam;
}

ORC_API void orc_set_attachments(orc_sim* s, int n, const int* verts, const double* targets)
{
    free(s->att_v); free(s->att_p); free(s->att_lam);
    s->na = n;
    s->att_v = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    s->att_p = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    s->att_lam = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    if (n > 0) {
        memcpy(s->att_v, verts, sizeof(int) * n);
        memcpy(s->att_p, targets, sizeof(double) * 3 * n);
    }
}

/* Gauss-Seidel projection of element e, its 12 updates applied at once */
static void gs_element(orc_sim* s, int e)
{
    const int* t = s->tets + 4 * e;
    const double* p[4] = { s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3] };
    const double w[4] = { s->w[t[0]], s->w[t[1]], s->w[t[2]], s->w[t[3]] };
    double dx[12];
    project_combined(p, w, s->Q + 9 * e, s->aD[e], s->aH[e], s->gam[e], s->ts, s->lam + 2 * e, dx);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) s->x[3 * t[i] + k] += dx[3 * i + k];
}

static void iterate(orc_sim* s)
{
    const int nt = s->nt;
    if (s->mode == ORC_GS) {
        if (s->ncolor > 0 && s->nthreads > 1) {
            /* units of one colour share no vertex: projecting them concurrently,
             * each unit in order, is the sequential pass over the same order,
             * bit for bit */
            for (int c = 0; c < s->ncolor; c++) {
                const int b = s->color_off[c], en = s->color_off[c + 1];
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
                for (int u = b; u < en; u++)
                    for (int oi = s->unit_off[u]; oi < s->unit_off[u + 1]; oi++) gs_element(s, s->order[oi]);
            }
        } else {
            for (int oi = 0; oi < nt; oi++) gs_element(s, s->order[oi]);
        }
        for (int ci = 0; ci < s->nc; ci++) {
            const int c = s->corder[ci];
            const int* f = s->faces + 3 * s->c_face[c];
            const double* p[3] = { s->x + 3 * f[0], s->x + 3 * f[1], s->x + 3 * f[2] };
            const double w[3] = { s->w[f[0]], s->w[f[1]], s->w[f[2]] };
            double dx[9];
            if (project_contact(p, w, s->c_p + 3 * c, s->c_n + 3 * c, s->c_bary + 3 * c, s->c_lam + c, dx))
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++) s->x[3 * f[i] + k] += dx[3 * i + k];
        }
        /* attachments follow the collision projectors (type-list order, XPBDTypedefs.hpp:56) */
        for (int a = 0; a < s->na; a++) {
            double dx[3];
            project_attachment(s, a, dx);
            for (int k = 0; k < 3; k++) s->x[3 * s->att_v[a] + k] += dx[k];
        }
        return;
    }
    /* Jacobi: every element writes its own 12 buffer slots (in parallel when
     * threads are set); the sum below stays in buffer order */
    const size_t nu = 12 * (size_t)nt;
    double* cbuf = NULL;
    int* cidx = NULL;
#pragma omp parallel for if (s->nthreads > 1) num_threads(s->nthreads > 1 ? s->nthreads : 1) schedule(static)
    for (int e = 0; e < nt; e++) {
        const int* t = s->tets + 4 * e;
        const double* p[4] = { s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3] };
        const double w[4] = { s->w[t[0]], s->w[t[1]], s->w[t[2]], s->w[t[3]] };
        const size_t u0 = 12 * (size_t)e;
        project_combined(p, w, s->Q + 9 * e, s->aD[e], s->aH[e], s->gam[e], s->ts, s->lam + 2 * e, s->jbuf + u0);
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) s->jidx[u0 + 3 * i + k] = 3 * t[i] + k;
    }
    int ncu = 0;
    if (s->nc) {
        cbuf = (double*)malloc(sizeof(double) * 9 * s->nc);
        cidx = (int*)malloc(sizeof(int) * 9 * s->nc);
        for (int c = 0; c < s->nc; c++) {
            const int* f = s->faces + 3 * s->c_face[c];
            const double* p[3] = { s->x + 3 * f[0], s->x + 3 * f[1], s->x + 3 * f[2] };
            const double w[3] = { s->w[f[0]], s->w[f[1]], s->w[f[2]] };
            if (project_contact(p, w, s->c_p + 3 * c, s->c_n + 3 * c, s->c_bary + 3 * c, s->c_lam + c, cbuf + ncu)) {
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++) cidx[ncu + 3 * i + k] = 3 * f[i] + k;
                ncu += 9;
            }
        }
    }
    double* abuf = NULL;
    if (s->na) {
        abuf = (double*)malloc(sizeof(double) * 3 * s->na);
        for (int a = 0; a < s->na; a++) project_attachment(s, a, abuf + 3 * a);
    }
    for (size_t u = 0; u < nu; u++) s->x[s->jidx[u]] += s->jbuf[u];
    for (int u = 0; u < ncu; u++) s->x[cidx[u]] += cbuf[u];
    for (int a = 0; a < s->na; a++)
        for (int k = 0; k < 3; k++) s->x[3 * s->att_v[a] + k] += abuf[3 * a + k];
    free(cbuf);
    free(cidx);
    free(abuf);
}

/* Local collision iterations (XPBDMeshObject.cpp:476-483 with the gather at
 * XPBDMeshObject.cpp:764-825 and XPBDSolver.hpp:106-121), after the global solve:
 * the projector list holds, for every contact whose multiplier is non-zero
 * (insertion order), for each of its three face vertices, every element attached
 * to that vertex (ascending element index) -- an element appears once per such
 * (contact, vertex) pair -- followed by every contact.  Each of local_iters
 * passes projects that list without resetting any multiplier: Gauss-Seidel
 * applies each projection at once (XPBDGaussSeidelSolver.hpp:44-53), Jacobi
 * buffers all of them and sums them afterwards (XPBDJacobiSolver.hpp:119-145).
 * The reference visits the element list in gather order; when a device order
 * is set (orc_set_order) the list is visited in that order instead, repeats of
 * an element back to back, which is how the colour-ordered device sequences it. */
static __thread const int* g_elem_rank;
static int cmp_elem(const void* a, const void* b)
{
    const int ra = g_elem_rank[*(const int*)a], rb = g_elem_rank[*(const int*)b];
    return (ra > rb) - (ra < rb);
}

static void project_list(orc_sim* s, int* list, size_t n, int iters, int with_contacts);

static void local_collision_iterations(orc_sim* s)
{
    s->last_local_n = 0;
    if (s->local_iters <= 0 || s->nc == 0) return;
    const int nv = s->nv, nt = s->nt;
    if (!s->v2e_off) {
        s->v2e_off = (int*)calloc((size_t)nv + 1, sizeof(int));
        s->v2e = (int*)malloc(sizeof(int) * 4 * (size_t)(nt > 0 ? nt : 1));
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) s->v2e_off[s->tets[4 * e + j] + 1]++;
        for (int v = 0; v < nv; v++) s->v2e_off[v + 1] += s->v2e_off[v];
        int* fill = (int*)malloc(sizeof(int) * (size_t)(nv > 0 ? nv : 1));
        memcpy(fill, s->v2e_off, sizeof(int) * (size_t)nv);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) s->v2e[fill[s->tets[4 * e + j]]++] = e;
        free(fill);
    }
    size_t n = 0, cap = 64;
    int* list = (int*)malloc(sizeof(int) * cap);
    for (int c = 0; c < s->nc; c++) {
        if (s->c_lam[c] == 0) continue;
        const int* f = s->faces + 3 * s->c_face[c];
        for (int i = 0; i < 3; i++)
            for (int k = s->v2e_off[f[i]]; k < s->v2e_off[f[i] + 1]; k++) {
                if (n == cap) {
                    cap *= 2;
                    list = (int*)realloc(list, sizeof(int) * cap);
                }
                list[n++] = s->v2e[k];
            }
    }
    s->last_local_n = (int)n;
    project_list(s, list, n, s->local_iters, 1);
    free(list);
}

/* Project the element list (repeats allowed) iters times, each pass followed
 * by every contact when with_contacts, with the multipliers kept: Gauss-Seidel
 * applies each projection at once (XPBDGaussSeidelSolver.hpp:44-53), Jacobi
 * buffers all of a pass and sums it afterwards (XPBDJacobiSolver.hpp:119-145).
 * With a device order set (orc_set_order) the list is sorted into it first,
 * repeats of an element back to back. */
static void project_list(orc_sim* s, int* list, size_t n, int iters, int with_contacts)
{
    const int nt = s->nt;
    if (s->order_set && n > 1) {
        int* rank = (int*)malloc(sizeof(int) * (size_t)nt);
        for (int i = 0; i < nt; i++) rank[s->order[i]] = i;
        g_elem_rank = rank;
        qsort(list, n, sizeof(int), cmp_elem);  /* equal keys are the same element */
        free(rank);
    }
    double* buf = s->mode == ORC_JACOBI ? (double*)malloc(sizeof(double) * 12 * (n > 0 ? n : 1)) : NULL;
    double* cbuf = s->mode == ORC_JACOBI ? (double*)malloc(sizeof(double) * 9 * (size_t)s->nc) : NULL;
    int* cact = s->mode == ORC_JACOBI ? (int*)malloc(sizeof(int) * (size_t)s->nc) : NULL;
    for (int it = 0; it < iters; it++) {
        for (size_t u = 0; u < n; u++) {
            const int e = list[u];
            const int* t = s->tets + 4 * e;
            const double* p[4] = { s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3] };
            const double w[4] = { s->w[t[0]], s->w[t[1]], s->w[t[2]], s->w[t[3]] };
            double dx[12];
            project_combined(p, w, s->Q + 9 * e, s->aD[e], s->aH[e], s->gam[e], s->ts, s->lam + 2 * e,
                             buf ? buf + 12 * u : dx);
            if (!buf)
                for (int i = 0; i < 4; i++)
                    for (int k = 0; k < 3; k++) s->x[3 * t[i] + k] += dx[3 * i + k];
        }
        for (int ci = 0; with_contacts && ci < s->nc; ci++) {
            const int c = buf ? ci : s->corder[ci];
            const int* f = s->faces + 3 * s->c_face[c];
            const double* p[3] = { s->x + 3 * f[0], s->x + 3 * f[1], s->x + 3 * f[2] };
            const double w[3] = { s->w[f[0]], s->w[f[1]], s->w[f[2]] };
            double dx[9];
            const int act = project_contact(p, w, s->c_p + 3 * c, s->c_n + 3 * c, s->c_bary + 3 * c, s->c_lam + c,
                                            buf ? cbuf + 9 * c : dx);
            if (buf) cact[c] = act;
            else if (act)
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++) s->x[3 * f[i] + k] += dx[3 * i + k];
        }
        if (buf) {
            for (size_t u = 0; u < n; u++) {
                const int* t = s->tets + 4 * list[u];
                for (int i = 0; i < 4; i++)
                    for (int k = 0; k < 3; k++) s->x[3 * t[i] + k] += buf[12 * u + 3 * i + k];
            }
            for (int c = 0; with_contacts && c < s->nc; c++) {
                if (!cact[c]) continue;
                const int* f = s->faces + 3 * s->c_face[c];
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++) s->x[3 * f[i] + k] += cbuf[9 * c + 3 * i + k];
            }
        }
    }
    free(buf);
    free(cbuf);
    free(cact);
}

/* XPBDSolver::solve(projector_references, num_iter, initialize)
 * (include/solver/xpbd_solver/XPBDSolver.hpp:106-125) on a list of n elements
 * (repeats allowed), followed by every contact when with_contacts: with
 * initialize the listed projectors' multipliers (and the contacts') start at
 * zero, then num_iter passes (project_list).  Positions change in place; no
 * velocity update and no pin reset happen here. */
ORC_API void orc_solve_elements(orc_sim* s, int n, const int* elements, int num_iter, int initialize,
                                int with_contacts)
{
    int* list = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    memcpy(list, elements, sizeof(int) * (size_t)(n > 0 ? n : 0));
    if (initialize) {
        for (int u = 0; u < n; u++) s->lam[2 * list[u]] = s->lam[2 * list[u] + 1] = 0;
        for (int c = 0; with_contacts && c < s->nc; c++) s->c_lam[c] = 0;
    }
    project_list(s, list, (size_t)n, num_iter, with_contacts);
    free(list);
}

ORC_API void orc_set_local_collision_iters(orc_sim* s, int n) { s->local_iters = n > 0 ? n : 0; }
ORC_API int orc_last_local_list_size(const orc_sim* s) { return s->last_local_n; }
ORC_API void orc_get_contact_lambdas(const orc_sim* s, double* out)
{
    memcpy(out, s->c_lam, sizeof(double) * (size_t)s->nc);
}

/* Number of element projections in one local pass (the gathered list length,
 * repeats included), counted by brute force over all elements, for checks of
 * the gather. */
ORC_API int orc_local_list_size(const orc_sim* s)
{
    int n = 0;
    for (int c = 0; c < s->nc; c++) {
        if (s->c_lam[c] == 0) continue;
        const int* f = s->faces + 3 * s->c_face[c];
        for (int i = 0; i < 3; i++) {
            int cnt = 0;
            for (int e = 0; e < s->nt; e++)
                for (int j = 0; j < 4; j++) cnt += s->tets[4 * e + j] == f[i];
            n += cnt;
        }
    }
    return n;
}

/* The per-object part of one time step, with contacts optionally re-detected
 * first (Simulation.cpp:202-241 runs detection before update()):
 *   XPBDMeshObject.cpp:426-444  update: x_prev = x, predict, project, pin reset
 *   XPBDMeshObject.cpp:459-469  predict (2nd order): x += dt v ; z -= g dt^2
 *   XPBDSolver.hpp:66-97        solve: lambda = 0, iters x iterate
 *   XPBDMeshObject.cpp:497-503  pinned vertices reset to x_prev
 *   XPBDMeshObject.cpp:507-535  v = (x - x_prev) / dt                       */
ORC_API double orc_residual(const orc_sim* s, double* vec);

static void record_primary(orc_sim* s);

ORC_API void orc_step(orc_sim* s, int detect)
{
    if (detect) orc_detect(s);
    const int nv = s->nv;
    const double dt = s->dt;
    move_inertially(s);
    memset(s->lam, 0, sizeof(double) * 2 * (size_t)s->nt);
    for (int a = 0; a < s->na; a++) s->att_lam[a] = 0;
    for (int c = 0; c < s->nc; c++) s->c_lam[c] = 0;
    for (int it = 0; it < s->iters; it++) iterate(s);
    /* EVERY_SUBSTEP residual is taken here, at the end of the global solve,
     * before the local collision iterations and the pin reset (XPBDSolver.hpp:91-96,
     * XPBDMeshObject.cpp:473-503) */
    s->last_res = orc_residual(s, NULL); record_primary(s);
    local_collision_iterations(s);
    for (int i = 0; i < nv; i++)
        if (s->pin[i])
            for (int k = 0; k < 3; k++) s->x[3 * i + k] = s->xp[3 * i + k];
    for (int i = 0; i < 3 * nv; i++) s->v[i] = (s->x[i] - s->xp[i]) / dt;
}

/* Warm-started step: like orc_step, but corr (3*nv, may be NULL) is added to
 * the predicted positions before the solve.  This is the reference semantics
 * of the learned step (the network supplies corr); with corr = NULL it is
 * exactly orc_step. */
ORC_API void orc_step_corrected(orc_sim* s, int detect, const double* corr, int iters)
{
    if (detect) orc_detect(s);
    const int nv = s->nv;
    const double dt = s->dt;
    memcpy(s->xp, s->x, sizeof(double) * 3 * nv);
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) s->x[3 * i + k] += dt * s->v[3 * i + k];
    const double dz = -s->g * dt * dt;
    for (int i = 0; i < nv; i++) s->x[3 * i + 2] += dz;
    keep_prediction(s);
    if (corr)
        for (int i = 0; i < 3 * nv; i++) s->x[i] += corr[i];
    memset(s->lam, 0, sizeof(double) * 2 * (size_t)s->nt);
    for (int c = 0; c < s->nc; c++) s->c_lam[c] = 0;
    for (int it = 0; it < iters; it++) iterate(s);
    s->last_res = orc_residual(s, NULL); record_primary(s);
    local_collision_iterations(s);
    for (int i = 0; i < nv; i++)
        if (s->pin[i])
            for (int k = 0; k < 3; k++) s->x[3 * i + k] = s->xp[3 * i + k];
    for (int i = 0; i < 3 * nv; i++) s->v[i] = (s->x[i] - s->xp[i]) / dt;
}

/* Dual warm-started step: the multipliers start at lam0 (2*nt) instead of 0,
 * and the predicted positions are moved to the consistent start
 *     x0 = x~ + M^-1 sum_e grad C_e(x~)^T lam0_e
 * so that the momentum equation M (x - x~) = grad C^T lambda, which the XPBD
 * update preserves, holds at the first iteration.  The fixed point of the
 * iteration is unchanged; only the starting point moves.  lam0 = NULL gives
 * exactly orc_step. */
ORC_API void orc_step_dual_warm(orc_sim* s, int detect, const double* lam0, int iters)
{
    if (detect) orc_detect(s);
    const int nv = s->nv;
    const double dt = s->dt;
    move_inertially(s);
    if (lam0) {
        memcpy(s->lam, lam0, sizeof(double) * 2 * (size_t)s->nt);
        double* dx = (double*)calloc(3 * (size_t)nv, sizeof(double));
        for (int e = 0; e < s->nt; e++) {
            const int* t = s->tets + 4 * e;
            double x12[12], C[2], g[24];
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 3; k++) x12[3 * j + k] = s->x[3 * t[j] + k];
            orc_eval_tet(x12, s->Q + 9 * e, s->gam[e], C, g);
            for (int j = 0; j < 4; j++)
                for (int k = 0; k < 3; k++)
                    dx[3 * t[j] + k] += s->w[t[j]] * (g[3 * j + k] * lam0[2 * e] + g[12 + 3 * j + k] * lam0[2 * e + 1]);
        }
        for (int i = 0; i < 3 * nv; i++) s->x[i] += dx[i];
        free(dx);
    } else {
        memset(s->lam, 0, sizeof(double) * 2 * (size_t)s->nt);
    }
    for (int c = 0; c < s->nc; c++) s->c_lam[c] = 0;
    for (int a = 0; a < s->na; a++) s->att_lam[a] = 0;
    for (int it = 0; it < iters; it++) iterate(s);
    s->last_res = orc_residual(s, NULL); record_primary(s);
    local_collision_iterations(s);
    for (int i = 0; i < nv; i++)
        if (s->pin[i])
            for (int k = 0; k < 3; k++) s->x[3 * i + k] = s->xp[3 * i + k];
    for (int i = 0; i < 3 * nv; i++) s->v[i] = (s->x[i] - s->xp[i]) / dt;
}

/* Learned-warm-start step (the reference semantics of the device's learned
 * step).  Given a predicted position correction corr (3*nv; NULL = zero), the
 * solve starts from the multipliers that satisfy the constraint equations at
 * the predicted point xl = x~ + corr,
 *     lam0_e = -C_e(xl) / alpha~_e,
 * and from the positions that satisfy the momentum equation for them,
 *     x0 = x~ + M^-1 sum_e grad C_e(xl)^T lam0_e.
 * If corr is the exact correction this start is the solution of the step; for
 * any corr the start obeys M (x0 - x~) = grad C^T lam0, which the XPBD updates
 * preserve, so the iteration converges to the same fixed point as a cold start. */
ORC_API void orc_step_learned(orc_sim* s, int detect, const double* corr, int iters)
{
    if (detect) orc_detect(s);
    const int nv = s->nv;
    const double dt = s->dt, dt2 = s->ts;   /* alpha~ = alpha / dt^2, or alpha / dt for first order */
    move_inertially(s);
    double* dx = (double*)calloc(3 * (size_t)nv, sizeof(double));
    for (int e = 0; e < s->nt; e++) {
        const int* t = s->tets + 4 * e;
        double x12[12], C[2], g[24];
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++)
                x12[3 * j + k] = s->x[3 * t[j] + k] + (corr ? corr[3 * t[j] + k] : 0.0);
        orc_eval_tet(x12, s->Q + 9 * e, s->gam[e], C, g);
        const double l0 = -C[0] / (s->aD[e] / dt2), l1 = -C[1] / (s->aH[e] / dt2);
        s->lam[2 * e] = l0;
        s->lam[2 * e + 1] = l1;
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) dx[3 * t[j] + k] += s->w[t[j]] * (g[3 * j + k] * l0 + g[12 + 3 * j + k] * l1);
    }
    for (int i = 0; i < 3 * nv; i++) s->x[i] += dx[i];
    free(dx);
    for (int c = 0; c < s->nc; c++) s->c_lam[c] = 0;
    for (int a = 0; a < s->na; a++) s->att_lam[a] = 0;
    for (int it = 0; it < iters; it++) iterate(s);
    s->last_res = orc_residual(s, NULL); record_primary(s);
    local_collision_iterations(s);
    for (int i = 0; i < nv; i++)
        if (s->pin[i])
            for (int k = 0; k < 3; k++) s->x[3 * i + k] = s->xp[3 * i + k];
    for (int i = 0; i < 3 * nv; i++) s->v[i] = (s->x[i] - s->xp[i]) / dt;
}

/* Prediction half of a step only (used to build learned-warm-start features). */
ORC_API void orc_predict(orc_sim* s) { move_inertially(s); }

ORC_API void orc_get_x(const orc_sim* s, double* out) { memcpy(out, s->x, sizeof(double) * 3 * s->nv); }
ORC_API void orc_get_xprev(const orc_sim* s, double* out) { memcpy(out, s->xp, sizeof(double) * 3 * s->nv); }
ORC_API void orc_get_v(const orc_sim* s, double* out) { memcpy(out, s->v, sizeof(double) * 3 * s->nv); }
ORC_API void orc_get_lambda(const orc_sim* s, double* out) { memcpy(out, s->lam, sizeof(double) * 2 * s->nt); }
ORC_API void orc_set_lambdas(orc_sim* s, const double* lam) { memcpy(s->lam, lam, sizeof(double) * 2 * s->nt); }
ORC_API void orc_get_masses(const orc_sim* s, double* out) { memcpy(out, s->m, sizeof(double) * s->nv); }
ORC_API void orc_get_rest(const orc_sim* s, double* Q, double* vol)
{
    if (Q) memcpy(Q, s->Q, sizeof(double) * 9 * s->nt);
    if (vol) memcpy(vol, s->vol, sizeof(double) * s->nt);
}
ORC_API void orc_get_compliance(const orc_sim* s, double* aD, double* aH, double* gam)
{
    if (aD) memcpy(aD, s->aD, sizeof(double) * s->nt);
    if (aH) memcpy(aH, s->aH, sizeof(double) * s->nt);
    if (gam) memcpy(gam, s->gam, sizeof(double) * s->nt);
}

/* a11 constraint residual, projector order (Dev, Hyd per tet):
 *     include/solver/xpbd_solver/XPBDSolver.hpp:327-391.  Inequality (contact)
 *     constraints are skipped there, so only elastic entries appear.
 *     Returns ||r||_2; vec (2*nt, optional) receives r. */
ORC_API double orc_residual(const orc_sim* s, double* vec)
{
    const double dt2 = s->ts;  /* dt^2, or dt for first order */
    double acc = 0;
    for (int e = 0; e < s->nt; e++) {
        const int* t = s->tets + 4 * e;
        double F[9];
        compute_F(s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3], s->Q + 9 * e, F);
        const double C1 = dev_eval(F), C2 = hyd_eval(F, s->gam[e]);
        const double r1 = C1 + s->aD[e] / dt2 * s->lam[2 * e];
        const double r2 = C2 + s->aH[e] / dt2 * s->lam[2 * e + 1];
        if (vec) { vec[2 * e] = r1; vec[2 * e + 1] = r2; }
        acc += r1 * r1 + r2 * r2;
    }
    return sqrt(acc);
}

/* Primary residual (Equation (8) of the XPBD paper; the reference declares it
 * as XPBDSolver::primaryResidual, XPBDSolver.hpp:127-130, and leaves
 * _calculatePrimaryResidual empty, :322): for every free vertex i,
 *     R_i = (x_i - xt_i) / w_i - sum_c grad_i C_c(x)^T lambda_c
 * over the constraints c that hold multipliers: every element's deviatoric and
 * hydrostatic constraint, every contact (grad = barycentric * normal,
 * StaticDeformableCollisionConstraint.cpp:20-55).  xt is the
 * step's inertial prediction (3*nv, given), w the inverse inertia the
 * projections use (1/m, or the first-order damping's), lambda the
 * multipliers left by the solve (alpha~ lambda is the constraint force times
 * dt^2, XPBDSolver.hpp:66-97 with CombinedNeohookeanConstraintProjector.cpp:98-179).
 * Pinned and attached vertices are left out: a pin, and an attachment with
 * its zero compliance (AttachmentConstraint.hpp:33-88), hold the vertex where
 * it must be, so their force is whatever the balance needs (the gradient
 * (x - p) / |x - p| of a met attachment has no direction to test either).
 * Returns ||R||_2; vec (3*nv, optional) receives R (zero on the vertices left
 * out).
 * Gauss-Seidel XPBD keeps R = 0 only up to the change of the gradients along
 * the iterates; a start that sets lambda without moving x by M^-1 grad C^T
 * lambda violates it by the difference. */
ORC_API double orc_momentum_residual(const orc_sim* s, const double* xt, double* vec)
{
    const int nv = s->nv;
    double* R = (double*)calloc(3 * (size_t)nv, sizeof(double));
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) R[3 * i + k] = (s->x[3 * i + k] - xt[3 * i + k]) / s->w[i];
    for (int e = 0; e < s->nt; e++) {
        const int* t = s->tets + 4 * e;
        double x12[12], C[2], g[24];
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) x12[3 * j + k] = s->x[3 * t[j] + k];
        orc_eval_tet(x12, s->Q + 9 * e, s->gam[e], C, g);
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++)
                R[3 * t[j] + k] -= g[3 * j + k] * s->lam[2 * e] + g[12 + 3 * j + k] * s->lam[2 * e + 1];
    }
    for (int c = 0; c < s->nc; c++) {
        const int* f = s->faces + 3 * s->c_face[c];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) R[3 * f[i] + k] -= s->c_bary[3 * c + i] * s->c_n[3 * c + k] * s->c_lam[c];
    }
    for (int a = 0; a < s->na; a++)
        for (int k = 0; k < 3; k++) R[3 * s->att_v[a] + k] = 0;
    double acc = 0;
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) {
            if (s->pin[i]) R[3 * i + k] = 0;
            acc += R[3 * i + k] * R[3 * i + k];
        }
    if (vec) memcpy(vec, R, sizeof(double) * 3 * (size_t)nv);
    free(R);
    return sqrt(acc);
}

/* Record the primary residual where the reference's EVERY_SUBSTEP policy
 * evaluates its residuals: at the end of the global solve, before the local
 * collision iterations and the pin reset (XPBDSolver.hpp:91-96). */
static void record_primary(orc_sim* s)
{
    if (s->want_pres) s->last_pres = orc_momentum_residual(s, s->xt, s->pres);
}

ORC_API void orc_set_primary_residual(orc_sim* s, int on)
{
    s->want_pres = on != 0;
    if (s->want_pres && !s->xt) {
        s->xt = (double*)calloc(3 * (size_t)s->nv, sizeof(double));
        s->pres = (double*)calloc(3 * (size_t)s->nv, sizeof(double));
    }
}

/* the inertial prediction kept for the primary residual (3*nv) */
ORC_API void orc_last_prediction(const orc_sim* s, double* out)
{
    if (s->xt) memcpy(out, s->xt, sizeof(double) * 3 * (size_t)s->nv);
}

ORC_API double orc_last_primary_residual(const orc_sim* s, double* vec)
{
    if (vec && s->pres) memcpy(vec, s->pres, sizeof(double) * 3 * (size_t)s->nv);
    return s->last_pres;
}

/* One contact pass from the current positions, as in iterate(): Gauss-Seidel
 * applies each contact's correction at once (corder), Jacobi sums all of them
 * from the same positions afterwards. */
static void contact_pass(orc_sim* s)
{
    if (s->nc <= 0) return;
    double* buf = s->mode == ORC_JACOBI ? (double*)calloc(9 * (size_t)s->nc, sizeof(double)) : NULL;
    for (int ci = 0; ci < s->nc; ci++) {
        const int c = buf ? ci : s->corder[ci];
        const int* f = s->faces + 3 * s->c_face[c];
        const double* p[3] = { s->x + 3 * f[0], s->x + 3 * f[1], s->x + 3 * f[2] };
        const double w[3] = { s->w[f[0]], s->w[f[1]], s->w[f[2]] };
        double dx[9];
        if (project_contact(p, w, s->c_p + 3 * c, s->c_n + 3 * c, s->c_bary + 3 * c, s->c_lam + c, buf ? buf + 9 * c : dx) &&
            !buf)
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) s->x[3 * f[i] + k] += dx[3 * i + k];
    }
    if (buf) {
        for (int c = 0; c < s->nc; c++) {
            const int* f = s->faces + 3 * s->c_face[c];
            for (int i = 0; i < 3; i++)
                for (int k = 0; k < 3; k++) s->x[3 * f[i] + k] += buf[9 * c + 3 * i + k];
        }
        free(buf);
    }
}

/* Multiplier warm start from predicted positions (the device's learned step):
 * the positions start at the prediction xl = xt + corr (corr: 3*nv, NULL =
 * zero; pinned vertices take corr as given), the contacts are projected once
 * there from zero multipliers (contact_pass), then every element's
 * multipliers are set to the values that satisfy its constraints,
 *     lambda0_e = -C_e(x) / alpha~_e,
 * attachment multipliers to 0, and iters iterations follow as in orc_step
 * without a reset.  The start has zero elastic residual.  Unlike
 * orc_step_learned it does not move the positions by M^-1 grad C^T lambda0, so
 * the momentum balance holds only as well as corr predicts the solved step
 * (see orc_momentum_residual). */
ORC_API void orc_step_warm(orc_sim* s, int detect, const double* corr, int iters)
{
    if (detect) orc_detect(s);
    const int nv = s->nv;
    const double dt = s->dt;
    move_inertially(s);
    if (corr)
        for (int i = 0; i < 3 * nv; i++) s->x[i] += corr[i];
    for (int c = 0; c < s->nc; c++) s->c_lam[c] = 0;
    contact_pass(s);
    for (int e = 0; e < s->nt; e++) {
        const int* t = s->tets + 4 * e;
        double F[9];
        compute_F(s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3], s->Q + 9 * e, F);
        s->lam[2 * e] = -dev_eval(F) / (s->aD[e] / s->ts);
        s->lam[2 * e + 1] = -hyd_eval(F, s->gam[e]) / (s->aH[e] / s->ts);
    }
    for (int a = 0; a < s->na; a++) s->att_lam[a] = 0;
    for (int it = 0; it < iters; it++) iterate(s);
    s->last_res = orc_residual(s, NULL); record_primary(s);
    local_collision_iterations(s);
    for (int i = 0; i < nv; i++)
        if (s->pin[i])
            for (int k = 0; k < 3; k++) s->x[3 * i + k] = s->xp[3 * i + k];
    for (int i = 0; i < 3 * nv; i++) s->v[i] = (s->x[i] - s->xp[i]) / dt;
}

/* a12 total strain energy: XPBDMeshObject.cpp:537-549 (all Dev, then all Hyd,
 *     sum of C^2 / alpha, no 1/2 factor). */
ORC_API double orc_energy(const orc_sim* s)
{
    double E = 0;
    for (int pass = 0; pass < 2; pass++)
        for (int e = 0; e < s->nt; e++) {
            const int* t = s->tets + 4 * e;
            double F[9];
            compute_F(s->x + 3 * t[0], s->x + 3 * t[1], s->x + 3 * t[2], s->x + 3 * t[3], s->Q + 9 * e, F);
            if (pass == 0) { const double C = dev_eval(F); E += C * C / s->aD[e]; }
            else { const double C = hyd_eval(F, s->gam[e]); E += C * C / s->aH[e]; }
        }
    return E;
}
