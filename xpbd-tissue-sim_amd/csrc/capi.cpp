// C-ABI of the XPBD solver: set-up (FP64 on the host), device buffers, the
// graph-captured time step, diagnostics and the batched path.
#include "xpbd_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.h"
#include "kernels.h"
#include "mesh_tools.h"

using namespace xpbd;

namespace {
thread_local std::string g_create_error;
// largest patch count for which a patch Jacobi handle starts on the fused
// one-launch iteration before any autotuning
constexpr int kFusedDefaultMaxPatches = 256;
}

namespace xpbd {

int set_err(xpbd_handle* h, int code, const std::string& msg)
{
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_err(xpbd_handle* h, hipError_t e, const char* what)
{
    return set_err(h, XPBD_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
static int dalloc(xpbd_handle* h, T** p, size_t n)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
    if (e != hipSuccess) return set_err(h, XPBD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    h->allocs.push_back(*p);
    return XPBD_OK;
}

static void release_graphs(xpbd_handle* h)
{
    for (int i = 0; i < 2; i++) {
        if (h->graph_exec[i]) (void)hipGraphExecDestroy(h->graph_exec[i]);
        h->graph_exec[i] = nullptr;
    }
    h->graph_gen++;
}

void invalidate_graphs(xpbd_handle* h) { release_graphs(h); }

TetArrays tet_arrays(xpbd_handle* h)
{
    TetArrays ta;
    ta.tet = h->d_tet;
    ta.Q = h->d_Q;
    ta.qstride = h->nt;
    ta.at = h->d_at;
    ta.gamma = h->d_gamma;
    ta.lam = h->d_lam;
    return ta;
}

// contacts are projected when a collider detects them or the caller gave them
bool contacts_enabled(const xpbd_handle* h) { return h->nf > 0 && (h->sdf_active || h->contacts_given); }

PatchArrays patch_arrays(xpbd_handle* h)
{
    PatchArrays pa;
    pa.tloc = reinterpret_cast<const uint2*>(h->d_ptloc);
    pa.pvoff = h->d_pvoff;
    pa.pverts = h->d_pverts;
    pa.pslot = h->d_pslot;
    pa.inc_off = h->d_inc_off;
    pa.inc = h->d_inc;
    pa.partial = h->d_partial;
    pa.shared = h->d_shared;
    pa.soff = h->d_soff;
    pa.npatch = h->patches.npatch;
    pa.maxl = h->patches.maxl;
    pa.nshared = (int)h->patches.shared.size();
    return pa;
}

}  // namespace xpbd

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
// Waits for the stream when it goes out of scope: host vectors that feed
// asynchronous copies stay alive until the copies have run, on every return path.
struct StreamDrain {
    hipStream_t s;
    ~StreamDrain() { (void)hipStreamSynchronize(s); }
};

static bool finite_all(const double* p, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(p[i])) return false;
    return true;
}

static int validate_desc(const xpbd_mesh_desc* m)
{
    if (!m || m->num_vertices <= 0 || m->num_tets <= 0 || !m->x0 || !m->tets || !m->E || !m->nu || !m->rho)
        return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh: missing arrays or non-positive sizes");
    if (m->num_faces < 0 || (m->num_faces > 0 && !m->faces))
        return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh: bad face list");
    const size_t V = m->num_vertices, T = m->num_tets;
    for (size_t i = 0; i < 4 * T; i++)
        if (m->tets[i] < 0 || m->tets[i] >= m->num_vertices)
            return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh: tet vertex index out of range");
    for (size_t i = 0; i < 3 * (size_t)m->num_faces; i++)
        if (m->faces[i] < 0 || m->faces[i] >= m->num_vertices)
            return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh: face vertex index out of range");
    if (!finite_all(m->x0, 3 * V) || (m->v0 && !finite_all(m->v0, 3 * V)) || !finite_all(m->E, T) ||
        !finite_all(m->nu, T) || !finite_all(m->rho, T))
        return set_err(nullptr, XPBD_E_NONFINITE, "mesh: non-finite value in an input array");
    for (size_t e = 0; e < T; e++)
        if (!(m->E[e] > 0) || !(m->nu[e] > -1.0 && m->nu[e] < 0.5) || !(m->rho[e] > 0))
            return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh: material out of range (E>0, -1<nu<0.5, rho>0)");
    return XPBD_OK;
}

// Rest state of one element in FP64: Q = Dm^-1 (adjugate / det), V = |det|/6.
// include/solver/constraint/ElementConstraint.hpp:36-50
static bool rest_tet(const double* X0, const double* X1, const double* X2, const double* X3, double* Q,
                     double* vol)
{
    double D[9];
    for (int k = 0; k < 3; k++) {
        D[k] = X0[k] - X3[k];
        D[3 + k] = X1[k] - X3[k];
        D[6 + k] = X2[k] - X3[k];
    }
    auto d = [&](int r, int c) { return D[3 * c + r]; };
    const double c00 = d(1, 1) * d(2, 2) - d(1, 2) * d(2, 1);
    const double c01 = d(1, 2) * d(2, 0) - d(1, 0) * d(2, 2);
    const double c02 = d(1, 0) * d(2, 1) - d(1, 1) * d(2, 0);
    const double det = d(0, 0) * c00 + d(0, 1) * c01 + d(0, 2) * c02;
    if (!(std::fabs(det) > 0) || !std::isfinite(det)) return false;
    const double inv[9] = {
        c00 / det, c01 / det, c02 / det,                                        // column 0
        (d(0, 2) * d(2, 1) - d(0, 1) * d(2, 2)) / det,                          // (0,1)
        (d(0, 0) * d(2, 2) - d(0, 2) * d(2, 0)) / det,                          // (1,1)
        (d(0, 1) * d(2, 0) - d(0, 0) * d(2, 1)) / det,                          // (2,1)
        (d(0, 1) * d(1, 2) - d(0, 2) * d(1, 1)) / det,                          // (0,2)
        (d(0, 2) * d(1, 0) - d(0, 0) * d(1, 2)) / det,                          // (1,2)
        (d(0, 0) * d(1, 1) - d(0, 1) * d(1, 0)) / det,                          // (2,2)
    };
    std::memcpy(Q, inv, sizeof(inv));
    *vol = std::fabs(det / 6);
    return true;
}

static int build(int32_t n, const xpbd_mesh_desc* meshes, const xpbd_params* P, xpbd_handle** out)
{
    g_create_error.clear();
    if (!out) return set_err(nullptr, XPBD_E_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (n <= 0 || !meshes || !P) return set_err(nullptr, XPBD_E_INVALID_ARG, "need n > 0 meshes and params");
    if (!(P->dt > 0) || !std::isfinite(P->dt) || !std::isfinite(P->g) || P->iters < 0 ||
        (P->solver != XPBD_SOLVER_JACOBI && P->solver != XPBD_SOLVER_GS_COLORED) || P->collision_every < 0 ||
        (P->first_order != 0 && P->first_order != 1) ||
        (P->first_order && !(P->damping_multiplier > 0 && std::isfinite(P->damping_multiplier))) ||
        P->local_collision_iters < 0)
        return set_err(nullptr, XPBD_E_INVALID_ARG,
                       "params: need dt > 0, finite g, iters >= 0, local_collision_iters >= 0, known solver");
    for (int32_t i = 0; i < n; i++) {
        int rc = validate_desc(&meshes[i]);
        if (rc) return rc;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_err(nullptr, XPBD_E_HIP, "no HIP device");

    xpbd_handle* h = new (std::nothrow) xpbd_handle();
    if (!h) return set_err(nullptr, XPBD_E_NOMEM, "handle allocation failed");
    h->device = dev;
    h->params = *P;
    h->ninst = n;
    // ---- union of instances
    h->v_off.resize(n + 1);
    h->t_off.resize(n + 1);
    h->f_off.resize(n + 1);
    h->v_off[0] = h->t_off[0] = h->f_off[0] = 0;
    for (int32_t i = 0; i < n; i++) {
        h->v_off[i + 1] = h->v_off[i] + meshes[i].num_vertices;
        h->t_off[i + 1] = h->t_off[i] + meshes[i].num_tets;
        h->f_off[i + 1] = h->f_off[i] + meshes[i].num_faces;
    }
    if (h->v_off[n] > INT32_MAX / 4 || h->t_off[n] > INT32_MAX / 12) {
        delete h;
        return set_err(nullptr, XPBD_E_INVALID_ARG, "mesh too large for 32-bit indexing");
    }
    const int nv = h->nv = (int)h->v_off[n];
    const int nt = h->nt = (int)h->t_off[n];
    const int nf = h->nf = (int)h->f_off[n];
    h->rest.resize(3 * (size_t)nv);
    std::vector<double> vel(3 * (size_t)nv, 0.0);
    std::vector<uint8_t> pin(nv, 0);
    std::vector<int32_t> tets(4 * (size_t)nt), faces(4 * (size_t)(nf > 0 ? nf : 0));
    std::vector<double> E(nt), nu(nt), rho(nt);
    for (int32_t i = 0; i < n; i++) {
        const xpbd_mesh_desc& m = meshes[i];
        const size_t vo = h->v_off[i], to = h->t_off[i], fo = h->f_off[i];
        std::memcpy(&h->rest[3 * vo], m.x0, sizeof(double) * 3 * m.num_vertices);
        if (m.v0) std::memcpy(&vel[3 * vo], m.v0, sizeof(double) * 3 * m.num_vertices);
        if (m.pinned) std::memcpy(&pin[vo], m.pinned, m.num_vertices);
        for (size_t k = 0; k < 4 * (size_t)m.num_tets; k++) tets[4 * to + k] = m.tets[k] + (int32_t)vo;
        std::memcpy(&E[to], m.E, sizeof(double) * m.num_tets);
        std::memcpy(&nu[to], m.nu, sizeof(double) * m.num_tets);
        std::memcpy(&rho[to], m.rho, sizeof(double) * m.num_tets);
        for (int32_t f = 0; f < m.num_faces; f++) {
            faces[4 * (fo + f) + 0] = m.faces[3 * f + 0] + (int32_t)vo;
            faces[4 * (fo + f) + 1] = m.faces[3 * f + 1] + (int32_t)vo;
            faces[4 * (fo + f) + 2] = m.faces[3 * f + 2] + (int32_t)vo;
            faces[4 * (fo + f) + 3] = i;
        }
    }
    h->pinned = pin;
    // ---- FP64 rest-state precompute
    std::vector<double> Q(9 * (size_t)nt), vol(nt);
    for (int e = 0; e < nt; e++) {
        const int32_t* t = &tets[4 * (size_t)e];
        if (!rest_tet(&h->rest[3 * (size_t)t[0]], &h->rest[3 * (size_t)t[1]], &h->rest[3 * (size_t)t[2]],
                      &h->rest[3 * (size_t)t[3]], &Q[9 * (size_t)e], &vol[e])) {
            delete h;
            return set_err(nullptr, XPBD_E_DEGENERATE, "element " + std::to_string(e) + " has zero volume");
        }
    }
    // lumped masses in element order (XPBDMeshObject.cpp:288-313)
    h->mass.assign(nv, 0.0);
    for (int e = 0; e < nt; e++) {
        const double em = vol[e] * rho[e];
        for (int j = 0; j < 4; j++) h->mass[tets[4 * (size_t)e + j]] += em / 4.0;
    }
    // per-vertex material summary (learned warm-start features)
    h->vmat.assign(2 * (size_t)nv, 0.f);
    {
        std::vector<double> sE(nv, 0.0), sN(nv, 0.0);
        std::vector<int32_t> cnt(nv, 0);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) {
                const int32_t v = tets[4 * (size_t)e + j];
                sE[v] += std::log10(E[e]);
                sN[v] += nu[e];
                cnt[v]++;
            }
        for (int v = 0; v < nv; v++)
            if (cnt[v]) {
                h->vmat[2 * (size_t)v] = (float)(sE[v] / cnt[v]);
                h->vmat[2 * (size_t)v + 1] = (float)(sN[v] / cnt[v]);
            }
    }
    h->tets_host = tets;
    // constraint inertia: the lumped mass, or for first-order objects the vertex
    // damping B_i = V_i b [E_i / (1 + nu_i)] with V_i = sum vol / 4 and E_i, nu_i
    // the means over attached elements (XPBDMeshObject.cpp:277-348,
    // XPBDMeshObjectBase.hpp:125-136); first order also moves z by -g m dt / B
    h->inertia = h->mass;
    if (P->first_order) {
        std::vector<double> V(nv, 0.0), sE(nv, 0.0), sN(nv, 0.0);
        std::vector<int32_t> cnt(nv, 0);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) cnt[tets[4 * (size_t)e + j]]++;
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) {
                const int32_t v = tets[4 * (size_t)e + j];
                V[v] += vol[e] / 4.0;
                sE[v] += E[e] / cnt[v];
                sN[v] += nu[e] / cnt[v];
            }
        h->dz1.assign(nv, 0.f);
        for (int v = 0; v < nv; v++) {
            const double B = P->adjust_damping ? V[v] * P->damping_multiplier * sE[v] / (1 + sN[v])
                                               : V[v] * P->damping_multiplier;
            h->inertia[v] = B;
            h->dz1[v] = (float)(-P->g * h->mass[v] * P->dt / B);
        }
    }
    // compliances (DeviatoricConstraint.cpp:13, HydrostaticConstraint.cpp:15-16,
    // ElasticMaterial.hpp:31-32) and alpha tilde = alpha / dt^2 (first order: alpha / dt)
    const double dt2 = P->first_order ? P->dt : P->dt * P->dt;
    std::vector<float> atD(nt), atH(nt), gam(nt);
    for (int e = 0; e < nt; e++) {
        const double mu = E[e] / (2 * (1 + nu[e]));
        const double lm = (E[e] * nu[e]) / ((1 + nu[e]) * (1 - 2 * nu[e]));
        atD[e] = (float)(1 / (mu * vol[e]) / dt2);
        atH[e] = (float)(1 / (lm * vol[e]) / dt2);
        gam[e] = (float)(mu / lm);
    }
    // ---- solver order of elements
    h->order.resize(nt);
    int variant = P->variant;
    // untuned default: the quad cluster kernel (faster than one thread per
    // cluster on every measured workload) or patch Jacobi
    if (variant == XPBD_VARIANT_AUTO)
        variant = P->solver == XPBD_SOLVER_GS_COLORED ? XPBD_VARIANT_CLUSTER_QUAD : XPBD_VARIANT_PATCH;
    const bool cluster = variant == XPBD_VARIANT_CLUSTER || variant == XPBD_VARIANT_CLUSTER_QUAD;
    if ((P->solver == XPBD_SOLVER_GS_COLORED && variant == XPBD_VARIANT_PATCH) ||
        (P->solver == XPBD_SOLVER_JACOBI && cluster) || variant < 0 || variant > XPBD_VARIANT_CLUSTER_QUAD) {
        delete h;
        return set_err(nullptr, XPBD_E_INVALID_ARG, "variant does not apply to this solver");
    }
    h->variant = variant;
    std::vector<uint32_t> tloc_solver;
    std::vector<int32_t> cverts_solver;
    if (P->solver == XPBD_SOLVER_GS_COLORED && cluster) {
        // clusters of consecutive tets, coloured; solver order = colour, cluster, tet
        std::vector<int32_t> cst, cv, ccol;
        std::vector<uint32_t> tl;
        const int ncl = build_clusters(nt, tets.data(), cst, cv, tl);
        h->cluster_max_tets = 1;
        for (int c = 0; c < ncl; c++) h->cluster_max_tets = std::max(h->cluster_max_tets, cst[c + 1] - cst[c]);
        const int nc = greedy_color(nv, ncl, kClusterVerts, cv.data(), ccol);
        if (nc <= 0) {
            delete h;
            return set_err(nullptr, XPBD_E_INVALID_ARG, "cluster colouring failed");
        }
        std::vector<int32_t> cnt(nc + 1, 0);
        for (int c = 0; c < ncl; c++) cnt[ccol[c] + 1]++;
        for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
        h->color_off.assign(cnt.begin(), cnt.end());  // offsets over clusters
        std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1), corder(ncl);
        for (int c = 0; c < ncl; c++) corder[pos[ccol[c]]++] = c;
        h->cl_start.assign(1, 0);
        cverts_solver.reserve((size_t)kClusterVerts * ncl);
        tloc_solver.reserve(nt);
        int at_pos = 0;
        for (int i = 0; i < ncl; i++) {
            const int c = corder[i];
            for (int e = cst[c]; e < cst[c + 1]; e++) {
                h->order[at_pos++] = e;
                tloc_solver.push_back(tl[e]);
            }
            h->cl_start.push_back(at_pos);
            for (int j = 0; j < kClusterVerts; j++) cverts_solver.push_back(cv[(size_t)kClusterVerts * c + j]);
        }
        h->ncolors = nc;
        h->nclusters = ncl;
    } else if (P->solver == XPBD_SOLVER_GS_COLORED) {
        std::vector<int32_t> col;
        int nc = greedy_color(nv, nt, 4, tets.data(), col);
        if (nc <= 0) {
            delete h;
            return set_err(nullptr, XPBD_E_INVALID_ARG, "tet colouring failed");
        }
        std::vector<int32_t> cnt(nc + 1, 0);
        for (int e = 0; e < nt; e++) cnt[col[e] + 1]++;
        for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
        h->color_off.assign(cnt.begin(), cnt.end());
        std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < nt; e++) h->order[pos[col[e]]++] = e;
        h->ncolors = nc;
    } else if (variant == XPBD_VARIANT_PATCH) {
        std::vector<int32_t> tinst(nt);
        for (int32_t i = 0; i < n; i++)
            for (int64_t e = h->t_off[i]; e < h->t_off[i + 1]; e++) tinst[e] = i;
        const int maxl = build_patches(nv, nt, tets.data(), h->rest.data(), tinst.data(), jacobi_patch_size(),
                                       h->patches);
        if (maxl < 0 || jacobi_patch_smem(maxl, true) > 64 * 1024) {
            // pathological vertex spread: fall back to one thread per tet
            h->variant = XPBD_VARIANT_SIMPLE;
            std::iota(h->order.begin(), h->order.end(), 0);
        } else {
            h->order = h->patches.order;
        }
        h->color_off = { 0, nt };
        h->ncolors = 1;
    } else {
        std::iota(h->order.begin(), h->order.end(), 0);
        h->color_off = { 0, nt };
        h->ncolors = 1;
    }
    // ---- contact face order: colour-sorted so contacts of one colour share no vertex
    h->fperm.resize(nf);
    h->fcol_off = { 0, nf };
    if (nf > 0) {
        std::vector<int32_t> f3(3 * (size_t)nf), col;
        for (int f = 0; f < nf; f++)
            for (int j = 0; j < 3; j++) f3[3 * (size_t)f + j] = faces[4 * (size_t)f + j];
        int nc = greedy_color(nv, nf, 3, f3.data(), col);
        if (nc <= 0) {
            delete h;
            return set_err(nullptr, XPBD_E_INVALID_ARG, "face colouring failed");
        }
        std::vector<int32_t> cnt(nc + 1, 0);
        for (int f = 0; f < nf; f++) cnt[col[f] + 1]++;
        for (int c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
        h->fcol_off.assign(cnt.begin(), cnt.end());
        std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
        for (int f = 0; f < nf; f++) h->fperm[pos[col[f]]++] = f;
    }
    h->nfcolors = (int)h->fcol_off.size() - 1;
    h->fci_off.assign((size_t)h->nfcolors * (h->ninst + 1), 0);
    for (int c = 0; c < h->nfcolors; c++) {
        int j = h->fcol_off[c];
        for (int i = 0; i <= h->ninst; i++) {
            while (j < h->fcol_off[c + 1] && faces[4 * (size_t)h->fperm[j] + 3] < i) j++;
            h->fci_off[(size_t)c * (h->ninst + 1) + i] = j;
        }
    }

    // ---- device buffers
    hipError_t he = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (he != hipSuccess) {
        int rc = hip_err(nullptr, he, "hipStreamCreate");
        delete h;
        return rc;
    }
    int rc = 0;
    rc |= dalloc(h, &h->d_pos, 2 * (size_t)nv);
    rc |= dalloc(h, &h->d_lo, nv);
    rc |= dalloc(h, &h->d_vel, nv);
    if (P->first_order) rc |= dalloc(h, &h->d_dz1, nv);
    rc |= dalloc(h, &h->d_step_res, 1);
    rc |= dalloc(h, &h->d_acc, 4 * (size_t)nv);
    rc |= dalloc(h, &h->d_rest, 3 * (size_t)nv);
    rc |= dalloc(h, &h->d_tet, nt);
    rc |= dalloc(h, &h->d_Q, 9 * (size_t)nt);
    rc |= dalloc(h, &h->d_at, nt);
    rc |= dalloc(h, &h->d_gamma, nt);
    rc |= dalloc(h, &h->d_lam, nt);
    if (h->variant == XPBD_VARIANT_PATCH) {
        const Patches& Pt = h->patches;
        rc |= dalloc(h, &h->d_ptloc, Pt.tloc.size());  // 2 words (4 x 16-bit slots) per tet
        rc |= dalloc(h, &h->d_pvoff, Pt.pvoff.size());
        rc |= dalloc(h, &h->d_pverts, Pt.pverts.size());
        rc |= dalloc(h, &h->d_pslot, Pt.pslot.size());
        rc |= dalloc(h, &h->d_inc_off, Pt.inc_off.size());
        rc |= dalloc(h, &h->d_inc, Pt.inc.size());
        rc |= dalloc(h, &h->d_partial, 2 * (size_t)Pt.nslots);
        rc |= dalloc(h, &h->d_shared, Pt.shared.size());
        rc |= dalloc(h, &h->d_soff, Pt.soff.size());
        rc |= dalloc(h, &h->d_pshare, Pt.pverts.size());
        rc |= dalloc(h, &h->d_sbuf, 2 * Pt.shared.size());
    }
    if (h->nclusters > 0) {
        rc |= dalloc(h, &h->d_cverts, 2 * (size_t)h->nclusters);
        rc |= dalloc(h, &h->d_dsep, 2 * (size_t)nv);   // deltas, then bases
        rc |= dalloc(h, &h->d_cstart, (size_t)h->nclusters + 1);
        rc |= dalloc(h, &h->d_tloc, nt);
    }
    h->red_blocks = std::max(1, std::min(1024, (nt + 255) / 256));
    rc |= dalloc(h, &h->d_part, h->red_blocks);
    rc |= dalloc(h, &h->d_scalar, 2);
    rc |= dalloc(h, &h->d_step, 1);
    rc |= dalloc(h, &h->d_sdf, n);
    rc |= dalloc(h, &h->d_faces, nf);
    rc |= dalloc(h, &h->d_fci_off, h->fci_off.size());
    rc |= dalloc(h, &h->d_flag, nf);
    rc |= dalloc(h, &h->d_cpos, (size_t)nf + 1);
    rc |= dalloc(h, &h->d_clist, nf);
    rc |= dalloc(h, &h->d_bsum, (size_t)compact_blocks(nf));
    rc |= dalloc(h, &h->d_cp, 3 * (size_t)nf);
    rc |= dalloc(h, &h->d_cn, 3 * (size_t)nf);
    rc |= dalloc(h, &h->d_cb, 3 * (size_t)nf);
    rc |= dalloc(h, &h->d_clam, nf);
    // deterministic patch Jacobi: contact corrections per face slot, gathered per
    // vertex over (face slot * 3 + corner) in ascending face slot
    std::vector<int32_t> vf_off, vf;
    if (P->solver == XPBD_SOLVER_JACOBI && nf > 0) {
        vf_off.assign((size_t)nv + 1, 0);
        vf.resize(3 * (size_t)nf);
        // solver face slot j holds caller face fperm[j] (4 ints per face: a, b, c, instance)
        auto corner = [&](int j, int c) { return faces[4 * (size_t)h->fperm[j] + c]; };
        for (int j = 0; j < nf; j++)
            for (int c = 0; c < 3; c++) vf_off[corner(j, c) + 1]++;
        for (int v = 0; v < nv; v++) vf_off[v + 1] += vf_off[v];
        std::vector<int32_t> fill(vf_off.begin(), vf_off.end() - 1);
        for (int j = 0; j < nf; j++)
            for (int c = 0; c < 3; c++) vf[fill[corner(j, c)]++] = 3 * j + c;
        rc |= dalloc(h, &h->d_cdx, 9 * (size_t)nf);
        rc |= dalloc(h, &h->d_vf_off, vf_off.size());
        rc |= dalloc(h, &h->d_vf, vf.size());
        if (P->local_collision_iters > 0) rc |= dalloc(h, &h->d_edx, 12 * (size_t)nt);
    }
    // local collision iterations: elements attached to each vertex, as solver positions
    std::vector<int32_t> vt_off, vt_pos;
    if (P->local_collision_iters > 0 && nf > 0) {
        vt_off.assign((size_t)nv + 1, 0);
        vt_pos.resize(4 * (size_t)nt);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) vt_off[tets[4 * (size_t)e + j] + 1]++;
        for (int v = 0; v < nv; v++) vt_off[v + 1] += vt_off[v];
        std::vector<int32_t> fill(vt_off.begin(), vt_off.end() - 1);
        for (int i = 0; i < nt; i++) {
            const int e = h->order[i];
            for (int j = 0; j < 4; j++) vt_pos[fill[tets[4 * (size_t)e + j]]++] = i;
        }
        rc |= dalloc(h, &h->d_vt_off, vt_off.size());
        rc |= dalloc(h, &h->d_vt_pos, vt_pos.size());
        rc |= dalloc(h, &h->d_mult, nt);
    }
    if (rc) {
        std::string msg = h->err;
        xpbd_destroy(h);
        return set_err(nullptr, XPBD_E_NOMEM, msg);
    }
    // host staging in solver order
    std::vector<float4> hu(2 * (size_t)nv), hv(nv);
    for (int i = 0; i < nv; i++) {
        const double w = h->inertia[i] > 0 ? 1.0 / h->inertia[i] : 0.0;  // PositionReference.hpp:30-33
        hu[2 * (size_t)i] = make_float4(0.f, 0.f, 0.f, (float)w);
        hu[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        hv[i] = make_float4((float)vel[3 * (size_t)i], (float)vel[3 * (size_t)i + 1], (float)vel[3 * (size_t)i + 2],
                            pin[i] ? 1.f : 0.f);
    }
    std::vector<int4> ht(nt);
    std::vector<float> hQ(9 * (size_t)nt);
    std::vector<float2> hat(nt);
    std::vector<float> hg(nt);
    for (int i = 0; i < nt; i++) {
        const int e = h->order[i];
        ht[i] = make_int4(tets[4 * (size_t)e], tets[4 * (size_t)e + 1], tets[4 * (size_t)e + 2], tets[4 * (size_t)e + 3]);
        for (int k = 0; k < 9; k++) hQ[(size_t)k * nt + i] = (float)Q[9 * (size_t)e + k];
        hat[i] = make_float2(atD[e], atH[e]);
        hg[i] = gam[e];
    }
    std::vector<int4> hf(nf);
    for (int j = 0; j < nf; j++) {
        const int f = h->fperm[j];
        hf[j] = make_int4(faces[4 * (size_t)f], faces[4 * (size_t)f + 1], faces[4 * (size_t)f + 2], faces[4 * (size_t)f + 3]);
    }
    // fused patch Jacobi: every face goes to the patch of an element that has it,
    // with the local slots of its three vertices there
    std::vector<int32_t> pf_off;
    std::vector<int4> pfv;
    if (h->variant == XPBD_VARIANT_PATCH) {
        const Patches& Pt = h->patches;
        bool ok = Pt.fusable && jacobi_fused_smem(Pt.maxl, true) <= 64 * 1024;
        const int psz = jacobi_patch_size();
        pf_off.assign((size_t)Pt.npatch + 1, 0);
        std::vector<int32_t> fpatch(nf, -1);
        std::vector<int4> floc(nf);
        if (ok && nf > 0) {
            std::vector<int32_t> at(nt);  // element -> solver position
            for (int i = 0; i < nt; i++) at[h->order[i]] = i;
            std::vector<int32_t> voff((size_t)nv + 1, 0), vel(4 * (size_t)nt);
            for (int e = 0; e < nt; e++)
                for (int k = 0; k < 4; k++) voff[tets[4 * (size_t)e + k] + 1]++;
            for (int v = 0; v < nv; v++) voff[v + 1] += voff[v];
            std::vector<int32_t> fill(voff.begin(), voff.end() - 1);
            for (int e = 0; e < nt; e++)
                for (int k = 0; k < 4; k++) vel[fill[tets[4 * (size_t)e + k]]++] = e;
            for (int j = 0; j < nf && ok; j++) {
                const int a = hf[j].x, b = hf[j].y, c = hf[j].z;
                int found = -1;
                for (int k = voff[a]; k < voff[a + 1] && found < 0; k++) {
                    const int32_t* t = &tets[4 * (size_t)vel[k]];
                    int hits = 0;
                    for (int m = 0; m < 4; m++) hits += (t[m] == b) + (t[m] == c);
                    if (hits == 2) found = vel[k];
                }
                if (found < 0) {
                    ok = false;  // a face that is no element's face: contacts stay in their own pass
                    break;
                }
                const int p = at[found] / psz;
                auto lb_ = Pt.pverts.begin() + Pt.pvoff[p], le_ = Pt.pverts.begin() + Pt.pvoff[p + 1];
                auto slot = [&](int v) { return (int)(std::lower_bound(lb_, le_, v) - lb_); };
                fpatch[j] = p;
                floc[j] = make_int4(j, slot(a), slot(b), slot(c));
                pf_off[p + 1]++;
            }
        }
        if (ok) {
            for (int p = 0; p < Pt.npatch; p++) pf_off[p + 1] += pf_off[p];
            pfv.resize(nf);
            std::vector<int32_t> fill(pf_off.begin(), pf_off.end() - 1);
            for (int j = 0; j < nf; j++) pfv[fill[fpatch[j]]++] = floc[j];  // face order within a patch
            int rc2 = dalloc(h, &h->d_pf_off, pf_off.size());
            if (!pfv.empty()) rc2 |= dalloc(h, &h->d_pf, pfv.size());
            if (rc2) ok = false;
        }
        h->fused_ok = ok;
        // untuned default: the one-launch iteration wins where launches dominate
        // (measured: 24 patches), the three-launch one on large meshes (1,875 and
        // 4,116 patches); xpbd_autotune times both
        h->patch_fused = ok && Pt.npatch <= kFusedDefaultMaxPatches;
    }
    h->sdf_host.assign(n, SdfDev{});
    for (auto& s : h->sdf_host) s.quat[3] = 1.0;
#define UP(dst, src, bytes)                                                      \
    do {                                                                         \
        hipError_t e_ = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);      \
        if (e_ != hipSuccess) {                                                  \
            int r_ = hip_err(nullptr, e_, "hipMemcpy");                          \
            xpbd_destroy(h);                                                     \
            return r_;                                                           \
        }                                                                        \
    } while (0)
    UP(h->d_pos, hu.data(), sizeof(float4) * 2 * (size_t)nv);
    UP(h->d_vel, hv.data(), sizeof(float4) * nv);
    if (P->first_order) UP(h->d_dz1, h->dz1.data(), sizeof(float) * nv);
    UP(h->d_rest, h->rest.data(), sizeof(double) * 3 * nv);
    UP(h->d_tet, ht.data(), sizeof(int4) * nt);
    UP(h->d_Q, hQ.data(), sizeof(float) * 9 * (size_t)nt);
    UP(h->d_at, hat.data(), sizeof(float2) * nt);
    UP(h->d_gamma, hg.data(), sizeof(float) * nt);
    UP(h->d_sdf, h->sdf_host.data(), sizeof(SdfDev) * n);
    if (h->variant == XPBD_VARIANT_PATCH) {
        const Patches& Pt = h->patches;
        UP(h->d_ptloc, Pt.tloc.data(), sizeof(uint32_t) * Pt.tloc.size());
        UP(h->d_pvoff, Pt.pvoff.data(), sizeof(int32_t) * Pt.pvoff.size());
        UP(h->d_pverts, Pt.pverts.data(), sizeof(int32_t) * Pt.pverts.size());
        UP(h->d_pslot, Pt.pslot.data(), sizeof(int32_t) * Pt.pslot.size());
        UP(h->d_inc_off, Pt.inc_off.data(), sizeof(int32_t) * Pt.inc_off.size());
        UP(h->d_inc, Pt.inc.data(), sizeof(uint16_t) * Pt.inc.size());
        if (!Pt.shared.empty()) {
            UP(h->d_shared, Pt.shared.data(), sizeof(int32_t) * Pt.shared.size());
        }
        UP(h->d_soff, Pt.soff.data(), sizeof(int32_t) * Pt.soff.size());
        UP(h->d_pshare, Pt.pshare.data(), sizeof(int32_t) * Pt.pshare.size());
        if (h->fused_ok) {
            UP(h->d_pf_off, pf_off.data(), sizeof(int32_t) * pf_off.size());
            if (!pfv.empty()) UP(h->d_pf, pfv.data(), sizeof(int4) * pfv.size());
        }
    }
    if (h->nclusters > 0) {
        UP(h->d_cverts, cverts_solver.data(), sizeof(int32_t) * cverts_solver.size());
        UP(h->d_cstart, h->cl_start.data(), sizeof(int32_t) * h->cl_start.size());
        UP(h->d_tloc, tloc_solver.data(), sizeof(uint32_t) * tloc_solver.size());
    }
    if (nf > 0) UP(h->d_faces, hf.data(), sizeof(int4) * nf);
    UP(h->d_fci_off, h->fci_off.data(), sizeof(int32_t) * h->fci_off.size());
    if (h->d_vf_off) {
        UP(h->d_vf_off, vf_off.data(), sizeof(int32_t) * vf_off.size());
        UP(h->d_vf, vf.data(), sizeof(int32_t) * vf.size());
    }
    if (h->d_vt_off) {
        UP(h->d_vt_off, vt_off.data(), sizeof(int32_t) * vt_off.size());
        UP(h->d_vt_pos, vt_pos.data(), sizeof(int32_t) * vt_pos.size());
    }
#undef UP
    // zero-initialised device state (fill kernels, then wait)
    hipError_t fe = hipSuccess;
    auto zero = [&](void* p, size_t bytes) {
        if (fe == hipSuccess) fe = launch_fill_zero(h->stream, p, bytes);
    };
    zero(h->d_acc, sizeof(float) * 4 * (size_t)nv);
    zero(h->d_lo, sizeof(float4) * (size_t)nv);
    zero(h->d_step_res, sizeof(double));
    zero(h->d_lam, sizeof(float2) * nt);
    zero(h->d_step, sizeof(unsigned long long));
    zero(h->d_cpos, sizeof(int32_t) * ((size_t)nf + 1));
    zero(h->d_clam, sizeof(double) * (nf > 0 ? nf : 1));
    zero(h->d_flag, sizeof(int32_t) * (nf > 0 ? nf : 1));   // no face in contact before the first detection
    he = fe != hipSuccess ? fe : hipStreamSynchronize(h->stream);
    if (he != hipSuccess) {
        int r = hip_err(nullptr, he, "initialisation");
        xpbd_destroy(h);
        return r;
    }
    *out = h;
    return XPBD_OK;
}

extern "C" int xpbd_create(const xpbd_mesh_desc* mesh, const xpbd_params* params, xpbd_handle** out)
{
    return build(1, mesh, params, out);
}

extern "C" int xpbd_batch_create(int32_t n, const xpbd_mesh_desc* meshes, const xpbd_params* params,
                                 xpbd_batch** out)
{
    return build(n, meshes, params, out);
}

extern "C" int xpbd_destroy(xpbd_handle* h)
{
    if (!h) return XPBD_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    release_graphs(h);
    for (auto& ev : h->events) {
        if (ev.first) (void)hipEventDestroy(ev.first);
        if (ev.second) (void)hipEventDestroy(ev.second);
    }
    for (void* p : h->allocs) (void)hipFree(p);
    destroy_learned(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return XPBD_OK;
}

extern "C" const char* xpbd_last_error(const xpbd_handle* h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

// ---------------------------------------------------------------------------
// time step
// ---------------------------------------------------------------------------
namespace xpbd {

// Patch Jacobi in its fixed-order mode: every accumulation (elements, contacts,
// attachments, local passes) is summed in an order fixed at set-up.
static bool jacobi_fixed_order(const xpbd_handle* h)
{
    return h->params.solver == XPBD_SOLVER_JACOBI && h->variant == XPBD_VARIANT_PATCH &&
           h->params.deterministic != 0 && h->d_cdx != nullptr;
}

// The vertex -> solver positions CSR of the elements attached to each vertex
// (the fixed-order Jacobi local pass and the primary residual gather over it)
// and, with slots, the per-element correction slots of the local pass; built
// at set-up when local collision passes are configured, else on first use.
static int ensure_vertex_elements(xpbd_handle* h, bool slots)
{
    const int nv = h->nv, nt = h->nt;
    if (!h->d_vt_off) {
        std::vector<int32_t> vt_off((size_t)nv + 1, 0), vt_pos(4 * (size_t)nt);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) vt_off[h->tets_host[4 * (size_t)e + j] + 1]++;
        for (int v = 0; v < nv; v++) vt_off[v + 1] += vt_off[v];
        std::vector<int32_t> fill(vt_off.begin(), vt_off.end() - 1);
        for (int i = 0; i < nt; i++) {
            const int e = h->order[i];
            for (int j = 0; j < 4; j++) vt_pos[fill[h->tets_host[4 * (size_t)e + j]]++] = i;
        }
        int32_t *off = nullptr, *pos = nullptr;
        int rc = dalloc(h, &off, vt_off.size());
        if (!rc) rc = dalloc(h, &pos, vt_pos.size());
        if (rc) return rc;
        hipError_t e = hipMemcpy(off, vt_off.data(), sizeof(int32_t) * vt_off.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(pos, vt_pos.data(), sizeof(int32_t) * vt_pos.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_err(h, e, "vertex-element lists");
        h->d_vt_off = off;
        h->d_vt_pos = pos;
    }
    if (slots && !h->d_edx) return dalloc(h, &h->d_edx, 12 * (size_t)nt);
    return XPBD_OK;
}

// The vertex -> (face slot * 3 + corner) CSR over the contact faces, in
// ascending face slot; built at set-up for Jacobi, else on first use (from
// the device's face list).
static int ensure_vertex_faces(xpbd_handle* h)
{
    const int nv = h->nv, nf = h->nf;
    if (h->d_vf_off || nf <= 0) return XPBD_OK;
    std::vector<int4> hf(nf);
    hipError_t e = hipMemcpy(hf.data(), h->d_faces, sizeof(int4) * nf, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(h, e, "vertex-face lists");
    std::vector<int32_t> vf_off((size_t)nv + 1, 0), vf(3 * (size_t)nf);
    auto corner = [&](int j, int c) { return c == 0 ? hf[j].x : c == 1 ? hf[j].y : hf[j].z; };
    for (int j = 0; j < nf; j++)
        for (int c = 0; c < 3; c++) vf_off[corner(j, c) + 1]++;
    for (int v = 0; v < nv; v++) vf_off[v + 1] += vf_off[v];
    std::vector<int32_t> fill(vf_off.begin(), vf_off.end() - 1);
    for (int j = 0; j < nf; j++)
        for (int c = 0; c < 3; c++) vf[fill[corner(j, c)]++] = 3 * j + c;
    int32_t *off = nullptr, *lst = nullptr;
    int rc = dalloc(h, &off, vf_off.size());
    if (!rc) rc = dalloc(h, &lst, vf.size());
    if (rc) return rc;
    e = hipMemcpy(off, vf_off.data(), sizeof(int32_t) * vf_off.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(lst, vf.data(), sizeof(int32_t) * vf.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err(h, e, "vertex-face lists");
    h->d_vf_off = off;
    h->d_vf = lst;
    return XPBD_OK;
}

int enqueue_contacts_jacobi(xpbd_handle* h, bool first)
{
    hipError_t e;
    if (jacobi_fixed_order(h)) {
        e = launch_contacts_jacobi_det(h->stream, h->d_faces, h->nf, h->d_cpos, h->d_clist, h->d_rest, h->d_pos,
                                       h->d_lo, h->d_cdx, h->d_cp, h->d_cn, h->d_cb, h->d_clam, first);
        if (e == hipSuccess)
            e = launch_contact_gather(h->stream, h->nv, h->d_vf_off, h->d_vf, h->d_flag, h->d_cdx, h->d_acc);
    } else {
        e = launch_contacts_jacobi(h->stream, h->d_faces, h->nf, h->d_cpos, h->d_clist, h->d_rest, h->d_pos, h->d_lo,
                                   h->d_acc, h->d_cp, h->d_cn, h->d_cb, h->d_clam, first);
    }
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "contacts_jacobi launch");
}

// Enqueue the solver iterations of one step (no predict / finish).  With
// reset_lambda (the reference solve, XPBDSolver.hpp:66-77) the first iteration
// starts every multiplier at zero; without it every multiplier (element,
// contact, attachment) keeps the value the caller left (a warm start).
int enqueue_iterations(xpbd_handle* h, int iters, bool timed, bool reset_lambda)
{
    hipStream_t s = h->stream;
    TetArrays ta = tet_arrays(h);
    const bool con = contacts_enabled(h);
    hipError_t e = hipSuccess;
    if (h->variant == XPBD_VARIANT_PATCH && h->patch_fused && h->fused_ok && h->params.deterministic != 0 &&
        h->natt == 0) {
        // fused patch Jacobi: one launch per iteration, the shared vertices' last
        // sums after the loop
        const Patches& Pt = h->patches;
        const size_t nsl = (size_t)Pt.nslots, nsh = Pt.shared.size();
        FusedPatchArgs A{};
        A.tloc = reinterpret_cast<const uint2*>(h->d_ptloc);
        A.pvoff = h->d_pvoff;
        A.pverts = h->d_pverts;
        A.pslot = h->d_pslot;
        A.pshare = h->d_pshare;
        A.inc_off = h->d_inc_off;
        A.inc = h->d_inc;
        A.pos = h->d_pos;
        A.npatch = Pt.npatch;
        A.maxl = Pt.maxl;
        A.nt = h->nt;
        A.xcd = h->xcd_map ? 1 : 0;
        A.pf_off = h->d_pf_off;
        A.pf = h->d_pf;
        A.flag = h->d_flag;
        A.rest = h->d_rest;
        A.lo = h->d_lo;
        A.cp = h->d_cp;
        A.cn = h->d_cn;
        A.cb = h->d_cb;
        A.clam = h->d_clam;
        for (int it = 0; it < iters; it++) {
            const int par = it & 1;
            A.prd = h->d_partial + (1 - par) * nsl;
            A.pwr = h->d_partial + par * nsl;
            A.brd = h->d_sbuf + par * nsh;
            A.bwr = h->d_sbuf + (1 - par) * nsh;
            A.first_iter = it == 0;
            if (timed) timed_begin(h);
            e = launch_jacobi_fused(s, ta, A, it == 0 && reset_lambda, con);
            if (e != hipSuccess) return hip_err(h, e, "jacobi_fused launch");
            if (timed) {
                timed_end(h);
                h->timed_iterations++;
            }
        }
        if (iters > 0) {
            // after iteration n-1: values before it in half n % 2, its partials in half (n-1) % 2
            e = launch_jacobi_fused_final(s, h->d_shared, h->d_soff, h->d_partial + ((iters - 1) & 1) * nsl,
                                          h->d_sbuf + (iters & 1) * nsh, h->d_pos, (int)nsh);
            if (e != hipSuccess) return hip_err(h, e, "jacobi_fused_final launch");
        }
        return XPBD_OK;
    }
    // Quad Gauss-Seidel in the split layout: the deltas leave the vertex records
    // for their own array before the loop and return after it, and the bases,
    // which the loop only reads, are copied into one (the attachment pass reads
    // the records, so handles with attachments stay interleaved).
    float4* dsep = h->params.solver == XPBD_SOLVER_GS_COLORED && h->variant == XPBD_VARIANT_CLUSTER_QUAD &&
                           quad_split(h->quad_block) && h->natt == 0 && iters > 0
                       ? h->d_dsep
                       : nullptr;
    if (dsep) {
        e = launch_split_deltas(s, h->d_pos, dsep, h->nv);
        if (e != hipSuccess) return hip_err(h, e, "split_deltas launch");
    }
    for (int it = 0; it < iters; it++) {
        const bool first = it == 0 && reset_lambda;
        if (h->params.solver == XPBD_SOLVER_GS_COLORED) {
            // one timed span covers the colour launches of the iteration (back to back on the stream)
            if (timed) timed_begin(h);
            for (int c = 0; c < h->ncolors; c++) {
                if (h->variant == XPBD_VARIANT_CLUSTER)
                    e = launch_gs_cluster(s, ta, h->d_cverts, h->d_cstart, h->d_tloc, h->d_pos, h->color_off[c],
                                          h->color_off[c + 1], first, h->xcd_map);
                else if (h->variant == XPBD_VARIANT_CLUSTER_QUAD)
                    e = launch_gs_cluster_quad(s, ta, h->d_cverts, h->d_cstart, h->d_tloc, h->d_pos,
                                               h->color_off[c], h->color_off[c + 1], first, h->quad_block,
                                               h->xcd_map, h->cluster_max_tets, dsep, dsep ? dsep + h->nv : nullptr);
                else
                    e = launch_gs_color(s, ta, h->d_pos, h->color_off[c], h->color_off[c + 1], first);
                if (e != hipSuccess) return hip_err(h, e, "gs_color launch");
            }
            if (timed) timed_end(h);
            if (con) {
                e = launch_contacts_gs(s, h->d_faces, h->d_fci_off, h->nfcolors, h->ninst, h->d_cpos, h->d_clist, h->d_rest,
                                       h->d_pos, h->d_lo, h->d_cp, h->d_cn, h->d_cb, h->d_clam, first, dsep,
                                       dsep ? dsep + h->nv : nullptr);
                if (e != hipSuccess) return hip_err(h, e, "contacts_gs launch");
            }
            // attachments follow the collision projectors (type-list order, XPBDTypedefs.hpp:56)
            if (h->natt > 0) {
                e = launch_attach_gs(s, h->d_att_v, h->d_att_p, h->d_att_lam, h->d_rest, h->d_pos, h->d_lo, h->natt,
                                     first);
                if (e != hipSuccess) return hip_err(h, e, "attach_gs launch");
            }
        } else if (h->variant == XPBD_VARIANT_PATCH) {
            if (h->natt > 0) {
                e = launch_attach_jacobi(s, h->d_att_v, h->d_att_p, h->d_att_lam, h->d_rest, h->d_pos, h->d_lo,
                                         h->d_acc, h->natt, first);
                if (e != hipSuccess) return hip_err(h, e, "attach_jacobi launch");
            }
            if (con) {
                // contact corrections read the pre-iteration positions, so they run first
                int rc = enqueue_contacts_jacobi(h, first);
                if (rc) return rc;
            }
            if (timed) timed_begin(h);
            e = launch_jacobi_patch(s, ta, patch_arrays(h), h->d_pos, reinterpret_cast<float4*>(h->d_acc), h->nt,
                                    first, h->params.deterministic != 0, con || h->natt > 0, h->xcd_map);
            if (e != hipSuccess) return hip_err(h, e, "jacobi_patch launch");
            if (timed) timed_end(h);
        } else {
            if (timed) timed_begin(h);
            e = launch_jacobi_project(s, ta, h->d_pos, h->d_acc, h->nt, first);
            if (e != hipSuccess) return hip_err(h, e, "jacobi_project launch");
            if (timed) timed_end(h);
            if (con) {
                int rc = enqueue_contacts_jacobi(h, first);
                if (rc) return rc;
            }
            if (h->natt > 0) {
                e = launch_attach_jacobi(s, h->d_att_v, h->d_att_p, h->d_att_lam, h->d_rest, h->d_pos, h->d_lo,
                                         h->d_acc, h->natt, first);
                if (e != hipSuccess) return hip_err(h, e, "attach_jacobi launch");
            }
            e = launch_jacobi_apply(s, h->d_pos, h->d_acc, h->nv);
            if (e != hipSuccess) return hip_err(h, e, "jacobi_apply launch");
        }
        if (timed) h->timed_iterations++;
    }
    if (dsep) {
        e = launch_join_deltas(s, h->d_pos, dsep, h->nv);
        if (e != hipSuccess) return hip_err(h, e, "join_deltas launch");
    }
    return XPBD_OK;
}

// Local collision iterations after the global solve (XPBDMeshObject.cpp:476-483,
// 764-825): count the element projections of a pass from the contacts whose
// multiplier is non-zero, then local_collision_iters passes of (elements,
// contacts) with the multipliers kept.  Gauss-Seidel visits the elements in the
// global solve's order, an element's repeats back to back; Jacobi sums every
// projection of a pass from the same positions.  Attachments are not part of it.
int enqueue_local_collision(xpbd_handle* h)
{
    const int n = h->params.local_collision_iters;
    if (n <= 0 || !h->d_mult || !contacts_enabled(h)) return XPBD_OK;
    hipStream_t s = h->stream;
    const TetArrays ta = tet_arrays(h);
    hipError_t e = launch_fill_zero(s, h->d_mult, sizeof(int32_t) * (size_t)h->nt);
    if (e == hipSuccess)
        e = launch_local_count(s, h->d_faces, h->nf, h->d_cpos, h->d_clist, h->d_clam, h->d_vt_off, h->d_vt_pos,
                               h->d_mult);
    if (e != hipSuccess) return hip_err(h, e, "local collision count launch");
    const bool gs = h->params.solver == XPBD_SOLVER_GS_COLORED;
    const bool cluster = h->variant == XPBD_VARIANT_CLUSTER || h->variant == XPBD_VARIANT_CLUSTER_QUAD;
    for (int it = 0; it < n; it++) {
        if (gs) {
            for (int c = 0; c < h->ncolors; c++) {
                e = cluster ? launch_gs_local_cluster(s, ta, h->d_cstart, h->d_mult, h->d_pos, h->color_off[c],
                                                      h->color_off[c + 1])
                            : launch_gs_local_color(s, ta, h->d_mult, h->d_pos, h->color_off[c], h->color_off[c + 1]);
                if (e != hipSuccess) return hip_err(h, e, "local gs launch");
            }
            e = launch_contacts_gs(s, h->d_faces, h->d_fci_off, h->nfcolors, h->ninst, h->d_cpos, h->d_clist, h->d_rest,
                                   h->d_pos, h->d_lo, h->d_cp, h->d_cn, h->d_cb, h->d_clam, false);
            if (e != hipSuccess) return hip_err(h, e, "local contacts_gs launch");
        } else {
            if (jacobi_fixed_order(h) && h->d_edx) {
                e = launch_jacobi_local_det(s, ta, h->d_mult, h->d_pos, h->d_edx, h->nt);
                if (e == hipSuccess)
                    e = launch_local_gather(s, ta, h->nv, h->d_vt_off, h->d_vt_pos, h->d_mult, h->d_edx, h->d_acc);
            } else {
                e = launch_jacobi_local(s, ta, h->d_mult, h->d_pos, h->d_acc, h->nt);
            }
            if (e != hipSuccess) return hip_err(h, e, "local jacobi launch");
            int rc = enqueue_contacts_jacobi(h, false);
            if (rc) return rc;
            e = launch_jacobi_apply(s, h->d_pos, h->d_acc, h->nv);
            if (e != hipSuccess) return hip_err(h, e, "local jacobi apply launch");
        }
    }
    return XPBD_OK;
}

int enqueue_detect(xpbd_handle* h)
{
    hipError_t e = launch_detect(h->stream, h->d_faces, h->nf, h->d_rest, h->d_pos, h->d_lo, h->d_sdf, h->d_step,
                                 h->params.dt, h->d_flag, h->d_cp, h->d_cn, h->d_cb);
    if (e != hipSuccess) return hip_err(h, e, "detect launch");
    e = launch_compact(h->stream, h->d_flag, h->nf, h->d_cpos, h->d_clist, h->d_bsum);
    if (e != hipSuccess) return hip_err(h, e, "compact launch");
    return XPBD_OK;
}

int enqueue_predict(xpbd_handle* h)
{
    const double dt = h->params.dt;
    hipError_t e = h->params.first_order
                       ? launch_predict_first_order(h->stream, h->d_pos, h->d_dz1, h->nv)
                       : launch_predict(h->stream, h->d_pos, h->d_vel, h->nv, (float)dt, (float)(-h->params.g * dt * dt));
    if (e == hipSuccess && h->primary_residual) e = launch_keep_prediction(h->stream, h->d_pos, h->d_pred, h->nv);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "predict launch");
}

int enqueue_finish(xpbd_handle* h)
{
    hipError_t e = launch_finish(h->stream, h->d_pos, h->d_lo, h->d_vel, h->nv, (float)h->params.dt, h->d_step);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "finish launch");
}

// Whole step: [detect] predict iterations finish
static int enqueue_step(xpbd_handle* h, bool detect, bool timed)
{
    int rc;
    if (detect && contacts_enabled(h) && (rc = enqueue_detect(h))) return rc;
    if ((rc = enqueue_predict(h))) return rc;
    if ((rc = enqueue_iterations(h, h->params.iters, timed))) return rc;
    if (records_residuals(h) && (rc = enqueue_step_residual(h))) return rc;
    if ((rc = enqueue_local_collision(h))) return rc;
    return enqueue_finish(h);
}

// Kernel timing spans.  An event that could not be created stays null and its
// span is left out of the statistics.
void timed_begin(xpbd_handle* h)
{
    if (h->ev_used == h->events.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess) a = nullptr;
        if (hipEventCreate(&b) != hipSuccess) b = nullptr;
        h->events.push_back({ a, b });
    }
    if (h->events[h->ev_used].first) (void)hipEventRecord(h->events[h->ev_used].first, h->stream);
}

void timed_end(xpbd_handle* h)
{
    if (h->events[h->ev_used].second) (void)hipEventRecord(h->events[h->ev_used].second, h->stream);
    h->ev_used++;
}

int collect_timing(xpbd_handle* h)
{
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "step");
    int64_t spans = 0;
    for (size_t i = 0; i < h->ev_used; i++) {
        float ms = 0.f;
        if (!h->events[i].first || !h->events[i].second ||
            hipEventElapsedTime(&ms, h->events[i].first, h->events[i].second) != hipSuccess)
            continue;
        h->timed_ms += ms;
        spans++;
    }
    h->timed_launches += spans;
    h->ev_used = 0;
    return XPBD_OK;
}

bool detect_this_step(const xpbd_handle* h)
{
    const int k = h->params.collision_every;
    return h->nf > 0 && h->sdf_active && k > 0 && (h->host_step % k) == 0;
}

int run_steps(xpbd_handle* h, int32_t nsteps)
{
    for (int32_t n = 0; n < nsteps; n++) {
        const bool det = detect_this_step(h);
        int rc;
        if (h->timing) {
            if ((rc = enqueue_step(h, det, true))) return rc;
            if ((rc = collect_timing(h))) return rc;
        } else {
            const int gi = det ? 1 : 0;
            if (!h->graph_exec[gi]) {
                hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_err(h, e, "hipStreamBeginCapture");
                rc = enqueue_step(h, det, false);
                hipGraph_t g = nullptr;
                e = hipStreamEndCapture(h->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) return hip_err(h, e, "hipStreamEndCapture");
                e = hipGraphInstantiate(&h->graph_exec[gi], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) return hip_err(h, e, "hipGraphInstantiate");
            }
            hipError_t e = hipGraphLaunch(h->graph_exec[gi], h->stream);
            if (e != hipSuccess) return hip_err(h, e, "hipGraphLaunch");
        }
        h->host_step++;
    }
    return XPBD_OK;
}

}  // namespace xpbd

extern "C" int xpbd_step(xpbd_handle* h, int32_t nsteps)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (nsteps < 0) return set_err(h, XPBD_E_INVALID_ARG, "nsteps < 0");
    return run_steps(h, nsteps);
}

extern "C" int xpbd_set_iterations(xpbd_handle* h, int32_t iters)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (iters < 0) return set_err(h, XPBD_E_INVALID_ARG, "iters < 0");
    if (iters != h->params.iters) {
        h->params.iters = iters;
        invalidate_graphs(h);
    }
    return XPBD_OK;
}

extern "C" int xpbd_synchronize(xpbd_handle* h)
{
    if (!h) return XPBD_E_INVALID_ARG;
    hipError_t e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "hipStreamSynchronize");
}

// ---------------------------------------------------------------------------
// state access
// ---------------------------------------------------------------------------
namespace xpbd {
int download_f4(xpbd_handle* h, std::vector<float4>& hu, const float4* src, size_t n)
{
    hu.resize(n);
    hipError_t e = hipMemcpyAsync(hu.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "download");
}
}  // namespace xpbd

// position = rest + base + low part + in-step motion (zero between steps)
static int positions_range(xpbd_handle* h, int64_t v0, int64_t v1, double* out)
{
    std::vector<float4> hp, hl;
    int rc = download_f4(h, hp, h->d_pos, 2 * (size_t)h->nv);
    if (!rc) rc = download_f4(h, hl, h->d_lo, h->nv);
    if (rc) return rc;
    for (int64_t i = v0; i < v1; i++) {
        const float4 b = hp[2 * i], d = hp[2 * i + 1], l = hl[i];
        out[3 * (i - v0) + 0] = h->rest[3 * i + 0] + ((double)b.x + (double)l.x + (double)d.x);
        out[3 * (i - v0) + 1] = h->rest[3 * i + 1] + ((double)b.y + (double)l.y + (double)d.y);
        out[3 * (i - v0) + 2] = h->rest[3 * i + 2] + ((double)b.z + (double)l.z + (double)d.z);
    }
    return XPBD_OK;
}

extern "C" int xpbd_get_positions(xpbd_handle* h, double* out)
{
    if (!h || !out) return h ? set_err(h, XPBD_E_INVALID_ARG, "out is NULL") : XPBD_E_INVALID_ARG;
    return positions_range(h, 0, h->nv, out);
}

extern "C" int xpbd_get_positions_f32(xpbd_handle* h, float* out)
{
    if (!h || !out) return h ? set_err(h, XPBD_E_INVALID_ARG, "out is NULL") : XPBD_E_INVALID_ARG;
    std::vector<double> tmp(3 * (size_t)h->nv);
    int rc = positions_range(h, 0, h->nv, tmp.data());
    if (rc) return rc;
    for (size_t i = 0; i < tmp.size(); i++) out[i] = (float)tmp[i];
    return XPBD_OK;
}

extern "C" int xpbd_batch_get_positions(xpbd_batch* b, int32_t instance, double* out)
{
    if (!b || !out) return b ? set_err(b, XPBD_E_INVALID_ARG, "out is NULL") : XPBD_E_INVALID_ARG;
    if (instance < 0 || instance >= b->ninst) return set_err(b, XPBD_E_INVALID_ARG, "instance out of range");
    return positions_range(b, b->v_off[instance], b->v_off[instance + 1], out);
}

extern "C" int xpbd_batch_step(xpbd_batch* b, int32_t nsteps)
{
    return xpbd_step(b, nsteps);
}

extern "C" int xpbd_batch_num_instances(xpbd_batch* b, int32_t* out)
{
    if (!b || !out) return XPBD_E_INVALID_ARG;
    *out = b->ninst;
    return XPBD_OK;
}

extern "C" int xpbd_get_velocities(xpbd_handle* h, double* out)
{
    if (!h || !out) return h ? set_err(h, XPBD_E_INVALID_ARG, "out is NULL") : XPBD_E_INVALID_ARG;
    std::vector<float4> hv;
    int rc = download_f4(h, hv, h->d_vel, h->nv);
    if (rc) return rc;
    for (int i = 0; i < h->nv; i++) {
        out[3 * (size_t)i] = hv[i].x;
        out[3 * (size_t)i + 1] = hv[i].y;
        out[3 * (size_t)i + 2] = hv[i].z;
    }
    return XPBD_OK;
}

extern "C" int xpbd_set_state(xpbd_handle* h, const double* x, const double* v)
{
    if (!h) return XPBD_E_INVALID_ARG;
    const size_t n3 = 3 * (size_t)h->nv;
    if ((x && !finite_all(x, n3)) || (v && !finite_all(v, n3)))
        return set_err(h, XPBD_E_NONFINITE, "set_state: non-finite input");
    int rc;
    if (x) {
        // base = nearest float to the displacement, low part = remainder; no in-step motion
        std::vector<float4> hp, hl(h->nv);
        if ((rc = download_f4(h, hp, h->d_pos, 2 * (size_t)h->nv))) return rc;
        for (int i = 0; i < h->nv; i++) {
            float hi[3], lw[3];
            for (int k = 0; k < 3; k++) {
                const double u = x[3 * (size_t)i + k] - h->rest[3 * (size_t)i + k];
                hi[k] = (float)u;
                lw[k] = (float)(u - (double)hi[k]);
            }
            hp[2 * (size_t)i] = make_float4(hi[0], hi[1], hi[2], hp[2 * (size_t)i].w);
            hp[2 * (size_t)i + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            hl[i] = make_float4(lw[0], lw[1], lw[2], 0.f);
        }
        hipError_t e = hipMemcpy(h->d_pos, hp.data(), sizeof(float4) * hp.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->d_lo, hl.data(), sizeof(float4) * hl.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_err(h, e, "set_state");
    }
    if (v) {
        std::vector<float4> hv;
        if ((rc = download_f4(h, hv, h->d_vel, h->nv))) return rc;
        for (int i = 0; i < h->nv; i++) {
            hv[i].x = (float)v[3 * (size_t)i];
            hv[i].y = (float)v[3 * (size_t)i + 1];
            hv[i].z = (float)v[3 * (size_t)i + 2];
        }
        hipError_t e = hipMemcpy(h->d_vel, hv.data(), sizeof(float4) * h->nv, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_err(h, e, "set_state");
    }
    return learned_reset_history(h);
}

extern "C" int xpbd_get_masses(xpbd_handle* h, double* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    std::memcpy(out, h->mass.data(), sizeof(double) * h->nv);
    return XPBD_OK;
}

extern "C" int xpbd_get_lambdas(xpbd_handle* h, double* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    std::vector<float2> hl(h->nt);
    hipError_t e = hipMemcpyAsync(hl.data(), h->d_lam, sizeof(float2) * h->nt, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "get_lambdas");
    for (int i = 0; i < h->nt; i++) {
        const int t = h->order[i];
        out[2 * (size_t)t] = hl[i].x;
        out[2 * (size_t)t + 1] = hl[i].y;
    }
    return XPBD_OK;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
namespace xpbd {
int device_residual(xpbd_handle* h, double* norm, double* vec_solver_order_dev)
{
    hipError_t e = launch_residual(h->stream, tet_arrays(h), h->d_pos, h->d_lo, true, h->nt, h->d_part,
                                   h->red_blocks, vec_solver_order_dev);
    if (e == hipSuccess) e = launch_sum_partials(h->stream, h->d_part, h->red_blocks, h->d_scalar);
    double s = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&s, h->d_scalar, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "residual");
    *norm = std::sqrt(s);
    return XPBD_OK;
}
// ||r||^2 (residual policy) and the primary residual (when on) at the end of
// the solve, before the local collision passes and the pin reset (where the
// reference's EVERY_SUBSTEP policy evaluates both, XPBDSolver.hpp:91-96), into
// d_step_res, d_pres and d_step_pres.
int enqueue_step_residual(xpbd_handle* h)
{
    hipError_t e = hipSuccess;
    if (h->residual_policy) {
        e = launch_residual(h->stream, tet_arrays(h), h->d_pos, h->d_lo, true, h->nt, h->d_part, h->red_blocks,
                            nullptr);
        if (e == hipSuccess) e = launch_sum_partials(h->stream, h->d_part, h->red_blocks, h->d_step_res);
    }
    if (e == hipSuccess && h->primary_residual) {
        PrimaryArgs A = element_force_args(h);
        A.vel = h->d_vel;
        A.pred = h->d_pred;
        A.vf_off = h->nf > 0 ? h->d_vf_off : nullptr;
        A.vf = h->d_vf;
        A.flag = h->d_flag;
        A.cn = h->d_cn;
        A.cb = h->d_cb;
        A.clam = h->d_clam;
        A.att_v = h->d_att_v;
        A.natt = h->natt;
        A.R = h->d_pres;
        A.part = h->d_part;
        A.nblocks = h->red_blocks;
        e = launch_primary_residual(h->stream, tet_arrays(h), A, h->d_step_pres);
    }
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "step residual");
}

// The vertex -> element list with a slot per (element, vertex) pair and the
// forces buffer over those slots: what k_element_forces writes and the
// per-vertex gathers (primary residual, consistent learned start) read.
int ensure_vertex_slots(xpbd_handle* h)
{
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "vertex slots");
    int rc = ensure_vertex_elements(h, false);
    if (rc) return rc;
    if (!h->d_vt_slot) {
        // the slot of (solver position i, corner j) in the vertex -> element list,
        // filled in the order that list was built (positions ascending, corners in order)
        const int nv = h->nv, nt = h->nt;
        std::vector<int32_t> fill((size_t)nv + 1, 0), slot(4 * (size_t)nt);
        for (int t = 0; t < nt; t++)
            for (int j = 0; j < 4; j++) fill[h->tets_host[4 * (size_t)t + j] + 1]++;
        for (int v = 0; v < nv; v++) fill[v + 1] += fill[v];
        for (int i = 0; i < nt; i++) {
            const int t = h->order[i];
            for (int j = 0; j < 4; j++) slot[4 * (size_t)i + j] = fill[h->tets_host[4 * (size_t)t + j]]++;
        }
        int32_t* d = nullptr;
        if ((rc = dalloc(h, &d, slot.size()))) return rc;
        e = hipMemcpy(d, slot.data(), sizeof(int32_t) * slot.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_err(h, e, "vertex-element slots");
        h->d_vt_slot = d;
    }
    if (!h->d_eforce) return dalloc(h, &h->d_eforce, 3 * 4 * (size_t)h->nt);
    return XPBD_OK;
}

PrimaryArgs element_force_args(xpbd_handle* h)
{
    PrimaryArgs A{};
    A.pos = h->d_pos;
    A.lo = h->d_lo;
    A.vt_off = h->d_vt_off;
    A.vt_slot = h->d_vt_slot;
    A.nt = h->nt;
    A.nv = h->nv;
    A.ef = h->d_eforce;
    return A;
}

// Buffers of the primary residual: the vertex lists it gathers over, the kept
// prediction, the residual vector and its norm (zero until the first step).
static int ensure_primary_residual(xpbd_handle* h)
{
    int rc = ensure_vertex_slots(h);
    if (!rc) rc = ensure_vertex_faces(h);
    if (rc) return rc;
    if (!h->d_pred) {
        float4* pred = nullptr;
        double *pres = nullptr, *norm = nullptr;
        rc = dalloc(h, &pred, (size_t)h->nv);
        if (!rc) rc = dalloc(h, &pres, 3 * (size_t)h->nv);
        if (!rc) rc = dalloc(h, &norm, 1);
        if (rc) return rc;
        hipError_t e = launch_fill_zero(h->stream, pred, sizeof(float4) * (size_t)h->nv);
        if (e == hipSuccess) e = launch_fill_zero(h->stream, pres, sizeof(double) * 3 * (size_t)h->nv);
        if (e == hipSuccess) e = launch_fill_zero(h->stream, norm, sizeof(double));
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_err(h, e, "primary residual buffers");
        h->d_pred = pred;
        h->d_pres = pres;
        h->d_step_pres = norm;
    }
    return XPBD_OK;
}
}  // namespace xpbd

extern "C" int xpbd_set_residual_policy(xpbd_handle* h, int32_t policy)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (policy != 0 && policy != 1) return set_err(h, XPBD_E_INVALID_ARG, "policy must be 0 or 1");
    if (policy != h->residual_policy) {
        h->residual_policy = policy;
        invalidate_graphs(h);
    }
    return XPBD_OK;
}

extern "C" int xpbd_set_primary_residual(xpbd_handle* h, int32_t on)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (on != 0 && on != 1) return set_err(h, XPBD_E_INVALID_ARG, "on must be 0 or 1");
    if (on != h->primary_residual) {
        if (on) {
            const int rc = ensure_primary_residual(h);
            if (rc) return rc;
        }
        h->primary_residual = on;
        invalidate_graphs(h);
    }
    return XPBD_OK;
}

extern "C" int xpbd_last_step_primary_residual(xpbd_handle* h, double* norm_out, double* vec_out)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (!norm_out) return set_err(h, XPBD_E_INVALID_ARG, "norm_out is NULL");
    if (!h->primary_residual) return set_err(h, XPBD_E_STATE, "primary residual is off");
    double s = 0;
    hipError_t e = hipMemcpyAsync(&s, h->d_step_pres, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && vec_out)
        e = hipMemcpyAsync(vec_out, h->d_pres, sizeof(double) * 3 * (size_t)h->nv, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "last_step_primary_residual");
    *norm_out = std::sqrt(s);
    return XPBD_OK;
}

extern "C" int xpbd_last_step_residual(xpbd_handle* h, double* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    if (!h->residual_policy) return set_err(h, XPBD_E_STATE, "residual policy is off");
    double s = 0;
    hipError_t e = hipMemcpyAsync(&s, h->d_step_res, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "last_step_residual");
    *out = std::sqrt(s);
    return XPBD_OK;
}

extern "C" int xpbd_constraint_residual(xpbd_handle* h, double* norm_out, double* vec_out)
{
    if (!h || !norm_out) return h ? set_err(h, XPBD_E_INVALID_ARG, "norm_out is NULL") : XPBD_E_INVALID_ARG;
    double* dvec = nullptr;
    if (vec_out) {
        if (!h->d_resvec) {
            if (hipMalloc(reinterpret_cast<void**>(&h->d_resvec), sizeof(double) * 2 * (size_t)h->nt) != hipSuccess)
                return set_err(h, XPBD_E_NOMEM, "residual vector allocation failed");
            h->allocs.push_back(h->d_resvec);
        }
        dvec = h->d_resvec;
    }
    int rc = device_residual(h, norm_out, dvec);
    if (rc || !vec_out) return rc;
    std::vector<double> tmp(2 * (size_t)h->nt);
    hipError_t e = hipMemcpy(tmp.data(), dvec, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(h, e, "residual vector");
    for (int i = 0; i < h->nt; i++) {
        const int t = h->order[i];
        vec_out[2 * (size_t)t] = tmp[2 * (size_t)i];
        vec_out[2 * (size_t)t + 1] = tmp[2 * (size_t)i + 1];
    }
    return XPBD_OK;
}

extern "C" int xpbd_strain_energy(xpbd_handle* h, double* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    // alpha = alpha~ * dt^2 (first order: alpha~ * dt)
    const double dt2 = h->params.first_order ? h->params.dt : h->params.dt * h->params.dt;
    hipError_t e = launch_energy(h->stream, tet_arrays(h), h->d_pos, h->d_lo, h->nt, dt2, h->d_part, h->red_blocks);
    if (e == hipSuccess) e = launch_sum_partials(h->stream, h->d_part, h->red_blocks, h->d_scalar);
    if (e == hipSuccess) e = hipMemcpyAsync(out, h->d_scalar, sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "energy");
}

extern "C" int xpbd_num_colors(xpbd_handle* h, int32_t* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    *out = h->ncolors;
    return XPBD_OK;
}

extern "C" int xpbd_get_tet_order(xpbd_handle* h, int32_t* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    std::memcpy(out, h->order.data(), sizeof(int32_t) * h->nt);
    return XPBD_OK;
}

extern "C" int xpbd_get_color_offsets(xpbd_handle* h, int32_t* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    for (int c = 0; c <= h->ncolors; c++)
        out[c] = h->nclusters > 0 ? h->cl_start[h->color_off[c]] : h->color_off[c];
    return XPBD_OK;
}

extern "C" int xpbd_get_face_rank(xpbd_handle* h, int32_t* out)
{
    if (!h || !out) return XPBD_E_INVALID_ARG;
    for (int j = 0; j < h->nf; j++) out[h->fperm[j]] = j;
    return XPBD_OK;
}

// ---------------------------------------------------------------------------
// contacts
// ---------------------------------------------------------------------------
static int put_sdf(xpbd_handle* h, int32_t inst, int32_t kind, const double* pose7, const double* params3)
{
    if (kind != XPBD_SDF_NONE && kind != XPBD_SDF_SPHERE && kind != XPBD_SDF_BOX)
        return set_err(h, XPBD_E_INVALID_ARG, "unknown SDF kind");
    if (kind != XPBD_SDF_NONE && (!pose7 || !params3)) return set_err(h, XPBD_E_INVALID_ARG, "pose/params NULL");
    if (kind != XPBD_SDF_NONE && (!finite_all(pose7, 7) || !finite_all(params3, 3)))
        return set_err(h, XPBD_E_NONFINITE, "SDF pose/params not finite");
    SdfDev& s = h->sdf_host[inst];
    const double vel[3] = { s.vel[0], s.vel[1], s.vel[2] };
    s = SdfDev{};
    s.kind = kind;
    for (int k = 0; k < 3; k++) s.vel[k] = vel[k];
    if (kind != XPBD_SDF_NONE) {
        for (int k = 0; k < 3; k++) s.pos[k] = pose7[k];
        for (int k = 0; k < 4; k++) s.quat[k] = pose7[3 + k];
        for (int k = 0; k < 3; k++) s.prm[k] = params3[k];
        if (kind == XPBD_SDF_SPHERE && !(params3[0] > 0)) return set_err(h, XPBD_E_INVALID_ARG, "radius <= 0");
        const double qn = s.quat[0] * s.quat[0] + s.quat[1] * s.quat[1] + s.quat[2] * s.quat[2] + s.quat[3] * s.quat[3];
        if (!(qn > 0)) return set_err(h, XPBD_E_INVALID_ARG, "zero quaternion");
    } else {
        s.quat[3] = 1.0;
    }
    hipError_t e = hipMemcpyAsync(h->d_sdf + inst, &s, sizeof(SdfDev), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = launch_fill_zero(h->stream, h->d_step, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "set_sdf");
    bool any = false;
    for (auto& q : h->sdf_host) any = any || q.kind != XPBD_SDF_NONE;
    if (any != h->sdf_active) {
        h->sdf_active = any;
        invalidate_graphs(h);
    }
    h->host_step = 0;
    return XPBD_OK;
}

extern "C" int xpbd_set_sdf(xpbd_handle* h, int32_t kind, const double* pose7, const double* params3)
{
    if (!h) return XPBD_E_INVALID_ARG;
    for (int i = 0; i < h->ninst; i++) {
        int rc = put_sdf(h, i, kind, pose7, params3);
        if (rc) return rc;
    }
    return XPBD_OK;
}

extern "C" int xpbd_batch_set_sdf(xpbd_batch* b, int32_t instance, int32_t kind, const double* pose7,
                                  const double* params3)
{
    if (!b) return XPBD_E_INVALID_ARG;
    if (instance < 0 || instance >= b->ninst) return set_err(b, XPBD_E_INVALID_ARG, "instance out of range");
    return put_sdf(b, instance, kind, pose7, params3);
}

// Attachments (XPBDMeshObject_::addAttachmentConstraint / clearAttachmentConstraints,
// XPBDMeshObject.cpp:255-275): replaces the whole list; n = 0 clears it.
extern "C" int xpbd_set_attachments(xpbd_handle* h, int32_t n, const int32_t* vertices, const double* targets)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (n < 0 || (n > 0 && (!vertices || !targets))) return set_err(h, XPBD_E_INVALID_ARG, "attachments: bad arguments");
    std::vector<uint8_t> seen(h->nv, 0);
    for (int32_t a = 0; a < n; a++) {
        if (vertices[a] < 0 || vertices[a] >= h->nv) return set_err(h, XPBD_E_INVALID_ARG, "attachment vertex out of range");
        if (seen[vertices[a]]) return set_err(h, XPBD_E_INVALID_ARG, "a vertex may carry one attachment");
        seen[vertices[a]] = 1;
        for (int k = 0; k < 3; k++)
            if (!std::isfinite(targets[3 * (size_t)a + k])) return set_err(h, XPBD_E_NONFINITE, "attachment target");
    }
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "attachments");
    if (n > h->att_cap) {
        // grow: the old buffers are released with the handle
        if (dalloc(h, &h->d_att_v, n) || dalloc(h, &h->d_att_p, 3 * (size_t)n) || dalloc(h, &h->d_att_lam, n))
            return XPBD_E_NOMEM;
        h->att_cap = n;
    }
    if (n > 0) {
        e = hipMemcpy(h->d_att_v, vertices, sizeof(int32_t) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->d_att_p, targets, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_err(h, e, "attachments upload");
    }
    if (n != h->natt) invalidate_graphs(h);  // launches capture the count
    h->natt = n;
    return XPBD_OK;
}

extern "C" int xpbd_set_sdf_velocity(xpbd_handle* h, const double* vel3)
{
    if (!h || !vel3) return XPBD_E_INVALID_ARG;
    if (!finite_all(vel3, 3)) return set_err(h, XPBD_E_NONFINITE, "velocity not finite");
    for (int i = 0; i < h->ninst; i++) {
        for (int k = 0; k < 3; k++) h->sdf_host[i].vel[k] = vel3[k];
    }
    hipError_t e = hipMemcpyAsync(h->d_sdf, h->sdf_host.data(), sizeof(SdfDev) * h->ninst, hipMemcpyHostToDevice,
                                  h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "set_sdf_velocity");
}

extern "C" int xpbd_detect_contacts(xpbd_handle* h)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (!(h->nf > 0 && h->sdf_active)) return XPBD_OK;
    int rc = enqueue_detect(h);
    if (rc) return rc;
    return xpbd_synchronize(h);
}

// XPBDMeshObject_::clearCollisionConstraints + addStaticCollisionConstraint
// (XPBDMeshObject.cpp:180-251): the contact set replaced by the caller's.
extern "C" int xpbd_set_contacts(xpbd_handle* h, int32_t n, const int32_t* faces, const double* p,
                                 const double* nrm, const double* bary)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (n < 0 || (n > 0 && (!faces || !p || !nrm || !bary)))
        return set_err(h, XPBD_E_INVALID_ARG, "set_contacts: n >= 0 and all arrays required");
    if (n > 0 && h->nf == 0) return set_err(h, XPBD_E_STATE, "set_contacts: the object has no surface faces");
    if (n > 0 && (!finite_all(p, 3 * (size_t)n) || !finite_all(nrm, 3 * (size_t)n) || !finite_all(bary, 3 * (size_t)n)))
        return set_err(h, XPBD_E_NONFINITE, "set_contacts: non-finite input");
    const int nf = h->nf;
    std::vector<int32_t> slot(nf), flag(std::max(nf, 1), 0);
    for (int j = 0; j < nf; j++) slot[h->fperm[j]] = j;
    std::vector<double> cp(3 * (size_t)std::max(nf, 1), 0.0), cn(cp), cb(cp);
    const StreamDrain drain{ h->stream };  // declared after the vectors: runs before they are freed
    for (int c = 0; c < n; c++) {
        const int f = faces[c];
        if (f < 0 || f >= nf) return set_err(h, XPBD_E_INVALID_ARG, "set_contacts: face index out of range");
        const int j = slot[f];
        if (flag[j]) return set_err(h, XPBD_E_INVALID_ARG, "set_contacts: a face may carry one contact");
        flag[j] = 1;
        for (int k = 0; k < 3; k++) {
            cp[3 * (size_t)j + k] = p[3 * (size_t)c + k];
            cn[3 * (size_t)j + k] = nrm[3 * (size_t)c + k];
            cb[3 * (size_t)j + k] = bary[3 * (size_t)c + k];
        }
    }
    if (nf > 0) {
        hipError_t e = hipMemcpyAsync(h->d_flag, flag.data(), sizeof(int32_t) * nf, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h->d_cp, cp.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h->d_cn, cn.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h->d_cb, cb.data(), sizeof(double) * 3 * nf, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = launch_fill_zero(h->stream, h->d_clam, sizeof(double) * nf);
        if (e == hipSuccess) e = launch_compact(h->stream, h->d_flag, nf, h->d_cpos, h->d_clist, h->d_bsum);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return hip_err(h, e, "set_contacts");
    }
    const bool given = n > 0;
    if (given != h->contacts_given) {
        h->contacts_given = given;
        invalidate_graphs(h);  // the captured steps depend on whether contacts are projected
    }
    return XPBD_OK;
}

// XPBDSolver::solve(projector_references, num_iter, initialize)
// (XPBDSolver.hpp:106-125) on a list of elements, optionally followed by
// every contact, between steps.
extern "C" int xpbd_solve_elements(xpbd_handle* h, int32_t n, const int32_t* elements, int32_t num_iter,
                                   int32_t initialize, int32_t with_contacts)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (n < 0 || num_iter < 0 || (n > 0 && !elements))
        return set_err(h, XPBD_E_INVALID_ARG, "solve_elements: n >= 0, num_iter >= 0 and an element list required");
    const int nt = h->nt;
    std::vector<int32_t> rank(nt), mult(std::max(nt, 1), 0);
    for (int i = 0; i < nt; i++) rank[h->order[i]] = i;
    for (int u = 0; u < n; u++) {
        const int e = elements[u];
        if (e < 0 || e >= nt) return set_err(h, XPBD_E_INVALID_ARG, "solve_elements: element index out of range");
        mult[rank[e]]++;
    }
    hipStream_t s = h->stream;
    hipError_t e = hipSuccess;
    const StreamDrain drain{ s };  // mult feeds an asynchronous copy
    if (!h->d_mult) {
        if (hipMalloc(reinterpret_cast<void**>(&h->d_mult), sizeof(int32_t) * (size_t)std::max(nt, 1)) != hipSuccess)
            return set_err(h, XPBD_E_NOMEM, "solve_elements: count buffer allocation failed");
        h->allocs.push_back(h->d_mult);
    }
    e = hipMemcpyAsync(h->d_mult, mult.data(), sizeof(int32_t) * nt, hipMemcpyHostToDevice, s);
    const TetArrays ta = tet_arrays(h);
    if (e == hipSuccess && initialize) e = launch_reset_listed(s, ta.lam, h->d_mult, nt);
    if (e != hipSuccess) return hip_err(h, e, "solve_elements set-up");
    const bool con = with_contacts && contacts_enabled(h);
    const bool gs = h->params.solver == XPBD_SOLVER_GS_COLORED;
    const bool cluster = h->variant == XPBD_VARIANT_CLUSTER || h->variant == XPBD_VARIANT_CLUSTER_QUAD;
    // deterministic patch Jacobi: each element's pass correction to its own
    // slot, summed per vertex in a fixed order (as the local collision passes)
    const bool fixed = !gs && h->variant == XPBD_VARIANT_PATCH && h->params.deterministic != 0;
    if (fixed) {
        int rc = ensure_vertex_elements(h, true);
        if (rc) return rc;
    }
    for (int it = 0; it < num_iter; it++) {
        const bool first = initialize && it == 0;  // contacts restart from zero multipliers
        if (gs) {
            for (int c = 0; c < h->ncolors && e == hipSuccess; c++)
                e = cluster ? launch_gs_local_cluster(s, ta, h->d_cstart, h->d_mult, h->d_pos, h->color_off[c],
                                                      h->color_off[c + 1])
                            : launch_gs_local_color(s, ta, h->d_mult, h->d_pos, h->color_off[c], h->color_off[c + 1]);
            if (e == hipSuccess && con)
                e = launch_contacts_gs(s, h->d_faces, h->d_fci_off, h->nfcolors, h->ninst, h->d_cpos, h->d_clist,
                                       h->d_rest, h->d_pos, h->d_lo, h->d_cp, h->d_cn, h->d_cb, h->d_clam, first);
            if (e != hipSuccess) return hip_err(h, e, "solve_elements launch");
        } else {
            if (fixed) {
                e = launch_jacobi_local_det(s, ta, h->d_mult, h->d_pos, h->d_edx, nt);
                if (e == hipSuccess)
                    e = launch_local_gather(s, ta, h->nv, h->d_vt_off, h->d_vt_pos, h->d_mult, h->d_edx, h->d_acc);
            } else {
                e = launch_jacobi_local(s, ta, h->d_mult, h->d_pos, h->d_acc, nt);
            }
            if (e != hipSuccess) return hip_err(h, e, "solve_elements launch");
            if (con) {
                int rc = enqueue_contacts_jacobi(h, first);
                if (rc) return rc;
            }
            e = launch_jacobi_apply(s, h->d_pos, h->d_acc, h->nv);
            if (e != hipSuccess) return hip_err(h, e, "solve_elements apply launch");
        }
    }
    e = launch_commit(s, h->d_pos, h->d_lo, h->nv);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? XPBD_OK : hip_err(h, e, "solve_elements");
}

extern "C" int xpbd_get_contacts(xpbd_handle* h, int32_t cap, int32_t* count, int32_t* face, double* p, double* n,
                                 double* bary)
{
    if (!h || !count) return XPBD_E_INVALID_ARG;
    *count = 0;
    if (h->nf == 0) return XPBD_OK;
    std::vector<int32_t> cpos(h->nf + 1), clist(h->nf);
    std::vector<double> cp(3 * (size_t)h->nf), cn(3 * (size_t)h->nf), cb(3 * (size_t)h->nf);
    hipError_t e = hipMemcpyAsync(cpos.data(), h->d_cpos, sizeof(int32_t) * (h->nf + 1), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(clist.data(), h->d_clist, sizeof(int32_t) * h->nf, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cp.data(), h->d_cp, sizeof(double) * cp.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cn.data(), h->d_cn, sizeof(double) * cn.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(cb.data(), h->d_cb, sizeof(double) * cb.size(), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return hip_err(h, e, "get_contacts");
    const int total = cpos[h->nf];
    *count = total;
    for (int k = 0; k < total && k < cap; k++) {
        const int j = clist[k];
        if (face) face[k] = h->fperm[j];
        for (int c = 0; c < 3; c++) {
            if (p) p[3 * k + c] = cp[3 * (size_t)j + c];
            if (n) n[3 * k + c] = cn[3 * (size_t)j + c];
            if (bary) bary[3 * k + c] = cb[3 * (size_t)j + c];
        }
    }
    return XPBD_OK;
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
extern "C" int xpbd_set_kernel_timing(xpbd_handle* h, int32_t on)
{
    if (!h) return XPBD_E_INVALID_ARG;
    h->timing = on != 0;
    h->timed_ms = 0;
    h->timed_iterations = 0;
    h->timed_launches = 0;
    return XPBD_OK;
}

extern "C" int xpbd_kernel_stats(xpbd_handle* h, double* proj_ms, int64_t* iterations, int64_t* launches)
{
    if (!h) return XPBD_E_INVALID_ARG;
    if (proj_ms) *proj_ms = h->timed_ms;
    if (iterations) *iterations = h->timed_iterations;
    if (launches) *launches = h->timed_launches;
    return XPBD_OK;
}

// ---------------------------------------------------------------------------
// host-side mesh utilities
// ---------------------------------------------------------------------------
extern "C" int xpbd_mesh_grid(int32_t nx, int32_t ny, int32_t nz, double hx, double hy, double hz,
                              const double* origin3, double* x, int32_t* tets)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(hx > 0) || !(hy > 0) || !(hz > 0) || !x || !tets)
        return XPBD_E_INVALID_ARG;
    if ((int64_t)nx * ny * nz * 24 > INT32_MAX) return XPBD_E_INVALID_ARG;
    const double o[3] = { origin3 ? origin3[0] : 0.0, origin3 ? origin3[1] : 0.0, origin3 ? origin3[2] : 0.0 };
    std::vector<double> vx;
    std::vector<int32_t> vt;
    grid_mesh(nx, ny, nz, hx, hy, hz, o, vx, vt);
    std::memcpy(x, vx.data(), sizeof(double) * vx.size());
    std::memcpy(tets, vt.data(), sizeof(int32_t) * vt.size());
    return XPBD_OK;
}

extern "C" int xpbd_mesh_surface(int32_t num_vertices, int32_t num_tets, const int32_t* tets, int32_t* num_faces,
                                 int32_t* faces)
{
    if (num_vertices <= 0 || num_tets <= 0 || !tets || !num_faces || !faces) return XPBD_E_INVALID_ARG;
    std::vector<int32_t> f;
    if (surface_faces(num_vertices, num_tets, tets, f) < 0) return XPBD_E_INVALID_ARG;
    *num_faces = (int32_t)(f.size() / 3);
    std::memcpy(faces, f.data(), sizeof(int32_t) * f.size());
    return XPBD_OK;
}

extern "C" int xpbd_color_items(int32_t num_vertices, int32_t num_items, int32_t verts_per_item,
                                const int32_t* items, int32_t* colors, int32_t* num_colors)
{
    if (num_vertices <= 0 || num_items < 0 || verts_per_item <= 0 || !items || !colors || !num_colors)
        return XPBD_E_INVALID_ARG;
    std::vector<int32_t> c;
    const int nc = greedy_color(num_vertices, num_items, verts_per_item, items, c);
    if (nc < 0) return XPBD_E_INVALID_ARG;
    std::memcpy(colors, c.data(), sizeof(int32_t) * num_items);
    *num_colors = nc;
    return XPBD_OK;
}

extern "C" int xpbd_read_msh(const char* path, int32_t* num_vertices, int32_t* num_tets, int32_t* num_tris,
                             double* x, int32_t* tets, int32_t* tris)
{
    if (!path || !num_vertices || !num_tets || !num_tris) return XPBD_E_INVALID_ARG;
    std::vector<double> vx;
    std::vector<int32_t> vt, vf;
    std::string err;
    if (read_msh41(path, vx, vt, vf, err)) return set_err(nullptr, XPBD_E_IO, err);
    *num_vertices = (int32_t)(vx.size() / 3);
    *num_tets = (int32_t)(vt.size() / 4);
    *num_tris = (int32_t)(vf.size() / 3);
    if (x) std::memcpy(x, vx.data(), sizeof(double) * vx.size());
    if (tets) std::memcpy(tets, vt.data(), sizeof(int32_t) * vt.size());
    if (tris) std::memcpy(tris, vf.data(), sizeof(int32_t) * vf.size());
    return XPBD_OK;
}

extern "C" int xpbd_read_element_classes(const char* path, int32_t cap, int32_t* classes, int32_t* count)
{
    if (!path || !count || cap < 0 || (cap > 0 && !classes)) return XPBD_E_INVALID_ARG;
    std::vector<int32_t> v;
    std::string err;
    if (read_int_lines(path, v, err)) return set_err(nullptr, XPBD_E_IO, err);
    *count = (int32_t)v.size();
    if (classes) std::memcpy(classes, v.data(), sizeof(int32_t) * std::min<size_t>((size_t)cap, v.size()));
    return XPBD_OK;
}
