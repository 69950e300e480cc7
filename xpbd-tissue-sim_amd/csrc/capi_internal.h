// Internal state behind the C-ABI handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "kernels.h"
#include "mesh_tools.h"
#include "xpbd_hip.h"

namespace xpbd {
struct Learned;
}

struct xpbd_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    xpbd_params params{};
    int ninst = 1;
    int nv = 0, nt = 0, nf = 0;
    std::vector<int64_t> v_off, t_off, f_off;  // per-instance offsets in the concatenated arrays
    std::vector<double> rest;                  // 3*nv rest positions (FP64, host copy)
    std::vector<double> mass;                  // nv lumped masses
    std::vector<double> inertia;               // constraint inertia: mass, or damping B (first order)
    std::vector<float> dz1;                    // first order: inertial move -g m dt / B per vertex
    std::vector<uint8_t> pinned;
    std::vector<int32_t> tets_host;            // 4*nt vertex ids in the caller's element order
    std::vector<float> vmat;                   // 2*nv: mean log10(E) and mean nu of incident elements
    std::vector<int32_t> order;                // solver position -> caller's tet index
    std::vector<int32_t> color_off;            // colour c occupies [color_off[c], color_off[c+1])
                                               // (tets, or clusters for the cluster variant)
    int ncolors = 1;
    int variant = XPBD_VARIANT_SIMPLE;         // resolved kernel variant
    int quad_block = 256;                      // CLUSTER_QUAD workgroup size (autotuned)
    int cluster_max_tets = xpbd::kClusterTets; // most elements in one cluster
    bool xcd_map = true;                       // XCD-contiguous workgroup mapping (autotuned)
    std::string tuned;                         // last autotuner choice, "" if none
    std::vector<int32_t> cl_start;             // cluster variant: tet range of each cluster (solver order)
    int nclusters = 0;
    xpbd::Patches patches;                     // patch variant (host copy)
    std::vector<int32_t> fperm;                // solver face position -> caller's face index
    std::vector<int32_t> fcol_off;
    int nfcolors = 1;
    // within a face colour the faces follow face order, so each instance's faces
    // are one run: fci_off[c * (ninst + 1) + i] is where instance i's faces of
    // colour c start (i = ninst: the colour's end)
    std::vector<int32_t> fci_off;
    std::vector<xpbd::SdfDev> sdf_host;
    bool sdf_active = false;
    bool contacts_given = false;  // a contact set from xpbd_set_contacts is in use

    // device state
    float4* d_pos = nullptr;     // 2 per vertex: (base displacement, inverse mass), (in-step motion, 0)
    float4* d_lo = nullptr;      // low-order part of the base displacement
    float4* d_vel = nullptr;     // velocity xyz, pinned flag w
    float* d_dz1 = nullptr;      // first order: inertial move per vertex
    float* d_acc = nullptr;      // Jacobi accumulator (4 per vertex)
    double* d_rest = nullptr;
    int4* d_tet = nullptr;
    float* d_Q = nullptr;
    float2* d_at = nullptr;
    float* d_gamma = nullptr;
    float2* d_lam = nullptr;
    int4* d_cverts = nullptr;    // cluster variant: 8 vertex ids per cluster (2 x int4)
    float4* d_dsep = nullptr;    // quad variant, split layout: vertex deltas, then bases, during an iteration loop
    int32_t* d_cstart = nullptr;
    uint32_t* d_tloc = nullptr;
    uint32_t* d_ptloc = nullptr; // patch variant (see PatchArrays)
    int32_t* d_pvoff = nullptr;
    int32_t* d_pverts = nullptr;
    int32_t* d_pslot = nullptr;
    int32_t* d_inc_off = nullptr;
    uint16_t* d_inc = nullptr;
    float4* d_partial = nullptr;     // 2 x nslots: partial sums, by iteration parity in the fused variant
    int32_t* d_shared = nullptr;
    int32_t* d_soff = nullptr;
    // fused patch Jacobi (one launch per iteration, see FusedPatchArgs)
    bool fused_ok = false;           // the patch layout and the faces allow it
    bool patch_fused = false;        // in use (autotuned)
    int2* d_pshare = nullptr;
    int32_t* d_pf_off = nullptr;
    int4* d_pf = nullptr;
    float4* d_sbuf = nullptr;        // 2 x nshared shared-vertex values, by iteration parity
    double* d_part = nullptr;
    int red_blocks = 1;
    double* d_scalar = nullptr;
    double* d_resvec = nullptr;
    unsigned long long* d_step = nullptr;
    xpbd::SdfDev* d_sdf = nullptr;
    int4* d_faces = nullptr;
    int32_t* d_fci_off = nullptr;    // see fci_off
    int32_t* d_flag = nullptr;
    int32_t* d_cpos = nullptr;
    int32_t* d_clist = nullptr;
    int32_t* d_bsum = nullptr;   // compaction scratch (per-block hit counts)
    double* d_cp = nullptr;
    double* d_cn = nullptr;
    double* d_cb = nullptr;
    double* d_clam = nullptr;
    int natt = 0, att_cap = 0;   // attachments (vertex, FP64 target point, multiplier)
    int32_t* d_att_v = nullptr;
    double* d_att_p = nullptr;
    double* d_att_lam = nullptr;
    // local collision iterations: CSR vertex -> solver positions of its elements,
    // per-element projection counts of a local pass
    int32_t* d_vt_off = nullptr;
    int32_t* d_vt_pos = nullptr;
    int32_t* d_mult = nullptr;
    // deterministic patch Jacobi: per-face contact corrections and the vertex ->
    // (face slot * 3 + corner) CSR that gathers them; per-element local-pass sums
    float* d_cdx = nullptr;
    int32_t* d_vf_off = nullptr;
    int32_t* d_vf = nullptr;
    float* d_edx = nullptr;
    std::vector<void*> allocs;

    // stepping
    int64_t host_step = 0;
    int residual_policy = 0;     // 1: record ||r|| at the end of every solve (reference EVERY_SUBSTEP)
    double* d_step_res = nullptr;
    // primary residual, recorded at the same point when on: the prediction's
    // in-step motion, the constraint forces of every (element, vertex) pair in
    // vertex order, the per-vertex residual (3 nv) and its squared norm
    int primary_residual = 0;
    float4* d_pred = nullptr;
    int32_t* d_vt_slot = nullptr;   // 4 nt: slot of (solver position, corner) in the vertex -> element list
    double* d_eforce = nullptr;     // 3 per slot
    double* d_pres = nullptr;
    double* d_step_pres = nullptr;
    hipGraphExec_t graph_exec[2] = { nullptr, nullptr };  // [0] plain step, [1] step with detection
    uint64_t graph_gen = 0;      // bumped whenever captured steps go stale (other graphs compare against it)

    // kernel timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    size_t ev_used = 0;
    double timed_ms = 0;
    int64_t timed_iterations = 0;
    int64_t timed_launches = 0;

    // learned warm start (owned by learned.hip)
    xpbd::Learned* learned = nullptr;

    std::string err;
};

namespace xpbd {
int set_err(xpbd_handle* h, int code, const std::string& msg);
int hip_err(xpbd_handle* h, hipError_t e, const char* what);
void invalidate_graphs(xpbd_handle* h);
TetArrays tet_arrays(xpbd_handle* h);
PatchArrays patch_arrays(xpbd_handle* h);
bool contacts_enabled(const xpbd_handle* h);
int enqueue_iterations(xpbd_handle* h, int iters, bool timed, bool reset_lambda = true);
int enqueue_detect(xpbd_handle* h);
int enqueue_predict(xpbd_handle* h);
int enqueue_finish(xpbd_handle* h);
int enqueue_local_collision(xpbd_handle* h);
// Jacobi contact pass into the accumulator (fixed order in the deterministic patch mode).
int enqueue_contacts_jacobi(xpbd_handle* h, bool first);
bool detect_this_step(const xpbd_handle* h);
int run_steps(xpbd_handle* h, int32_t nsteps);
void timed_begin(xpbd_handle* h);
void timed_end(xpbd_handle* h);
int collect_timing(xpbd_handle* h);
int download_f4(xpbd_handle* h, std::vector<float4>& hu, const float4* src, size_t n);
int device_residual(xpbd_handle* h, double* norm, double* vec_solver_order_dev);
// residuals recorded at the end of every solve (constraint residual under the
// residual policy, primary residual when switched on)
int enqueue_step_residual(xpbd_handle* h);
inline bool records_residuals(const xpbd_handle* h) { return h->residual_policy || h->primary_residual; }
// the vertex -> element slots and the forces buffer over them (built on first use)
int ensure_vertex_slots(xpbd_handle* h);
PrimaryArgs element_force_args(xpbd_handle* h);
void destroy_learned(xpbd_handle* h);
// Clears the previous-step correction the learned warm start reads (state was replaced).
int learned_reset_history(xpbd_handle* h);
}  // namespace xpbd
