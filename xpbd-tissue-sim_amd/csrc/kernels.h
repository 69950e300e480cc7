// Launch wrappers for the solver kernels (implemented in solver_kernels.hip).
// Every wrapper enqueues on the given stream and returns the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "xpbd_hip.h"

namespace xpbd {

// Kinematic signed-distance field of one instance.  Pose at step n is
// pos + vel * (n * dt), orientation quat (x, y, z, w).
struct SdfDev {
    int32_t kind;  // 0 none, 1 sphere, 2 box
    int32_t pad;
    double pos[3];
    double vel[3];
    double quat[4];
    double prm[3];  // sphere: radius in prm[0]; box: edge lengths
};

// Structure-of-arrays element data in solver order.
struct TetArrays {
    const int4* tet;      // vertex ids
    const float* Q;       // 9 planes, Q[k * qstride + e]
    int qstride;
    const float2* at;     // (alpha_D, alpha_H) / dt^2
    const float* gamma;   // mu / lambda
    float2* lam;          // (lambda_D, lambda_H)
};

// Vertex records: pos[2i] = (base displacement xyz, inverse mass),
// pos[2i+1] = (in-step motion xyz, 0); lo[i] = low-order part of the base;
// vel[i] = (velocity xyz, pinned flag).
hipError_t launch_predict(hipStream_t s, float4* pos, const float4* vel, int nv, float dt, float dz);
// first-order inertial move: delta = (0, 0, dz1[i])
hipError_t launch_predict_first_order(hipStream_t s, float4* pos, const float* dz1, int nv);
hipError_t launch_finish(hipStream_t s, float4* pos, float4* lo, float4* vel, int nv, float dt,
                         unsigned long long* step_counter);
// Deterministic Jacobi: contacts into per-face slots, then a per-vertex gather in
// face-slot order; local passes into per-element slots, then a per-vertex gather
// in solver-position order.
hipError_t launch_contacts_jacobi_det(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                                      const double* rest, const float4* pos, const float4* lo, float* cdx,
                                      const double* cp, const double* cn, const double* cb, double* clam, bool first);
hipError_t launch_contact_gather(hipStream_t s, int nv, const int* vf_off, const int* vf, const int* flag,
                                 const float* cdx, float* acc);
hipError_t launch_jacobi_local_det(hipStream_t s, const TetArrays& ta, const int* mult, const float4* pos, float* edx,
                                   int nt);
hipError_t launch_local_gather(hipStream_t s, const TetArrays& ta, int nv, const int* vt_off, const int* vt_pos,
                               const int* mult, const float* edx, float* acc);
// in-step motion folded into the base, no velocity update (positions moved between steps)
hipError_t launch_commit(hipStream_t s, float4* pos, float4* lo, int nv);
// multipliers of the elements with mult > 0 set to zero
hipError_t launch_reset_listed(hipStream_t s, float2* lam, const int* mult, int nt);
// lambda = -C / alpha~ of every element at the current positions (learned step start)
hipError_t launch_lambda_start(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, int nt);
hipError_t launch_gs_color(hipStream_t s, const TetArrays& ta, float4* pos, int begin, int end, bool first);
// Cluster Gauss-Seidel over clusters [cbeg, cend) of one colour: cverts holds 8
// vertex ids per cluster (-1 padded), cstart the tet range of each cluster,
// tloc the 4-bit local slots of each tet's vertices.  xcd: give each XCD one
// contiguous eighth of the clusters (see xcd_item).
hipError_t launch_gs_cluster(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                             const uint32_t* tloc, float4* pos, int cbeg, int cend, bool first, bool xcd);
// Same clusters and order, four lanes per cluster (one per vertex of the element
// being projected); block = threads per workgroup (64, 128 or 256; negative:
// the same shape in the LDS-lean form, which fits more waves per CU; plus
// kQuadSplit in magnitude: the shape run with the split layout); max_tets = the
// most elements any cluster holds (at most kClusterTets).  dsep, dbase: the
// deltas and the bases in arrays of their own (one float4 per vertex each)
// while an iteration loop runs in the split layout, else null (both in the
// interleaved vertex records).
constexpr int kQuadSplit = XPBD_QUAD_SPLIT;
hipError_t launch_gs_cluster_quad(hipStream_t s, const TetArrays& ta, const int4* cverts, const int* cstart,
                                  const uint32_t* tloc, float4* pos, int cbeg, int cend, bool first, int block,
                                  bool xcd, int max_tets, float4* dsep = nullptr, const float4* dbase = nullptr);
bool quad_block_supported(int block);
int quad_shape(int block);   // the block value without the split flag
bool quad_split(int block);  // whether the block value asks for the split layout
// the deltas out of the vertex records into dsep[0, nv) and the bases into
// dsep[nv, 2 nv); the deltas back
hipError_t launch_split_deltas(hipStream_t s, const float4* pos, float4* dsep, int nv);
hipError_t launch_join_deltas(hipStream_t s, float4* pos, const float4* dsep, int nv);
hipError_t launch_jacobi_project(hipStream_t s, const TetArrays& ta, const float4* pos, float* acc, int nt,
                                 bool first);
hipError_t launch_jacobi_apply(hipStream_t s, float4* pos, float* acc, int nv);

// Patch Jacobi (see k_jacobi_patch): device views of the host Patches layout.
struct PatchArrays {
    const uint2* tloc;       // 4 x 16-bit local vertex ids per tet (solver order)
    const int* pvoff;        // patch -> range in pverts
    const int* pverts;       // distinct vertices of each patch
    const int* pslot;        // -1 interior, else partial slot
    const int* inc_off;      // per patch local incidence offsets (deterministic mode)
    const uint16_t* inc;     // per patch local incidence entries
    float4* partial;         // partial sums of shared vertices
    const int* shared;       // shared vertex ids
    const int* soff;         // shared vertex -> its contiguous range of partial slots
    int npatch, maxl, nshared;
};
int jacobi_patch_size();
size_t jacobi_patch_smem(int maxl, bool det);
// One Jacobi iteration: patch kernel + shared-vertex kernel.  cacc holds contact
// corrections (float4 per vertex) when con is set; it is consumed and zeroed.
hipError_t launch_jacobi_patch(hipStream_t s, const TetArrays& ta, const PatchArrays& pa, float4* pos, float4* cacc,
                               int nt, bool first, bool det, bool con, bool xcd);
// Fused patch Jacobi (fixed-order accumulation, no attachments): one launch per
// iteration; the shared vertices' sums are formed while they are staged in the
// next iteration, and contacts are projected by the patch that holds their
// face.  Per iteration i: prd/pwr are the partial buffers of iterations i-1 and i,
// brd/bwr the shared vertices' values before iterations i-1 and i (the two halves
// of two double buffers, by parity); first_iter stages every vertex from pos.
struct FusedPatchArgs {
    const uint2* tloc;
    const int* pvoff;
    const int* pverts;
    const int* pslot;
    const int2* pshare;      // per pverts entry: (shared index, first slot | count << 28) or (-1, 0)
    const int* inc_off;
    const uint16_t* inc;
    float4* pos;
    const float4* prd;
    float4* pwr;
    const float4* brd;
    float4* bwr;
    int first_iter;
    int npatch, maxl, nt, xcd;
    // contacts: per patch, its faces as (solver face index, three local slots)
    const int* pf_off;
    const int4* pf;
    const int* flag;
    const double* rest;
    const float4* lo;
    const double *cp, *cn, *cb;
    double* clam;
};
size_t jacobi_fused_smem(int maxl, bool con);
hipError_t launch_jacobi_fused(hipStream_t s, const TetArrays& ta, const FusedPatchArgs& A, bool first, bool con);
// shared vertex j = brd[j] + sum of prd over its slots, written to pos
hipError_t launch_jacobi_fused_final(hipStream_t s, const int* shared, const int* soff, const float4* prd,
                                     const float4* brd, float4* pos, int nshared);
hipError_t launch_fill_zero(hipStream_t s, void* p, size_t bytes);

// FP64 residual r = C + alpha_tilde * lambda per element, block partial sums of
// r^2 into part[nblocks]; vec (2*nt, optional) receives r in solver order.
// with_delta: evaluate at base + in-step motion (end of a solve) instead of base.
hipError_t launch_residual(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, bool with_delta,
                           int nt, double* part, int nblocks, double* vec);
// Strain energy sum C^2/alpha (alpha = alpha_tilde * dt^2), block partials.
hipError_t launch_energy(hipStream_t s, const TetArrays& ta, const float4* pos, const float4* lo, int nt, double dt2,
                         double* part, int nblocks);
hipError_t launch_sum_partials(hipStream_t s, const double* part, int n, double* out);

// Primary residual (momentum balance) of a solve, per vertex in FP64:
//   R_v = (x_v - xt_v) / w_v - sum_e grad_v C_e^T lambda_e - sum_c b_vc n_c lambda_c
// over the elements and active contacts around v; zero on pinned and attached
// vertices.  pred[v] is the in-step motion the inertial prediction set.
struct PrimaryArgs {
    const float4* pos;
    const float4* lo;
    const float4* vel;    // .w != 0 marks pins
    const float4* pred;
    const int* vt_off;    // vertex -> its range of element-corner slots
    const int* vt_slot;   // 4 * nt: slot of (element in solver position e, corner c) at 4 e + c
    const int* vf_off;    // vertex -> face slot * 3 + corner (null without faces)
    const int* vf;
    const int* flag;      // per face slot: contact active
    const double* cn;
    const double* cb;
    const double* clam;
    const int* att_v;
    int natt;
    int nv;
    int nt;
    double* ef;           // 3 per slot: grad C^T lambda of one element at one vertex, in vertex order
    double* R;            // 3 * nv
    double* part;         // nblocks partial sums of R^2
    int nblocks;
};
// keep the inertial prediction: pred[v] = in-step motion of v
hipError_t launch_keep_prediction(hipStream_t s, const float4* pos, float4* pred, int nv);
// R into A.R and ||R||^2 into *out
hipError_t launch_primary_residual(hipStream_t s, const TetArrays& ta, const PrimaryArgs& A, double* out);
// grad C^T lambda of every element at its four vertices into A.ef (the
// vertex-ordered slots of A.vt_slot), at the current positions and multipliers
hipError_t launch_element_forces(hipStream_t s, const TetArrays& ta, const PrimaryArgs& A);
// consistent start: in-step motion = pred + w * (sum of the vertex's slots of ef)
hipError_t launch_consistent_shift(hipStream_t s, float4* pos, const float4* pred, const int* vt_off,
                                   const double* ef, int nv);

// Contacts.  faces: (a, b, c, instance) in colour-sorted order.
hipError_t launch_detect(hipStream_t s, const int4* faces, int nf, const double* rest, const float4* pos,
                         const float4* lo, const SdfDev* sdf, const unsigned long long* step, double dt, int* flag,
                         double* cp, double* cn, double* cb);
// Ordered compaction of hit flags; bsum is scratch of compact_blocks(nf) ints.
int compact_blocks(int nf);
hipError_t launch_compact(hipStream_t s, const int* flag, int nf, int* cpos, int* clist, int* bsum);
// Gauss-Seidel contacts: one workgroup per instance (instances share no
// vertex), colour by colour; fci_off[c * (ninst + 1) + i] is the face position
// where instance i's faces of colour c start.
hipError_t launch_contacts_gs(hipStream_t s, const int4* faces, const int* fci_off, int ncol, int ninst,
                              const int* cpos, const int* clist, const double* rest, float4* pos, const float4* lo,
                              const double* cp, const double* cn, const double* cb, double* clam, bool first,
                              float4* dsep = nullptr, const float4* dbase = nullptr);
hipError_t launch_contacts_jacobi(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                                  const double* rest, const float4* pos, const float4* lo, float* acc,
                                  const double* cp, const double* cn, const double* cb, double* clam, bool first);

// Attachments: vertex av[a] pulled to the point ap[3a..3a+2] (alpha = 0).
// Gauss-Seidel applies in place; Jacobi accumulates into acc (4 floats per vertex).
hipError_t launch_attach_gs(hipStream_t s, const int* av, const double* ap, double* alam, const double* rest,
                            float4* pos, const float4* lo, int n, bool first);
hipError_t launch_attach_jacobi(hipStream_t s, const int* av, const double* ap, double* alam, const double* rest,
                                const float4* pos, const float4* lo, float* acc, int n, bool first);

// Local collision iterations.  vt_off / vt_pos: CSR of the solver positions of
// the elements attached to each vertex.  mult (nt ints, zeroed by the caller)
// receives, per solver position, how many times a local pass projects it.
hipError_t launch_local_count(hipStream_t s, const int4* faces, int nf, const int* cpos, const int* clist,
                              const double* clam, const int* vt_off, const int* vt_pos, int* mult);
// Gauss-Seidel local pass over one colour: elements [begin, end) (one element
// per thread) or clusters [cbeg, cend) (one cluster per thread).
hipError_t launch_gs_local_color(hipStream_t s, const TetArrays& ta, const int* mult, float4* pos, int begin, int end);
hipError_t launch_gs_local_cluster(hipStream_t s, const TetArrays& ta, const int* cstart, const int* mult, float4* pos,
                                   int cbeg, int cend);
// Jacobi local pass: corrections summed into acc (4 floats per vertex).
hipError_t launch_jacobi_local(hipStream_t s, const TetArrays& ta, const int* mult, const float4* pos, float* acc,
                               int nt);

}  // namespace xpbd
