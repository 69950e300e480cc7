/*
 * xpbd_hip.h -- C-ABI of the MI355X (gfx950) XPBD soft-tissue solver.
 *
 * Plain C types only: pointers to host arrays, sizes and int status codes.
 * Every call returns XPBD_OK (0) or a negative XPBD_E_* code; the message of
 * the last failure on a handle is xpbd_last_error(handle) (NULL handle: the
 * last failure of xpbd_create / xpbd_batch_create on this thread).
 *
 * Threading: one handle owns one HIP stream on the device that was current
 * when it was created.  Calls on one handle are not re-entrant; different
 * handles may be driven from different threads (the reference drives its
 * solver from a single physics thread, src/simulation/Simulation.cpp:331-336).
 *
 * Each entry point names the reference interface it replaces.  The reference
 * is C++ (MAVYWAVY902/xpbd-tissue-sim); paths are relative to its root.
 */
#ifndef XPBD_HIP_H
#define XPBD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XPBD_OK 0
#define XPBD_E_INVALID_ARG (-1)  /* null pointer, bad size, bad enum */
#define XPBD_E_HIP (-2)          /* HIP runtime failure (message has the HIP error) */
#define XPBD_E_NONFINITE (-3)    /* NaN or Inf in an input array */
#define XPBD_E_DEGENERATE (-4)   /* zero-volume element or zero-mass vertex */
#define XPBD_E_UNSUPPORTED (-5)  /* option not available in this build */
#define XPBD_E_NOMEM (-6)        /* host or device allocation failed */
#define XPBD_E_STATE (-7)        /* call not valid in the handle's current state */
#define XPBD_E_IO (-8)           /* file could not be read or written */

/* Solver types.  XPBD_SOLVER_JACOBI fills the reference's PARALLEL_JACOBI slot
 * (include/common/XPBDEnumTypes.hpp:6-11, "Parallel-Jacobi" in
 * include/config/simobject/XPBDMeshObjectConfig.hpp:26-34);
 * XPBD_SOLVER_GS_COLORED is the new "GPU-Gauss-Seidel" value beside it. */
#define XPBD_SOLVER_JACOBI 0
#define XPBD_SOLVER_GS_COLORED 1

/* Kernel variants (xpbd_params.variant).  AUTO picks CLUSTER_QUAD for
 * Gauss-Seidel and PATCH for Jacobi (xpbd_autotune then times the shapes).  SIMPLE: one thread per tet (Gauss-Seidel: one launch
 * per tet colour; Jacobi: float atomics).  CLUSTER (Gauss-Seidel): runs of
 * consecutive tets sharing at most 8 vertices form clusters, clusters are
 * coloured, one thread projects its cluster's tets in order.  PATCH (Jacobi):
 * workgroups stage a patch of tets and their vertices in LDS.  CLUSTER_QUAD
 * (Gauss-Seidel): the CLUSTER order, four lanes per cluster (one per vertex of
 * the element being projected). */
#define XPBD_VARIANT_AUTO 0
#define XPBD_VARIANT_SIMPLE 1
#define XPBD_VARIANT_CLUSTER 2
#define XPBD_VARIANT_PATCH 3
#define XPBD_VARIANT_CLUSTER_QUAD 4

/* Signed-distance field kinds (src/geometry/SphereSDF.cpp, BoxSDF.cpp). */
#define XPBD_SDF_NONE 0
#define XPBD_SDF_SPHERE 1
#define XPBD_SDF_BOX 2

typedef struct xpbd_handle xpbd_handle;
typedef struct xpbd_handle xpbd_batch;
typedef struct xpbd_mlp xpbd_mlp;

/* One tetrahedral object.  Replaces what XPBDMeshObject_::setup() reads from
 * the mesh file and config (src/simobject/XPBDMeshObject.cpp:102-160):
 * rest positions, elements, per-element material (the reference maps element
 * classes to ElasticMaterial, XPBDMeshObject.cpp:287-300), fixed vertices
 * (XPBDMeshObjectBase.hpp fixVertex) and the surface faces used by collision. */
typedef struct xpbd_mesh_desc {
    int32_t num_vertices;
    int32_t num_tets;
    const double* x0;        /* 3*V rest positions, xyz per vertex */
    const double* v0;        /* 3*V initial velocities, or NULL for zero */
    const int32_t* tets;     /* 4*T vertex indices */
    const double* E;         /* T Young's moduli [Pa] */
    const double* nu;        /* T Poisson ratios */
    const double* rho;       /* T densities [kg/m^3] */
    const uint8_t* pinned;   /* V flags (nonzero = fixed), or NULL */
    int32_t num_faces;       /* surface triangles for SDF contact, 0 = none */
    const int32_t* faces;    /* 3*F vertex indices */
} xpbd_mesh_desc;

/* Solver / step parameters: time-step, g-accel, num-solver-iters and
 * solver-type of the reference config (SimulationConfig.hpp:185-195,
 * XPBDMeshObjectConfig.hpp:26-125). */
typedef struct xpbd_params {
    double dt;                /* time step [s] */
    double g;                 /* gravity magnitude, acts along -z [m/s^2] */
    int32_t iters;            /* solver iterations per step */
    int32_t solver;           /* XPBD_SOLVER_* */
    int32_t deterministic;    /* Jacobi with the PATCH variant: 1 = every correction (elements,
                               * contacts, attachments, local passes, subset solves) summed in
                               * an order fixed at set-up, bitwise reproducible; 0 = LDS / global float
                               * atomics.  The SIMPLE variant always uses float atomics. */
    int32_t collision_every;  /* re-detect contacts every k steps; 0 = never */
    int32_t variant;          /* XPBD_VARIANT_* */
    /* First-order XPBD (FirstOrderXPBDMeshObject; XPBDMeshObject.cpp:277-348,
     * 450-457; CombinedNeohookeanConstraintProjector.cpp:16-96): the per-vertex
     * damping B_i = V_i b [E_i / (1 + nu_i)] replaces the mass as constraint
     * inertia, alpha~ = alpha / dt, and the inertial move is z -= g m_i dt / B_i.
     * Zero-initialised params give the second-order solver. */
    int32_t first_order;          /* 1 = first-order XPBD */
    double damping_multiplier;    /* b (config damping-multiplier) */
    int32_t adjust_damping;       /* 1 = scale B by E_i / (1 + nu_i) (adjust-damping-to-material) */
    /* Local collision iterations (num-local-collision-iters,
     * XPBDMeshObjectConfig.hpp:68; XPBDMeshObject.cpp:476-483, 764-825): after
     * the global solve, this many extra passes over (a) every element attached to
     * a vertex of a contact whose multiplier is non-zero, once per such
     * (contact, face vertex) pair, then (b) every contact, with the multipliers
     * kept.  0 = none. */
    int32_t local_collision_iters;
} xpbd_params;

/* ---- object life cycle ------------------------------------------------- */

/* Build the solver for one object.  Rest-state precompute (Dm^-1, volume,
 * lumped masses, Lame parameters, compliances) runs in FP64 on the host and is
 * stored in FP32 on the device.  Replaces XPBDMeshObject_::setup +
 * _calculatePerVertexQuantities + _createElasticConstraints
 * (src/simobject/XPBDMeshObject.cpp:102-160, 277-331, 352-417) together with
 * the XPBDSolver constructor (include/solver/xpbd_solver/XPBDSolver.hpp:47-61). */
int xpbd_create(const xpbd_mesh_desc* mesh, const xpbd_params* params, xpbd_handle** out);

/* Release all host and device resources of a handle (NULL is a no-op). */
int xpbd_destroy(xpbd_handle* h);

/* Message of the last failure on h (or of the last failed create on this
 * thread when h is NULL); never NULL, "" when there was none. */
const char* xpbd_last_error(const xpbd_handle* h);

/* Advance nsteps time steps on the device.  One step replaces
 * Simulation::_timeStep for this object (src/simulation/Simulation.cpp:197-276):
 * optional contact re-detection, XPBDMeshObject_::update (predict, solve with
 * lambda reset, pin reset; XPBDMeshObject.cpp:426-505), velocityUpdate
 * (XPBDMeshObject.cpp:507-535).  The solve replaces XPBDSolver::solve
 * (XPBDSolver.hpp:66-97) with XPBDGaussSeidelSolver / XPBDJacobiSolver
 * iteration semantics (XPBDGaussSeidelSolver.hpp:29-80,
 * XPBDJacobiSolver.hpp:79-114).  Steps are replayed from a captured HIP graph. */
int xpbd_step(xpbd_handle* h, int32_t nsteps);

/* Change the number of solver iterations (reference: num-solver-iters). */
int xpbd_set_iterations(xpbd_handle* h, int32_t iters);

/* Block until all work queued on the handle's stream has finished. */
int xpbd_synchronize(xpbd_handle* h);

/* ---- state ---------------------------------------------------------------- */

/* Current positions (3*V, xyz per vertex) in double / float.  Replaces reading
 * Geometry::Mesh::vertices() after the solve (the reference's device path
 * copies back at XPBDParallelJacobiSolver.hpp:111-113). */
int xpbd_get_positions(xpbd_handle* h, double* out);
int xpbd_get_positions_f32(xpbd_handle* h, float* out);
/* Current velocities (3*V) (XPBDMeshObjectBase.hpp vertexVelocity). */
int xpbd_get_velocities(xpbd_handle* h, double* out);
/* Overwrite positions and/or velocities (either may be NULL). */
int xpbd_set_state(xpbd_handle* h, const double* x, const double* v);
/* Lumped vertex masses (V) as computed at set-up (XPBDMeshObject.cpp:277-313). */
int xpbd_get_masses(xpbd_handle* h, double* out);
/* Lagrange multipliers (2*T, (lambda_D, lambda_H) per tet, in the caller's tet
 * order) left by the last solve (CombinedConstraintProjector::lambda()). */
int xpbd_get_lambdas(xpbd_handle* h, double* out);

/* ---- diagnostics: FP64 on the device ------------------------------------- */

/* Constraint residual r = C + alpha/dt^2 * lambda of every elastic constraint
 * after the last solve; ||r||_2 into *norm_out and, if vec_out != NULL, the
 * 2*T entries (Dev, Hyd per tet, caller's tet order).  Replaces
 * XPBDSolver::_calculateConstraintResidual + constraintResidual()
 * (XPBDSolver.hpp:130-135, 327-391). */
int xpbd_constraint_residual(xpbd_handle* h, double* norm_out, double* vec_out);

/* Residual policy (reference "residual-policy", XPBDMeshObjectConfig.hpp:45-53):
 * 0 = never (default), 1 = every step: ||r||_2 is evaluated on the device at the
 * end of every solve, before the pin reset, where XPBDSolver::solve computes it
 * for EVERY_SUBSTEP (XPBDSolver.hpp:91-96).  xpbd_last_step_residual returns
 * the value of the last step (XPBD_E_STATE when the policy is off). */
int xpbd_set_residual_policy(xpbd_handle* h, int32_t policy);
int xpbd_last_step_residual(xpbd_handle* h, double* out);

/* Primary residual of every step's solve (Equation (8) of the XPBD paper),
 * on = 1 / off = 0 (default), recorded at the point of the constraint
 * residual.  The reference's EVERY_SUBSTEP policy asks for both
 * (XPBDSolver.hpp:91-96); here it has its own switch because it costs an FP64
 * pass over every vertex's elements, which the constraint residual that the
 * learned step's fallback reads does not need.  Switching it on allocates 4
 * floats and 3 doubles per vertex, 12 doubles per element, and the vertex ->
 * element and vertex -> face lists.  xpbd_last_step_primary_residual returns, per vertex
 *     R_v = (x_v - xt_v) / w_v - sum grad_v C^T lambda
 * over both constraints of every element and every active contact around v,
 * with xt the step's inertial prediction and w the inverse mass (first order:
 * the inverse damping).  Pinned and attached vertices read zero.  ||R||_2 into
 * *norm_out and, if vec_out != NULL, the 3*V entries (caller's vertex order).
 * Replaces XPBDSolver::_calculatePrimaryResidual + primaryResidual()
 * (XPBDSolver.hpp:127-130; the reference's _calculatePrimaryResidual, :322,
 * is empty, so its vector stays zero).  XPBD_E_STATE when it is off; zero
 * before the first step. */
int xpbd_set_primary_residual(xpbd_handle* h, int32_t on);
int xpbd_last_step_primary_residual(xpbd_handle* h, double* norm_out, double* vec_out);

/* Total strain energy sum(C^2 / alpha) over all deviatoric and hydrostatic
 * constraints (XPBDMeshObject_::totalStrainEnergy, XPBDMeshObject.cpp:537-549). */
int xpbd_strain_energy(xpbd_handle* h, double* out);

/* ---- solver ordering (parity checks) ----------------------------------- */

/* Number of colours of the Gauss-Seidel tet colouring (1 for Jacobi). */
int xpbd_num_colors(xpbd_handle* h, int32_t* out);
/* Order in which the solver visits tets (T entries, caller's tet indices).
 * Colour-GS equals sequential Gauss-Seidel in exactly this order. */
int xpbd_get_tet_order(xpbd_handle* h, int32_t* out);
/* Rank of every surface face in the contact visiting order (F entries). */
int xpbd_get_face_rank(xpbd_handle* h, int32_t* out);
/* Colour boundaries of the visiting order: ncolors + 1 positions into the
 * xpbd_get_tet_order array (colour c visits positions [out[c], out[c+1])).
 * No two tets of one colour in different clusters share a vertex, so a
 * sequential Gauss-Seidel pass may run the clusters of one colour side by side
 * (XPBDGaussSeidelSolver.hpp:29-80 visits the same order one tet at a time).
 * Jacobi: {0, T}. */
int xpbd_get_color_offsets(xpbd_handle* h, int32_t* out);

/* ---- contacts ------------------------------------------------------------- */

/* Kinematic SDF collider for the object: pose = (px, py, pz, qx, qy, qz, qw),
 * params: sphere radius in params[0], box edge lengths in params[0..2].
 * Replaces the fixed RigidSphere / RigidBox colliders whose SDFs are queried by
 * CollisionScene::_collideObjectPair (src/collision/CollisionScene.cpp:171-214).
 * Resets the collider clock: the pose at step n after this call is
 * pose + lin_vel * n * dt (see xpbd_set_sdf_velocity). */
int xpbd_set_sdf(xpbd_handle* h, int32_t kind, const double* pose7, const double* params3);
/* Linear velocity of the collider (m/s), applied from the last xpbd_set_sdf. */
int xpbd_set_sdf_velocity(xpbd_handle* h, const double* vel3);
/* Run contact detection now on the current positions (the detection step of
 * Simulation::_timeStep, Simulation.cpp:202-241) and replace the contact set;
 * clearCollisionConstraints + addStaticCollisionConstraint
 * (XPBDMeshObject.cpp:180-251). */
int xpbd_detect_contacts(xpbd_handle* h);
/* Replace the contact set with the caller's n contacts: face index (caller's
 * face numbering), surface point p (3), normal nrm (3) and barycentrics bary
 * (3) each, multipliers zero; n = 0 clears it.  Replaces
 * XPBDMeshObject_::clearCollisionConstraints followed by one
 * addStaticCollisionConstraint per contact (XPBDMeshObject.cpp:180-251), so a
 * host-side collision scene (the reference's CollisionScene) can drive the
 * device solver.  A face carries at most one contact.  The contacts are
 * projected in every iteration until the next detection (xpbd_detect_contacts,
 * or a step that detects: collision_every > 0 with a collider set) or the next
 * call of this function. */
int xpbd_set_contacts(xpbd_handle* h, int32_t n, const int32_t* faces, const double* p, const double* nrm,
                      const double* bary);

/* XPBDSolver::solve(projector_references, num_iter, initialize)
 * (include/solver/xpbd_solver/XPBDSolver.hpp:106-125) between steps: the n
 * listed elements (caller's indices; an element listed k times is projected k
 * times per pass, back to back) are projected num_iter times in the solver's
 * visiting order, each pass followed by every contact when with_contacts;
 * initialize = 1 starts their multipliers (and the contacts') at zero first
 * (the projectors' initialize()).  Gauss-Seidel applies each projection at
 * once; Jacobi sums a pass from the same positions, per vertex in a fixed
 * order for the deterministic PATCH variant, with float atomics otherwise.  The
 * positions change in place; velocities are not updated (the next step's
 * velocity update sees the moved positions), and pinned vertices are not
 * reset. */
int xpbd_solve_elements(xpbd_handle* h, int32_t n, const int32_t* elements, int32_t num_iter, int32_t initialize,
                        int32_t with_contacts);

/* Copy out the current contact set (at most cap entries), in solver order:
 * face index, surface point p (3), normal n (3), barycentrics (3). */
int xpbd_get_contacts(xpbd_handle* h, int32_t cap, int32_t* count, int32_t* face, double* p,
                      double* n, double* bary);

/* ---- timing --------------------------------------------------------------- */

/* Attachment constraints: vertex vertices[a] is pulled to the world point
 * targets[3a..3a+2] every iteration (alpha = 0, after the elastic and contact
 * projectors).  Replaces XPBDMeshObject_::addAttachmentConstraint /
 * clearAttachmentConstraints (src/simobject/XPBDMeshObject.cpp:255-275) with
 * AttachmentConstraint (include/solver/constraint/AttachmentConstraint.hpp:31-71)
 * and ConstraintProjector::project (ConstraintProjector.hpp:111-172):
 * C = |x - p|, alpha = 0.  targets[3a..3a+2] is the reference's attached point
 * plus offset.  Below |x - p| = 1e-12 no correction is made (the reference's
 * substitute gradient there moves the vertex by ~3e11 |x - p|).  The call
 * replaces the whole list (n = 0 clears); targets may be updated every step
 * with the same vertices.  A vertex may carry one attachment. */
int xpbd_set_attachments(xpbd_handle* h, int32_t n, const int32_t* vertices, const double* targets);

/* Kernel timing mode: when on, xpbd_step launches without the HIP graph and
 * brackets the element-projection launches of every solver iteration (all
 * colour launches of a Gauss-Seidel iteration, back to back; the patch and
 * shared-vertex launches of a Jacobi iteration) with one pair of HIP events on
 * the handle's stream; xpbd_kernel_stats returns the summed milliseconds, the
 * number of solver iterations and the number of timed spans since the last
 * reset. */
int xpbd_set_kernel_timing(xpbd_handle* h, int32_t on);
int xpbd_kernel_stats(xpbd_handle* h, double* proj_ms, int64_t* iterations, int64_t* launches);

/* ---- batched objects -------------------------------------------------------- */

/* Many independent objects advanced together in one captured substep: every
 * launch covers all instances (element and vertex arrays are concatenated
 * with per-instance offsets; each instance keeps its own SDF collider).
 * The reference steps objects one after another (Simulation.cpp:246-255). */
int xpbd_batch_create(int32_t n, const xpbd_mesh_desc* meshes, const xpbd_params* params,
                      xpbd_batch** out);
int xpbd_batch_num_instances(xpbd_batch* b, int32_t* out);
int xpbd_batch_set_sdf(xpbd_batch* b, int32_t instance, int32_t kind, const double* pose7,
                       const double* params3);
int xpbd_batch_get_positions(xpbd_batch* b, int32_t instance, double* out);
/* Advance every instance by nsteps, one graph replay per step covering all of
 * them (the same as xpbd_step on the batch handle, which is also read and
 * destroyed with the single-object calls).  Replaces the reference's loop over
 * objects in Simulation::_timeStep (Simulation.cpp:246-255). */
int xpbd_batch_step(xpbd_batch* b, int32_t nsteps);
/* Time the kernel variants and launch shapes that share this handle's device
 * layout on its own problem (Gauss-Seidel with clusters: the one-thread
 * cluster kernel and the quad kernel at 64/128/256 threads per workgroup, each
 * in its full and its LDS-lean form (vertex 3's weights summed in the quad and
 * vertex bases read across lanes: more instructions, more waves per CU), each
 * with the deltas interleaved and split (XPBD_QUAD_SPLIT);
 * the patch Jacobi kernel with and without the fused one-launch iteration;
 * cluster and patch kernels each with and without the XCD-contiguous
 * workgroup mapping) as replays of a captured graph of the handle's
 * iterations, restore the state, and keep the fastest.  The choice is
 * appended to cache_path (may be NULL) as
 * "shape-key variant block xcd ms-per-iteration" and reused from there for the
 * same shape.  No reference counterpart: the reference has no kernel variants. */
int xpbd_autotune(xpbd_batch* b, const char* cache_path);
/* The kernel variant in use (XPBD_VARIANT_*); block: for CLUSTER_QUAD the
 * workgroup size (negative for the LDS-lean form; magnitude XPBD_QUAD_SPLIT +
 * size when the iterations keep the vertex deltas in their own array), for
 * PATCH 1 when the fused one-launch iteration is in use
 * (deterministic mode without attachments), else 0; xcd: whether the
 * XCD-contiguous workgroup mapping is on (block and xcd may be NULL). */
int xpbd_get_variant(xpbd_handle* h, int32_t* variant, int32_t* block, int32_t* xcd);
/* Added to a quad block size: the Gauss-Seidel iterations copy the vertex
 * deltas out of the interleaved vertex records into an array of their own
 * before the loop and back after it, so that the deltas of neighbouring
 * vertices share 32-byte write sectors (same results).  Handles with
 * attachments keep the interleaved layout. */
#define XPBD_QUAD_SPLIT 1000
/* Batch handles are also plain handles: xpbd_step, xpbd_destroy, ... apply. */

/* ---- host-side mesh utilities (no GPU needed) ------------------------------ */

/* Structured grid of nx*ny*nz cells split 6 tets per cell (conforming);
 * x: 3*(nx+1)*(ny+1)*(nz+1), tets: 24*nx*ny*nz.  Vertex (i,j,k) has index
 * i + (nx+1)*(j + (ny+1)*k). */
int xpbd_mesh_grid(int32_t nx, int32_t ny, int32_t nz, double hx, double hy, double hz,
                   const double* origin3, double* x, int32_t* tets);
/* Boundary triangles (faces of exactly one tet); faces must hold 12*T ints.
 * Replaces the surface-face search of TetMesh (src/geometry/TetMesh.cpp:123-141). */
int xpbd_mesh_surface(int32_t num_vertices, int32_t num_tets, const int32_t* tets,
                      int32_t* num_faces, int32_t* faces);
/* Greedy conflict colouring (items of k vertices; no shared vertex within a
 * colour).  Returns the colour count in *num_colors. */
int xpbd_color_items(int32_t num_vertices, int32_t num_items, int32_t verts_per_item,
                     const int32_t* items, int32_t* colors, int32_t* num_colors);
/* GMSH 4.1 ASCII reader (src/utils/MeshUtils.cpp:92-218): first call with
 * NULL arrays to get the counts, then again with arrays of that size. */
int xpbd_read_msh(const char* path, int32_t* num_vertices, int32_t* num_tets, int32_t* num_tris,
                  double* x, int32_t* tets, int32_t* tris);
/* Element classes file (element-classes-filename, XPBDMeshObjectConfig.hpp:63;
 * read by XPBDMeshObject.cpp:112-149 with FileUtils::readVectorFromFile<int>):
 * one integer per line, up to the first line without one.  *count receives
 * the number read; classes (may be NULL) the first min(cap, count) of them.
 * Each class indexes the object's material list (XPBDMeshObject.cpp:287-300). */
int xpbd_read_element_classes(const char* path, int32_t cap, int32_t* classes, int32_t* count);

/* ---- learned warm start ----------------------------------------------------- */

#define XPBD_MLP_IN 24
#define XPBD_MLP_HIDDEN 128
#define XPBD_MLP_OUT 3

/* A 24 -> 128 -> 128 -> 3 per-vertex MLP with SiLU, BF16 weights and
 * activations, FP32 accumulation on MFMA.  No reference counterpart: the
 * learned warm start is new in this solver.  Parameter layout (float32,
 * 20,099 values): W1[128][24], b1[128], W2[128][128], b2[128], W3[3][128], b3[3]. */
int xpbd_mlp_create(xpbd_mlp** out);
int xpbd_mlp_destroy(xpbd_mlp* m);
const char* xpbd_mlp_last_error(const xpbd_mlp* m);
int xpbd_mlp_num_params(int32_t* out);
int xpbd_mlp_set_params(xpbd_mlp* m, const float* params);
int xpbd_mlp_get_params(xpbd_mlp* m, float* params);
/* Feature / target scaling: y_model = scale_out * MLP((x - mean_in) / std_in). */
int xpbd_mlp_set_scaling(xpbd_mlp* m, const float* mean_in24, const float* std_in24, float scale_out);
/* Network file: "XPBDMLP1", int32 parameter and feature counts (20,099, 24),
 * the float32 parameters in the layout above, mean_in[24], std_in[24],
 * scale_out (little-endian).  Save needs the scaling set; load checks the tag
 * and the counts and sets parameters and scaling. */
int xpbd_mlp_save(xpbd_mlp* m, const char* path);
int xpbd_mlp_load(xpbd_mlp* m, const char* path);
/* Forward on n host rows (n x 24) -> out (n x 3). */
int xpbd_mlp_forward(xpbd_mlp* m, int32_t n, const float* x, float* out);
/* Forward without device allocation: the same inference as xpbd_mlp_forward
 * (the same kernel, the same results), with the rows staged through a fixed
 * file-scope device array of XPBD_MLP_NOALLOC_MAX_ROWS rows (8 MB, held on
 * every GPU the library runs on until the process ends) instead of a
 * workspace grown with hipMalloc.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (a serving loop, code under HIP
 * graph capture).  n > XPBD_MLP_NOALLOC_MAX_ROWS is refused with
 * XPBD_E_INVALID_ARG.  All calls on one GPU share the array: calls that
 * overlap in time on one GPU (from several threads or networks) race. */
#define XPBD_MLP_NOALLOC_MAX_ROWS 65536
int xpbd_mlp_forward_noalloc(xpbd_mlp* m, int32_t n, const float* x, float* out);
/* Loss and parameter gradient on n rows (host arrays): the mean over n*3 of
 * (MLP(x) - y / scale_out)^2, the squared error in units of scale_out; grad
 * receives 20,099 floats.  Used by the gradient checks; training minimises
 * the same loss. */
int xpbd_mlp_loss_grad(xpbd_mlp* m, int32_t n, const float* x, const float* y, double* loss,
                       float* grad);
/* Adam training on device-resident samples: epochs passes of minibatches. */
int xpbd_mlp_train(xpbd_mlp* m, int32_t n, const float* x, const float* y, int32_t batch,
                   int32_t epochs, float lr, uint64_t seed, double* final_loss);

/* Record training pairs from the plain solver on h: for nsteps steps, the 24
 * features of every free vertex after the prediction and the solver's
 * position correction (x_solved - x_predicted).  Arrays hold cap rows. */
int xpbd_collect_training(xpbd_handle* h, int32_t nsteps, int32_t cap, float* x, float* y,
                          int32_t* rows);
/* Features of every vertex for the current state (V x 24), for inspection. */
int xpbd_features(xpbd_handle* h, float* out);
/* Attach (or detach with NULL) a network to the handle for learned steps. */
int xpbd_attach_mlp(xpbd_handle* h, xpbd_mlp* m);
/* One learned step: predict, add the network's correction to the predicted
 * positions of the free vertices, start the solve as xpbd_set_learned_start
 * selects, then iters solver iterations.  If ||r||_2 (r = C + alpha~ lambda at
 * the end of the solve) exceeds fallback_ratio * ref_residual, the step is
 * redone from the saved start-of-step state with the handle's full iteration
 * count, which reproduces the plain step exactly; *fell_back reports it and
 * *residual the final ||r||_2.  ref_residual <= 0 disables the fallback.
 * Everything up to the fallback decision (and the rest of the step, which a
 * fallback redoes) replays one captured graph per network, iteration count,
 * start mode and detection; in kernel timing mode the step launches eagerly
 * and times the network with events (xpbd_learned_stats).  The same bits
 * either way. */
int xpbd_step_learned(xpbd_handle* h, int32_t iters, double ref_residual, double fallback_ratio,
                      int32_t* fell_back, double* residual);

/* Start of the learned step's solve.
 * POSITIONS: every multiplier from zero, as the reference solve
 *   (XPBDSolver.hpp:66-77); the network only moves the starting positions.
 * MULTIPLIERS (default): the contacts are projected once from the corrected
 *   positions (contact multipliers from zero), then every element's
 *   multipliers start at lambda = -C / alpha~ there (C in FP64), the values
 *   that satisfy its constraints; attachment multipliers start at zero and
 *   the iterations keep all of them.  The elastic residual then starts at the
 *   rounding of the FP32 multipliers whatever the correction, so the
 *   residual test no longer judges the correction: the trajectory is as
 *   accurate as the network (the positions are not moved by
 *   M^-1 grad C^T lambda, the implicit-Euler momentum balance).
 * CONSISTENT: every element's multipliers start at lambda0 = -C / alpha~ at
 *   the corrected positions xt + corr (C and its gradients in FP64), and the
 *   positions at xt + M^-1 sum grad C(xt + corr)^T lambda0, so the momentum
 *   balance M (x - xt) = grad C^T lambda that the XPBD updates preserve holds
 *   from the first iteration; contact and attachment multipliers start at
 *   zero.  From the exact correction this start is the step's solution.  The
 *   first call with this mode builds the vertex -> element slot lists (16
 *   bytes and 12 doubles per element). */
#define XPBD_LEARNED_START_POSITIONS 0
#define XPBD_LEARNED_START_MULTIPLIERS 1
#define XPBD_LEARNED_START_CONSISTENT 2
int xpbd_set_learned_start(xpbd_handle* h, int32_t mode);
/* Counters of the learned steps on h: steps taken, fallbacks, and (with kernel
 * timing on) the summed milliseconds and number of network inferences. */
int xpbd_learned_stats(xpbd_handle* h, int64_t* steps, int64_t* fallbacks, double* mlp_ms,
                       int64_t* mlp_calls);

#ifdef __cplusplus
}
#endif

#endif /* XPBD_HIP_H */
