/*
 * hfpx3d.h -- C ABI of the MI355X (gfx950) HFPx3D hot path.
 *
 * Plain C types only: device pointers are `void*`/`double*`/`float*`
 * allocated by the caller on the context's device, host pointers are marked
 * `_host`.  Every function returns an hfpx status code (HFPX_OK = 0) and never
 * aborts; hfpx_last_error() gives the message of the last failure on a
 * context.  Arrays are column-major with explicit leading dimensions, the
 * layout of il::Array2D (il/container/2d/Array2D.h:233-263 in the
 * reference).  A context is not thread-safe; distinct contexts are
 * independent.  All work is enqueued on the context's stream; functions that
 * return host-visible results (info, reports) synchronise that stream.
 */
#ifndef HFPX3D_H
#define HFPX3D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define HFPX_OK 0
#define HFPX_ERR_INVALID 1        /* invalid argument / shape              */
#define HFPX_ERR_NOMEM 2          /* device or host allocation failed      */
#define HFPX_ERR_SINGULAR 3       /* exact zero pivot (info = column, 1-based) */
#define HFPX_ERR_NOT_CONVERGED 4  /* iterative refinement did not converge */
#define HFPX_ERR_HIP 5            /* HIP runtime error                     */

#define HFPX_F64 0
#define HFPX_F32 1

typedef struct hfpx_ctx hfpx_ctx;
typedef struct hfpx_mesh hfpx_mesh;

/* Outcome of the mixed-precision solve. */
typedef struct hfpx_ir_report {
  int iterations;        /* outer refinement steps taken (classic + GMRES) */
  int inner_iterations;  /* GMRES steps summed over the GMRES outer steps */
  double backward_error; /* max over columns of |b-Ax|/(|A||x|+|b|), inf-norms */
  int converged;         /* 1 if backward_error <= tol                     */
  int fell_back;         /* 1 if refinement failed and the solution came
                            from an FP64 LU of a copy of A64               */
} hfpx_ir_report;

/* ---- context ---------------------------------------------------------- */
/* Create a context on `device`; `stream` is a hipStream_t to use (NULL: the
 * library creates its own default-priority stream). */
int hfpx_create(hfpx_ctx** ctx, int device, void* stream);
/* Free the context's device memory.  With a caller's stream, the caller
 * finishes the work it enqueued (the library does not touch a stream it
 * does not own here); call it while the HIP runtime is still up, not from
 * exit-time destructors. */
int hfpx_destroy(hfpx_ctx* ctx);
const char* hfpx_status_string(int status);
const char* hfpx_last_error(const hfpx_ctx* ctx);
/* Stream the context enqueues on (a hipStream_t). */
void* hfpx_stream(const hfpx_ctx* ctx);
int hfpx_synchronize(hfpx_ctx* ctx);

/* ---- mesh ------------------------------------------------------------- */
/* Upload a crack mesh.  Replaces the mesh/DoF containers consumed by
 * hfp3d::make_3dbem_matrix_s (src/Core/surface_mesh_utilities.h:25-47):
 *   nods_host  nods_rows x n_nodes, rows 0..2 = x, y, z (extra rows ignored)
 *   conn_host  conn_rows x n_el, rows 0..2 = 0-based vertex ids
 *   dof_h_host n_el x 18 (col-major, -1 = fixed DoF) as built by
 *              hfp3d::make_dof_h_crack (surface_mesh_utilities.cpp:28-112),
 *              or NULL for the identity map (fix_tip = 0, n_dof = 18 n_el)
 *   beta       collocation offset (Num_Param_T::beta, model_parameters.h:78)
 */
int hfpx_mesh_create(hfpx_ctx* ctx, const double* nods_host, int64_t nods_rows,
                     int64_t n_nodes, const int64_t* conn_host, int64_t conn_rows,
                     int64_t n_el, const int64_t* dof_h_host, int64_t n_dof,
                     double beta, hfpx_mesh** mesh);
int hfpx_mesh_destroy(hfpx_mesh* mesh);
/* A device mesh straight from the reference's npy mesh files: nodes_npy
 * ((3 or 6) x n_nodes, <f8 or <f4, either memory order) and elems_npy
 * ((3 or 5) x n_el, <i4 or <i8); index_origin is subtracted from the
 * vertex rows (1 for the reference's Mesh_Files); the DoF handle is built
 * for tip_type 0..2 (make_dof_h_crack) and beta is the collocation offset.
 * hfpx_mesh_n_dof gives the DoF count.  Replaces load_mesh_from_numpy_32 /
 * _64 (src/IO/mesh_file_io.cpp:186-225, declared mesh_file_io.h:53-70) followed
 * by make_dof_h_crack and the upload; a missing or malformed file is
 * HFPX_ERR_INVALID with the reason in hfpx_last_error. */
int hfpx_mesh_create_npy(hfpx_ctx* ctx, const char* nodes_npy, const char* elems_npy, int64_t index_origin,
                         int tip_type, double beta, hfpx_mesh** mesh);
/* DD in each element's own frame (1) or global (0, the default): the
 * assembly's columns, the volume-control row and the DD given to
 * hfpx_stress follow it.  Replaces Num_Param_T::is_dd_local
 * (src/Core/model_parameters.h:89; branches at system_assembly.cpp:186-195,
 * 358-387, 529-536). */
int hfpx_mesh_set_dd_local(hfpx_mesh* mesh, int is_dd_local);
/* Non-uniform edge partition: per-node vertex weights wts_host (n_nodes,
 * host, every weight finite and > 0).  Element e takes the weights w of its
 * three vertex nodes; its edge node opposite vertex v then sits at
 * (w_a x_a + w_b x_b) / (w_a + w_b), (a, b) = (v+1, v+2), with the quadratic
 * Lagrange shape functions of those six nodes and collocation points pulled
 * from them towards the centroid by beta.  Assembly, right-hand sides, the
 * volume-control border and hfpx_stress follow it.  wts_host = NULL restores
 * the uniform element.  Replaces the vertex-weight hooks of
 * hfp3d::make_el_sfm_nonuniform / el_cp_nonuniform
 * (src/Core/element_utilities.cpp:240-322, 397-431, declared
 * element_utilities.h:68-71, 87-90; call sites left for them at
 * system_assembly.cpp:67-81, 92-115). */
int hfpx_mesh_set_vertex_weights(hfpx_ctx* ctx, hfpx_mesh* mesh, const double* wts_host, int64_t n_nodes);
int64_t hfpx_mesh_n_dof(const hfpx_mesh* mesh);

/* ---- assembly ---------------------------------------------------------- */
/* Dense influence matrix A (n_dof x n_dof, FP64, device, lda >= n_dof):
 * element geometry + every (collocation point, source element) traction
 * block, written once per entry.  Replaces hfp3d::make_3dbem_matrix_s
 * (src/Solvers/system_assembly.cpp:30-219, declared system_assembly.h:35-39). */
int hfpx_assemble(hfpx_ctx* ctx, hfpx_mesh* mesh, double shear_m, double poiss_r,
                  double* A, int64_t lda);
/* Same matrix rounded to FP32 (device, lda32), optionally together with the
 * FP64 copy in the same pass (A64 may be NULL).  No reference counterpart:
 * the FP32 copy is the factorisation operand of the mixed-precision solve. */
int hfpx_assemble_mixed(hfpx_ctx* ctx, hfpx_mesh* mesh, double shear_m, double poiss_r,
                        double* A64, int64_t lda64, float* A32, int64_t lda32);

/* The element integral on the reference's own arguments, for n_cases
 * independent cases (device arrays): h[c] (minus the point's coordinate
 * along the element normal), z[2c..2c+1] (the point's in-plane coordinates,
 * re / im), tau[6c..6c+5] (the three vertices' in-plane coordinates, re / im
 * pairs) and sfm[72c..72c+71] (the 6 x 6 complex shape-function coefficients,
 * column-major, re / im interleaved; row = node, column = monomial 1, tau,
 * conj(tau), tau^2, conj(tau)^2, tau conj(tau)), all in the element's own
 * frame.  out[108c..108c+107] receives the 6 x 18 stress block, column-major:
 * rows S11 S22 S33 S12 S13 S23, column 3 node + DD component.  kernel_id
 * must be 1 (the hypersingular traction kernel; the reference's 0 is an
 * empty branch).  Replaces hfp3d::make_local_3dbem_submatrix
 * (src/Elasticity/elasticity_kernel_integration.cpp:34-205, declared
 * elasticity_kernel_integration.h:25-29), one case per thread. */
int hfpx_local_submatrix(hfpx_ctx* ctx, int kernel_id, double shear_m, double poiss_r, int64_t n_cases,
                         const double* h, const double* z, const double* tau, const double* sfm, double* out);

/* Right-hand sides for nrhs far-field stress tensors sinf_host (6 x nrhs,
 * order xx yy zz xy xz yz, tension positive): B(:, r) -= traction of
 * sinf(:, r) on every free DoF (B device, n_dof x nrhs, ldb).  Replaces
 * hfp3d::add_s_inf_to_3dbem_rhs (src/Solvers/system_assembly.cpp:222-274). */
int hfpx_rhs_sinf(hfpx_ctx* ctx, hfpx_mesh* mesh, const double* sinf_host, int nrhs,
                  double* B, int64_t ldb);

/* ---- volume control -------------------------------------------------------- */
/* Volume-control matrix (device, (n_dof+1) x (n_dof+1), lda >= n_dof + 1):
 * the influence matrix of hfpx_assemble in rows/columns 0..n_dof-1, row
 * n_dof = opening volume per unit DD (area integral of the DoF's shape
 * function times the element normal's component), column n_dof = -(element
 * normal component) (traction per unit pressure), corner 0.  FP64 (A64)
 * and/or its FP32 rounding (A32) in one pass; either may be NULL.  With
 * hfpx_mesh_set_dd_local(mesh, 1) the row holds the area integral on the
 * opening (local component 2) only.  Replaces
 * hfp3d::make_3dbem_matrix_vc (src/Solvers/system_assembly.cpp:458-557,
 * declared system_assembly.h:59-64); the reference leaves the corner unset. */
int hfpx_assemble_vc(hfpx_ctx* ctx, hfpx_mesh* mesh, double shear_m, double poiss_r, double* A64,
                     int64_t lda64, float* A32, int64_t lda32);
/* Truncated volume-control system from a resident one.  A (device, lda):
 * the (orig_ndof+1)^2 volume-control matrix of the original DoF handle
 * orig_dof_h_host (n_el x ndpe, col-major, -1 = fixed); dof_h_host (same
 * shape) numbers the used DoF 0..used_ndof-1 (-1 = not used).  Writes
 *   M(t, s) = sum of A(o_a, o_b) over used slots a -> t, b -> s,
 *   M(used, s) / M(s, used) = sums of the border row / column, M(used, used) = 0
 * to M64 (ldm64) and/or its FP32 rounding M32 (ldm32), and rhs (device,
 * used_ndof + 1): rhs[s] = delta_t_host[...] (indexed by slot e * ndpe + j,
 * original DoF or used DoF when tsize is n_el * ndpe, orig_ndof or
 * used_ndof; the last slot of a DoF wins), rhs[used_ndof] = delta_v.
 * Synchronises the stream.  Replaces hfp3d::mod_3dbem_system_vc
 * (src/Solvers/system_assembly.cpp:560-629); the border is read at index
 * orig_ndof, which is the reference's n_el * ndpe whenever the original
 * handle has no fixed DoF (the only case the reference keeps in bounds). */
int hfpx_vc_truncate(hfpx_ctx* ctx, const double* A, int64_t lda, int64_t orig_ndof,
                     const int64_t* orig_dof_h_host, const int64_t* dof_h_host, int64_t n_el, int64_t ndpe,
                     int64_t used_ndof, const double* delta_t_host, int64_t tsize, double delta_v, double* M64,
                     int64_t ldm64, float* M32, int64_t ldm32, double* rhs);
/* ---- cohesion / friction state of the crack surface ---------------------- */
/* Friction / cohesion parameters of one surface material (F_C_Param_T,
 * src/Core/model_parameters.h:22-30). */
typedef struct hfpx_fc_param {
  double cr_open;  /* critical opening */
  double peak_ts;  /* peak tensile stress */
  double cr_slip;  /* critical slip */
  double peak_sc;  /* peak shear cohesion */
  double peak_sf;  /* peak friction coefficient */
  double res_sf;   /* residual friction coefficient */
} hfpx_fc_param;
/* Friction coefficient, shear cohesion and opening cohesion of n_nod nodes
 * (device arrays of n_nod) from their DD (device, n_nod x 3 col-major, ldd;
 * components used as given: shear 0, 1 and opening 2), material ids mat_id
 * (device, 0 <= id < n_mat), the largest relative opening / slip reached
 * (mr_open, mr_slip, device) and, per material, the parameters
 * f_c_param_host (n_mat) and model ids c_model_id_host (host, n_mat x 2
 * col-major: cohesion 0 box / 1 linear softening, friction 0 constant /
 * 1 linear slip weakening; other ids give no cohesion / residual
 * friction).  Synchronous.  A material id outside [0, n_mat) or a zero
 * critical opening or slip is HFPX_ERR_INVALID.  Replaces hfp3d::match_f_c
 * (src/Development/cohesion_friction.cpp:22-127, declared
 * cohesion_friction.h:22-29). */
int hfpx_match_f_c(hfpx_ctx* ctx, int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                   const int32_t* c_model_id_host, const hfpx_fc_param* f_c_param_host, int64_t n_mat,
                   const double* mr_open, const double* mr_slip, double* friction_coef,
                   double* slip_cohesion, double* open_cohesion);
/* Largest relative opening (mr_open) and slip (mr_slip) reached, updated in
 * place from the DD (same layouts as hfpx_match_f_c).  Synchronous.
 * Replaces hfp3d::update_mrd (src/Development/cohesion_friction.cpp:131-189,
 * declared cohesion_friction.h:34-41). */
int hfpx_update_mrd(hfpx_ctx* ctx, int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                    const hfpx_fc_param* f_c_param_host, int64_t n_mat, double* mr_open, double* mr_slip);
/* ---- cohesion / friction iteration on a resident volume-control matrix ---- */
typedef struct hfpx_cf_sys hfpx_cf_sys;
/* Outcome of one step of hfpx_vc_cf_iteration. */
typedef struct hfpx_cf_report {
  int64_t n_dof;      /* DoF active after the step (0: nothing moved)          */
  int refactored;     /* 1 if the truncated system was gathered and factorised
                         (its active set differs from the cached factors')    */
  double pressure;    /* pressure after the step                              */
  double residual;    /* |increments and pressure increment|_1 +
                         |traction adjustments|_1                             */
  hfpx_ir_report ir;  /* refinement of the truncated system                   */
} hfpx_cf_report;
/* Iteration state on the volume-control matrix A_vc (device, (n_dof+1)^2,
 * lda >= n_dof + 1, as written by hfpx_assemble_vc for `mesh`, whose DoF
 * map is the original handle; every DoF on one slot).  A_vc and mesh must
 * outlive it.  Holds the truncated system (FP64 and FP32, (n_dof+1)^2
 * each, allocated here) and the FP32 factors of the last active set. */
int hfpx_cf_create(hfpx_ctx* ctx, hfpx_mesh* mesh, const double* A_vc, int64_t lda, hfpx_cf_sys** out);
int hfpx_cf_destroy(hfpx_cf_sys* sys);
/* One step of the volume-control cohesion / friction iteration.  Device
 * arrays of the 6 n_el nodes: mat_id, dd (6 n_el x 3, ldd; global, or
 * element frames with hfpx_mesh_set_dd_local) and pp (nodal pressure) are
 * updated in place, mr_open / mr_slip read, friction_coef / slip_cohesion /
 * open_cohesion written (hfpx_match_f_c of the entry DD).  Host: the model
 * ids and parameters as for hfpx_match_f_c, sinf_host (6, xx yy zz xy xz yz,
 * tension positive), the injection node (element inj_el, local node
 * inj_node), the injected volume new_vol, and the current DoF handle
 * dof_h_host (n_el x 18 col-major, a subset of the mesh's original DoF,
 * -1 = not active) with *n_dof, replaced by the new handle (the active
 * original DoF renumbered in their original order).  The truncated system
 * is solved by FP32 LU + FP64 refinement (tol, max_outer as for
 * hfpx_solve_ir); its FP32 factors are reused while the active set stays
 * the same.  Synchronous.  Replaces hfp3d::vc_cf_iteration
 * (src/Development/c_f_iteration.cpp:27-436, declared c_f_iteration.h:24-38)
 * with the traction adjustment entering the system as -R^T dt in all three
 * components (the reference adds +R^T dt in two, :279-285) and the opening
 * of a collocation point taken in the element frame (the reference rotates
 * it twice, :404-409). */
int hfpx_vc_cf_iteration(hfpx_ctx* ctx, hfpx_cf_sys* sys, const int32_t* mat_id, const int32_t* c_model_id_host,
                         const hfpx_fc_param* f_c_param_host, int64_t n_mat, const double* sinf_host,
                         int64_t inj_el, int64_t inj_node, double new_vol, double* dd, int64_t ldd, double* pp,
                         const double* mr_open, const double* mr_slip, double* friction_coef,
                         double* slip_cohesion, double* open_cohesion, int64_t* dof_h_host, int64_t* n_dof,
                         double tol, int max_outer, hfpx_cf_report* rep);
/* ---- postprocessing ------------------------------------------------------ */
/* Stress at n_pts observation points: S (device, n_pts x 6, col-major, ldS,
 * order Sxx Syy Szz Sxy Sxz Syz, tension positive) = sinf_host (6 values) +
 * the stress of the nodal DD dd (device, 6 n_el x 3 col-major, ldd, global
 * components, row 6 e + node) at pts (device, 3 x n_pts col-major).
 * Deterministic (fixed summation order).  Replaces
 * hfp3d::make_3dbem_stress_f_s (src/Solvers/system_assembly.cpp:279-455,
 * declared system_assembly.h:49-54). */
int hfpx_stress(hfpx_ctx* ctx, hfpx_mesh* mesh, double shear_m, double poiss_r, const double* dd,
                int64_t ldd, const double* pts, int64_t n_pts, const double* sinf_host, double* S,
                int64_t ldS);

/* hfpx_stress without any device allocation (same arguments and result).
 * No counterpart in the reference: for callers that cannot allocate device
 * memory -- code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc.  Its workspace (the DD in element frames and per-chunk
 * partial sums) is a 64 MiB device array of the library, one per GPU, held
 * until the process ends: 18 n_el + 6 n_pts must not exceed 8,388,608
 * doubles (else HFPX_ERR_INVALID; fewer element chunks are used where the
 * allocating form would need more).  The far-field stress is passed by
 * value.  Enqueues on the context's stream and returns without
 * synchronising.  Calls on one GPU share the array: calls on one stream are
 * ordered, calls that overlap in time (other streams or contexts on the same
 * GPU) race. */
int hfpx_stress_noalloc(hfpx_ctx* ctx, hfpx_mesh* mesh, double shear_m, double poiss_r, const double* dd,
                        int64_t ldd, const double* pts, int64_t n_pts, const double* sinf_host, double* S,
                        int64_t ldS);

/* ---- dense solve -------------------------------------------------------- */
/* In-place LU with partial pivoting of the n x n matrix A (device, lda),
 * dtype HFPX_F64 or HFPX_F32: P A = L U, ipiv (device, n int32) 1-based
 * sequential row interchanges, *info_host = 0 or the first column (1-based)
 * with an exactly zero pivot, in which case HFPX_ERR_SINGULAR is returned and
 * the factorisation is still completed, as LAPACK does.  Replaces the
 * il::LU<il::Array2D<double>> constructor
 * (il/linear_algebra/dense/factorization/LU.h:79-99, LAPACKE_dgetrf). */
int hfpx_getrf(hfpx_ctx* ctx, int dtype, void* A, int64_t lda, int64_t n, int32_t* ipiv,
               int64_t* info_host);
/* B (n x nrhs, device, ldb) := A^{-1} B from hfpx_getrf's factors.  Replaces
 * il::LU::solve (LU.h:106-138, LAPACKE_dgetrs 'N'). */
int hfpx_getrs(hfpx_ctx* ctx, int dtype, const void* LU, int64_t lda, const int32_t* ipiv,
               int64_t n, void* B, int64_t ldb, int64_t nrhs);
/* Mixed-precision solve of A X = B for nrhs columns: FP32 LU factors of A
 * (LU32, ipiv from hfpx_getrf with HFPX_F32) precondition FP64 iterative
 * refinement whose residuals use the FP64 matrix A64 (left untouched):
 * classic steps x += (LU)^-1 r while each lowers the backward error at least
 * 100-fold, GMRES(m) steps from the first that does not (HFPX_KNOB_IR_MODE).
 * X (device, ldx) receives the solution; refinement stops when every
 * column's normwise backward error is <= tol or after max_outer steps
 * (then HFPX_ERR_NOT_CONVERGED, X holding the last iterate and rep the
 * achieved error).  Refinement that stagnates or meets a non-finite iterate
 * is replaced by an FP64 LU solve of a copy of A64 (HFPX_KNOB_IR_FALLBACK;
 * rep->fell_back = 1); the backward error is NaN when A64 or B holds a
 * non-finite value.  No reference counterpart: it stands in for il::LU::solve
 * (LU.h:106-138) on the FP64 matrix. */
int hfpx_solve_ir(hfpx_ctx* ctx, const double* A64, int64_t lda64, const float* LU32,
                  int64_t ldlu32, const int32_t* ipiv, int64_t n, const double* B, int64_t ldb,
                  double* X, int64_t ldx, int64_t nrhs, double tol, int max_outer, int gmres_m,
                  hfpx_ir_report* rep);
/* C (M x N) -= A (M x K) B (K x N), column-major, on MFMA: the
 * Schur-complement update of hfpx_getrf, exposed for testing and reuse.
 * No reference counterpart (LAPACK-internal). */
int hfpx_gemm_sub(hfpx_ctx* ctx, int dtype, void* C, int64_t ldc, const void* A, int64_t lda,
                  const void* B, int64_t ldb, int64_t M, int64_t N, int64_t K);

/* ---- measurement --------------------------------------------------------- */
/* Kernel classes timed by the opt-in profiler. */
#define HFPX_KC_ASSEMBLE 0
#define HFPX_KC_GEMM_F64 1
#define HFPX_KC_GEMM_F32 2
#define HFPX_KC_PANEL 3
#define HFPX_KC_LASWP 4
#define HFPX_KC_TRSM 5
#define HFPX_KC_STRESS 6
#define HFPX_KC_VC 7
#define HFPX_KC_COUNT 8
/* Start (enable = 1) or stop (0) recording a HIP event pair around every
 * launch of the classes above on the context's stream; starting clears the
 * previous records.  No reference counterpart (the reference times whole
 * stages with il::Timer, main.cpp:89-175). */
int hfpx_profile(hfpx_ctx* ctx, int enable);
/* Restrict recording to the classes whose bit (1 << HFPX_KC_*) is set in
 * mask (default: every class), so that a timed run carries events only
 * around the launches it reports.  No reference counterpart. */
int hfpx_profile_classes(hfpx_ctx* ctx, uint32_t mask);
/* Totals for one kernel class since the last start: launches, summed kernel
 * time (ms), algorithmic flops and algorithmic HBM bytes.  Synchronises. */
int hfpx_profile_read(hfpx_ctx* ctx, int kernel_class, int64_t* launches, double* ms, double* flops,
                      double* bytes);

/* ---- batches of independent cracks -------------------------------------- */
/* No reference counterpart (the reference solves one crack per run,
 * src/main.cpp:89-175): many cracks of uneven size assembled and solved
 * together, each as make_3dbem_matrix_s + add_s_inf_to_3dbem_rhs + il::LU
 * with fix_tip = 0 (all 18 n_el DoF free; at most 512 elements per crack).
 * Crack i has n_nodes[i] nodes and n_el[i] elements; nods_host (3 x sum
 * n_nodes, rows x y z) and conn_host (3 x sum n_el, crack-local 0-based node
 * ids) hold the cracks one after another; shear_m[i], poiss_r[i] and
 * sinf_host[6 i .. 6 i + 5] (xx yy zz xy xz yz, tension positive) are its
 * material and far-field load.  Cracks with n >= mixed_min_n are solved with
 * an FP32 LU and FP64 iterative refinement, the others with an FP64 LU.  The
 * batch owns its device memory (FP64 matrices, FP32 copies, vectors). */
typedef struct hfpx_batch hfpx_batch;
int hfpx_batch_create(hfpx_ctx* ctx, int n_cracks, const int64_t* n_nodes, const int64_t* n_el,
                      const double* nods_host, const int64_t* conn_host, double beta, const double* shear_m,
                      const double* poiss_r, const double* sinf_host, int64_t mixed_min_n, hfpx_batch** batch);
int hfpx_batch_destroy(hfpx_batch* batch);
/* Assemble every crack's matrix and right-hand side (no solve). */
int hfpx_batch_assemble(hfpx_ctx* ctx, hfpx_batch* batch);
/* Assemble and solve every crack: batched LU (FP64 group and FP32 group),
 * solves, n_refine refinement steps for the FP32 group, per-crack backward
 * errors.  use_graph = 1 captures the whole launch sequence into a hipGraph
 * on the first call (and whenever n_refine changes) and replays it. */
int hfpx_batch_solve(hfpx_ctx* ctx, hfpx_batch* batch, int n_refine, int use_graph);
/* Crack `crack` (input order): solution (n doubles, DoF order) to x_host (may
 * be NULL), *n_out = its DoF count, *berr = normwise backward error
 * max|b-Ax| / (|A|_inf max|x| + max|b|) against its FP64 matrix, *info = first
 * zero pivot column (1-based) or 0.  Any pointer may be NULL. */
int hfpx_batch_solution(hfpx_ctx* ctx, hfpx_batch* batch, int crack, double* x_host, int64_t* n_out,
                        double* berr, int64_t* info);
/* Crack `crack`'s right-hand side (n doubles) to host memory. */
int hfpx_batch_rhs(hfpx_ctx* ctx, hfpx_batch* batch, int crack, double* b_host);
/* Copy crack `crack`'s assembled FP64 matrix (n x n column-major) to A_host
 * (may be NULL; *n_out = n).  Solving leaves it unchanged (the FP64 group is
 * factorised in a working copy). */
int hfpx_batch_matrix(hfpx_ctx* ctx, hfpx_batch* batch, int crack, double* A_host, int64_t* n_out);

/* ---- tuning -------------------------------------------------------------- */
/* Kernel-variant knobs of a context (no reference counterpart).
 * HFPX_KNOB_GEMM_F32: FP32 trailing-update GEMM (products with more than 64
 * columns): 0 = automatic (15 when M, N >= 4096 and K >= 512; 17 when
 * M >= 256, N >= 128 and K >= 64; else 7 for K <= 384, 2 for longer K); 1 = 16x16x4 FP32 MFMA, 128 x 128 tile;
 * 2..12 = 32x32x2 FP32 MFMA, K chunk 16, with 2 = 128 x 128 tile, no
 * options, 3 = unguarded interior tiles (edge bands in separate launches),
 * 4 = 3 + pipelined LDS operands, 5 = 4 + XCD-grouped tile order, 6 = 3 +
 * XCD-grouped tile order, 7 = 2 and 8 = 6 with C loaded into the
 * accumulators first, 9 = 6 and 11 = 8 with 256 x 128 tiles, 10 = 6 with
 * 128 x 256 tiles, 12 = 9 + pipelined LDS operands; 13 = bf16 MFMA
 * (32x32x16) on three-way bf16 splits of both operands, six products per
 * pair, FP32 accumulation: FP32-accurate, not bitwise equal to the FP32
 * MFMA kernels; 14 = 13 with 128 x 256 tiles (waves of 64 x 128, unpadded
 * LDS rows); 15 = 14 with a software-pipelined loop (operand reads and LDS
 * stores between the products; bitwise the result of 14); 16 = 15 with the
 * split done in the product's load stage (FP32 operands read directly, no
 * split pass; bitwise the result of 14); 17 = 16 with 128 x 128 tiles
 * (bitwise the result of 13). */
#define HFPX_KNOB_GEMM_F32 0
/* HFPX_KNOB_LU_NB: outer block of the LU (columns per trailing update):
 * 0 = default (1024), or one of 256, 512, 1024, 2048, 4096. */
#define HFPX_KNOB_LU_NB 1
/* HFPX_KNOB_BGEMM_F32: FP32 Schur-complement updates of the batched LU:
 * 0 = default (3), 1 = FP32 MFMA with a read-modify-write epilogue, 2 =
 * FP32 MFMA with C loaded into the accumulators first, 3 = outer updates as
 * bf16 MFMA products of three-way splits (as HFPX_KNOB_GEMM_F32 13), the
 * updates inside an outer block as 2. */
#define HFPX_KNOB_BGEMM_F32 2
/* HFPX_KNOB_PANEL: LU panel kernel: 0 = default (1), 1 = grid-barrier
 * panel, 2 = exchange panel (per-column tagged words, no barrier; used for
 * grids up to 256 workgroups, i.e. n up to 262,144). */
#define HFPX_KNOB_PANEL 3
/* HFPX_KNOB_X3_BAND: row tiles per band of the bf16-split product's tile
 * walk: 0 = automatic (the band's split A panel in an XCD's L2),
 * 1 .. 64 = fixed. */
#define HFPX_KNOB_X3_BAND 4
/* HFPX_KNOB_IR_FALLBACK: what hfpx_solve_ir does when refinement stagnates,
 * meets a non-finite iterate or runs out of steps with finite data:
 * 0 = default (1), 1 = solve again by an FP64 LU of a copy of A64 (n^2
 * doubles allocated for the call; report fell_back = 1), 2 = return
 * HFPX_ERR_NOT_CONVERGED with the last iterate. */
#define HFPX_KNOB_IR_FALLBACK 5
/* HFPX_KNOB_IR_MODE: correction steps of hfpx_solve_ir: 0 = default
 * (classic steps while each contracts the backward error 100-fold, then
 * GMRES), 1 = GMRES steps only, 2 = classic steps only (a step that
 * lowers the error less than 10 % counts as stagnation). */
#define HFPX_KNOB_IR_MODE 9
/* HFPX_KNOB_TEST_PANEL_TIMEOUT: 1 = hfpx_getrf starts with its panel
 * time-out flag already raised, as if a grid barrier had timed out: every
 * panel returns at entry and getrf reports HFPX_ERR_HIP (a host-forced test
 * of the error path; every kernel still exits normally).  0 = off. */
#define HFPX_KNOB_TEST_PANEL_TIMEOUT 6
/* HFPX_KNOB_GEMM_F64: FP64 Schur-complement products with M >= 256 and
 * N >= 128: 0 = default (3); 1 = the generic 16x16x4 kernel (one wave per
 * SIMD); 2 = 128 x 128 tiles at two waves per SIMD, unguarded interior,
 * C read and written at the end; 3 = 2 with an XCD-grouped tile walk and C
 * preloaded into the accumulators; 4 = 3 without the XCD-grouped walk;
 * 5 = 3 with bands of 4 row tiles. */
#define HFPX_KNOB_GEMM_F64 7
/* HFPX_KNOB_LU_LOOKAHEAD: 0 = default (3 for FP32, 1 for FP64, where the
 * overlapped FP64 updates measured slower); 1 = off: each outer block's
 * trailing update follows its factorisation; 2 = look-ahead: the next
 * block's columns are updated first and that block is factorised on the
 * context's stream while the rest of the trailing update runs on a second
 * default-priority stream (the factors are bitwise those of 1); 3 = 2 with
 * the panel kernel launched without the cooperative launch. */
#define HFPX_KNOB_LU_LOOKAHEAD 8
#define HFPX_KNOB_COUNT 10
/* Set / read one knob; HFPX_ERR_INVALID for an unknown knob or value. */
int hfpx_tune_set(hfpx_ctx* ctx, int knob, int value);
int hfpx_tune_get(const hfpx_ctx* ctx, int knob, int* value);

#ifdef __cplusplus
}
#endif
#endif /* HFPX3D_H */
