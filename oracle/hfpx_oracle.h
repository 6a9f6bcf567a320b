/*
 * hfpx_oracle.h -- CPU restatement of the HFPx3D hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker, not the product.  Only tests/, the smoke()
 * entry point and the cpu_baseline leg of bench.py may load it.  The HIP path
 * never calls into it and fails loudly when its own extension is missing.
 *
 * Conventions (all follow the reference's il::Array2D semantics):
 *   - arrays are column-major; a "3x6" array has element (i, j) at i + 3*j;
 *   - complex values are stored as interleaved (re, im) doubles;
 *   - indices are int64_t (il::int_t), pivots are 1-based int32 (lapack_int).
 */
#ifndef HFPX_ORACLE_H
#define HFPX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element geometry.  vert: 3x3 (coordinate k of vertex j at k + 3*j).
 * Outputs rt (3x3 rotation, rows e1, e2, e3), tau (3 complex),
 * sfm (6x6 complex, rows = shape functions, cols = monomials
 * [1, t, tb, t^2, tb^2, t*tb]), cp (3x6 collocation points for beta). */
void hfpo_el_geom(const double* vert, double beta, double* rt, double* tau,
                  double* sfm, double* cp);
/* Same with a non-uniform edge partition: wts (3, all > 0) are the vertex
 * weights; the edge node opposite vertex v sits at
 * (w_a v_a + w_b v_b) / (w_a + w_b), (a, b) = (v+1, v+2)
 * (make_el_sfm_nonuniform / el_cp_nonuniform, element_utilities.cpp:240-322,
 * 397-431).  wts = NULL is the uniform element. */
void hfpo_el_geom_w(const double* vert, const double* wts, double beta, double* rt,
                    double* tau, double* sfm, double* cp);

/* 6x18 stress influence (rows S11 S22 S33 S12 S13 S23, cols 3*node + DD
 * component, all in the source element's local frame) of one element at the
 * point (h, z) given in that frame. */
void hfpo_local_submatrix(double shear_m, double poiss_r, double h, double z_re,
                          double z_im, const double* tau, const double* sfm,
                          double* out);

/* DoF handle (n_el x 18, col-major, -1 = fixed).  Returns n_dof. */
int64_t hfpo_dof_h(const double* nods, int64_t nods_rows, int64_t n_nodes,
                   const int64_t* conn, int64_t conn_rows, int64_t n_el,
                   int tip_type, int64_t* dof_h);

/* Dense influence matrix, n_dof x n_dof col-major (lda = n_dof), zero-filled
 * by the callee.  Target rows [t0, t1) of ELEMENTS only are computed when
 * t1 > t0 (rows of other targets are left zero); pass t0 = t1 = 0 for all.
 * is_dd_local: columns are DD in each source element's own frame
 * (Num_Param_T::is_dd_local). */
int64_t hfpo_assemble(const double* nods, int64_t nods_rows, int64_t n_nodes,
                      const int64_t* conn, int64_t conn_rows, int64_t n_el,
                      const int64_t* dof_h, int64_t n_dof, double shear_m,
                      double poiss_r, double beta, int64_t t0, int64_t t1,
                      int is_dd_local, int nthreads, double* A);
/* Same with per-node vertex weights node_wts (n_nodes, NULL = uniform):
 * element e takes the weights of its three vertex nodes. */
int64_t hfpo_assemble_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                        const int64_t* conn, int64_t conn_rows, int64_t n_el,
                        const int64_t* dof_h, int64_t n_dof, double shear_m,
                        double poiss_r, double beta, const double* node_wts, int64_t t0,
                        int64_t t1, int is_dd_local, int nthreads, double* A);
/* Rows of target elements [t0, t1) only (global frame, uniform elements):
 * R (18 (t1 - t0)) x n_dof, row 18 (et - t0) + r = local traction slot r of
 * target et (0 where fixed), column = source DoF (system_assembly.cpp:61-213
 * restricted to those targets). */
void hfpo_assemble_rows(const double* nods, int64_t nods_rows, int64_t n_nodes, const int64_t* conn,
                        int64_t conn_rows, int64_t n_el, const int64_t* dof_h, int64_t n_dof, double shear_m,
                        double poiss_r, double beta, int64_t t0, int64_t t1, int nthreads, double* R);

/* rhs[dof] -= (n . s_inf)_l for every free DoF (caller zero-fills rhs). */
void hfpo_rhs(const double* nods, int64_t nods_rows, int64_t n_nodes,
              const int64_t* conn, int64_t conn_rows, int64_t n_el,
              const int64_t* dof_h, const double* s_inf, double* rhs);

/* Solution vector -> nodal DD array (6*n_el x 3, col-major); fixed DoF -> 0. */
void hfpo_write_dd(const double* x, const int64_t* dof_h, int64_t n_el,
                   double* dd);

/* Stress at n_pts points (pts 3 x n_pts), out n_pts x 6 col-major
 * [Sxx Syy Szz Sxy Sxz Syz], s_inf added, deterministic; is_dd_local: dd are
 * in each element's own frame. */
void hfpo_stress(const double* nods, int64_t nods_rows, int64_t n_nodes,
                 const int64_t* conn, int64_t conn_rows, int64_t n_el,
                 double shear_m, double poiss_r, const double* dd,
                 const double* pts, int64_t n_pts, const double* s_inf,
                 int is_dd_local, int nthreads, double* out);
void hfpo_stress_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                   const int64_t* conn, int64_t conn_rows, int64_t n_el,
                   double shear_m, double poiss_r, const double* dd,
                   const double* pts, int64_t n_pts, const double* s_inf,
                   const double* node_wts, int is_dd_local, int nthreads, double* out);

/* LU with partial pivoting (LAPACK dgetrf semantics): A n x n col-major
 * (lda), ipiv 1-based.  Returns info (0 ok, k > 0: U(k,k) exactly zero). */
int64_t hfpo_getrf(int64_t n, double* A, int64_t lda, int32_t* ipiv,
                   int nthreads);
/* Solve with the factors (dgetrs 'N'): B n x nrhs col-major (ldb). */
void hfpo_getrs(int64_t n, int64_t nrhs, const double* LU, int64_t lda,
                const int32_t* ipiv, double* B, int64_t ldb);

/* Sneddon penny crack (radius a, pressure p): opening DD at a point of the
 * crack plane and the stress field (including sigma_inf_zz = p). */
void hfpo_sneddon_dd(double shear_m, double poiss_r, double a, double p,
                     const double* xyz, double* dd);

/* Volume control (system_assembly.cpp:458-629).  Area integrals of the six
 * shape functions (el_p2_sf_integral, element_utilities.cpp:573-595; tau and
 * sfm as returned by hfpo_el_geom). */
void hfpo_sf_integral(const double* tau, const double* sfm, double* out);
/* Border of make_3dbem_matrix_vc: row[d] = volume per unit DD of DoF d,
 * col[d] = -normal component (both n_dof long, written for free DoF). */
void hfpo_vc_border(const double* nods, int64_t nods_rows, int64_t n_nodes,
                    const int64_t* conn, int64_t conn_rows, int64_t n_el,
                    const int64_t* dof_h, int is_dd_local, double* row, double* col);
void hfpo_vc_border_w(const double* nods, int64_t nods_rows, int64_t n_nodes,
                      const int64_t* conn, int64_t conn_rows, int64_t n_el,
                      const int64_t* dof_h, const double* node_wts, int is_dd_local,
                      double* row, double* col);
/* mod_3dbem_system_vc: M ((used+1)^2 col-major) and rhs (used+1) of the
 * truncated system; dof handles n_el x ndpe col-major.  0 or -1 (invalid). */
int hfpo_mod_system_vc(const double* A, int64_t lda, int64_t orig_ndof,
                       const int64_t* orig_dof_h, const int64_t* dof_h,
                       int64_t n_el, int64_t ndpe, int64_t used_ndof,
                       const double* delta_t, int64_t tsize, double delta_v,
                       double* M, double* rhs);

/* Cohesion / friction state per node (src/Development/cohesion_friction.cpp).
 * dd n_nod x 3 col-major (ldd); mat_id n_nod; cmid n_mat x 2 col-major
 * (cohesion model, friction model); par 6 per material: cr_open, peak_ts,
 * cr_slip, peak_sc, peak_sf, res_sf. */
void hfpo_match_f_c(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                    const int32_t* cmid, int64_t n_mat, const double* par,
                    const double* mr_open, const double* mr_slip, double* friction_coef,
                    double* slip_cohesion, double* open_cohesion);
void hfpo_update_mrd(int64_t n_nod, const double* dd, int64_t ldd, const int32_t* mat_id,
                     int64_t n_mat, const double* par, double* mr_open, double* mr_slip);

/* One step of the volume-control cohesion / friction iteration
 * (vc_cf_iteration, src/Development/c_f_iteration.cpp:27-436), uniform
 * elements.  A: the (orig_ndof+1)^2 volume-control matrix of the original
 * handle orig_dof_h (n_el x 18 col-major, one slot per DoF).  In/out: dd
 * (6 n_el x 3 col-major, ld 6 n_el), pp (6 n_el), dof_h / *n_dof (the
 * current handle in, the new one out).  Out: friction_coef, slip_cohesion,
 * open_cohesion (match_f_c of the entry DD), delta_trac (orig_ndof, the
 * right-hand side by original DoF), incr (n_dof + 1 after the call: DD
 * increments by new DoF, then the pressure increment), *pressure, *residual
 * (|incr|_1 + |delta_trac|_1).  Returns 0; 1 when no DoF stays active (zero
 * increments); 2 when the truncated system is singular; -1 on bad input. */
int hfpo_vc_cf_iteration(const double* nods, int64_t nods_rows, int64_t n_nodes, const int64_t* conn,
                         int64_t conn_rows, int64_t n_el, double beta, int is_dd_local,
                         const int64_t* orig_dof_h, int64_t orig_ndof, const double* A, int64_t lda,
                         const int32_t* mat_id, const int32_t* cmid, int64_t n_mat, const double* par,
                         const double* mr_open, const double* mr_slip, const double* s_inf, int64_t inj_el,
                         int64_t inj_node, double new_vol, double* dd, double* pp, int64_t* dof_h,
                         int64_t* n_dof, double* friction_coef, double* slip_cohesion, double* open_cohesion,
                         double* delta_trac, double* incr, double* pressure, double* residual);

#ifdef __cplusplus
}
#endif
#endif
