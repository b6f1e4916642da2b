/*
 * dgtd_oracle.h - CPU FP64 restatement of rapidfem-td's nodal DGTD path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product is
 * compared against. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product (rapidfem_amd/) never links,
 * loads or calls it.
 *
 * Each function restates one reference function; the citation is on the
 * definition in dgtd_oracle.c. Paths are relative to the reference root.
 *
 * Parity pins (what checks each part, all in tests/):
 *   orc_refelem          sympy goldens p=1..3 (reference tests), exact
 *                        rationals p=1..4 incl. LIFT (own sympy script),
 *                        invariants of dg_basis.rs tests.
 *   orc_element_curl     rhs_curl_golden_test.rs vector (p=2 sheared tet),
 *                        exact curl of polynomial fields p=1..4.
 *   orc_op_* / apply     central-flux skew-symmetry of M~A, upwind
 *                        dissipation sign, PEC cavity eigenfrequency,
 *                        sigma damping rate, brute-force dense rules.
 *   orc_lserk_*          lserk4_convergence_test.rs (6x6 skew operator).
 *   orc_kcl_*            4th-order rate + embedded 3rd-order estimate on
 *                        a closed-form linear ODE.
 *   orc_expm/expmv/etd   propagator_expm_test.rs (sympy 30-digit values),
 *                        closed forms, scipy.linalg.expm, ETD analytic.
 *   orc_cfl_dt           the LSERK4 stability boundary along the dominant
 *                        eigenvalue (dense eigensolve) within the
 *                        bisection's resolution; session.rs:1352-1364.
 *   orc_assemble_sparse  dense A of unit-vector applies bit for bit;
 *                        rhs.rs:2259-2330, 3133-3255.
 *   orc_op_set_dispersive  n_dof and P rows (rhs.rs:3077-3121), analytic
 *                        Debye permittivity, exact energy-dissipation
 *                        identity of the central-flux Debye system.
 *   orc_apply_ghost      A y + b for the numpy port source b, any y.
 */
#ifndef DGTD_ORACLE_H
#define DGTD_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference element (dg_basis.rs) ---------------------------------- */
int orc_np(int p);
int orc_nfp(int p);
/* Fills caller arrays: nodes np*3, mass/dr/ds/dt np*np, face_nodes 4*nfp,
 * lift np*4nfp, fweights 4*nfp. Returns 0. */
int orc_refelem(int p, double *nodes, double *mass, double *dr, double *ds,
                double *dt, int *face_nodes, double *lift, double *fweights);

/* ---- structured mesher (mesh_gen.rs) ----------------------------------- */
void orc_box_nodes(int nx, int ny, int nz, double lx, double ly, double lz,
                   double amplitude, uint64_t seed, int jittered, double *xyz);
void orc_kuhn_tets(int nx, int ny, int nz, int64_t *tets);
void orc_jitter_offset(int64_t i, int64_t j, int64_t k, uint64_t seed,
                       double *off3);

/* ---- one-element curl (rhs.rs element_curl) ----------------------------- */
void orc_element_curl(int p, const double *tet_xyz /*4x3*/,
                      const double *field /*np*3*/, double *out /*np*3*/);
/* Geometric factors of one tet: jac 9, jinv 9 ([i][k] row major), det. */
void orc_geom_factors(const double *tet_xyz, double *jac, double *jinv,
                      double *det);

/* ---- operator (rhs.rs MaxwellOperator) ---------------------------------- */
typedef struct orc_op orc_op;
/* face_port: optional 4K array (elem*4+lf -> port id or -1), NULL = none.
 * eps, mu: K*3 (NULL = vacuum); sigma, sigma_m: K (NULL = 0). */
orc_op *orc_op_new(int64_t n_nodes, const double *xyz, int64_t n_tets,
                   const int64_t *tets, int order, double flux_alpha,
                   const double *eps, const double *mu, const double *sigma,
                   const double *sigma_m, const int *face_port);
void orc_op_free(orc_op *op);
int64_t orc_op_n_dof(const orc_op *op);
int64_t orc_op_n_elem(const orc_op *op);
int orc_op_order(const orc_op *op);
/* Copy out the arrays GpuOperator::new reads (gpu/operator.rs:216-292). */
void orc_op_export(const orc_op *op, double *jinv, double *det,
                   double *normal, double *fscale, int *nbr, int *nbr_lf,
                   int *port, int *perm, double *inv_eps, double *inv_mu,
                   double *sig_eps, double *sig_mu);
void orc_op_set_threads(int nthreads);
/* Debye slots in list order (rhs.rs:963-985); -1 for a bad element list. */
int orc_op_set_dispersive(orc_op *op, int64_t n, const int64_t *elem,
                          const double *eps_inf, const double *eps_static,
                          const double *tau);
int64_t orc_op_n_dispersive(const orc_op *op);
void orc_apply(const orc_op *op, const double *y, double *dy);
/* A y with ghost[(4e + f)*Nfp + m][6] as the port faces' incident field. */
void orc_apply_ghost(const orc_op *op, const double *y, const double *ghost,
                     double *dy);
/* MaxwellOperator::assemble_sparse: col_idx == NULL fills row_ptr
 * (n_dof + 1) and returns nnz; a second call fills col_idx/values. */
int64_t orc_assemble_sparse(const orc_op *op, int64_t *row_ptr,
                            int64_t *col_idx, double *values);
double orc_field_energy(const orc_op *op, const double *y);

/* ---- LSERK4 (explicit.rs) ---------------------------------------------- */
/* src_kind 0 free, 1 point (src_dof), 2 vector (src_vec, n). g: one sample
 * per step (held across the step), may be NULL when free. */
void orc_lserk_run(const orc_op *op, double *y, double dt, int64_t nsteps,
                   int src_kind, int64_t src_dof, const double *src_vec,
                   const double *g);
void orc_lserk_run_dense(const double *a, int n, double *y, double dt,
                         int64_t nsteps);

/* ---- KCL RK4(3)5[2R+]C + PI controller (explicit_adaptive.rs, session.rs) */
typedef struct {
  double atol, rtol, safety, growth_limit, shrink_limit, pi_alpha, pi_beta,
      min_step_factor;
} orc_controller;
void orc_controller_default(orc_controller *c);
void orc_kcl_step(const orc_op *op, double *y, double *err, double dt,
                  int src_kind, int64_t src_dof, const double *src_vec,
                  double g);
void orc_kcl_step_dense(const double *a, int n, double *y, double *err,
                        double dt);
double orc_err_norm(const orc_controller *c, const double *y_old,
                    const double *y_new, const double *err, int64_t n);
double orc_factor(const orc_controller *c, double err_norm, double prev,
                  int reject);
/* Adaptive transient, steps frames of dt; out_traj (steps+1)*n or NULL.
 * stats: accepted, rejected; hminmax: h_min, h_max. Returns 0, or -1 when
 * the step collapses (message in the reference). */
int orc_adaptive_run(const orc_op *op, double *y, double dt, int64_t steps,
                     const orc_controller *c, int src_kind, int64_t src_dof,
                     const double *src_vec, const double *g,
                     int64_t *stats, double *hminmax);

/* ---- CFL (session.rs cfl_dt, normal_vector) ----------------------------- */
void orc_normal_vector(int64_t n, uint64_t seed, double *out);
double orc_cfl_dt(const orc_op *op);

/* ---- Krylov exponential (propagator.rs) --------------------------------- */
void orc_expm(const double *a, int n, double *out);
int orc_expmv_dense(const double *a, int n, const double *v, double t,
                    int max_dim, double tol, double *out);
int orc_expmv_op(const orc_op *op, const double *v, double t, int max_dim,
                 double tol, double *out);
void orc_etd_dense(const double *a, int n, const double *y, const double *b,
                   double h, int max_dim, double tol, double *out);
void orc_etd_op(const orc_op *op, const double *y, const double *b, double h,
                int max_dim, double tol, double *out);

#ifdef __cplusplus
}
#endif
#endif
