/*
 * rfhip.h - MI355X (gfx950) HIP backend for rapidfem-td's DGTD time-stepper.
 *
 * Drop-in C ABI for the reference's optional GPU backend
 * `rapidfem_td::gpu::{GpuContext, GpuOperator}` (crates/rapidfem-td/src/gpu/
 * mod.rs:25-26), which session.rs calls at :266-277, :589-612 and :866-970.
 * Each entry point below names the reference method it replaces. Plain
 * pointers and sizes only; no torch or HIP types cross this boundary.
 *
 * Conventions
 *  - Every function returns an int status, RFHIP_OK (0) on success, and never
 *    aborts. The message of the last failure on the calling thread is read
 *    with rfhip_last_error(). The reference returns Result<_, String>; its
 *    length-mismatch asserts (gpu/operator.rs:518,552) become RFHIP_EINVAL.
 *  - The caller owns every host buffer: inputs are read during the call,
 *    outputs are preallocated ((steps+1)*n_dof when recording, else n_dof).
 *  - The library owns device memory for the lifetime of a handle.
 *  - One handle = one host thread = one HIP stream (the reference's &mut self
 *    plus an in-order queue, gpu/operator.rs:436-438).
 *  - State layout at the boundary is the reference's node-major AoS,
 *    y[(e*Np + node)*6 + c], c = Ex,Ey,Ez,Hx,Hy,Hz (rhs.rs:492-497).
 */
#ifndef RFHIP_H
#define RFHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RFHIP_OK 0
#define RFHIP_EINVAL 1      /* bad argument or length mismatch */
#define RFHIP_ENODEV 2      /* no usable HIP device ("no OpenCL GPU device found", gpu/mod.rs:48-50) */
#define RFHIP_ENOMEM 3      /* device allocation failed */
#define RFHIP_EHIP 4        /* a HIP runtime call failed */
#define RFHIP_ECOLLAPSE 5   /* adaptive step collapsed below min_step_factor*dt (gpu/operator.rs:881-888) */
#define RFHIP_EUNSUPPORTED 6 /* operator data this form cannot run (e.g. non-conforming faces for a fused form) */

/* Kernel forms of one operator.
 *  PLAIN_F32  the reference GPU structure and precision: one work-item per
 *             node apply kernel + a separate stage kernel, f32 state
 *             (gpu/apply.cl, gpu/lserk.cl).
 *  PLAIN_F64  the same structure in FP64: the baseline "plain FP64 GPU form".
 *  FUSED_F64  one fused kernel per RK stage (volume curl + upwind flux + LIFT
 *             + materials + stage update), FP64 state and arithmetic.
 *  FUSED_F64C the fused kernel with FP64 arithmetic and compressed
 *             registers: y stored as f32 hi + bf16 lo, p as f32.
 *  FUSED_BF16X3, FUSED_BF16X2, FUSED_BF16X4  the precision forms:
 *             FUSED_F64C's registers, with the volume contraction D_k u on the
 *             matrix cores in BF16 split into 3 (2: hi + lo; 4) pieces per
 *             operand, products accumulated in FP32; J^-1, curl, flux, LIFT
 *             and the update stay FP64. Orders 2..4. */
#define RFHIP_FORM_PLAIN_F32 0
#define RFHIP_FORM_PLAIN_F64 1
#define RFHIP_FORM_FUSED_F64 2
#define RFHIP_FORM_FUSED_F64C 3
#define RFHIP_FORM_FUSED_BF16X3 4
#define RFHIP_FORM_FUSED_BF16X2 5
#define RFHIP_FORM_FUSED_BF16X4 6

/* Drive of dy/dt = A*y + b*g(t) (session.rs:107-114 Drive). */
#define RFHIP_DRIVE_FREE 0
#define RFHIP_DRIVE_POINT 1
#define RFHIP_DRIVE_VECTOR 2

typedef struct rfhip_ctx rfhip_ctx;
typedef struct rfhip_op rfhip_op;

/* Copies the calling thread's last error message (NUL-terminated, truncated
 * to len). Returns the full message length. */
size_t rfhip_last_error(char *buf, size_t len);

/* ---- GpuContext (gpu/mod.rs:33-135) ------------------------------------- */

/* GpuContext::new (gpu/mod.rs:46-65): first HIP device. */
int rfhip_ctx_create(rfhip_ctx **out);
void rfhip_ctx_destroy(rfhip_ctx *ctx);
/* GpuContext::device_name (gpu/mod.rs:35). */
int rfhip_ctx_device_name(const rfhip_ctx *ctx, char *buf, size_t len);
/* GpuContext::fp64 (gpu/mod.rs:38): always 1 on gfx950. */
int rfhip_ctx_has_fp64(const rfhip_ctx *ctx);

/* ---- GpuOperator (gpu/operator.rs:78-1486) ------------------------------ */

/* Everything GpuOperator::new reads from MaxwellOperator's fields
 * (gpu/operator.rs:216-292), as the f64 host data MaxwellOperator holds.
 * face arrays are flattened over e*4 + local_face. */
typedef struct {
  int32_t np, nfp;                /* Np, Nfp */
  int64_t n_elem;                 /* K */
  const double *diff_r;           /* Np*Np row-major (re.diff_r) */
  const double *diff_s;
  const double *diff_t;
  const double *lift;             /* Np*(4*Nfp) row-major (re.lift) */
  const int32_t *face_nodes;      /* 4*Nfp (re.face_nodes) */
  const double *jinv;             /* 9*K, jacobian_inv[i][k] row-major */
  const double *inv_eps;          /* 3*K */
  const double *inv_mu;           /* 3*K */
  const double *sigma_eps;        /* 3*K */
  const double *sigma_mu;         /* 3*K */
  const double *face_normal;      /* 12*K */
  const double *face_fscale;      /* 4*K */
  const int32_t *face_neighbor;   /* 4*K, -1 on a boundary face */
  const int32_t *face_nbr_local;  /* 4*K */
  const int32_t *face_port;       /* 4*K, -1 for none */
  const int32_t *face_perm;       /* 4*K*Nfp, 0-padded on boundary faces */
  double flux_alpha;              /* 0 central .. 1 upwind */
} rfhip_op_desc;

/* GpuOperator::new (gpu/operator.rs:207-396). form: RFHIP_FORM_*. */
int rfhip_op_create(rfhip_ctx *ctx, const rfhip_op_desc *desc, int form,
                    rfhip_op **out);
/* GpuOperator::new for an operator with Debye dispersive elements, which the
 * reference GPU path refuses (gpu/operator.rs:211-215): MaxwellOperator's
 * DispersiveElem list (rhs.rs:469-487, resolved at rhs.rs:963-985), slot q
 * being element disp_elem[q] with Pdot = a*P + g*E (a = -1/tau,
 * g = (eps_s - eps_inf)/tau, dispersive.rs:50-53) and 1/eps_inf. The state
 * then appends P[(q*Np + node)*3 + c] after the 6*Np*K [E,H] values
 * (rhs.rs:490-500) and n_dof grows by 3*Np per slot; every form, apply,
 * transient, KCL and Krylov path carries it (rhs.rs:1295-1324,1456-1472).
 * The element's own inv_eps in desc is used as given: the reference expects
 * 1/eps_inf there. RFHIP_EINVAL for an element out of range or listed twice
 * (rhs.rs:970-974). n_disp = 0 is rfhip_op_create. */
int rfhip_op_create_dispersive(rfhip_ctx *ctx, const rfhip_op_desc *desc,
                               int form, int64_t n_disp,
                               const int64_t *disp_elem, const double *disp_a,
                               const double *disp_g,
                               const double *disp_inv_eps_inf, rfhip_op **out);
void rfhip_op_destroy(rfhip_op *op);
/* GpuOperator::n_dof (gpu/operator.rs:430); MaxwellOperator::n_dof
 * (rhs.rs:1003-1010) with the appended P block. */
int64_t rfhip_op_n_dof(const rfhip_op *op);
/* MaxwellOperator::n_dispersive (rhs.rs:1021-1024). */
int64_t rfhip_op_n_dispersive(const rfhip_op *op);
int rfhip_op_form(const rfhip_op *op);

/* GpuOperator::apply (gpu/operator.rs:513-523): dy = A*y, f32 in/out. */
int rfhip_op_apply(rfhip_op *op, const float *y, float *dy);
/* FP64 variant (no reference counterpart: the reference GPU apply is f32
 * only). */
int rfhip_op_apply_f64(rfhip_op *op, const double *y, double *dy);

/* GpuOperator::transient (gpu/operator.rs:906-953): LSERK4, `steps` frames
 * of `substeps` steps of dt/substeps, source b*g held per substep
 * (g has steps*substeps samples unless free). record != 0 writes
 * (steps+1)*n_dof rows (row 0 = y0), else the final state (n_dof). */
int rfhip_op_transient(rfhip_op *op, const float *y0, float dt, int64_t steps,
                       int64_t substeps, int drive_kind, int64_t point_dof,
                       const double *src_vec, const float *g, int64_t n_g,
                       int record, float *out);
/* FP64 boundary variant: the session's f64<->f32 round trip
 * (session.rs:912-916) would otherwise cap accuracy at f32. */
int rfhip_op_transient_f64(rfhip_op *op, const double *y0, double dt,
                           int64_t steps, int64_t substeps, int drive_kind,
                           int64_t point_dof, const double *src_vec,
                           const double *g, int64_t n_g, int record,
                           double *out);

/* Explicit transient that records probe DOFs on the device instead of
 * whole states (session.rs:118-123 Record::Dofs; the reference GPU path
 * downloads the full state every frame, session.rs:914-921, and picks the
 * DOFs on the host). rows receives (steps+1)*n_probes values, row 0 taken
 * from y0, one device-to-host copy at the end; y_final (may be NULL) the
 * final state. Arguments otherwise as rfhip_op_transient_f64. */
int rfhip_op_transient_probes_f64(rfhip_op *op, const double *y0, double dt,
                                  int64_t steps, int64_t substeps,
                                  int drive_kind, int64_t point_dof,
                                  const double *src_vec, const double *g,
                                  int64_t n_g, const int64_t *probes,
                                  int64_t n_probes, double *rows,
                                  double *y_final);

/* Port modal extraction on the device (rhs.rs:1211-1242
 * port_modal_projections, recorded per frame as TdSession::port_signals,
 * session.rs:1041-1054, does from full downloaded states). Port `port` is
 * given by its n_faces (element, local face) pairs, the mode's transverse
 * profiles at each face node (profiles: n_faces*Nfp*6 doubles, e_t xyz then
 * h_t xyz, face-node order of face_nodes) and the quadrature weights
 * 2*area*face_node_weight (weights: n_faces*Nfp). Setting a port again
 * replaces it; ports are numbered 0..rfhip_op_n_ports()-1 (unset ones are
 * empty and record 0). */
int rfhip_op_set_port(rfhip_op *op, int32_t port, int64_t n_faces,
                      const int32_t *elem, const int32_t *local_face,
                      const double *profiles, const double *weights);
int rfhip_op_n_ports(const rfhip_op *op);
/* MaxwellOperator::port_source (rhs.rs:1157-1199), assembled on the device
 * from the faces and mode profiles given to rfhip_op_set_port: the n_dof
 * drive pattern of the port's incident mode (the characteristic flux of the
 * jumps [E] = -e_t, [H] = -h_t, lifted and scaled by the element's 1/eps,
 * 1/mu). It is written to b (n_dof doubles) when b is not NULL, and it stays
 * on the device as the operator's vector source: a following transient with
 * RFHIP_DRIVE_VECTOR and src = NULL drives with it, without a host copy
 * (the reference builds it on the host and uploads it per run). Passing a
 * src pattern again replaces it. */
int rfhip_op_port_source_f64(rfhip_op *op, int32_t port, double *b);
/* Explicit transient recording (P_e, P_h) of every port per frame on the
 * device: signals receives (steps+1)*n_ports*2 doubles (row 0 from y0),
 * y_final (may be NULL) the final state. Arguments otherwise as
 * rfhip_op_transient_f64. */
int rfhip_op_transient_ports_f64(rfhip_op *op, const double *y0, double dt,
                                 int64_t steps, int64_t substeps,
                                 int drive_kind, int64_t point_dof,
                                 const double *src_vec, const double *g,
                                 int64_t n_g, double *signals,
                                 double *y_final);

/* session.rs:130-140 Controller, all f64. */
typedef struct {
  double atol, rtol, safety, growth_limit, shrink_limit, pi_alpha, pi_beta,
      min_step_factor;
} rfhip_controller;

/* KclStats (session.rs:218-223); h in operator time. */
typedef struct {
  int64_t accepted, rejected;
  double h_min, h_max;
} rfhip_kcl_stats;

/* GpuOperator::transient_kcl (gpu/operator.rs:823-893): KCL RK4(3)5[2R+]C
 * with the PI controller, g[k] held across frame k (steps samples unless
 * free). out receives (steps+1)*n_dof rows. RFHIP_ECOLLAPSE when the step
 * collapses. */
int rfhip_op_transient_kcl(rfhip_op *op, const float *y0, float dt,
                           int64_t steps, int drive_kind, int64_t point_dof,
                           const double *src_vec, const float *g, int64_t n_g,
                           const rfhip_controller *ctl, float *out,
                           rfhip_kcl_stats *stats);
int rfhip_op_transient_kcl_f64(rfhip_op *op, const double *y0, double dt,
                               int64_t steps, int drive_kind,
                               int64_t point_dof, const double *src_vec,
                               const double *g, int64_t n_g,
                               const rfhip_controller *ctl, double *out,
                               rfhip_kcl_stats *stats);

/* KCL registers (no counterpart in the reference, whose KCL keeps f32
 * registers, gpu/operator.rs:823-893). RFHIP_KCL_REG_F64 (default): y, err,
 * the stage input and the substep backup in FP64, 48 bytes per entry per
 * stage update; on a fused form without Debye elements whose mesh has the
 * neighbour-trace table the update runs in the stage kernel's epilogue (one
 * launch per stage, k never stored); RFHIP_KCL_REG_F64_SEPARATE keeps it as
 * a pass of its own after each apply. RFHIP_KCL_REG_COMPRESSED (forms with compressed state
 * registers only, else RFHIP_EUNSUPPORTED): y and its backup as f32 hi +
 * bf16 lo, err as f32, the stage input FP64, 36 bytes; the controller and
 * every apply still in FP64. On an operator without Debye elements whose
 * mesh has the neighbour-trace table (conforming) and whose form keeps the
 * volume in FP64, the update runs in the stage kernel's epilogue, so k is
 * never stored (one launch per stage instead of two), and the stage input
 * is held as f32 hi + bf16 lo too.
 * RFHIP_KCL_REG_COMPRESSED_SEPARATE: the same registers with the update
 * always in a pass of its own after each apply. Applies to later
 * transient_kcl calls. */
#define RFHIP_KCL_REG_F64 0
#define RFHIP_KCL_REG_COMPRESSED 1
#define RFHIP_KCL_REG_COMPRESSED_SEPARATE 2
#define RFHIP_KCL_REG_F64_SEPARATE 3
int rfhip_op_set_kcl_registers(rfhip_op *op, int32_t mode);

/* GpuOperator::expmv (gpu/operator.rs:1175-1197): out = exp(t*A)*v, FP64
 * Arnoldi (CGS2) with the Hessenberg exponential, following
 * propagator.rs:186-326: max_dim caps the dimension, tol > 0 enables the
 * a-posteriori stop (0 = fixed dimension). The basis is held whole in HBM,
 * so no KRYLOV_CHUNK sub-stepping. dim_used may be NULL. */
int rfhip_op_expmv(rfhip_op *op, const double *v, double t, int32_t max_dim,
                   double tol, double *out, int32_t *dim_used);
/* GpuOperator::etd_step (gpu/operator.rs:1267-1298): out = exp(hA)y +
 * h*phi1(hA)b via the augmented [[A,b],[0,0]] system
 * (propagator.rs:338-373). */
int rfhip_op_etd_step(rfhip_op *op, const double *y, const double *b,
                      double h, int32_t max_dim, double tol, double *out,
                      int32_t *dim_used);
/* ---- adaptive Krylov exponential and its learned (h, m) proposer ---------
 * No reference counterpart: the reference's exponential route takes the
 * frame step with a fixed (or KRYLOV_TOL-stopped) dimension
 * (session.rs:700-726, propagator.rs:186-326). These calls integrate
 * y' = A y over [0, t_end] with a step h and Krylov dimension m chosen per
 * step. Every step is judged by the a-posteriori estimate of
 * propagator.rs:278-294, scaled by the propagated norm beta:
 *   est(h, m) = beta * H[m][m-1] * |[exp(h H_m)]_{m-1,0}|,
 * accepted when est <= tol * |v| * h / t_end (so the estimates of a run add
 * up to at most tol * |v|). Hessenberg exponentials: degree-13 Padé with
 * scaling and squaring, FP64; Arnoldi: classical Gram-Schmidt with delayed
 * reorthogonalisation (DCGS2) on the device, FP64. */
typedef struct rfhip_mlp rfhip_mlp;

typedef struct {
  double tol;              /* relative error target of the whole run */
  int32_t m;               /* the plain controller's Krylov dimension (2..64) */
  int32_t m_min, m_max;    /* range of the proposer's dimensions (2..64) */
  double h0;               /* first step (operator time); <= 0: the largest
                              admissible step on the first basis */
  double safety;           /* step growth factor safety * (bound/est)^(1/m) */
  double growth_limit;     /* clamps of that factor */
  double shrink_limit;
  int64_t max_steps;       /* accepted steps before RFHIP_EINVAL */
  double h_scale;          /* characteristic step of the proposer's features
                              and outputs (the CFL step); <= 0: 1/|H[1][0]|
                              of the first basis */
} rfhip_expctl;

/* One accepted step of an adaptive run (the accept/reject log). */
typedef struct {
  double t;                /* operator time at the step's start */
  double h;                /* accepted step */
  double h_first;          /* first proposal (plain: == h when accepted) */
  double est_first, bound_first;  /* its estimate and bound */
  double est, bound;       /* estimate and bound of the accepted step */
  double beta;             /* norm of the propagated vector */
  double t_left;           /* t_end - t */
  double ms;               /* wall time of the step: Arnoldi, exponentials */
  int32_t m;               /* Krylov dimension built */
  int32_t proposer;        /* 0 plain controller, 1 network */
  int32_t n_judged;        /* proposals judged (1: the first was admissible) */
  int32_t breakdown;       /* 1: the Krylov subspace was invariant (exact) */
  float feat[12];          /* the proposer's features for this step */
} rfhip_exp_step;

/* out = exp(t_end A) v. proposer NULL: the plain controller (dimension
 * ctl->m every step, the next step grown from the last estimate, a rejected
 * proposal shrunk on the same basis, which costs no operator applies).
 * Otherwise, after two Arnoldi columns of each step the network proposes
 * (h, m), m clamped to [m_min, m_max]; the step taken is the largest one
 * the estimate admits on that basis (searched from the proposed h, up or
 * down, to 2^-10 in log h). steps (capacity
 * steps_cap, may be NULL) receives each step's record; hess (may be NULL)
 * receives each logged step's Hessenberg matrix as a zero-padded
 * (hess_dim+1) x hess_dim row-major block. n_steps (may be NULL) gets the
 * number of steps taken. */
int rfhip_op_expmv_adaptive(rfhip_op *op, const double *v, double t_end,
                            const rfhip_expctl *ctl, rfhip_mlp *proposer,
                            rfhip_exp_step *steps, int64_t steps_cap,
                            double *hess, int32_t hess_dim, int64_t *n_steps,
                            double *out);

/* Basis storage of the adaptive exponential (no counterpart in the
 * reference; for callers that trade the basis' precision for HBM traffic).
 * RFHIP_KRYLOV_BASIS_F64 (default): every column FP64.
 * RFHIP_KRYLOV_BASIS_COMPRESSED: the finished columns as f32 hi + bf16 lo
 * (6 bytes per entry, about 2^-32 relative rounding), the state, the
 * provisional vector and the operator product FP64; the Gram and update
 * sweeps, which read every finished column, move a quarter fewer bytes.
 * Meant for tolerances well above 1e-9. Applies to later
 * rfhip_op_expmv_adaptive calls on op. */
#define RFHIP_KRYLOV_BASIS_F64 0
#define RFHIP_KRYLOV_BASIS_COMPRESSED 1
int rfhip_op_set_krylov_basis(rfhip_op *op, int32_t mode);

/* Hindsight labels for the proposer (host computation, nthreads threads):
 * for each of count logged steps (Hessenberg blocks as written by
 * rfhip_op_expmv_adaptive with hess_dim = m_log, dim_built[q] columns
 * valid), the largest admissible step h_j <= h_cap[q] at each grid
 * dimension j (estimate as above with beta[q] and bound rate[q] * h, found
 * by bracketing from h_ref[q] and 12 bisections of log h), and the j of the
 * highest h_j / cost(j), cost(j) = c0 + c1 j + c2 j (j+1) / 2
 * (cost_model = {c0, c1, c2}: per step, per operator apply, per basis
 * column swept). eff_best may be NULL. */
int rfhip_krylov_labels(const double *hess, int32_t m_log, int64_t count,
                        const int32_t *dim_built, const double *beta,
                        const double *rate, const double *h_ref,
                        const double *h_cap, const int32_t *m_grid,
                        int32_t n_grid, const double *cost_model,
                        int32_t nthreads, double *h_best, int32_t *m_best,
                        double *eff_best);
/* exp(a) for a row-major n x n matrix (n <= 4096), degree-13 Padé with
 * scaling and squaring (Higham 2005), FP64, on the host: the Hessenberg
 * exponential of the adaptive path. (The fixed-step rfhip_op_expmv keeps
 * the reference's Taylor-18, propagator.rs:69-114.) */
int rfhip_expm_pade(const double *a, int32_t n, double *out);

/* The proposer: a 12 -> 64 -> 64 -> 2 perceptron, tanh hidden layers,
 * FP32, run and trained by the library's own HIP kernels. Parameters
 * (5122 floats): W1[64][12] b1[64] W2[64][64] b2[64] W3[2][64] b3[2],
 * row-major; normalisation (28 floats): in_mean[12] in_inv_scale[12]
 * out_mean[2] out_scale[2]. Outputs are (log(h / h_scale), m). */
int rfhip_mlp_create(rfhip_ctx *ctx, uint64_t seed, rfhip_mlp **out);
void rfhip_mlp_destroy(rfhip_mlp *mlp);
int rfhip_mlp_n_params(void);
/* Either pointer may be NULL. */
int rfhip_mlp_get_params(rfhip_mlp *mlp, float *params, float *norm);
/* Also resets the optimiser's moments. */
int rfhip_mlp_set_params(rfhip_mlp *mlp, const float *params,
                         const float *norm);
/* y[s*2 + o] for the n feature rows x[s*12 + i]. */
int rfhip_mlp_forward(rfhip_mlp *mlp, const float *x, int64_t n, float *y);
/* Loss sum_s w_s |yn_s - tn_s|^2 / sum_s w_s in normalised outputs (w NULL:
 * equal weights) and its gradient with respect to the parameters (grad may
 * be NULL): the backward kernel's check. */
int rfhip_mlp_loss_grad(rfhip_mlp *mlp, const float *x, const float *t,
                        const float *w, int64_t n, double *loss, float *grad);
/* Adam training (beta1 0.9, beta2 0.999, eps 1e-8) over epochs shuffled
 * passes (seeded) in minibatches; epoch_loss (may be NULL) gets each
 * epoch's mean batch loss. */
int rfhip_mlp_train(rfhip_mlp *mlp, const float *x, const float *t,
                    const float *w, int64_t n, int32_t epochs, int32_t batch,
                    float lr, uint64_t seed, double *epoch_loss);

/* Wall time, in milliseconds, of the device part of the operator's last
 * rfhip_op_expmv, rfhip_op_etd_step or rfhip_op_transient_kcl(_f64) call:
 * from its input resident on the device to its output ready, the host
 * transfers at the boundary left out. (No reference counterpart: the
 * benchmark's Krylov and KCL measurements read it.) */
int rfhip_op_last_call_ms(const rfhip_op *op, double *ms);

/* ---- kernel variants (no reference counterpart) --------------------------
 * The fused forms have several kernel variants per order (work per thread,
 * block shape). rfhip_op_autotune times each on the operator's own mesh with
 * scratch registers (the state is not modified) and keeps the fastest;
 * chosen/ms_best may be NULL. rfhip_op_set_variant pins one. A fused
 * operator starts on the variant measured fastest for its order and register
 * form on MI355X. Plain forms have a single variant (0). */
int rfhip_op_n_variants(const rfhip_op *op);
int rfhip_op_autotune(rfhip_op *op, int *chosen, double *ms_best);
int rfhip_op_set_variant(rfhip_op *op, int variant);
int rfhip_op_variant(const rfhip_op *op);
/* 1 when the variant runs as itself on this operator (some variants need the
 * neighbour-trace table of a conforming mesh, one needs Np <= 20);
 * rfhip_op_set_variant refuses the others with RFHIP_EUNSUPPORTED and the
 * autotuner skips them. */
int rfhip_op_variant_available(const rfhip_op *op, int variant);
/* Test aid (no reference counterpart): fills the shared memory (LDS) of
 * every CU with NaN on the operator's stream and waits, so a kernel variant
 * that reads LDS it did not write fails its parity check instead of
 * passing on values an earlier launch left behind. */
int rfhip_op_poison_lds(rfhip_op *op);

/* ---- device-resident runs (no reference counterpart) -------------------
 * For callers that keep the state in HBM between calls (benchmarks, long
 * runs without per-frame recording): load once, step, read back. */
int rfhip_op_state_load(rfhip_op *op, const double *y0);
int rfhip_op_state_store(rfhip_op *op, double *y);
/* nsteps free LSERK4 steps of dt on the resident state. */
int rfhip_op_state_lserk(rfhip_op *op, double dt, int64_t nsteps);
/* The same steps replayed from a captured HIP graph of two steps (ten stage
 * launches), cached per (dt, kernel variant) and per starting state buffer:
 * for small meshes, whose stage kernels run for microseconds, launch
 * overhead is what it removes. Results equal rfhip_op_state_lserk's bit for
 * bit. (No reference counterpart: the reference has no graph capture.) */
int rfhip_op_state_lserk_graph(rfhip_op *op, double dt, int64_t nsteps);
/* Several operators' resident states stepped together (the batching form;
 * no reference counterpart): nsteps free LSERK4 steps of dt[i] on ops[i],
 * the stage launches of two steps of every operator captured into one graph
 * that forks from ops[0]'s stream to every operator's stream and joins back,
 * so independent meshes (or several right-hand sides of one mesh, one
 * operator each) run concurrently. Each state ends as rfhip_op_state_lserk
 * would leave it. The operators must share a context and be distinct. */
int rfhip_batch_lserk(rfhip_op *const *ops, int32_t n_ops, const double *dt,
                      int64_t nsteps);
/* The batching form with port excitations (no reference counterpart): as
 * rfhip_batch_lserk, with ops[i] driven, when g[i] is not NULL, by its port
 * source as last assembled on the device (rfhip_op_port_source_f64) times
 * the waveform g[i][k] held over step k (nsteps samples; the reference
 * samples the waveform once per substep, session.rs:673-678, 885-927).
 * g[i] == NULL steps ops[i] free. The samples are uploaded once and indexed
 * on the device, so one captured graph of two steps of every operator serves
 * the whole run. RFHIP_EINVAL when a driven operator has no resident port
 * source. */
int rfhip_batch_lserk_ports(rfhip_op *const *ops, int32_t n_ops,
                            const double *dt, int64_t nsteps,
                            const double *const *g);
/* As rfhip_op_state_lserk, timed with HIP events on the operator's stream:
 * total_ms for the whole run and stage_ms_avg, the mean duration of one
 * launch of the dominant kernel (the fused stage kernel, or the plain form's
 * apply kernel). */
int rfhip_op_state_lserk_timed(rfhip_op *op, double dt, int64_t nsteps,
                               double *total_ms, double *stage_ms_avg);
/* Energy 0.5*y^T M~ y of the resident state (rhs.rs:1509-1544), FP64. */
int rfhip_op_state_energy(rfhip_op *op, const double *mass_ref,
                          const double *det, double *energy);

/* Allocation-free variant of rfhip_op_state_energy. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Every pointer is device memory the caller owns: y the f64 AoS
 * state (6*Np*K values of [E,H]; an appended Debye block is not read),
 * mass_ref (Np*Np), det (K), inv_eps and inv_mu (3K each, per element and
 * component); *energy receives 0.5 y^T M~ y and is ready when the call
 * returns. The reduction's partial sums use a fixed device array of 4096
 * doubles per GPU, so K is limited to 4096 * 256 = 1,048,576 elements
 * (RFHIP_EINVAL beyond). Calls that overlap on one GPU race on that array. */
int rfhip_op_energy_noalloc(rfhip_op *op, const double *y,
                            const double *mass_ref, const double *det,
                            const double *inv_eps, const double *inv_mu,
                            double *energy);

/* ---- sparse state-space export (rhs.rs:1575-1730) ------------------------ */

/* CsrMatrix (rhs.rs:1575-1591), held in device memory. */
typedef struct rfhip_csr rfhip_csr;

/* MaxwellOperator::assemble_sparse (rhs.rs:1621-1730) on the device: the
 * operator desc describes (with n_disp Debye slots as in
 * rfhip_op_create_dispersive; 0 for none) as an explicit CSR matrix, row
 * block e holding the nonzeros of e's stencil columns (e and its distinct
 * face neighbours, rhs.rs:1597-1610, then e's own P columns), each row's
 * columns ascending, then the appended P rows with exactly g and a. Values
 * and pattern are those of the reference's unit-column probe, bit for bit.
 * Orders 1..4. RFHIP_EINVAL for a malformed description or Debye list,
 * RFHIP_ENOMEM when the matrix does not fit in device memory. */
int rfhip_assemble_sparse(rfhip_ctx *ctx, const rfhip_op_desc *desc,
                          int64_t n_disp, const int64_t *disp_elem,
                          const double *disp_a, const double *disp_g,
                          const double *disp_inv_eps_inf, rfhip_csr **out);
void rfhip_csr_destroy(rfhip_csr *csr);
/* CsrMatrix::n and CsrMatrix::nnz (rhs.rs:1588-1590). */
int rfhip_csr_shape(const rfhip_csr *csr, int64_t *n, int64_t *nnz);
/* Copies row_ptr (n+1), col_idx (nnz) and values (nnz) to host memory; any
 * of the three may be NULL. */
int rfhip_csr_download(const rfhip_csr *csr, int64_t *row_ptr,
                       int64_t *col_idx, double *values);
/* The device arrays themselves, owned by csr (valid until
 * rfhip_csr_destroy), for consumers that stay on the GPU. */
int rfhip_csr_device(const rfhip_csr *csr, const int64_t **row_ptr,
                     const int64_t **col_idx, const double **values);
/* Durations of the counting and the filling pass over the rows (HIP events
 * on the assembly stream), in milliseconds. (No reference counterpart.) */
int rfhip_csr_timing(const rfhip_csr *csr, double *count_ms, double *fill_ms);

#ifdef __cplusplus
}
#endif
#endif
