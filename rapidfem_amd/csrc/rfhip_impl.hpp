// Internal declarations shared by the rfhip translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rfhip.h"

namespace rfhip {

// ---- errors ------------------------------------------------------------

void set_error(const std::string &msg);

struct Error {
  int code;
  std::string msg;
};

#define RF_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::rfhip::Error{RFHIP_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define RF_REQUIRE(cond, msg)                                \
  do {                                                       \
    if (!(cond)) throw ::rfhip::Error{RFHIP_EINVAL, (msg)}; \
  } while (0)

// ---- device memory -----------------------------------------------------

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t nbytes) {
    if (nbytes <= bytes && p) return;
    release();
    if (nbytes == 0) nbytes = 8;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{RFHIP_ENOMEM, "hipMalloc of " + std::to_string(nbytes) + " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = nbytes;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Page-locked host memory (hipHostMalloc), for device-to-host copies that
// must stay asynchronous.
struct HostPinned {
  void *p = nullptr;
  size_t bytes = 0;
  HostPinned() = default;
  HostPinned(const HostPinned &) = delete;
  HostPinned &operator=(const HostPinned &) = delete;
  ~HostPinned() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t nbytes) {
    if (nbytes <= bytes && p) return;
    release();
    hipError_t e = hipHostMalloc(&p, nbytes, hipHostMallocDefault);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error{RFHIP_ENOMEM, "hipHostMalloc of " + std::to_string(nbytes) + " bytes failed: " + hipGetErrorString(e)};
    }
    bytes = nbytes;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

// Copies of large pageable host buffers through a pinned staging ring
// (rfhip_xfer.hip). copy_d2h returns with the data in dst; copy_h2d returns
// once src may be reused, the device copy ordered on s.
void copy_d2h(void *dst, const void *src, size_t bytes, hipStream_t s);
void copy_h2d(void *dst, const void *src, size_t bytes, hipStream_t s);

template <class T>
void upload(DevBuf &b, const T *host, size_t count, hipStream_t s) {
  b.alloc(sizeof(T) * count);
  RF_HIP(hipMemcpyAsync(b.p, host, sizeof(T) * count, hipMemcpyHostToDevice, s));
}

// ---- operator ----------------------------------------------------------

// Forms with the fused stage kernel; with compressed state registers (y f32
// hi + bf16 lo, p f32); and the BF16 pieces of the precision forms' volume
// contraction (0: FP64).
inline bool form_fused(int f) { return f >= RFHIP_FORM_FUSED_F64 && f <= RFHIP_FORM_FUSED_BF16X4; }
inline bool form_compressed(int f) { return f >= RFHIP_FORM_FUSED_F64C && f <= RFHIP_FORM_FUSED_BF16X4; }
inline int form_vmf(int f) {
  return f == RFHIP_FORM_FUSED_BF16X3 ? 3 : f == RFHIP_FORM_FUSED_BF16X2 ? 2 : f == RFHIP_FORM_FUSED_BF16X4 ? 4 : 0;
}

// Per-face code of the fused kernels, one int32 per (element, face):
//   interior: (neighbour << 5) | (neighbour_local_face << 3) | orientation
//   PEC wall: -1;  port p: -2 - p.
// orientation indexes the six vertex correspondences between the two
// copies of the shared triangle (see the neighbour-node table).
constexpr int kCodePec = -1;

// Device pointers of the plain form, passed by value.
struct DevOp {
  int np, nfp;
  int64_t n_elem;
  double alpha;
  const void *dr, *ds, *dt, *lift;  // Np*Np, Np*4Nfp (row-major), form precision
  const int *face_nodes;            // 4*Nfp
  const void *jinv, *inv_eps, *inv_mu, *sig_eps, *sig_mu, *normal, *fscale;
  const int *nbr, *nbr_lf, *port, *perm;
};

// Arguments of the fused stage kernel (rfhip_fused.hip), passed by value.
// Vectors use the reference's AoS layout y[(e*Np + node)*6 + comp].
struct StageArgs {
  int64_t n_elem;
  double alpha;
  int variant;            // kernel variant (rfhip_fused.hip)
  const double *dT;       // 3*Np*Np: dT[(k*Np + j)*Np + i] = D_k[i][j]
  const double *lift;     // LIFT column-major: lift[fm*Np + i]
  const void *afragb;     // split-BF16 A fragments of [Dr; Ds; Dt] (precision forms, k_fused11 VMF)
  const double *lfrag;    // FP64 MFMA A fragments of LIFT (k_fused8)
  const double *afrag7;   // node-interleaved A fragments of [Dr; Ds; Dt] (k_fused8)
  const double *jinv;     // 9K
  const int *fcode;       // 4K
  const int *mat;         // K
  const double *mat_tab;  // n_mat * 12: inv_eps3, inv_mu3, sig_eps3, sig_mu3
  const int *face_nodes;  // 4*Nfp
  const int *nbr_node;    // [(f*4 + nlf)*6 + orientation][Nfp] -> neighbour volume node
  const int *nbase;       // [e*4Nfp + fm]: AoS index of the neighbour trace, or the (negative) face code
  const double *clane;    // nonzeros of D_k then LIFT, in k_clane's order (clane_pack), or null
  const void *yin;        // double, or float hi (compressed)
  const void *yin_lo;     // bf16 lo (compressed) or null
  void *yout, *yout_lo;
  void *p;                // double, or float (compressed)
  double *dy;             // apply output
  double sa, sb, dt;      // stage coefficients
  int src_kind;
  int64_t src_dof;
  const double *src_vec;
  double g;
  // KCL stage with compressed registers (launch_kcl_stage): yin / yin_lo
  // is the stage input and yout / yout_lo the next one, f32 hi + bf16 lo
  // like y and the substep backup; err f32.
  float *ky, *kerr, *kbak;
  uint16_t *ky_lo, *kbak_lo;
  double kb, kew, kamb;   // b_s, bhat_s - b_s, a_s - b_s (0 on the last stage)
  double katol, krtol;
  int kfirst, klast;
  double *kpart;          // last stage: one partial sum of the error norm per wave of the grid
  // the same with FP64 registers (launch_kcl_stage, r64): y, err and the
  // backup FP64, the stage inputs yin / yout FP64
  double *kyd, *kerrd, *kbakd;
};

}  // namespace rfhip

struct rfhip_ctx {
  int device = 0;
  std::string name;
};

struct rfhip_op {
  rfhip_ctx *ctx = nullptr;
  int form = 0;
  int np = 0, nfp = 0;
  int64_t n_elem = 0, n_dof = 0;
  double alpha = 1.0;
  hipStream_t stream = nullptr;
  // host copies used by the host-side logic
  std::vector<int32_t> port_host;  // 4K
  std::vector<int32_t> face_nodes_host;  // 4*Nfp
  // device operator data
  rfhip::DevBuf dr, ds, dt, lift, face_nodes, jinv, inv_eps, inv_mu, sig_eps, sig_mu, normal, fscale, nbr,
      nbr_lf, port, perm;
  rfhip::DevBuf dT, liftc, lfrag, afrag7, afragb, clane, jinv64, fcode, mat, mat_tab, nbr_node, nbase;
  int64_t n_vec = 0;  // length of device vectors (n_dof)
  // Debye dispersive block: slot -> element and (a, g, 1/eps_inf); the P
  // block follows the 6 Np K values of [E,H] in every state vector
  int64_t n_disp = 0;
  int kcl_reg = 0;  // KCL registers: RFHIP_KCL_REG_* (rfhip_op_set_kcl_registers)
  int kb_mode = 0;  // adaptive Arnoldi basis: RFHIP_KRYLOV_BASIS_* (rfhip_op_set_krylov_basis)
  rfhip::DevBuf disp_elem, disp_coef;
  int order = 0;
  // state registers (precision depends on form)
  rfhip::DevBuf y[2], ylo[2], p, k, src;
  int cur = 0;  // index of the current state buffer for fused forms
  // KCL / Krylov / scratch buffers
  rfhip::DevBuf kst, kerr, kbak, kyc, red, aos;
  rfhip::DevBuf kb_basis, kb_w, kb_proj, kb_part, kb_out, kb_x;
  rfhip::DevBuf kb_cbasis;  // compressed finished columns of the adaptive Arnoldi
  rfhip::HostPinned kb_host;  // Krylov projections and norm, copied without a sync per pass
  rfhip::DevBuf probe_idx, probe_rows;  // device-side probe recording
  // port modal extraction: host-side per-port entries, device copies
  struct PortEntries {
    std::vector<int64_t> dof;
    std::vector<int32_t> face_elem, face_lf;  // the port faces (element, local face)
    std::vector<double> prof, w;
    double e_norm = 0.0, h_norm = 0.0;
  };
  std::vector<PortEntries> ports;
  rfhip::DevBuf port_dof, port_prof, port_w, port_off, port_norm, port_rows;
  rfhip::DevBuf ps_elem, ps_off, ps_lf, ps_prof;  // port-source assembly input
  int64_t ps_count = 0;  // elements with port faces of the assembled source (ps_elem)
  // waveform samples of a batched port drive and the step index into them
  rfhip::DevBuf gseq, gctr;
  bool src_resident = false;  // src holds a device-assembled port source
  // wall time of the device part of the last Krylov or KCL call (input
  // resident to output ready, host transfers excluded), milliseconds
  double last_call_ms = 0.0;
  // captured two-step LSERK4 graphs of the resident stepper, one per starting
  // state buffer, with the (dt, variant) they were captured for
  struct StepGraph {
    hipGraphExec_t exec = nullptr;
    double dt = 0.0;
    int variant = -1;
  };
  StepGraph step_graph[2];
  bool ports_dirty = true;
  int kb_cap = 0;
  rfhip::DevOp dev{};
  rfhip::StageArgs sa{};  // fused forms: geometry part filled at creation
};

// CsrMatrix (rhs.rs:1575-1591) held in HBM: row_ptr n+1, col_idx and values
// nnz, column indices as int64 (the reference's usize).
struct rfhip_csr {
  rfhip_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  int64_t n = 0, nnz = 0;
  float count_ms = 0.0f, fill_ms = 0.0f;  // the two row passes, HIP events
  rfhip::DevBuf row_ptr, col_idx, values;
  ~rfhip_csr() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace rfhip {

// Kernel launchers (implemented in the .hip files).
void launch_plain_apply(const rfhip_op *op, const void *y, void *dy, hipStream_t s);
void launch_plain_stage(const rfhip_op *op, void *p, const void *k, void *y, double a, double b, double dt,
                        hipStream_t s);
void launch_add_point(const rfhip_op *op, void *k, int64_t dof, double val, hipStream_t s);
void launch_add_vec(const rfhip_op *op, void *k, const void *src, double g, hipStream_t s);

// Fused forms: epi 0 = apply (dy), 1 = LSERK stage (FP64 or compressed
// registers by form). Stage-kernel variants 0 .. kStageVariants-1 (the
// precision forms: 0 .. kStageVariantsBf16-1).
constexpr int kStageVariants = 9;
constexpr int kStageVariantsBf16 = 2;
void launch_stage(int order, int form, int epi, const StageArgs &a, hipStream_t s);
// Whether stage-kernel variant `variant` runs as itself on this operator's
// data (some need the neighbour-trace table or Np <= 20; launch_stage runs
// the base variant for them otherwise).
bool variant_available(int order, int form, int variant, const StageArgs &a);
// One KCL stage with the register update in the stage kernel's epilogue
// (k_fused11): with compressed registers (EPI 3) the apply of a.yin /
// a.yin_lo, the register update and the next stage's input into a.yout /
// a.yout_lo; with FP64 registers (r64, EPI 4) the same on a.kyd, a.kerrd,
// a.kbakd and FP64 stage inputs. Needs the neighbour-trace table and a form
// without the matrix-core volume; runs variant 6 when a.variant is 6 and
// variant 7 otherwise. kcl_stage_slots: the number of partial sums the last
// stage leaves in a.kpart.
bool kcl_stage_fusable(int form, const StageArgs &a);
int64_t kcl_stage_slots(int order, const StageArgs &a, bool r64);
void launch_kcl_stage(int order, const StageArgs &a, bool r64, hipStream_t s);
// The table of the component-lane stage kernel (variant 8): the nonzeros of
// D_r, D_s, D_t (row-major D_k[i][j], the order of (i, j, k)) and then LIFT
// ((face, node, face node) order). False when the operator's face nodes or
// the zeros of its D_k differ from the compiled equispaced tables.
bool clane_pack(int np, const int *face_nodes, const double *const D[3], const double *lift, std::vector<double> &out);
// Debye block after an applied derivative (dy = A y; f32: plain FP32 form),
// with the drive of the P rows, and after a fused LSERK4 stage
// (rfhip_debye.hip). No-ops without Debye elements.
void launch_debye_apply(const rfhip_op *op, const void *y, void *dy, bool f32, int kind, int64_t dof,
                        const double *vec, double g, hipStream_t s);
void launch_debye_stage(const rfhip_op *op, const StageArgs &a, hipStream_t s);

// The held port drive of one LSERK4 stage read from op->gseq[*op->gctr]
// over the port elements of the assembled source (rfhip_fused.hip): before
// (plain forms) k += b g; after a fused stage p += dt b g, y_out += sb dt b g.
void launch_port_stage(const rfhip_op *op, bool before, const StageArgs &a, void *k, double dt, double sb,
                       hipStream_t s);
// *op->gctr += 1: the next step's waveform sample.
void launch_step_advance(const rfhip_op *op, hipStream_t s);

// Port source b (rhs.rs:1157-1199) into the zeroed f64 vector b for the port
// faces grouped by element: element elem[u] has faces [off[u], off[u+1]) with
// local faces lf[q] and profiles prof[q*Nfp*6 ...] (rfhip_fused.hip).
void launch_port_source(const rfhip_op *op, int64_t n_elems, const int64_t *elem, const int64_t *off,
                        const int32_t *lf, const double *prof, double *b, hipStream_t s);

// MaxwellOperator::assemble_sparse (rhs.rs:1621-1730) of the operator desc
// describes, with n_disp Debye slots (elements disp_elem, coefficients
// disp_coef[3q..] = a, g, 1/eps_inf), into csr (rfhip_sparse.hip). The
// description's indices must already be validated.
void assemble_sparse(rfhip_csr *csr, const rfhip_op_desc &d, int64_t n_disp, const int64_t *disp_elem,
                     const double *disp_coef);


// Conversions between boundary AoS and the form's storage.
void launch_f64_to_c(const double *src, float *hi, uint16_t *lo, int64_t n, hipStream_t s);
void launch_c_to_f64(const float *hi, const uint16_t *lo, double *dst, int64_t n, hipStream_t s);
void launch_cast_d2f(const double *src, float *dst, int64_t n, hipStream_t s);
void launch_cast_f2d(const float *src, double *dst, int64_t n, hipStream_t s);
void launch_fill(void *dst, int64_t nbytes, hipStream_t s);
void launch_poison_lds(hipStream_t s);
// Port modal projections (P_e, P_h) of every port into out[2p], out[2p+1].
void launch_port_proj(const void *y, const uint16_t *lo, int kind, const int64_t *dof, const double *prof,
                      const double *w, const int64_t *off, const double *norm, int n_ports, double *out,
                      hipStream_t s);
// out[q] = y[idx[q]] as f64; kind 0 f64 state, 1 f32, 2 f32 hi + bf16 lo.
void launch_gather(const void *y, const uint16_t *lo, int kind, const int64_t *idx, int64_t n, double *out,
                   hipStream_t s);

// One KCL stage's register update on T = float or double (f64 selects):
// f = dt k, y += b f, err = ew f (first) or err += ew f, st = y + amb f
// (next). The first stage saves the starting y into bak; the last (no next)
// reduces sum((err/(atol + rtol*max(|bak|,|y|)))^2) into *err_sum (device
// double, +inf for a non-finite y) through nblk partials.
struct KclStage {
  void *y, *err, *st, *bak;
  uint16_t *y_lo, *bak_lo;  // compressed registers: bf16 lo parts (launch_kcl_update_c)
  const void *k;
  double dt, b, ew, amb;
  bool first, next;
  int64_t n;
  double atol, rtol;
  double *partials;
  int nblk;
  double *err_sum;
};
void launch_kcl_update(bool f64, const KclStage &a, hipStream_t s);
// Compressed registers (y, bak f32 + bf16, err f32, st f64); the first
// stage is never the last.
void launch_kcl_update_c(const KclStage &a, hipStream_t s);
void launch_copy(void *dst, const void *src, int64_t nbytes, hipStream_t s);
int weighted_err_blocks(int64_t n);
// out[0] = the sum of part[0 .. n) in a fixed order (one block).
void launch_sum_partials(const double *part, int n, double *out, hipStream_t s);

// Krylov (FP64).
int krylov_dot_blocks(int64_t n);
void launch_multi_dot(const double *basis, int64_t n, int cols, const double *w, double *partials, int nblk,
                      double *proj, hipStream_t s);
void launch_axpy_basis_norm(double *w, const double *basis, int64_t n, int cols, const double *proj,
                            double *partials, int nblk, double *norm2, hipStream_t s);
// CGS2 with delayed reorthogonalisation (rfhip_vec.hip): out[2c + r] =
// b_c . b_{cols-2+r} for c < cols; the update sweep of column j (coef =
// [s (j), a (j), bq, 1/r]) and |b_{j+1}|^2; out = sum coef_c b_c with out
// allowed to be a column.
void launch_gram2(const double *basis, int64_t n, int cols, double *partials, int nblk, double *out, hipStream_t s);
void launch_dcgs_update(double *basis, int64_t n, int j, const double *coef, double *partials, int nblk,
                        double *norm2, hipStream_t s);
void launch_lincomb_alias(const double *basis, int64_t n, int dim, const double *coef, double *out, hipStream_t s);
// The same three on a compressed basis (rfhip_vec.hip): finished columns as
// f32 hi[c n + i] + bf16 lo[c nl + i] (nl = n rounded up to even), the
// provisional q and the product z FP64 vectors of their own. gram2c: out[2c
// + r] = b_c . (q, z)_r for c <= j, b_j = q; dcgs_updatec: writes v_j as
// column j and the next provisional vector into q, |q|^2 to norm2;
// lincombc: out = sum_{c<dim} coef_c v_c.
void launch_gram2c(const float *hi, const uint16_t *lo, int64_t n, int64_t nl, int j, const double *q,
                   const double *z, double *partials, int nblk, double *out, hipStream_t s);
void launch_dcgs_updatec(float *hi, uint16_t *lo, int64_t n, int64_t nl, int j, const double *coef, double *q,
                         const double *z, double *partials, int nblk, double *norm2, hipStream_t s);
void launch_lincombc(const float *hi, const uint16_t *lo, int64_t n, int64_t nl, int dim, const double *coef,
                     double *out, hipStream_t s);
// w -= basis[:, :cols] proj, then proj_out = basis[:, :cols]^T w (the updated
// w): a CGS2 first-pass update fused with the second pass's projections.
// proj_out may be proj: it is written after every block has read proj.
void launch_axpy_dot(double *w, const double *basis, int64_t n, int cols, const double *proj, double *partials,
                     int nblk, double *proj_out, hipStream_t s);
void launch_scale(const double *w, double *dst, int64_t n, double inv, hipStream_t s);
void launch_lincomb(const double *basis, int64_t n, int dim, const double *coef, double *out, hipStream_t s);
void launch_axpy(double *w, const double *x, double a, int64_t n, hipStream_t s);
void launch_energy(const double *y, const double *mass, const double *det, const double *inv_eps,
                   const double *inv_mu, int np, int64_t K, double *partials, int nblk, double *out,
                   hipStream_t s);

// The same energy with a fixed grid of nblk blocks (each of
// energy_elems_per_block() elements covers at most one element per thread
// when nblk * that >= K), written as scale * sum to out on the device.
int energy_elems_per_block();
void launch_energy_scaled(const double *y, const double *mass, const double *det, const double *inv_eps,
                          const double *inv_mu, int np, int64_t K, double *partials, int nblk, double scale,
                          double *out, hipStream_t s);
// The allocation-free entry points' scratch array (rfhip_noalloc.hip): its
// device address on `device`, looked up once with that device current.
void noalloc_init(int device);

// Host-side dense exponential (propagator.rs:69-114).
void expm_host(const double *a, int n, double *out);

}  // namespace rfhip
