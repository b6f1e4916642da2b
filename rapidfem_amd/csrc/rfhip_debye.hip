// Debye dispersive block (rhs.rs:469-500 DispersiveElem and the appended
// polarisation block, rhs.rs:1295-1324 its rows, rhs.rs:1456-1472 the
// polarisation current on the E rows; dispersive.rs:50-53 the relaxation
// coefficients). The state carries, after the [E,H] block of 6 Np K values,
// 3 Np values P[(slot*Np + node)*3 + c] per Debye element; with
// Pdot = a*P + g*E (a = -1/tau, g = (eps_s - eps_inf)/tau) the E rows of that
// element gain -Pdot/eps_inf and the P rows are Pdot.
//
// The block is element-local and linear, so it runs as a small kernel after
// the main operator kernel, one thread per (slot, node, component):
//  * k_debye_apply adds it to an applied derivative dy = A y (the plain forms,
//    the fused apply, KCL and Krylov);
//  * k_debye_stage completes a fused LSERK4 stage. The stage kernel has
//    already written p_E = sa p_E + dt k_main and y_E' = y_E + sb p_E for
//    every node; with k = k_main + k_d this kernel adds dt k_d to p_E and
//    sb dt k_d to y_E', reading E and P from the stage's input buffer (the
//    stage writes the other one), and steps P itself.
#include "rfhip_impl.hpp"
#include "rfhip_regs.hpp"

namespace rfhip {
namespace {

constexpr int kDB = 256;


// Register formats: 0 = f64, 1 = f32, 2 = f32 hi + bf16 lo.
template <int FMT>
__device__ __forceinline__ double reg_load(const void *y, const uint16_t *lo, int64_t i) {
  if constexpr (FMT == 0) return static_cast<const double *>(y)[i];
  else if constexpr (FMT == 1) return (double)static_cast<const float *>(y)[i];
  else return (double)static_cast<const float *>(y)[i] + (double)bf16_to_f(lo[i]);
}
template <int FMT>
__device__ __forceinline__ void reg_store(void *y, uint16_t *lo, int64_t i, double v) {
  if constexpr (FMT == 0) {
    static_cast<double *>(y)[i] = v;
  } else if constexpr (FMT == 1) {
    static_cast<float *>(y)[i] = (float)v;
  } else {
    const float h = (float)v;
    static_cast<float *>(y)[i] = h;
    lo[i] = bf16_rne((float)(v - (double)h));
  }
}

struct DebyeArgs {
  const int64_t *elem;   // slot -> element
  const double *coef;    // slot -> (a, g, 1/eps_inf)
  int64_t n_disp;
  int np;
  int64_t eh;            // 6 Np K, the offset of the P block
  int src_kind;          // drive of the P rows (the main kernel drives [E,H])
  int64_t src_dof;
  const double *src_vec;
  double g;
};

__device__ __forceinline__ double p_source(const DebyeArgs &d, int64_t ip) {
  if (d.src_kind == RFHIP_DRIVE_POINT) return ip == d.src_dof ? d.g : 0.0;
  if (d.src_kind == RFHIP_DRIVE_VECTOR) return d.src_vec[ip] * d.g;
  return 0.0;
}

// dy of the Debye rows; FMT is the format of y and dy (0 f64, 1 f32).
template <int FMT>
__global__ void __launch_bounds__(kDB) k_debye_apply(const void *__restrict__ y, void *__restrict__ dy, DebyeArgs d) {
  const int64_t n = d.n_disp * d.np * 3;
  for (int64_t t = (int64_t)blockIdx.x * kDB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kDB) {
    const int64_t slot = t / (3 * d.np);
    const int r = (int)(t - slot * 3 * d.np), node = r / 3, c = r - node * 3;
    const double a = d.coef[3 * slot], g = d.coef[3 * slot + 1], iei = d.coef[3 * slot + 2];
    const int64_t ie = (d.elem[slot] * d.np + node) * 6 + c, ip = d.eh + t;
    const double pd = a * reg_load<FMT>(y, nullptr, ip) + g * reg_load<FMT>(y, nullptr, ie);
    reg_store<FMT>(dy, nullptr, ie, reg_load<FMT>(dy, nullptr, ie) - iei * pd);
    reg_store<FMT>(dy, nullptr, ip, pd + p_source(d, ip));
  }
}

// Completion of a fused LSERK4 stage. YF: state format (0 f64, 2 f32+bf16);
// PF: format of the p register (0 f64, 1 f32).
template <int YF, int PF>
__global__ void __launch_bounds__(kDB) k_debye_stage(const void *__restrict__ yin, const uint16_t *__restrict__ yin_lo,
                                                     void *__restrict__ yout, uint16_t *__restrict__ yout_lo,
                                                     void *__restrict__ p, double sa, double sb, double dt,
                                                     DebyeArgs d) {
  const int64_t n = d.n_disp * d.np * 3;
  for (int64_t t = (int64_t)blockIdx.x * kDB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kDB) {
    const int64_t slot = t / (3 * d.np);
    const int r = (int)(t - slot * 3 * d.np), node = r / 3, c = r - node * 3;
    const double a = d.coef[3 * slot], g = d.coef[3 * slot + 1], iei = d.coef[3 * slot + 2];
    const int64_t ie = (d.elem[slot] * d.np + node) * 6 + c, ip = d.eh + t;
    const double P = reg_load<YF>(yin, yin_lo, ip);
    const double pd = a * P + g * reg_load<YF>(yin, yin_lo, ie);
    // E: the stage kernel's p_E and y_E' plus the polarisation current
    const double kd = -iei * pd;
    reg_store<PF>(p, nullptr, ie, reg_load<PF>(p, nullptr, ie) + dt * kd);
    reg_store<YF>(yout, yout_lo, ie, reg_load<YF>(yout, yout_lo, ie) + sb * dt * kd);
    // P: its own low-storage step
    const double pp = sa * reg_load<PF>(p, nullptr, ip) + dt * (pd + p_source(d, ip));
    reg_store<PF>(p, nullptr, ip, pp);
    reg_store<YF>(yout, yout_lo, ip, P + sb * pp);
  }
}

DebyeArgs debye_args(const rfhip_op *op, int kind, int64_t dof, const double *vec, double g) {
  DebyeArgs d;
  d.elem = op->disp_elem.as<int64_t>();
  d.coef = op->disp_coef.as<double>();
  d.n_disp = op->n_disp;
  d.np = op->np;
  d.eh = 6 * (int64_t)op->np * op->n_elem;
  d.src_kind = kind;
  d.src_dof = dof;
  d.src_vec = vec;
  d.g = g;
  return d;
}

unsigned grid_of(int64_t n) {
  int64_t b = (n + kDB - 1) / kDB;
  if (b > 8192) b = 8192;
  return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

void launch_debye_apply(const rfhip_op *op, const void *y, void *dy, bool f32, int kind, int64_t dof,
                        const double *vec, double g, hipStream_t s) {
  if (op->n_disp == 0) return;
  const DebyeArgs d = debye_args(op, kind, dof, vec, g);
  const unsigned grid = grid_of(d.n_disp * d.np * 3);
  if (f32)
    hipLaunchKernelGGL(k_debye_apply<1>, dim3(grid), dim3(kDB), 0, s, y, dy, d);
  else
    hipLaunchKernelGGL(k_debye_apply<0>, dim3(grid), dim3(kDB), 0, s, y, dy, d);
  RF_HIP(hipGetLastError());
}

void launch_debye_stage(const rfhip_op *op, const StageArgs &a, hipStream_t s) {
  if (op->n_disp == 0) return;
  const DebyeArgs d = debye_args(op, a.src_kind, a.src_dof, a.src_vec, a.g);
  const unsigned grid = grid_of(d.n_disp * d.np * 3);
  if (form_compressed(op->form))
    hipLaunchKernelGGL((k_debye_stage<2, 1>), dim3(grid), dim3(kDB), 0, s, a.yin,
                       static_cast<const uint16_t *>(a.yin_lo), a.yout, static_cast<uint16_t *>(a.yout_lo), a.p,
                       a.sa, a.sb, a.dt, d);
  else
    hipLaunchKernelGGL((k_debye_stage<0, 0>), dim3(grid), dim3(kDB), 0, s, a.yin, nullptr, a.yout, nullptr, a.p,
                       a.sa, a.sb, a.dt, d);
  RF_HIP(hipGetLastError());
}

}  // namespace rfhip
