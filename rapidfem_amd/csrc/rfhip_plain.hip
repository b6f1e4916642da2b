// Plain form: the reference GPU backend's kernel structure (one work-item per
// DG node of a block of EPG = 128/Np elements, an apply kernel followed by a
// separate stage kernel; crates/rapidfem-td/src/gpu/apply.cl:17-209 and
// lserk.cl:8-40), in f32 or f64. It is the "plain GPU form" the fused
// kernels are measured against, so it keeps that structure on purpose.
#include "rfhip_impl.hpp"

namespace rfhip {
namespace {

template <typename T>
__device__ __forceinline__ void cross3(const T *a, const T *b, T *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

constexpr int kPlainWG = 128;  // APPLY_TARGET_WG, gpu/operator.rs:68

// dy = A*y, one work-item per node. Phase 0 stages the element state in
// LDS, phase 1 is the volume curl with the D rows read per node, phase 2
// the upwind flux per face node (PEC ghost / port / neighbour through perm),
// phase 3 the LIFT sum and the material epilogue (rhs.rs:1341-1454).
template <typename T, int NP, int NFP>
__global__ void __launch_bounds__(kPlainWG) k_plain_apply(const T *__restrict__ y, T *__restrict__ dy, DevOp d) {
  constexpr int EPG = (kPlainWG / NP) > 0 ? kPlainWG / NP : 1;
  constexpr int NF4 = 4 * NFP;
  __shared__ T su[EPG][6 * NP];
  __shared__ T sf[EPG][6 * NF4];
  const int lid = threadIdx.x;
  const int le = lid / NP, node = lid % NP;
  const int64_t e = (int64_t)blockIdx.x * EPG + le;
  const bool active = (le < EPG) && (e < d.n_elem);
  const T *Dr = static_cast<const T *>(d.dr), *Ds = static_cast<const T *>(d.ds), *Dt = static_cast<const T *>(d.dt);
  const T *lift = static_cast<const T *>(d.lift);
  const T *jinv = static_cast<const T *>(d.jinv);
  if (active) {
    const int64_t base = (e * NP + node) * 6;
#pragma unroll
    for (int c = 0; c < 6; ++c) su[le][node * 6 + c] = y[base + c];
  }
  __syncthreads();
  T acc[6] = {0, 0, 0, 0, 0, 0};
  if (active) {
    T ji[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) ji[q] = jinv[e * 9 + q];
    // Volume term: dE = curl H, dH = -curl E (rhs.rs:1359-1368, element_curl :39-85).
    // The reference derivatives of all six components are accumulated in
    // one pass over the node's D rows (each entry read once, each sum in
    // the reference's j order), then mapped through J^-1 into the curls.
    T r[3][6];
#pragma unroll
    for (int c = 0; c < 6; ++c) r[0][c] = r[1][c] = r[2][c] = 0;
#pragma unroll 5
    for (int j = 0; j < NP; ++j) {
      const T dr = Dr[node * NP + j], ds = Ds[node * NP + j], dt = Dt[node * NP + j];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const T u = su[le][j * 6 + c];
        r[0][c] += dr * u;
        r[1][c] += ds * u;
        r[2][c] += dt * u;
      }
    }
#pragma unroll
    for (int fld = 0; fld < 2; ++fld) {
      const int off = fld == 0 ? 3 : 0;
      T pd[3][3];
#pragma unroll
      for (int comp = 0; comp < 3; ++comp)
#pragma unroll
        for (int ph = 0; ph < 3; ++ph)
          pd[ph][comp] = ji[0 * 3 + ph] * r[0][off + comp] + ji[1 * 3 + ph] * r[1][off + comp] +
                         ji[2 * 3 + ph] * r[2][off + comp];
      const T cx = pd[1][2] - pd[2][1], cy = pd[2][0] - pd[0][2], cz = pd[0][1] - pd[1][0];
      if (fld == 0) {
        acc[0] = cx;
        acc[1] = cy;
        acc[2] = cz;
      } else {
        acc[3] = -cx;
        acc[4] = -cy;
        acc[5] = -cz;
      }
    }
    // Surface term per face node (rhs.rs:1370-1441).
    const T *nrm_all = static_cast<const T *>(d.normal);
    const T *fsc = static_cast<const T *>(d.fscale);
    for (int fm = node; fm < NF4; fm += NP) {
      const int f = fm / NFP, m = fm - f * NFP;
      const int64_t q = e * 4 + f;
      const int vi = d.face_nodes[f * NFP + m];
      const T n[3] = {nrm_all[q * 3], nrm_all[q * 3 + 1], nrm_all[q * 3 + 2]};
      const T coef = T(0.5) * fsc[q];
      const int port = d.port[q];
      const int nb = d.nbr[q];
      T em[3], hm[3], je[3], jh[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        em[k] = su[le][vi * 6 + k];
        hm[k] = su[le][vi * 6 + 3 + k];
      }
      if (port >= 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          je[k] = em[k];
          jh[k] = hm[k];
        }
      } else if (nb < 0) {
        const T edn = n[0] * em[0] + n[1] * em[1] + n[2] * em[2];
        const T hdn = n[0] * hm[0] + n[1] * hm[1] + n[2] * hm[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          je[k] = T(2) * (em[k] - edn * n[k]);
          jh[k] = T(2) * hdn * n[k];
        }
      } else {
        const int vj = d.face_nodes[d.nbr_lf[q] * NFP + d.perm[q * NFP + m]];
        const int64_t nbb = ((int64_t)nb * NP + vj) * 6;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          je[k] = em[k] - y[nbb + k];
          jh[k] = hm[k] - y[nbb + 3 + k];
        }
      }
      T nxjh[3], nxje[3], pe[3], ph[3];
      cross3(n, jh, nxjh);
      cross3(n, je, nxje);
      cross3(n, nxje, pe);
      cross3(n, nxjh, ph);
      const T a = port >= 0 ? T(1) : T(d.alpha);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        sf[le][fm * 6 + k] = coef * (-nxjh[k] + a * pe[k]);
        sf[le][fm * 6 + 3 + k] = coef * (nxje[k] + a * ph[k]);
      }
    }
  }
  __syncthreads();
  if (active) {
    for (int fm = 0; fm < NF4; ++fm) {
      const T w = lift[node * NF4 + fm];
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[c] += w * sf[le][fm * 6 + c];
    }
    const T *ie = static_cast<const T *>(d.inv_eps) + e * 3, *im = static_cast<const T *>(d.inv_mu) + e * 3;
    const T *se = static_cast<const T *>(d.sig_eps) + e * 3, *sm = static_cast<const T *>(d.sig_mu) + e * 3;
    const int64_t base = (e * NP + node) * 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      dy[base + c] = ie[c] * acc[c] - se[c] * su[le][node * 6 + c];
      dy[base + 3 + c] = im[c] * acc[3 + c] - sm[c] * su[le][node * 6 + 3 + c];
    }
  }
}

// p = a*p + dt*k; y += b*p (lserk.cl:8-20, explicit.rs:126-133).
template <typename T>
__global__ void __launch_bounds__(256) k_stage(T *__restrict__ p, const T *__restrict__ k, T *__restrict__ y, T a,
                                               T b, T dt, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T pi = a * p[i] + dt * k[i];
  p[i] = pi;
  y[i] += b * pi;
}

template <typename T>
__global__ void k_add_point(T *k, int64_t dof, T val) {
  if (threadIdx.x == 0 && blockIdx.x == 0) k[dof] += val;
}

template <typename T>
__global__ void __launch_bounds__(256) k_add_vec(T *__restrict__ k, const double *__restrict__ src, double g,
                                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) k[i] += T(g * src[i]);
}

template <typename T, int NP>
void plain_apply_np(const rfhip_op *op, const void *y, void *dy, hipStream_t s) {
  constexpr int EPG = (kPlainWG / NP) > 0 ? kPlainWG / NP : 1;
  const int64_t groups = (op->n_elem + EPG - 1) / EPG;
  constexpr int NFP = NP == 4 ? 3 : NP == 10 ? 6 : NP == 20 ? 10 : 15;
  hipLaunchKernelGGL((k_plain_apply<T, NP, NFP>), dim3((unsigned)groups), dim3(EPG * NP), 0, s,
                     static_cast<const T *>(y), static_cast<T *>(dy), op->dev);
}

inline unsigned blocks_for(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_plain_apply(const rfhip_op *op, const void *y, void *dy, hipStream_t s) {
  const bool f64 = op->form == RFHIP_FORM_PLAIN_F64;
  switch (op->np) {
    case 4: f64 ? plain_apply_np<double, 4>(op, y, dy, s) : plain_apply_np<float, 4>(op, y, dy, s); break;
    case 10: f64 ? plain_apply_np<double, 10>(op, y, dy, s) : plain_apply_np<float, 10>(op, y, dy, s); break;
    case 20: f64 ? plain_apply_np<double, 20>(op, y, dy, s) : plain_apply_np<float, 20>(op, y, dy, s); break;
    case 35: f64 ? plain_apply_np<double, 35>(op, y, dy, s) : plain_apply_np<float, 35>(op, y, dy, s); break;
    default: throw Error{RFHIP_EUNSUPPORTED, "order not compiled (Np must be 4, 10, 20 or 35)"};
  }
  RF_HIP(hipGetLastError());
}

void launch_plain_stage(const rfhip_op *op, void *p, const void *k, void *y, double a, double b, double dt,
                        hipStream_t s) {
  const int64_t n = op->n_dof;
  if (op->form == RFHIP_FORM_PLAIN_F64)
    hipLaunchKernelGGL(k_stage<double>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<double *>(p),
                       static_cast<const double *>(k), static_cast<double *>(y), a, b, dt, n);
  else
    hipLaunchKernelGGL(k_stage<float>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<float *>(p),
                       static_cast<const float *>(k), static_cast<float *>(y), (float)a, (float)b, (float)dt, n);
  RF_HIP(hipGetLastError());
}

void launch_add_point(const rfhip_op *op, void *k, int64_t dof, double val, hipStream_t s) {
  if (op->form == RFHIP_FORM_PLAIN_F32)
    hipLaunchKernelGGL(k_add_point<float>, dim3(1), dim3(1), 0, s, static_cast<float *>(k), dof, (float)val);
  else
    hipLaunchKernelGGL(k_add_point<double>, dim3(1), dim3(1), 0, s, static_cast<double *>(k), dof, val);
  RF_HIP(hipGetLastError());
}

void launch_add_vec(const rfhip_op *op, void *k, const void *src, double g, hipStream_t s) {
  const int64_t n = op->n_dof;
  if (op->form == RFHIP_FORM_PLAIN_F32)
    hipLaunchKernelGGL(k_add_vec<float>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<float *>(k),
                       static_cast<const double *>(src), g, n);
  else
    hipLaunchKernelGGL(k_add_vec<double>, dim3(blocks_for(n)), dim3(256), 0, s, static_cast<double *>(k),
                       static_cast<const double *>(src), g, n);
  RF_HIP(hipGetLastError());
}

}  // namespace rfhip
