// Streaming kernels around the operator: register conversions, the KCL
// stage updates and error norm (gpu/kcl.cl:12-112), the Krylov CGS2
// building blocks (gpu/expmv.cl:34-165 restructured: one multi-dot pass per
// CGS2 sweep instead of one work-group per basis row) and the field energy
// (rhs.rs:1509-1544). All HBM-bound; grid-stride loops over 256-thread
// blocks.
#include <type_traits>

#include "rfhip_impl.hpp"
#include "rfhip_regs.hpp"

namespace rfhip {
namespace {

constexpr int kBS = 256;

inline unsigned grid_for(int64_t n, int64_t cap = 8192) {
  int64_t g = (n + kBS - 1) / kBS;
  if (g > cap) g = cap;
  if (g < 1) g = 1;
  return (unsigned)g;
}


__global__ void __launch_bounds__(kBS) k_f64_to_c(const double *__restrict__ s, float *__restrict__ hi,
                                                  uint16_t *__restrict__ lo, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double v = s[i];
    const float h = (float)v;
    hi[i] = h;
    lo[i] = bf16_rne((float)(v - (double)h));
  }
}
// Probe rows (session.rs:118-123 Record::Dofs, recorded on the device):
// out[q] = y[idx[q]] for a state stored as f64 (kind 0), f32 (1) or
// f32 hi + bf16 lo (2).
__global__ void __launch_bounds__(kBS) k_gather(const void *__restrict__ y, const uint16_t *__restrict__ lo, int kind,
                                                const int64_t *__restrict__ idx, int64_t n, double *__restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * kBS + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBS) {
    const int64_t i = idx[q];
    double v;
    if (kind == 0)
      v = static_cast<const double *>(y)[i];
    else if (kind == 1)
      v = (double)static_cast<const float *>(y)[i];
    else
      v = (double)static_cast<const float *>(y)[i] + (double)bf16_to_f(lo[i]);
    out[q] = v;
  }
}

// Port modal projections (rhs.rs:1211-1242) of the current state, one block
// per port: out[2p] = sum_w w (E.e_t) / e_norm, out[2p+1] = sum_w w (H.h_t) /
// h_norm over the port's face-node entries [off[p], off[p+1]). Entry q has
// the AoS index of its node's first field value, the profile (e_t, h_t) and
// the weight 2 * area * face-node weight.
__global__ void __launch_bounds__(kBS) k_port_proj(const void *__restrict__ y, const uint16_t *__restrict__ lo,
                                                   int kind, const int64_t *__restrict__ dof,
                                                   const double *__restrict__ prof, const double *__restrict__ w,
                                                   const int64_t *__restrict__ off, const double *__restrict__ norm,
                                                   double *__restrict__ out) {
  __shared__ double red[2][kBS];
  const int p = blockIdx.x;
  double se = 0.0, sh = 0.0;
  for (int64_t q = off[p] + threadIdx.x; q < off[p + 1]; q += kBS) {
    const int64_t b = dof[q];
    double v[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (kind == 0)
        v[c] = static_cast<const double *>(y)[b + c];
      else if (kind == 1)
        v[c] = (double)static_cast<const float *>(y)[b + c];
      else
        v[c] = (double)static_cast<const float *>(y)[b + c] + (double)bf16_to_f(lo[b + c]);
    }
    const double *pr = prof + 6 * q;
    se += w[q] * (v[0] * pr[0] + v[1] * pr[1] + v[2] * pr[2]);
    sh += w[q] * (v[3] * pr[3] + v[4] * pr[4] + v[5] * pr[5]);
  }
  red[0][threadIdx.x] = se;
  red[1][threadIdx.x] = sh;
  __syncthreads();
  for (int st = kBS / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      red[0][threadIdx.x] += red[0][threadIdx.x + st];
      red[1][threadIdx.x] += red[1][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * p] = norm[2 * p] > 0.0 ? red[0][0] / norm[2 * p] : 0.0;
    out[2 * p + 1] = norm[2 * p + 1] > 0.0 ? red[1][0] / norm[2 * p + 1] : 0.0;
  }
}

__global__ void __launch_bounds__(kBS) k_c_to_f64(const float *__restrict__ hi, const uint16_t *__restrict__ lo,
                                                  double *__restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS)
    d[i] = (double)hi[i] + (double)bf16_to_f(lo[i]);
}
__global__ void __launch_bounds__(kBS) k_d2f(const double *__restrict__ s, float *__restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) d[i] = (float)s[i];
}
__global__ void __launch_bounds__(kBS) k_f2d(const float *__restrict__ s, double *__restrict__ d, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) d[i] = (double)s[i];
}
__global__ void __launch_bounds__(kBS) k_fill0(uint32_t *__restrict__ d, int64_t nwords) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * kBS) d[i] = 0u;
}
__global__ void __launch_bounds__(kBS) k_copy(uint32_t *__restrict__ d, const uint32_t *__restrict__ s, int64_t nw) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < nw; i += (int64_t)gridDim.x * kBS) d[i] = s[i];
}

// ---- KCL RK4(3)5[2R+]C (explicit_adaptive.rs:84-152) -------------------

__device__ __forceinline__ double block_sum(double v, double *sh);

// After stage s of a KCL substep: f = h k, y += b f, err = e f on the first
// stage and err += e f after it, and, when NEXT, the next stage's input
// st = y + (a - b) f with the updated y (the 2R update and the next stage's
// build in one pass; f itself is not stored). The first stage also saves
// the substep's starting y into bak (restored on a rejection); the last
// (no NEXT) also reads bak and leaves the block's partial sum of the error
// norm's entries in part[blockIdx.x].
template <typename T, bool FIRST, bool NEXT>
__global__ void __launch_bounds__(kBS) k_kcl_update(T *__restrict__ y, T *__restrict__ err, T *__restrict__ st,
                                                    const T *__restrict__ k, T *__restrict__ bak, T dt, T b, T ew,
                                                    T amb, double atol, double rtol, int64_t n,
                                                    double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const T yo = y[i];
    const T f = dt * k[i];
    const T yn = yo + b * f;
    y[i] = yn;
    if (FIRST) bak[i] = yo;
    const T en = FIRST ? ew * f : err[i] + ew * f;
    err[i] = en;
    if (NEXT)
      st[i] = yn + amb * f;
    else
      acc += werr_term((double)bak[i], (double)yn, (double)en, atol, rtol);
  }
  if (!NEXT) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// The same pass with compressed registers: y and the substep's starting y
// (bak) as f32 hi + bf16 lo, err as f32, the next stage's input st in FP64
// (what the apply reads). 6 + 6 + 4 + 4 + 8 (+ 8 for k) bytes per entry
// against 48 with FP64 registers. LAST: no next stage; st = y_new, the next
// substep's first input, and the block's partial sum of the error norm.
template <bool FIRST, bool LAST>
__global__ void __launch_bounds__(kBS) k_kcl_update_c(float *__restrict__ yh, uint16_t *__restrict__ yl,
                                                      float *__restrict__ err, double *__restrict__ st,
                                                      const double *__restrict__ k, float *__restrict__ bh,
                                                      uint16_t *__restrict__ bl, double dt, double b, double ew,
                                                      double amb, double atol, double rtol, int64_t n,
                                                      double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const float h0 = yh[i];
    const uint16_t l0 = yl[i];
    const double f = dt * k[i];
    const double yn = ((double)h0 + (double)bf16_to_f(l0)) + b * f;
    const float hn = (float)yn;
    yh[i] = hn;
    yl[i] = bf16_rne((float)(yn - (double)hn));
    if (FIRST) {
      bh[i] = h0;
      bl[i] = l0;
    }
    const double en = FIRST ? ew * f : (double)err[i] + ew * f;
    err[i] = (float)en;
    st[i] = LAST ? yn : yn + amb * f;
    if (LAST) acc += werr_term((double)bh[i] + (double)bf16_to_f(bl[i]), yn, en, atol, rtol);
  }
  if (LAST) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
  // wave64 shuffle reduction, then across the block's waves.
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh[q];
  __syncthreads();
  return t;  // valid on thread 0
}

// Sums nblk partials into out[0] with one block (fixed order: deterministic).
__global__ void __launch_bounds__(kBS) k_finish_sum(const double *__restrict__ part, int nblk, double *out) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kBS) acc += part[i];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = t;
}

// out[0] = scale * sum of part[0..nblk).
__global__ void __launch_bounds__(kBS) k_finish_sum_scaled(const double *__restrict__ part, int nblk, double scale,
                                                           double *out) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += kBS) acc += part[i];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = scale * t;
}

// ---- Krylov CGS2 pieces (propagator.rs:231-271) -------------------------

constexpr int kMaxCols = 64;

// partials[b*cols + i] = sum over block b's slice of basis_i . w. Columns
// are taken sixteen at a time so w is read ceil(cols/16) times, the basis
// once. The streaming updates take their columns eight at a time.
constexpr int kDotTile = 8;
constexpr int kDotTileM = 16;
// Largest column count k_axpy_dot holds in registers; above it the update
// and the products run as two sweeps.
constexpr int kAxpyDotMaxCols = 48;
// Per-block partial dot products of w with basis columns 0..cols-1, kDotTileM
// columns per sweep over the block's contiguous range. With VEC (n even, so
// every column starts 16-byte aligned) each lane reads two consecutive
// entries per 16-byte load.
template <bool VEC>
__global__ void __launch_bounds__(kBS) k_multi_dot(const double *__restrict__ basis, int64_t n, int cols,
                                                   const double *__restrict__ w, double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  const int64_t m = VEC ? n / 2 : n;  // items: pairs or single entries
  const int64_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < m ? i0 + chunk : m;
  for (int c0 = 0; c0 < cols; c0 += kDotTileM) {
    double acc[kDotTileM];
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) acc[t] = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kBS) {
      if constexpr (VEC) {
        const double2 wi = reinterpret_cast<const double2 *>(w)[i];
#pragma unroll
        for (int t = 0; t < kDotTileM; ++t)
          if (c0 + t < cols) {
            const double2 b = reinterpret_cast<const double2 *>(basis + (int64_t)(c0 + t) * n)[i];
            acc[t] += b.x * wi.x + b.y * wi.y;
          }
      } else {
        const double wi = w[i];
#pragma unroll
        for (int t = 0; t < kDotTileM; ++t)
          if (c0 + t < cols) acc[t] += basis[(int64_t)(c0 + t) * n + i] * wi;
      }
    }
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) {
      if (c0 + t >= cols) break;
      const double s = block_sum(acc[t], sh);
      if (threadIdx.x == 0) part[(int64_t)blockIdx.x * cols + c0 + t] = s;
    }
  }
}
__global__ void __launch_bounds__(kBS) k_dot_finish(const double *__restrict__ part, int nblk, int cols,
                                                    double *__restrict__ proj) {
  __shared__ double sh[kBS / 64];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kBS) acc += part[(int64_t)b * cols + c];
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) proj[c] = t;
}
// w -= sum_i proj[i]*basis_i, then per-block partial ||w||^2; VEC as in
// k_multi_dot.
template <bool VEC>
__global__ void __launch_bounds__(kBS) k_axpy_basis_norm(double *__restrict__ w, const double *__restrict__ basis,
                                                         int64_t n, int cols, const double *__restrict__ proj,
                                                         double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  __shared__ double sp[kMaxCols];
  for (int c = threadIdx.x; c < cols; c += kBS) sp[c] = proj[c];
  __syncthreads();
  double acc = 0.0;
  const int64_t m = VEC ? n / 2 : n;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBS) {
    // Columns in tiles of kDotTile: the tile's loads issue together, then
    // the subtractions run in column order (the same order as one at a time).
    if constexpr (VEC) {
      double2 wi = reinterpret_cast<const double2 *>(w)[i];
      for (int c0 = 0; c0 < cols; c0 += kDotTile) {
        double2 b[kDotTile];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < cols) b[t] = reinterpret_cast<const double2 *>(basis + (int64_t)(c0 + t) * n)[i];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < cols) {
            wi.x -= sp[c0 + t] * b[t].x;
            wi.y -= sp[c0 + t] * b[t].y;
          }
      }
      reinterpret_cast<double2 *>(w)[i] = wi;
      acc += wi.x * wi.x + wi.y * wi.y;
    } else {
      double wi = w[i];
      for (int c0 = 0; c0 < cols; c0 += kDotTile) {
        double b[kDotTile];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < cols) b[t] = basis[(int64_t)(c0 + t) * n + i];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < cols) wi -= sp[c0 + t] * b[t];
      }
      w[i] = wi;
      acc += wi * wi;
    }
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// The first pass's update and the second pass's projections of classical
// Gram-Schmidt with reorthogonalisation in one sweep over the basis:
// w -= sum_c proj[c] basis_c, then the per-block partials of basis_c . w
// (the updated w) over contiguous per-block slices as in k_multi_dot. Each
// lane keeps its entry of every column in registers between the update and
// the products, so the basis is read once instead of twice.
// MAXC >= cols is a multiple of 4 and the body has no branch per column:
// columns cols..MAXC-1 reload column cols-1 (a cache hit) with coefficient
// 0, so every load issues before the first use and the waits count down.
template <int MAXC>
__global__ void __launch_bounds__(kBS) k_axpy_dot(double *__restrict__ w, const double *__restrict__ basis,
                                                  int64_t n, int cols, const double *__restrict__ proj,
                                                  double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  __shared__ double sp[MAXC];
  for (int c = threadIdx.x; c < MAXC; c += kBS) sp[c] = c < cols ? proj[c] : 0.0;
  __syncthreads();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < n ? i0 + chunk : n;
  double acc[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) acc[c] = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBS) {
    double b[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) b[c] = basis[(int64_t)(c < cols ? c : cols - 1) * n + i];
    double wi = w[i];
    // z is 0 (i < n < 2^62) but depends on i, so the coefficients are read
    // from LDS (a broadcast) at each use instead of being held in two VGPRs
    // per column across the loop.
    const int z = (int)(i >> 62);
#pragma unroll
    for (int c = 0; c < MAXC; ++c) wi -= sp[c + z] * b[c];
    w[i] = wi;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] += b[c] * wi;
  }
  // A guard, not a break: with a constant trip count the loop unrolls fully
  // and acc stays in registers (an early exit left it in scratch memory
  // above 16 columns).
#pragma unroll
  for (int c = 0; c < MAXC; ++c) {
    if (c < cols) {
      const double t = block_sum(acc[c], sh);
      if (threadIdx.x == 0) part[(int64_t)blockIdx.x * cols + c] = t;
    }
  }
}
// ---- CGS2 with delayed reorthogonalisation (the adaptive Arnoldi) --------
//
// Per-block partials of the Gram products of basis columns 0..cols-1 with
// the last two, b_{cols-2} and b_{cols-1} (the provisional vector q and the
// new product z = A q): part[b*2cols + 2c + r] = b_c . b_{cols-2+r} over the
// block's contiguous slice, kDotTileM columns per pass. One sweep gives the
// reorthogonalisation coefficients of q, the projections of z, |q|^2 and
// q . z.
template <bool VEC>
__global__ void __launch_bounds__(kBS) k_gram2(const double *__restrict__ basis, int64_t n, int cols,
                                               double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  const int64_t m = VEC ? n / 2 : n;
  const int64_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < m ? i0 + chunk : m;
  const double *q = basis + (int64_t)(cols - 2) * n, *z = basis + (int64_t)(cols - 1) * n;
  for (int c0 = 0; c0 < cols; c0 += kDotTileM) {
    double aq[kDotTileM], az[kDotTileM];
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) aq[t] = az[t] = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kBS) {
      if constexpr (VEC) {
        const double2 qi = reinterpret_cast<const double2 *>(q)[i], zi = reinterpret_cast<const double2 *>(z)[i];
#pragma unroll
        for (int t = 0; t < kDotTileM; ++t)
          if (c0 + t < cols) {
            const double2 b = reinterpret_cast<const double2 *>(basis + (int64_t)(c0 + t) * n)[i];
            aq[t] += b.x * qi.x + b.y * qi.y;
            az[t] += b.x * zi.x + b.y * zi.y;
          }
      } else {
        const double qi = q[i], zi = z[i];
#pragma unroll
        for (int t = 0; t < kDotTileM; ++t)
          if (c0 + t < cols) {
            const double b = basis[(int64_t)(c0 + t) * n + i];
            aq[t] += b * qi;
            az[t] += b * zi;
          }
      }
    }
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) {
      if (c0 + t >= cols) break;
      const double sq = block_sum(aq[t], sh);
      const double sz = block_sum(az[t], sh);
      if (threadIdx.x == 0) {
        part[(int64_t)blockIdx.x * 2 * cols + 2 * (c0 + t)] = sq;
        part[(int64_t)blockIdx.x * 2 * cols + 2 * (c0 + t) + 1] = sz;
      }
    }
  }
}
// The update sweep: with q = b_j and z = b_{j+1},
//   b_j     <- (q - sum_{c<j} s_c b_c) / r          (the finished v_j)
//   b_{j+1} <- (z - sum_{c<j} a_c b_c - bq q) / r    (the next provisional vector)
// coef = [s_0..s_{j-1}, a_0..a_{j-1}, bq, 1/r]; the block's partial
// |b_{j+1}|^2 goes to part[blockIdx.x].
template <bool VEC>
__global__ void __launch_bounds__(kBS) k_dcgs_update(double *__restrict__ basis, int64_t n, int j,
                                                     const double *__restrict__ coef, double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  __shared__ double sc[2 * kMaxCols + 2];
  for (int c = threadIdx.x; c < 2 * j + 2; c += kBS) sc[c] = coef[c];
  __syncthreads();
  const double bq = sc[2 * j], ir = sc[2 * j + 1];
  double *q = basis + (int64_t)j * n, *z = basis + (int64_t)(j + 1) * n;
  double acc = 0.0;
  const int64_t m = VEC ? n / 2 : n;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBS) {
    if constexpr (VEC) {
      const double2 qi = reinterpret_cast<const double2 *>(q)[i];
      double2 zi = reinterpret_cast<const double2 *>(z)[i];
      zi.x -= bq * qi.x;
      zi.y -= bq * qi.y;
      double2 qn = qi;
      for (int c0 = 0; c0 < j; c0 += kDotTile) {
        double2 b[kDotTile];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < j) b[t] = reinterpret_cast<const double2 *>(basis + (int64_t)(c0 + t) * n)[i];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < j) {
            qn.x -= sc[c0 + t] * b[t].x;
            qn.y -= sc[c0 + t] * b[t].y;
            zi.x -= sc[j + c0 + t] * b[t].x;
            zi.y -= sc[j + c0 + t] * b[t].y;
          }
      }
      qn.x *= ir;
      qn.y *= ir;
      zi.x *= ir;
      zi.y *= ir;
      reinterpret_cast<double2 *>(q)[i] = qn;
      reinterpret_cast<double2 *>(z)[i] = zi;
      acc += zi.x * zi.x + zi.y * zi.y;
    } else {
      const double qi = q[i];
      double zi = z[i] - bq * qi, qn = qi;
      for (int c0 = 0; c0 < j; c0 += kDotTile) {
        double b[kDotTile];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < j) b[t] = basis[(int64_t)(c0 + t) * n + i];
#pragma unroll
        for (int t = 0; t < kDotTile; ++t)
          if (c0 + t < j) {
            qn -= sc[c0 + t] * b[t];
            zi -= sc[j + c0 + t] * b[t];
          }
      }
      const double zn = zi * ir;
      q[i] = qn * ir;
      z[i] = zn;
      acc += zn * zn;
    }
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// out = sum_c coef_c basis_c where out may be one of the columns (each
// entry is read by the thread that writes it, so in place is safe; no
// restrict on these pointers).
__global__ void __launch_bounds__(kBS) k_lincomb_alias(const double *basis, int64_t n, int dim,
                                                       const double *__restrict__ coef, double *out) {
  __shared__ double sc[kMaxCols];
  for (int c = threadIdx.x; c < dim; c += kBS) sc[c] = coef[c];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    double acc = 0.0;
    for (int c0 = 0; c0 < dim; c0 += kDotTile) {
      double b[kDotTile];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) b[t] = basis[(int64_t)(c0 + t) * n + i];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) acc += sc[c0 + t] * b[t];
    }
    out[i] = acc;
  }
}

// ---- the adaptive Arnoldi on a compressed basis ---------------------------
//
// Finished columns v_c held as f32 hi[c n + i] plus the bf16 of the
// remainder lo[c nl + i] (nl = n rounded up to even, so each column's lo
// part starts on a 4-byte word): 6 bytes per entry. The provisional vector
// q and the product z = A q stay FP64 in buffers of their own. Each lane
// takes V consecutive entries: V = 4 when n is a multiple of 4 (a 16-byte hi
// and an 8-byte lo load per column, two 16-byte loads of q and z), else 2
// (n even) or 1.
template <int V>
__device__ __forceinline__ void cb_load(const float *hi, const uint16_t *lo, int64_t n, int64_t nl, int c, int64_t i,
                                        double (&b)[V]) {
  if constexpr (V == 4) {
    const float4 h = reinterpret_cast<const float4 *>(hi + (int64_t)c * n)[i];
    const uint2 l = reinterpret_cast<const uint2 *>(lo + (int64_t)c * nl)[i];
    b[0] = (double)h.x + (double)__uint_as_float(l.x << 16);
    b[1] = (double)h.y + (double)__uint_as_float(l.x & 0xFFFF0000u);
    b[2] = (double)h.z + (double)__uint_as_float(l.y << 16);
    b[3] = (double)h.w + (double)__uint_as_float(l.y & 0xFFFF0000u);
  } else if constexpr (V == 2) {
    const float2 h = reinterpret_cast<const float2 *>(hi + (int64_t)c * n)[i];
    const uint32_t l = reinterpret_cast<const uint32_t *>(lo + (int64_t)c * nl)[i];
    b[0] = (double)h.x + (double)__uint_as_float(l << 16);
    b[1] = (double)h.y + (double)__uint_as_float(l & 0xFFFF0000u);
  } else {
    b[0] = (double)hi[(int64_t)c * n + i] + (double)bf16_to_f(lo[(int64_t)c * nl + i]);
  }
}
template <int V>
__device__ __forceinline__ void cb_store(float *hi, uint16_t *lo, int64_t n, int64_t nl, int c, int64_t i,
                                         const double (&v)[V]) {
  float h[V];
  uint32_t r[V];
#pragma unroll
  for (int e = 0; e < V; ++e) {
    h[e] = (float)v[e];
    r[e] = bf16_rne((float)(v[e] - (double)h[e]));
  }
  if constexpr (V == 4) {
    reinterpret_cast<float4 *>(hi + (int64_t)c * n)[i] = make_float4(h[0], h[1], h[2], h[3]);
    reinterpret_cast<uint2 *>(lo + (int64_t)c * nl)[i] = make_uint2(r[0] | (r[1] << 16), r[2] | (r[3] << 16));
  } else if constexpr (V == 2) {
    reinterpret_cast<float2 *>(hi + (int64_t)c * n)[i] = make_float2(h[0], h[1]);
    reinterpret_cast<uint32_t *>(lo + (int64_t)c * nl)[i] = r[0] | (r[1] << 16);
  } else {
    hi[(int64_t)c * n + i] = h[0];
    lo[(int64_t)c * nl + i] = (uint16_t)r[0];
  }
}
template <int V>
__device__ __forceinline__ void f64_load(const double *p, int64_t i, double (&v)[V]) {
  if constexpr (V == 1) {
    v[0] = p[i];
  } else {
#pragma unroll
    for (int e = 0; e < V; e += 2) {
      const double2 w = reinterpret_cast<const double2 *>(p)[i * (V / 2) + e / 2];
      v[e] = w.x;
      v[e + 1] = w.y;
    }
  }
}
template <int V>
__device__ __forceinline__ void f64_store(double *p, int64_t i, const double (&v)[V]) {
  if constexpr (V == 1) {
    p[i] = v[0];
  } else {
#pragma unroll
    for (int e = 0; e < V; e += 2) reinterpret_cast<double2 *>(p)[i * (V / 2) + e / 2] = make_double2(v[e], v[e + 1]);
  }
}

// k_gram2 on the compressed basis: "columns" 0..j-1 are v_0..v_{j-1} and
// column j is q; part[b*2(j+1) + 2c + r] = b_c . (q, z)_r over the block's
// slice, so c = j gives |q|^2 and q . z.
template <int V>
__global__ void __launch_bounds__(kBS) k_gram2c(const float *__restrict__ hi, const uint16_t *__restrict__ lo,
                                                int64_t n, int64_t nl, int j, const double *q, const double *z,
                                                double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  const int cols = j + 1;
  const int64_t m = n / V;
  const int64_t chunk = (m + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = i0 + chunk < m ? i0 + chunk : m;
  for (int c0 = 0; c0 < cols; c0 += kDotTileM) {
    double aq[kDotTileM], az[kDotTileM];
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) aq[t] = az[t] = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kBS) {
      double qi[V], zi[V];
      f64_load<V>(q, i, qi);
      f64_load<V>(z, i, zi);
#pragma unroll
      for (int t = 0; t < kDotTileM; ++t)
        if (c0 + t < cols) {
          double b[V];
          if (c0 + t < j) {
            cb_load<V>(hi, lo, n, nl, c0 + t, i, b);
          } else {
#pragma unroll
            for (int e = 0; e < V; ++e) b[e] = qi[e];
          }
#pragma unroll
          for (int e = 0; e < V; ++e) {
            aq[t] += b[e] * qi[e];
            az[t] += b[e] * zi[e];
          }
        }
    }
#pragma unroll
    for (int t = 0; t < kDotTileM; ++t) {
      if (c0 + t >= cols) break;
      const double sq = block_sum(aq[t], sh);
      const double sz = block_sum(az[t], sh);
      if (threadIdx.x == 0) {
        part[(int64_t)blockIdx.x * 2 * cols + 2 * (c0 + t)] = sq;
        part[(int64_t)blockIdx.x * 2 * cols + 2 * (c0 + t) + 1] = sz;
      }
    }
  }
}
// k_dcgs_update on the compressed basis: with coef = [s (j), a (j), bq, 1/r],
//   v_j <- (q - sum_{c<j} s_c v_c) / r   (stored compressed as column j)
//   q   <- (z - sum_{c<j} a_c v_c - bq q) / r   (the next provisional vector)
// and the block's partial |q|^2 in part[blockIdx.x]. Each entry of q is read
// and rewritten by the same thread.
template <int V>
__global__ void __launch_bounds__(kBS) k_dcgs_updatec(float *__restrict__ hi, uint16_t *__restrict__ lo, int64_t n,
                                                      int64_t nl, int j, const double *__restrict__ coef,
                                                      double *__restrict__ q, const double *__restrict__ z,
                                                      double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  __shared__ double sc[2 * kMaxCols + 2];
  for (int c = threadIdx.x; c < 2 * j + 2; c += kBS) sc[c] = coef[c];
  __syncthreads();
  const double bq = sc[2 * j], ir = sc[2 * j + 1];
  double acc = 0.0;
  const int64_t m = n / V;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBS) {
    double qn[V], zi[V];
    f64_load<V>(q, i, qn);
    f64_load<V>(z, i, zi);
#pragma unroll
    for (int e = 0; e < V; ++e) zi[e] -= bq * qn[e];
    for (int c0 = 0; c0 < j; c0 += kDotTile) {
      double b[kDotTile][V];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < j) cb_load<V>(hi, lo, n, nl, c0 + t, i, b[t]);
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < j) {
#pragma unroll
          for (int e = 0; e < V; ++e) {
            qn[e] -= sc[c0 + t] * b[t][e];
            zi[e] -= sc[j + c0 + t] * b[t][e];
          }
        }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
      qn[e] *= ir;
      zi[e] *= ir;
      acc += zi[e] * zi[e];
    }
    cb_store<V>(hi, lo, n, nl, j, i, qn);
    f64_store<V>(q, i, zi);
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
// out = sum_{c<dim} coef_c v_c over the compressed columns (FP64 out).
template <int V>
__global__ void __launch_bounds__(kBS) k_lincombc(const float *__restrict__ hi, const uint16_t *__restrict__ lo,
                                                  int64_t n, int64_t nl, int dim, const double *__restrict__ coef,
                                                  double *__restrict__ out) {
  __shared__ double sc[kMaxCols];
  for (int c = threadIdx.x; c < dim; c += kBS) sc[c] = coef[c];
  __syncthreads();
  const int64_t m = n / V;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBS) {
    double acc[V];
#pragma unroll
    for (int e = 0; e < V; ++e) acc[e] = 0.0;
    for (int c0 = 0; c0 < dim; c0 += kDotTile) {
      double b[kDotTile][V];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) cb_load<V>(hi, lo, n, nl, c0 + t, i, b[t]);
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) {
#pragma unroll
          for (int e = 0; e < V; ++e) acc[e] += sc[c0 + t] * b[t][e];
        }
    }
    f64_store<V>(out, i, acc);
  }
}

__global__ void __launch_bounds__(kBS) k_scale(const double *__restrict__ w, double *__restrict__ d, int64_t n,
                                               double inv) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) d[i] = w[i] * inv;
}
__global__ void __launch_bounds__(kBS) k_lincomb(const double *__restrict__ basis, int64_t n, int dim,
                                                 const double *__restrict__ coef, double *__restrict__ out) {
  __shared__ double sc[kMaxCols];
  for (int c = threadIdx.x; c < dim; c += kBS) sc[c] = coef[c];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    double acc = 0.0;
    for (int c0 = 0; c0 < dim; c0 += kDotTile) {
      double b[kDotTile];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) b[t] = basis[(int64_t)(c0 + t) * n + i];
#pragma unroll
      for (int t = 0; t < kDotTile; ++t)
        if (c0 + t < dim) acc += sc[c0 + t] * b[t];
    }
    out[i] = acc;
  }
}
__global__ void __launch_bounds__(kBS) k_axpy(double *__restrict__ w, const double *__restrict__ x, double a,
                                              int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) w[i] += a * x[i];
}

// 0.5 * sum_e |det J_e| sum_c w_c f_c^T M f_c, per-block partials
// (rhs.rs:1509-1544). One thread per element.
__global__ void __launch_bounds__(kBS) k_energy(const double *__restrict__ y, const double *__restrict__ mass,
                                                const double *__restrict__ det, const double *__restrict__ ie,
                                                const double *__restrict__ im, int np, int64_t K,
                                                double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * kBS + threadIdx.x; e < K; e += (int64_t)gridDim.x * kBS) {
    const double *ye = y + e * np * 6;
    double w[6];
    for (int c = 0; c < 3; ++c) {
      w[c] = 1.0 / ie[3 * e + c];
      w[3 + c] = 1.0 / im[3 * e + c];
    }
    double s = 0.0;
    for (int a = 0; a < np; ++a)
      for (int b = 0; b < np; ++b) {
        const double m = mass[a * np + b];
        for (int c = 0; c < 6; ++c) s += w[c] * m * ye[a * 6 + c] * ye[b * 6 + c];
      }
    acc += fabs(det[e]) * s;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

}  // namespace

void launch_f64_to_c(const double *src, float *hi, uint16_t *lo, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f64_to_c, dim3(grid_for(n)), dim3(kBS), 0, s, src, hi, lo, n);
  RF_HIP(hipGetLastError());
}
void launch_c_to_f64(const float *hi, const uint16_t *lo, double *dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_c_to_f64, dim3(grid_for(n)), dim3(kBS), 0, s, hi, lo, dst, n);
  RF_HIP(hipGetLastError());
}
void launch_cast_d2f(const double *src, float *dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(kBS), 0, s, src, dst, n);
  RF_HIP(hipGetLastError());
}
void launch_cast_f2d(const float *src, double *dst, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(kBS), 0, s, src, dst, n);
  RF_HIP(hipGetLastError());
}
void launch_gather(const void *y, const uint16_t *lo, int kind, const int64_t *idx, int64_t n, double *out,
                   hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 1024)), dim3(kBS), 0, s, y, lo, kind, idx, n, out);
  RF_HIP(hipGetLastError());
}

void launch_port_proj(const void *y, const uint16_t *lo, int kind, const int64_t *dof, const double *prof,
                      const double *w, const int64_t *off, const double *norm, int n_ports, double *out,
                      hipStream_t s) {
  if (n_ports <= 0) return;
  hipLaunchKernelGGL(k_port_proj, dim3((unsigned)n_ports), dim3(kBS), 0, s, y, lo, kind, dof, prof, w, off, norm,
                     out);
  RF_HIP(hipGetLastError());
}

// Fills the LDS of every resident block with a NaN pattern (volatile stores
// the compiler keeps). Launched with four 40 KiB blocks per CU it covers the
// whole 160 KiB of each CU, so a later kernel that reads LDS it did not
// write sees NaN instead of what an earlier kernel happened to leave there.
__global__ void __launch_bounds__(256) k_poison_lds() {
  __shared__ double pad[5120];
  volatile double *v = pad;
  const double nan = __longlong_as_double(0x7ff8dead7ff8deadll);
  for (int q = threadIdx.x; q < 5120; q += 256) v[q] = nan;
  __syncthreads();
}

void launch_poison_lds(hipStream_t s) {
  int dev = 0, cus = 0;
  RF_HIP(hipGetDevice(&dev));
  RF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  hipLaunchKernelGGL(k_poison_lds, dim3((unsigned)(8 * (cus > 0 ? cus : 1))), dim3(256), 0, s);
  RF_HIP(hipGetLastError());
}

void launch_fill(void *dst, int64_t nbytes, hipStream_t s) {
  const int64_t nw = nbytes / 4;
  hipLaunchKernelGGL(k_fill0, dim3(grid_for(nw)), dim3(kBS), 0, s, static_cast<uint32_t *>(dst), nw);
  RF_HIP(hipGetLastError());
}
void launch_copy(void *dst, const void *src, int64_t nbytes, hipStream_t s) {
  const int64_t nw = nbytes / 4;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(nw)), dim3(kBS), 0, s, static_cast<uint32_t *>(dst),
                     static_cast<const uint32_t *>(src), nw);
  RF_HIP(hipGetLastError());
}

// k_kcl_update on pairs of doubles (16-byte accesses; n even).
template <bool FIRST, bool NEXT>
__global__ void __launch_bounds__(kBS) k_kcl_update2(double2 *__restrict__ y, double2 *__restrict__ err,
                                                     double2 *__restrict__ st, const double2 *__restrict__ k,
                                                     double2 *__restrict__ bak, double dt, double b, double ew,
                                                     double amb, double atol, double rtol, int64_t m,
                                                     double *__restrict__ part) {
  __shared__ double sh[kBS / 64];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < m; i += (int64_t)gridDim.x * kBS) {
    const double2 kk = k[i], yo = y[i];
    const double f0 = dt * kk.x, f1 = dt * kk.y;
    const double2 yn = make_double2(yo.x + b * f0, yo.y + b * f1);
    y[i] = yn;
    if (FIRST) bak[i] = yo;
    double2 en;
    if (FIRST) {
      en = make_double2(ew * f0, ew * f1);
    } else {
      const double2 e = err[i];
      en = make_double2(e.x + ew * f0, e.y + ew * f1);
    }
    err[i] = en;
    if (NEXT) {
      st[i] = make_double2(yn.x + amb * f0, yn.y + amb * f1);
    } else {
      const double2 bo = bak[i];
      acc += werr_term(bo.x, yn.x, en.x, atol, rtol);
      acc += werr_term(bo.y, yn.y, en.y, atol, rtol);
    }
  }
  if (!NEXT) {
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

template <typename T>
void kcl_update_t(const KclStage &a, hipStream_t s) {
  // the last stage writes one error partial per block: its grid is nblk
  auto grid = [&](int64_t items) { return a.next ? dim3(grid_for(items)) : dim3((unsigned)a.nblk); };
  const dim3 bs(kBS);
  if (std::is_same<T, double>::value && a.n % 2 == 0) {
    double2 *yy = static_cast<double2 *>(a.y), *ee = static_cast<double2 *>(a.err);
    double2 *ss = static_cast<double2 *>(a.st), *bb = static_cast<double2 *>(a.bak);
    const double2 *kk = static_cast<const double2 *>(a.k);
    const int64_t m = a.n / 2;
    const dim3 g = grid(m);
    if (a.first && a.next)
      hipLaunchKernelGGL((k_kcl_update2<true, true>), g, bs, 0, s, yy, ee, ss, kk, bb, a.dt, a.b, a.ew, a.amb,
                         a.atol, a.rtol, m, a.partials);
    else if (a.next)
      hipLaunchKernelGGL((k_kcl_update2<false, true>), g, bs, 0, s, yy, ee, ss, kk, bb, a.dt, a.b, a.ew, a.amb,
                         a.atol, a.rtol, m, a.partials);
    else if (a.first)
      hipLaunchKernelGGL((k_kcl_update2<true, false>), g, bs, 0, s, yy, ee, ss, kk, bb, a.dt, a.b, a.ew, a.amb,
                         a.atol, a.rtol, m, a.partials);
    else
      hipLaunchKernelGGL((k_kcl_update2<false, false>), g, bs, 0, s, yy, ee, ss, kk, bb, a.dt, a.b, a.ew, a.amb,
                         a.atol, a.rtol, m, a.partials);
  } else {
    T *yy = static_cast<T *>(a.y), *ee = static_cast<T *>(a.err), *ss = static_cast<T *>(a.st);
    T *bb = static_cast<T *>(a.bak);
    const T *kk = static_cast<const T *>(a.k);
    const dim3 g = grid(a.n);
    const T dt = (T)a.dt, b = (T)a.b, ew = (T)a.ew, amb = (T)a.amb;
    if (a.first && a.next)
      hipLaunchKernelGGL((k_kcl_update<T, true, true>), g, bs, 0, s, yy, ee, ss, kk, bb, dt, b, ew, amb, a.atol,
                         a.rtol, a.n, a.partials);
    else if (a.next)
      hipLaunchKernelGGL((k_kcl_update<T, false, true>), g, bs, 0, s, yy, ee, ss, kk, bb, dt, b, ew, amb, a.atol,
                         a.rtol, a.n, a.partials);
    else if (a.first)
      hipLaunchKernelGGL((k_kcl_update<T, true, false>), g, bs, 0, s, yy, ee, ss, kk, bb, dt, b, ew, amb, a.atol,
                         a.rtol, a.n, a.partials);
    else
      hipLaunchKernelGGL((k_kcl_update<T, false, false>), g, bs, 0, s, yy, ee, ss, kk, bb, dt, b, ew, amb, a.atol,
                         a.rtol, a.n, a.partials);
  }
  if (!a.next) hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, a.partials, a.nblk, a.err_sum);
  RF_HIP(hipGetLastError());
}

void launch_kcl_update_c(const KclStage &a, hipStream_t s) {
  float *yh = static_cast<float *>(a.y), *err = static_cast<float *>(a.err), *bh = static_cast<float *>(a.bak);
  uint16_t *yl = a.y_lo, *bl = a.bak_lo;
  double *st = static_cast<double *>(a.st);
  const double *k = static_cast<const double *>(a.k);
  const dim3 g = a.next ? dim3(grid_for(a.n)) : dim3((unsigned)a.nblk), bs(kBS);
  if (a.first && a.next)
    hipLaunchKernelGGL((k_kcl_update_c<true, false>), g, bs, 0, s, yh, yl, err, st, k, bh, bl, a.dt, a.b, a.ew,
                       a.amb, a.atol, a.rtol, a.n, a.partials);
  else if (a.next)
    hipLaunchKernelGGL((k_kcl_update_c<false, false>), g, bs, 0, s, yh, yl, err, st, k, bh, bl, a.dt, a.b, a.ew,
                       a.amb, a.atol, a.rtol, a.n, a.partials);
  else
    hipLaunchKernelGGL((k_kcl_update_c<false, true>), g, bs, 0, s, yh, yl, err, st, k, bh, bl, a.dt, a.b, a.ew,
                       a.amb, a.atol, a.rtol, a.n, a.partials);
  if (!a.next) hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, a.partials, a.nblk, a.err_sum);
  RF_HIP(hipGetLastError());
}

void launch_kcl_update(bool f64, const KclStage &a, hipStream_t s) {
  RF_REQUIRE(a.next || (a.nblk >= 1 && a.partials && a.err_sum), "KCL last stage needs the error partials");
  if (f64)
    kcl_update_t<double>(a, s);
  else
    kcl_update_t<float>(a, s);
}

int weighted_err_blocks(int64_t n) { return (int)grid_for(n, 2048); }

void launch_sum_partials(const double *part, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, part, n, out);
  RF_HIP(hipGetLastError());
}


// 1024 blocks: k_multi_dot streams 5.7 TB/s over 1024 blocks and 5.3 over
// 2048 on a 24.2M-entry basis; the other sweeps are flat between the two.
int krylov_dot_blocks(int64_t n) { return (int)grid_for(n, 1024); }

void launch_multi_dot(const double *basis, int64_t n, int cols, const double *w, double *partials, int nblk,
                      double *proj, hipStream_t s) {
  RF_REQUIRE(cols <= kMaxCols, "Krylov dimension above 64");
  if (n % 2 == 0)
    hipLaunchKernelGGL(k_multi_dot<true>, dim3(nblk), dim3(kBS), 0, s, basis, n, cols, w, partials);
  else
    hipLaunchKernelGGL(k_multi_dot<false>, dim3(nblk), dim3(kBS), 0, s, basis, n, cols, w, partials);
  hipLaunchKernelGGL(k_dot_finish, dim3(cols), dim3(kBS), 0, s, partials, nblk, cols, proj);
  RF_HIP(hipGetLastError());
}
void launch_axpy_basis_norm(double *w, const double *basis, int64_t n, int cols, const double *proj,
                            double *partials, int nblk, double *norm2, hipStream_t s) {
  RF_REQUIRE(cols <= kMaxCols, "Krylov dimension above 64");
  if (n % 2 == 0)
    hipLaunchKernelGGL(k_axpy_basis_norm<true>, dim3(nblk), dim3(kBS), 0, s, w, basis, n, cols, proj, partials);
  else
    hipLaunchKernelGGL(k_axpy_basis_norm<false>, dim3(nblk), dim3(kBS), 0, s, w, basis, n, cols, proj, partials);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, partials, nblk, norm2);
  RF_HIP(hipGetLastError());
}
void launch_axpy_dot(double *w, const double *basis, int64_t n, int cols, const double *proj, double *partials,
                     int nblk, double *proj_out, hipStream_t s) {
  RF_REQUIRE(cols >= 1 && cols <= kMaxCols, "Krylov dimension above 64");
  if (cols > kAxpyDotMaxCols) {
    launch_axpy_basis_norm(w, basis, n, cols, proj, partials, nblk, proj_out, s);
    launch_multi_dot(basis, n, cols, w, partials, nblk, proj_out, s);
    return;
  }
  const dim3 g(nblk), b(kBS);
  switch ((cols + 3) / 4) {
#define RF_AXPY_DOT(q)                                                                          \
  case q:                                                                                       \
    hipLaunchKernelGGL(k_axpy_dot<4 * q>, g, b, 0, s, w, basis, n, cols, proj, partials);       \
    break;
    RF_AXPY_DOT(1) RF_AXPY_DOT(2) RF_AXPY_DOT(3) RF_AXPY_DOT(4) RF_AXPY_DOT(5) RF_AXPY_DOT(6)
    RF_AXPY_DOT(7) RF_AXPY_DOT(8) RF_AXPY_DOT(9) RF_AXPY_DOT(10) RF_AXPY_DOT(11) RF_AXPY_DOT(12)
#undef RF_AXPY_DOT
  }
  hipLaunchKernelGGL(k_dot_finish, dim3(cols), dim3(kBS), 0, s, partials, nblk, cols, proj_out);
  RF_HIP(hipGetLastError());
}
void launch_gram2(const double *basis, int64_t n, int cols, double *partials, int nblk, double *out,
                  hipStream_t s) {
  RF_REQUIRE(cols >= 2 && cols <= kMaxCols, "Krylov dimension above 64");
  if (n % 2 == 0)
    hipLaunchKernelGGL(k_gram2<true>, dim3(nblk), dim3(kBS), 0, s, basis, n, cols, partials);
  else
    hipLaunchKernelGGL(k_gram2<false>, dim3(nblk), dim3(kBS), 0, s, basis, n, cols, partials);
  hipLaunchKernelGGL(k_dot_finish, dim3(2 * cols), dim3(kBS), 0, s, partials, nblk, 2 * cols, out);
  RF_HIP(hipGetLastError());
}
void launch_dcgs_update(double *basis, int64_t n, int j, const double *coef, double *partials, int nblk,
                        double *norm2, hipStream_t s) {
  RF_REQUIRE(j >= 0 && j + 1 < kMaxCols, "Krylov dimension above 64");
  if (n % 2 == 0)
    hipLaunchKernelGGL(k_dcgs_update<true>, dim3(nblk), dim3(kBS), 0, s, basis, n, j, coef, partials);
  else
    hipLaunchKernelGGL(k_dcgs_update<false>, dim3(nblk), dim3(kBS), 0, s, basis, n, j, coef, partials);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, partials, nblk, norm2);
  RF_HIP(hipGetLastError());
}
void launch_lincomb_alias(const double *basis, int64_t n, int dim, const double *coef, double *out, hipStream_t s) {
  RF_REQUIRE(dim <= kMaxCols, "Krylov dimension above 64");
  hipLaunchKernelGGL(k_lincomb_alias, dim3(grid_for(n)), dim3(kBS), 0, s, basis, n, dim, coef, out);
  RF_HIP(hipGetLastError());
}
// Entries per lane of the compressed-basis sweeps (see k_gram2c).
inline int cb_vec(int64_t n) { return n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : 1; }
void launch_gram2c(const float *hi, const uint16_t *lo, int64_t n, int64_t nl, int j, const double *q,
                   const double *z, double *partials, int nblk, double *out, hipStream_t s) {
  RF_REQUIRE(j >= 0 && j + 1 <= kMaxCols, "Krylov dimension above 64");
  const int cols = j + 1;
  switch (cb_vec(n)) {
    case 4: hipLaunchKernelGGL(k_gram2c<4>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, q, z, partials); break;
    case 2: hipLaunchKernelGGL(k_gram2c<2>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, q, z, partials); break;
    default: hipLaunchKernelGGL(k_gram2c<1>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, q, z, partials); break;
  }
  hipLaunchKernelGGL(k_dot_finish, dim3(2 * cols), dim3(kBS), 0, s, partials, nblk, 2 * cols, out);
  RF_HIP(hipGetLastError());
}
void launch_dcgs_updatec(float *hi, uint16_t *lo, int64_t n, int64_t nl, int j, const double *coef, double *q,
                         const double *z, double *partials, int nblk, double *norm2, hipStream_t s) {
  RF_REQUIRE(j >= 0 && j + 1 < kMaxCols, "Krylov dimension above 64");
  switch (cb_vec(n)) {
    case 4:
      hipLaunchKernelGGL(k_dcgs_updatec<4>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, coef, q, z, partials);
      break;
    case 2:
      hipLaunchKernelGGL(k_dcgs_updatec<2>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, coef, q, z, partials);
      break;
    default:
      hipLaunchKernelGGL(k_dcgs_updatec<1>, dim3(nblk), dim3(kBS), 0, s, hi, lo, n, nl, j, coef, q, z, partials);
      break;
  }
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, partials, nblk, norm2);
  RF_HIP(hipGetLastError());
}
void launch_lincombc(const float *hi, const uint16_t *lo, int64_t n, int64_t nl, int dim, const double *coef,
                     double *out, hipStream_t s) {
  RF_REQUIRE(dim <= kMaxCols, "Krylov dimension above 64");
  const int v = cb_vec(n);
  const dim3 g(grid_for(n / v));
  switch (v) {
    case 4: hipLaunchKernelGGL(k_lincombc<4>, g, dim3(kBS), 0, s, hi, lo, n, nl, dim, coef, out); break;
    case 2: hipLaunchKernelGGL(k_lincombc<2>, g, dim3(kBS), 0, s, hi, lo, n, nl, dim, coef, out); break;
    default: hipLaunchKernelGGL(k_lincombc<1>, g, dim3(kBS), 0, s, hi, lo, n, nl, dim, coef, out); break;
  }
  RF_HIP(hipGetLastError());
}
void launch_scale(const double *w, double *dst, int64_t n, double inv, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBS), 0, s, w, dst, n, inv);
  RF_HIP(hipGetLastError());
}
void launch_lincomb(const double *basis, int64_t n, int dim, const double *coef, double *out, hipStream_t s) {
  RF_REQUIRE(dim <= kMaxCols, "Krylov dimension above 64");
  hipLaunchKernelGGL(k_lincomb, dim3(grid_for(n)), dim3(kBS), 0, s, basis, n, dim, coef, out);
  RF_HIP(hipGetLastError());
}
void launch_axpy(double *w, const double *x, double a, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kBS), 0, s, w, x, a, n);
  RF_HIP(hipGetLastError());
}
void launch_energy(const double *y, const double *mass, const double *det, const double *ie, const double *im, int np,
                   int64_t K, double *partials, int nblk, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy, dim3(nblk), dim3(kBS), 0, s, y, mass, det, ie, im, np, K, partials);
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(kBS), 0, s, partials, nblk, out);
  RF_HIP(hipGetLastError());
}

int energy_elems_per_block() { return kBS; }

void launch_energy_scaled(const double *y, const double *mass, const double *det, const double *ie, const double *im,
                          int np, int64_t K, double *partials, int nblk, double scale, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_energy, dim3(nblk), dim3(kBS), 0, s, y, mass, det, ie, im, np, K, partials);
  hipLaunchKernelGGL(k_finish_sum_scaled, dim3(1), dim3(kBS), 0, s, partials, nblk, scale, out);
  RF_HIP(hipGetLastError());
}

}  // namespace rfhip
