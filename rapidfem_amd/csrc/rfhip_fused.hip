// Fused DG stage kernel, FP64 arithmetic.
//
// One launch per LSERK4 stage does what the reference GPU path spreads over
// two kernels (gpu/apply.cl:17-209 + gpu/lserk.cl:8-20): volume curl,
// upwind flux, LIFT, materials, held source and the low-storage update
// p <- a p + dt k, y <- y + b p. y is double-buffered (read y_in, write
// y_out) because neighbouring elements read y_in's face traces during the
// same stage.
//
// Per-element HBM traffic is the state itself plus 92 bytes of geometry:
//  * J^-1 (9 doubles). Face normals and fscale are not stored: with
//    g_f = -grad(lambda_f) = {-row2, -row0, -row1, row0+row1+row2} of J^-1
//    for the four local faces, n = g/|g| and fscale = 2 area/|det J| = |g|,
//    so 0.5*fscale*n = 0.5*g (rhs.rs:829, topology.rs:56-73).
//  * one int32 per face: neighbour, its local face and the orientation of
//    the shared triangle, which index a small table of neighbour nodes
//    (replaces face_perm + face_nodes double indirection, apply.cl:151-152).
//  * a material index into a table of distinct (1/eps, 1/mu, sigma/eps,
//    sigma*/mu) rows.
//
// Work mapping: a block holds EPB = 256/Np elements, one thread per
// (element, node). The block's state is staged in LDS with coalesced loads;
// D^T and LIFT^T are read from L1 with consecutive nodes on consecutive
// lanes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "rfhip_elem_tables.hpp"
#include "rfhip_impl.hpp"
#include "rfhip_regs.hpp"

namespace rfhip {
namespace {

constexpr int kFusedWG = 256;

// Two floats to bf16 (round to nearest even, as bf16_rne for finite values),
// a in the low half: one v_cvt_pk_bf16_f32 on gfx950.
__device__ __forceinline__ uint32_t bf16_pair(float a, float b) {
  const __bf16 x = (__bf16)a, y = (__bf16)b;
  return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

template <bool CIN>
__device__ __forceinline__ double load_state(const StageArgs &io, int64_t i) {
  if constexpr (CIN) {
    return (double)static_cast<const float *>(io.yin)[i] + (double)bf16_to_f(static_cast<const uint16_t *>(io.yin_lo)[i]);
  } else {
    return static_cast<const double *>(io.yin)[i];
  }
}

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// The six field values of one node starting at element index nb (a multiple
// of 6): 8-byte f32 pairs + 4-byte bf16 pairs, or 16-byte f64 pairs.
template <bool CIN>
__device__ __forceinline__ void load_trace(const StageArgs &io, int64_t nb, double *v) {
  if constexpr (CIN) {
    const float2 *hp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + nb);
    const uint32_t *lp = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + nb);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float2 h = hp[q];
      const uint32_t l = lp[q];
      v[2 * q] = (double)h.x + (double)__uint_as_float(l << 16);
      v[2 * q + 1] = (double)h.y + (double)__uint_as_float(l & 0xFFFF0000u);
    }
  } else {
    const double2 *dp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + nb);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double2 w = dp[q];
      v[2 * q] = w.x;
      v[2 * q + 1] = w.y;
    }
  }
}

// EPI: 0 = apply (dy = A y, FP64), 1 = LSERK stage with FP64 registers,
// 2 = LSERK stage with compressed registers (y f32+bf16, p f32), 3 = KCL
// stage with compressed registers (k_fused11 only, node_kcl), 4 = KCL stage
// with FP64 registers (k_fused11 only, node_kcl64).
//
// KCL stage update of one node with compressed registers
// (explicit_adaptive.rs:107-149), the same arithmetic as the separate pass
// k_kcl_update_c: f = h k, y += b f, err = e f on the first stage and
// err += e f after it, the next stage's input y + (a - b) f (y itself after
// the last stage) into yout. The first stage saves the starting y (already
// in registers) to the backup; the last reads the backup and adds the
// node's error-norm terms (session.rs:160-171) to wsum. own: the stage
// input, yh / yl / ev: the node's y (f32 hi, bf16 lo pairs) and err. CIN:
// the next stage's input is written as f32 hi + bf16 lo (yout, yout_lo),
// else in FP64.
template <bool CIN>
__device__ __forceinline__ void node_kcl(const StageArgs &io, int64_t base, const double *mt, const double *acc,
                                         const double *own, const double *srcv, const float *yh,
                                         const uint32_t *yl, const float *ev, double &wsum) {
  // the node's right-hand side: materials and the held source
  // (rhs.rs:1443-1454), written out here as in node_update (a helper shared
  // by the two, with an output array, raised the N=3 stage kernels' spilled
  // VGPRs from 10 to 44-62)
  double kv[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kv[c] = mt[c] * acc[c] - mt[6 + c] * own[c];
    kv[3 + c] = mt[3 + c] * acc[3 + c] - mt[9 + c] * own[3 + c];
  }
  if (io.src_kind == RFHIP_DRIVE_POINT) {
    const int64_t off = io.src_dof - base;
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (off == c) kv[c] += io.g;
  } else if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
    for (int c = 0; c < 6; ++c) kv[c] += srcv[c] * io.g;
  }
  float hn[6], rn[6], en[6];
  double sn[6];
  float bh[6];
  uint32_t bl[3];
  if (io.klast) {
    const float2 *bp = reinterpret_cast<const float2 *>(io.kbak + base);
    const uint32_t *lp = reinterpret_cast<const uint32_t *>(io.kbak_lo + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float2 v = bp[q];
      bh[2 * q] = v.x;
      bh[2 * q + 1] = v.y;
      bl[q] = lp[q];
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const uint32_t lw = c & 1 ? yl[c >> 1] & 0xFFFF0000u : yl[c >> 1] << 16;
    const double f = io.dt * kv[c];
    const double yn = ((double)yh[c] + (double)__uint_as_float(lw)) + io.kb * f;
    hn[c] = (float)yn;
    rn[c] = (float)(yn - (double)hn[c]);
    const double e = io.kfirst ? io.kew * f : (double)ev[c] + io.kew * f;
    en[c] = (float)e;
    sn[c] = io.klast ? yn : yn + io.kamb * f;
    if (io.klast) {
      const uint32_t bw = c & 1 ? bl[c >> 1] & 0xFFFF0000u : bl[c >> 1] << 16;
      wsum += werr_term((double)bh[c] + (double)__uint_as_float(bw), yn, e, io.katol, io.krtol);
    }
  }
  float2 *yo = reinterpret_cast<float2 *>(io.ky + base);
  uint32_t *lo = reinterpret_cast<uint32_t *>(io.ky_lo + base);
  float2 *eo = reinterpret_cast<float2 *>(io.kerr + base);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    yo[q] = make_float2(hn[2 * q], hn[2 * q + 1]);
    lo[q] = bf16_pair(rn[2 * q], rn[2 * q + 1]);
    eo[q] = make_float2(en[2 * q], en[2 * q + 1]);
  }
  if constexpr (CIN) {
    float2 *sh = reinterpret_cast<float2 *>(static_cast<float *>(io.yout) + base);
    uint32_t *sl = reinterpret_cast<uint32_t *>(static_cast<uint16_t *>(io.yout_lo) + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const float h0 = (float)sn[2 * q], h1 = (float)sn[2 * q + 1];
      sh[q] = make_float2(h0, h1);
      sl[q] = bf16_pair((float)(sn[2 * q] - (double)h0), (float)(sn[2 * q + 1] - (double)h1));
    }
  } else {
    double2 *so = reinterpret_cast<double2 *>(static_cast<double *>(io.yout) + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) so[q] = make_double2(sn[2 * q], sn[2 * q + 1]);
  }
  if (io.kfirst) {
    float2 *bo = reinterpret_cast<float2 *>(io.kbak + base);
    uint32_t *blo = reinterpret_cast<uint32_t *>(io.kbak_lo + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      bo[q] = make_float2(yh[2 * q], yh[2 * q + 1]);
      blo[q] = yl[q];
    }
  }
}

// The same KCL stage update with FP64 registers, the arithmetic of the
// separate pass k_kcl_update2: yd / ed: the node's y and err; the next
// stage's input (FP64) into yout; the backup in kbakd.
__device__ __forceinline__ void node_kcl64(const StageArgs &io, int64_t base, const double *mt, const double *acc,
                                           const double *own, const double *srcv, const double *yd,
                                           const double *ed, double &wsum) {
  // the node's right-hand side (rhs.rs:1443-1454), written out as in
  // node_update
  double kv[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kv[c] = mt[c] * acc[c] - mt[6 + c] * own[c];
    kv[3 + c] = mt[3 + c] * acc[3 + c] - mt[9 + c] * own[3 + c];
  }
  if (io.src_kind == RFHIP_DRIVE_POINT) {
    const int64_t off = io.src_dof - base;
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (off == c) kv[c] += io.g;
  } else if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
    for (int c = 0; c < 6; ++c) kv[c] += srcv[c] * io.g;
  }
  double yn[6], en[6], sn[6], bo[6];
  if (io.klast) {
    const double2 *bp = reinterpret_cast<const double2 *>(io.kbakd + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double2 v = bp[q];
      bo[2 * q] = v.x;
      bo[2 * q + 1] = v.y;
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double f = io.dt * kv[c];
    yn[c] = yd[c] + io.kb * f;
    en[c] = io.kfirst ? io.kew * f : ed[c] + io.kew * f;
    sn[c] = io.klast ? yn[c] : yn[c] + io.kamb * f;
    if (io.klast) wsum += werr_term(bo[c], yn[c], en[c], io.katol, io.krtol);
  }
  double2 *yo = reinterpret_cast<double2 *>(io.kyd + base);
  double2 *eo = reinterpret_cast<double2 *>(io.kerrd + base);
  double2 *so = reinterpret_cast<double2 *>(static_cast<double *>(io.yout) + base);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    yo[q] = make_double2(yn[2 * q], yn[2 * q + 1]);
    eo[q] = make_double2(en[2 * q], en[2 * q + 1]);
    so[q] = make_double2(sn[2 * q], sn[2 * q + 1]);
  }
  if (io.kfirst) {
    double2 *bp = reinterpret_cast<double2 *>(io.kbakd + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) bp[q] = make_double2(yd[2 * q], yd[2 * q + 1]);
  }
}

// Stage update with the old p already in registers (prefetched by the
// kernels that use it): materials, held source, p and y.
template <int EPI>
__device__ __forceinline__ void node_update(const StageArgs &io, int64_t base, const double *mt, const double *acc,
                                            const double *own, const double *pold, const double *srcv) {
  double kv[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kv[c] = mt[c] * acc[c] - mt[6 + c] * own[c];
    kv[3 + c] = mt[3 + c] * acc[3 + c] - mt[9 + c] * own[3 + c];
  }
  if (io.src_kind == RFHIP_DRIVE_POINT) {
    const int64_t off = io.src_dof - base;
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (off == c) kv[c] += io.g;
  } else if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
    for (int c = 0; c < 6; ++c) kv[c] += srcv[c] * io.g;
  }
  if constexpr (EPI == 0) {
    double2 *dy = reinterpret_cast<double2 *>(io.dy + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) dy[q] = make_double2(kv[2 * q], kv[2 * q + 1]);
  } else if constexpr (EPI == 1) {
    double pn[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) pn[c] = io.sa * pold[c] + io.dt * kv[c];
    double2 *pp = reinterpret_cast<double2 *>(static_cast<double *>(io.p) + base);
    double2 *yo = reinterpret_cast<double2 *>(static_cast<double *>(io.yout) + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      pp[q] = make_double2(pn[2 * q], pn[2 * q + 1]);
      yo[q] = make_double2(own[2 * q] + io.sb * pn[2 * q], own[2 * q + 1] + io.sb * pn[2 * q + 1]);
    }
  } else {
    float pn[6];
    float yh[6];
    float yr[6];  // the part of y below f32, stored as bf16 (round to nearest even)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      pn[c] = (float)(io.sa * pold[c] + io.dt * kv[c]);
      const double yn = own[c] + io.sb * (double)pn[c];
      yh[c] = (float)yn;
      yr[c] = (float)(yn - (double)yh[c]);
    }
    float2 *pp = reinterpret_cast<float2 *>(static_cast<float *>(io.p) + base);
    float2 *yo = reinterpret_cast<float2 *>(static_cast<float *>(io.yout) + base);
    uint32_t *lo = reinterpret_cast<uint32_t *>(static_cast<uint16_t *>(io.yout_lo) + base);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      pp[q] = make_float2(pn[2 * q], pn[2 * q + 1]);
      yo[q] = make_float2(yh[2 * q], yh[2 * q + 1]);
      lo[q] = bf16_pair(yr[2 * q], yr[2 * q + 1]);
    }
  }
}

// Unit normal and 0.5*fscale of local face f from J^-1 (rows of ji):
// g_f = {-row2, -row0, -row1, row0+row1+row2}, n = g/|g|, fscale = |g|.
__device__ __forceinline__ void face_geom(const double *ji, int f, double *n, double &cf) {
  double g[3];
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    const double r0 = ji[x], r1 = ji[3 + x], r2 = ji[6 + x];
    g[x] = f == 0 ? -r2 : f == 1 ? -r0 : f == 2 ? -r1 : (r0 + r1 + r2);
  }
  // 1/|g| from the hardware reciprocal square root and two Newton steps
  // (y <- y (3/2 - s y^2 / 2), each squaring the relative error): a few ulp,
  // without the correctly rounded sqrt and divide sequences. |g| > 0 for a
  // non-degenerate element.
  const double s = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  const double hs = 0.5 * s;
  double y = __builtin_amdgcn_rsq(s);
  y = y * fma(-hs * y, y, 1.5);
  y = y * fma(-hs * y, y, 1.5);
  n[0] = g[0] * y;
  n[1] = g[1] * y;
  n[2] = g[2] * y;
  cf = hs * y;  // 0.5 |g|
}

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8).
// With a grid of whole rounds the blocks on one XCD are renumbered to
// consecutive indices, so that in each sweep of a grid-stride loop they take
// consecutive element groups: an element's face neighbours in the cells
// next to it along the two faster mesh directions are then mostly gathered
// on the same XCD at about the same time and come from its L2 rather than
// the fabric. Any renumbering visits every group once.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned nb = gridDim.x;
  return (nb & 7u) == 0 ? (int64_t)(blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3) : (int64_t)blockIdx.x;
}

// Order of element groups over the blocks. With xcd, blocks are taken to
// run round-robin over the 8 XCDs (block b on XCD b % 8): XCD x walks the
// contiguous groups [x*chunk, (x+1)*chunk), so the neighbours an element
// gathers from were mostly loaded by the same XCD and sit in its L2.
// Requires gridDim.x % 8 == 0 (else the plain order is used).
struct GroupRange {
  int64_t first, end, step;
};
__device__ __forceinline__ GroupRange group_range(bool xcd, int64_t ngroups) {
  if (xcd && (gridDim.x & 7u) == 0) {
    const int64_t chunk = (ngroups + 7) / 8;
    const int64_t beg = (int64_t)(blockIdx.x & 7u) * chunk;
    const int64_t end = beg + chunk < ngroups ? beg + chunk : ngroups;
    return {beg + (int64_t)(blockIdx.x >> 3), end, (int64_t)(gridDim.x >> 3)};
  }
  return {xcd_block(), ngroups, (int64_t)gridDim.x};
}

// One thread per (element, node), 256/Np elements per block:
//  * the face codes and neighbour-node indices of a thread's face nodes are
//    resolved before anything else (they are a chain of dependent loads),
//    and with PREV the neighbour traces themselves are requested too, so
//    their latency overlaps the state load and the volume contraction;
//  * the state is moved into LDS with 8/16-byte vector loads;
//  * the unit normal and 0.5*fscale of each face are computed once per
//    element (four threads) into LDS, not once per face node;
//  * DLDS: D_k^T and LIFT^T live in LDS and the block walks over element
//    groups (grid-stride), so the operator matrices are read from L2 once
//    per block instead of through L1 by every wave.
template <int NP, int NFP, bool CIN, int EPI, bool DLDS, bool PREV, bool XCD = false, bool SPF = false, int UNR = 4,
          bool LATE = false>
__global__ void __launch_bounds__(kFusedWG) __attribute__((amdgpu_waves_per_eu(3))) k_fused3(StageArgs io) {
  constexpr int EPB = kFusedWG / NP;
  constexpr int NF4 = 4 * NFP;
  constexpr int FPT = (NF4 + NP - 1) / NP;  // face nodes per thread
  constexpr int ND = 3 * NP * NP;
  constexpr int SD = DLDS ? ND + NF4 * NP : 1;
  __shared__ __align__(16) double su[EPB * NP * 6];
  // per-element flux stride padded by two doubles: the elements of a wave
  // then start on different LDS banks
  constexpr int SFS = NF4 * 6 + 2;
  __shared__ __align__(16) double sf[EPB * SFS];
  __shared__ __align__(16) double sg[EPB * 16];  // per face: n_x, n_y, n_z, 0.5*fscale
  __shared__ __align__(16) double sd[SD];
  __shared__ __align__(16) double sm[EPB * 12];  // material row per element
  const double *dT = io.dT;
  const double *lift = io.lift;
  if constexpr (DLDS) {
    for (int q = threadIdx.x; q < SD; q += kFusedWG) sd[q] = q < ND ? io.dT[q] : io.lift[q - ND];
    dT = sd;
    lift = sd + ND;
  }
  const int tid = threadIdx.x;
  const int le = tid / NP, i = tid - (tid / NP) * NP;
  const int64_t ngroups = (io.n_elem + EPB - 1) / EPB;
  const GroupRange gr = group_range(XCD, ngroups);
  // SPF: the state of a block's next group is read into registers while the
  // current group computes, and stored to LDS at the next group's start.
  constexpr int PPT = (EPB * NP * 3 + kFusedWG - 1) / kFusedWG;  // state pairs per thread
  using HT = typename std::conditional<CIN, float2, double2>::type;
  HT sh[SPF ? PPT : 1];
  uint32_t sl[SPF ? PPT : 1];
  auto prefetch_state = [&](int64_t g) {
    const int64_t rg = io.n_elem - g * EPB;
    const int cnt = (rg < EPB ? (int)rg : EPB) * NP * 3;
    const int64_t gb = g * EPB * NP * 6;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int q = tid + k * kFusedWG;
      if (q < cnt) {
        if constexpr (CIN) {
          sh[k] = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + gb)[q];
          sl[k] = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + gb)[q];
        } else {
          sh[k] = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + gb)[q];
        }
      }
    }
  };
  if constexpr (SPF) {
    if (gr.first < gr.end) prefetch_state(gr.first);
  }
  for (int64_t grp = gr.first; grp < gr.end; grp += gr.step) {
    const int64_t e0 = grp * EPB;
    const int64_t rem = io.n_elem - e0;
    const int nel = rem < EPB ? (int)rem : EPB;
    const bool active = tid < nel * NP;
    const int64_t e = e0 + le;
    if constexpr (SPF) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int q = tid + k * kFusedWG;
        if (q < nel * NP * 3) {
          double2 v;
          if constexpr (CIN) {
            v.x = (double)sh[k].x + (double)__uint_as_float(sl[k] << 16);
            v.y = (double)sh[k].y + (double)__uint_as_float(sl[k] & 0xFFFF0000u);
          } else {
            v = sh[k];
          }
          reinterpret_cast<double2 *>(su)[q] = v;
        }
      }
      // LATE: issued after this group's own reads instead (vector loads
      // retire in issue order, so a wait on J^-1 would otherwise wait on it)
      if (!LATE && grp + gr.step < gr.end) prefetch_state(grp + gr.step);
    }

    // 1. neighbour addresses (and with PREV the traces) of this thread's
    //    face nodes, in flight early.
    double nv[FPT][6];
    int64_t nbase[FPT];
    int fcodes[FPT];
    if (active) {
#pragma unroll
      for (int s = 0; s < FPT; ++s) {
        const int fm = i + s * NP;
        fcodes[s] = kCodePec;
        nbase[s] = 0;
        if (fm >= NF4) continue;
        if (io.nbase) {  // one read instead of the face code -> node table chain
          const int nbv = io.nbase[e * NF4 + fm];
          fcodes[s] = nbv >= 0 ? 0 : nbv;
          if (nbv < 0) continue;
          nbase[s] = nbv;
        } else {
          const int f = fm / NFP, m = fm - f * NFP;
          const int code = io.fcode[e * 4 + f];
          fcodes[s] = code;
          if (code < 0) continue;
          const int64_t nb = code >> 5;
          const int nlf = (code >> 3) & 3, o = code & 7;
          const int vj = io.nbr_node[((f * 4 + nlf) * 6 + o) * NFP + m];
          nbase[s] = (nb * NP + vj) * 6;
        }
        if constexpr (PREV) load_trace<CIN>(io, nbase[s], nv[s]);
      }
    }
    // old p and the material row, needed only by the epilogue
    using PT = typename std::conditional<EPI == 2, float, double>::type;
    PT pold[6];
    const int64_t base = (e * NP + i) * 6;
    if (active) {
      if constexpr (EPI == 1) {
        const double2 *pp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.p) + base);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double2 v = pp[q];
          pold[2 * q] = v.x;
          pold[2 * q + 1] = v.y;
        }
      } else if constexpr (EPI == 2) {
        const float2 *pp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.p) + base);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const float2 v = pp[q];
          pold[2 * q] = v.x;
          pold[2 * q + 1] = v.y;
        }
      }
      for (int q = i; q < 12; q += NP) sm[le * 12 + q] = io.mat_tab[12 * io.mat[e] + q];
    }
    // 2. the group's state into LDS, two values per lane and load.
    if constexpr (!SPF) {
      const int64_t g0 = e0 * NP * 6;
      const int cnt2 = nel * NP * 3;
      for (int q = tid; q < cnt2; q += kFusedWG) {
        double2 v;
        if constexpr (CIN) {
          const float2 h = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + g0)[q];
          const uint32_t l = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + g0)[q];
          v.x = (double)h.x + (double)__uint_as_float(l << 16);
          v.y = (double)h.y + (double)__uint_as_float(l & 0xFFFF0000u);
        } else {
          v = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + g0)[q];
        }
        reinterpret_cast<double2 *>(su)[q] = v;
      }
    }
    // 3. face geometry from J^-1: g_f = {-row2, -row0, -row1, row0+row1+row2},
    //    n = g/|g|, 0.5*fscale = 0.5*|g|.
    double ji[9];
    if (active) {
#pragma unroll
      for (int q = 0; q < 9; ++q) ji[q] = io.jinv[e * 9 + q];
    }
    if constexpr (SPF && LATE) {
      if (grp + gr.step < gr.end) prefetch_state(grp + gr.step);
    }
    if (active && !LATE) {
      if (i < 4) {
        double g[3];
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const double r0 = ji[x], r1 = ji[3 + x], r2 = ji[6 + x];
          g[x] = i == 0 ? -r2 : i == 1 ? -r0 : i == 2 ? -r1 : (r0 + r1 + r2);
        }
        const double len = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const double inv = 1.0 / len;
        double *o = sg + le * 16 + i * 4;
        o[0] = g[0] * inv;
        o[1] = g[1] * inv;
        o[2] = g[2] * inv;
        o[3] = 0.5 * len;
      }
    }
    __syncthreads();

    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (active) {
      const double *ub = su + le * NP * 6;
      double r[3][6];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) r[k][c] = 0.0;
#pragma unroll UNR
      for (int j = 0; j < NP; ++j) {
        double u[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = ub[j * 6 + c];
        const double d0 = dT[(0 * NP + j) * NP + i];
        const double d1 = dT[(1 * NP + j) * NP + i];
        const double d2 = dT[(2 * NP + j) * NP + i];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          r[0][c] = fma(d0, u[c], r[0][c]);
          r[1][c] = fma(d1, u[c], r[1][c]);
          r[2][c] = fma(d2, u[c], r[2][c]);
        }
      }
      auto pd = [&](int p, int c) { return ji[p] * r[0][c] + ji[3 + p] * r[1][c] + ji[6 + p] * r[2][c]; };
      acc[0] = pd(1, 5) - pd(2, 4);
      acc[1] = pd(2, 3) - pd(0, 5);
      acc[2] = pd(0, 4) - pd(1, 3);
      acc[3] = -(pd(1, 2) - pd(2, 1));
      acc[4] = -(pd(2, 0) - pd(0, 2));
      acc[5] = -(pd(0, 1) - pd(1, 0));

      // Upwind flux with unit normal n and c = 0.5*fscale:
      //   fE = c(-n x [H] + a((n.[E])n - [E])),  fH = c(n x [E] + a((n.[H])n - [H])).
      // PEC ([E] = 2E_t, [H] = 2(n.H)n): fE = -2ac(E - (n.E)n), fH = 2c n x E.
#pragma unroll
      for (int s = 0; s < FPT; ++s) {
        const int fm = i + s * NP;
        if (fm >= NF4) continue;
        const int f = fm / NFP, m = fm - f * NFP;
        double n0, n1, n2, cf;
        if constexpr (LATE) {  // geometry here, so the group start need not wait for J^-1
          double nn[3];
          face_geom(ji, f, nn, cf);
          n0 = nn[0];
          n1 = nn[1];
          n2 = nn[2];
        } else {
          const double *gq = sg + le * 16 + f * 4;
          n0 = gq[0];
          n1 = gq[1];
          n2 = gq[2];
          cf = gq[3];
        }
        const int vi = io.face_nodes[f * NFP + m];
        const double *uv = ub + vi * 6;
        const double e0v = uv[0], e1v = uv[1], e2v = uv[2], h0 = uv[3], h1 = uv[4], h2 = uv[5];
        double fe0, fe1, fe2, fh0, fh1, fh2;
        const int code = fcodes[s];
        if (code == kCodePec) {
          const double ne = n0 * e0v + n1 * e1v + n2 * e2v;
          const double ka = -2.0 * io.alpha * cf, k2 = 2.0 * cf;
          fe0 = ka * (e0v - ne * n0);
          fe1 = ka * (e1v - ne * n1);
          fe2 = ka * (e2v - ne * n2);
          fh0 = k2 * (n1 * e2v - n2 * e1v);
          fh1 = k2 * (n2 * e0v - n0 * e2v);
          fh2 = k2 * (n0 * e1v - n1 * e0v);
        } else {
          double je0, je1, je2, jh0, jh1, jh2, a;
          if (code >= 0) {
            if constexpr (!PREV) load_trace<CIN>(io, nbase[s], nv[s]);
            je0 = e0v - nv[s][0];
            je1 = e1v - nv[s][1];
            je2 = e2v - nv[s][2];
            jh0 = h0 - nv[s][3];
            jh1 = h1 - nv[s][4];
            jh2 = h2 - nv[s][5];
            a = io.alpha;
          } else {  // port: zero incident field, fully characteristic
            je0 = e0v, je1 = e1v, je2 = e2v, jh0 = h0, jh1 = h1, jh2 = h2;
            a = 1.0;
          }
          const double ne = n0 * je0 + n1 * je1 + n2 * je2;
          const double nh = n0 * jh0 + n1 * jh1 + n2 * jh2;
          fe0 = cf * (a * (ne * n0 - je0) - (n1 * jh2 - n2 * jh1));
          fe1 = cf * (a * (ne * n1 - je1) - (n2 * jh0 - n0 * jh2));
          fe2 = cf * (a * (ne * n2 - je2) - (n0 * jh1 - n1 * jh0));
          fh0 = cf * (a * (nh * n0 - jh0) + (n1 * je2 - n2 * je1));
          fh1 = cf * (a * (nh * n1 - jh1) + (n2 * je0 - n0 * je2));
          fh2 = cf * (a * (nh * n2 - jh2) + (n0 * je1 - n1 * je0));
        }
        double2 *out = reinterpret_cast<double2 *>(sf + le * SFS + fm * 6);
        out[0] = make_double2(fe0, fe1);
        out[1] = make_double2(fe2, fh0);
        out[2] = make_double2(fh1, fh2);
      }
    }
    __syncthreads();
    if (active) {
      const double *fb = sf + le * SFS;
#pragma unroll UNR
      for (int fm = 0; fm < NF4; ++fm) {
        const double w = lift[fm * NP + i];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(w, fb[fm * 6 + c], acc[c]);
      }
      const double *ub = su + (le * NP + i) * 6;
      double own[6], po[6], srcv[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        own[c] = ub[c];
        po[c] = (double)pold[c];
      }
      if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
        for (int c = 0; c < 6; ++c) srcv[c] = io.src_vec[base + c];
      }
      node_update<EPI>(io, base, sm + le * 12, acc, own, po, srcv);
    }
    __syncthreads();  // su, sf, sg and sm are rewritten by the next group
  }
}

// ---- shared pieces of the flux and the epilogue --------------------------
//
// Operand layout of v_mfma_f64_16x16x4f64 (checked on gfx950): lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and receives C[4r + l/16][l%16]
// in element r.
typedef double mf_d4 __attribute__((ext_vector_type(4)));
// Operands of v_mfma_f32_16x16x32_bf16: lane l supplies A[l%16][8(l/16) + t]
// and B[8(l/16) + t][l%16] (t = 0..7) and receives C[4(l/16) + r][l%16] in
// element r.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Upwind flux of one face node (rhs.rs:1372-1440) with unit normal n and
// cf = 0.5*fscale, own trace u and neighbour trace v (interior faces):
//   fE = cf(-n x [H] + a((n.[E])n - [E])),  fH = cf(n x [E] + a((n.[H])n - [H])),
// PEC ([E] = 2E_t, [H] = 2(n.H)n): fE = -2a cf(E - (n.E)n), fH = 2cf n x E;
// port: [.] = own trace, a = 1.
__device__ __forceinline__ void face_flux(int code, double alpha, const double *n, double cf, const double *u,
                                          const double *v, double *out) {
  const double n0 = n[0], n1 = n[1], n2 = n[2];
  if (code == kCodePec) {
    const double ne = n0 * u[0] + n1 * u[1] + n2 * u[2];
    const double ka = -2.0 * alpha * cf, k2 = 2.0 * cf;
    out[0] = ka * (u[0] - ne * n0);
    out[1] = ka * (u[1] - ne * n1);
    out[2] = ka * (u[2] - ne * n2);
    out[3] = k2 * (n1 * u[2] - n2 * u[1]);
    out[4] = k2 * (n2 * u[0] - n0 * u[2]);
    out[5] = k2 * (n0 * u[1] - n1 * u[0]);
    return;
  }
  double j[6], a;
  if (code >= 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) j[c] = u[c] - v[c];
    a = alpha;
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c) j[c] = u[c];
    a = 1.0;
  }
  const double ne = n0 * j[0] + n1 * j[1] + n2 * j[2];
  const double nh = n0 * j[3] + n1 * j[4] + n2 * j[5];
  out[0] = cf * (a * (ne * n0 - j[0]) - (n1 * j[5] - n2 * j[4]));
  out[1] = cf * (a * (ne * n1 - j[1]) - (n2 * j[3] - n0 * j[5]));
  out[2] = cf * (a * (ne * n2 - j[2]) - (n0 * j[4] - n1 * j[3]));
  out[3] = cf * (a * (nh * n0 - j[3]) + (n1 * j[2] - n2 * j[1]));
  out[4] = cf * (a * (nh * n1 - j[4]) + (n2 * j[0] - n0 * j[2]));
  out[5] = cf * (a * (nh * n2 - j[5]) + (n0 * j[1] - n1 * j[0]));
}

// Materials, held source and the stage update of one node (rhs.rs:1443-1454,
// explicit.rs:100-135): k = M^-1(acc) - sigma u (+ b g), then EPI.
template <int EPI>
__device__ __forceinline__ void node_epilogue(const StageArgs &io, int64_t base, const double *mt, const double *acc,
                                              const double *own) {
  double kv[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    kv[c] = mt[c] * acc[c] - mt[6 + c] * own[c];
    kv[3 + c] = mt[3 + c] * acc[3 + c] - mt[9 + c] * own[3 + c];
  }
  if (io.src_kind == RFHIP_DRIVE_POINT) {
    const int64_t off = io.src_dof - base;
#pragma unroll
    for (int c = 0; c < 6; ++c)
      if (off == c) kv[c] += io.g;
  } else if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
    for (int c = 0; c < 6; ++c) kv[c] += io.src_vec[base + c] * io.g;
  }
  if constexpr (EPI == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) io.dy[base + c] = kv[c];
  } else if constexpr (EPI == 1) {
    double *pp = static_cast<double *>(io.p);
    double *yo = static_cast<double *>(io.yout);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double pn = io.sa * pp[base + c] + io.dt * kv[c];
      pp[base + c] = pn;
      yo[base + c] = own[c] + io.sb * pn;
    }
  } else {
    float *pp = static_cast<float *>(io.p);
    float *yh = static_cast<float *>(io.yout);
    uint16_t *yl = static_cast<uint16_t *>(io.yout_lo);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const float pn = (float)(io.sa * (double)pp[base + c] + io.dt * kv[c]);
      pp[base + c] = pn;
      const double yn = own[c] + io.sb * (double)pn;
      const float h = (float)yn;
      yh[base + c] = h;
      yl[base + c] = bf16_rne((float)(yn - (double)h));
    }
  }
}

// ---- MFMA volume with the curl in registers -------------------------------
//
// k_fused8 maps wave w, lane l of a 16-element block to node 4w + l/16 of
// element l%16. Its volume A operand interleaves the stacked [Dr; Ds; Dt]
// by node (row 4k + q <-> derivative k of node 4w + q, a zero row for
// k = 3), so v_mfma_f64_16x16x4f64 hands each lane, in element r of its
// accumulator, the three reference derivatives of its own node for one
// component (one column tile per component, columns = elements). J^-1 and
// the curl are then applied in registers with no LDS round trip and no
// barrier between the two fields. The face flux goes straight to LDS and
// all six components are lifted in one MFMA phase (three barriers per
// element group in all).
template <int NP>
struct Mfma7Cfg {
  static constexpr int NW = (NP + 3) / 4;  // waves; wave w owns nodes 4w..4w+3
  static constexpr int NT = NW * 64;
  static constexpr int KS = (NP + 3) / 4;
};

// Raw field values of one node as stored (f32 pairs + bf16 pairs, or f64
// pairs), converted to f64 only where they are used, so a load needs no
// wait at the place it is issued.
template <bool CIN>
struct TraceRaw;
template <>
struct TraceRaw<true> {
  float2 h[3];
  uint32_t l[3];
};
template <>
struct TraceRaw<false> {
  double2 d[3];
};

template <bool CIN>
__device__ __forceinline__ void load_trace_raw(const StageArgs &io, int64_t nb, TraceRaw<CIN> &r) {
  if constexpr (CIN) {
    const float2 *hp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + nb);
    const uint32_t *lp = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + nb);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      r.h[q] = hp[q];
      r.l[q] = lp[q];
    }
  } else {
    const double2 *dp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + nb);
#pragma unroll
    for (int q = 0; q < 3; ++q) r.d[q] = dp[q];
  }
}

template <bool CIN>
__device__ __forceinline__ void trace_f64(const TraceRaw<CIN> &r, double *v) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if constexpr (CIN) {
      v[2 * q] = (double)r.h[q].x + (double)__uint_as_float(r.l[q] << 16);
      v[2 * q + 1] = (double)r.h[q].y + (double)__uint_as_float(r.l[q] & 0xFFFF0000u);
    } else {
      v[2 * q] = r.d[q].x;
      v[2 * q + 1] = r.d[q].y;
    }
  }
}

// Marks the registers of a raw trace as read here (an empty asm that takes
// them as inputs): forces the wait for their load at this point on every
// path, so no later write to those registers has to wait for it.
template <bool CIN>
__device__ __forceinline__ void touch_trace(const TraceRaw<CIN> &r) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if constexpr (CIN) {
      asm volatile("" ::"v"(r.h[q].x), "v"(r.h[q].y), "v"(r.l[q]));
    } else {
      asm volatile("" ::"v"(r.d[q].x), "v"(r.d[q].y));
    }
  }
}

#ifndef RF_K8_WPE
#define RF_K8_WPE 3  // waves per SIMD the persistent kernel is compiled for
#endif

// k_fused8 is persistent: each block walks over element groups and reads the
// next group's state, neighbour-table entries and material indices into
// registers while the current group computes, so a group starts without
// waiting on HBM. The group's own J^-1, old p, material rows and neighbour
// traces are requested at its start, ahead of the read-ahead (vector loads
// retire in issue order), and are first used after the barrier that
// publishes the state. All of these loads are unconditional (clamped to a
// valid index on inactive lanes and boundary faces) so that no conversion
// or wait is forced where they are issued. Face geometry is formed in the
// flux; the face-node table lives in LDS.
template <int NP, int NFP, bool CIN, int EPI>
__global__ void __launch_bounds__(Mfma7Cfg<NP>::NT) __attribute__((amdgpu_waves_per_eu(RF_K8_WPE)))
k_fused8(StageArgs io) {
  using Cfg = Mfma7Cfg<NP>;
  constexpr int EPB = 16, E2 = 17, NW = Cfg::NW, NT = Cfg::NT, KS = Cfg::KS;
  constexpr int NF4 = 4 * NFP;
  constexpr int FPT = (NF4 + NP - 1) / NP;
  constexpr int MTL = (NP + 15) / 16, KSL = NF4 / 4;
  constexpr int SS = 6 * NP * E2, SF = 6 * NF4 * E2;
  constexpr bool LFIX = MTL <= NW;
  constexpr int PPT = (EPB * NP * 3 + NT - 1) / NT;  // state pairs per thread
  constexpr int MPT = (EPB * 12 + NT - 1) / NT;  // material entries per thread
  __shared__ __align__(16) double sA[SS];  // state, then the lifted flux L[c][i][e]
  __shared__ __align__(16) double sF[SF];  // face flux F[c][fm][e]
  __shared__ __align__(16) double sm[EPB * 12];
  __shared__ __align__(16) double sL[MTL * KSL * 64];  // LIFT A fragments
  __shared__ int sfn[NF4];
  const int wave = threadIdx.x >> 6;
  const int64_t ngroups = (io.n_elem + EPB - 1) / EPB;
  for (int t = threadIdx.x; t < NF4; t += NT) sfn[t] = io.face_nodes[t];
  for (int t = threadIdx.x; t < MTL * KSL * 64; t += NT) sL[t] = io.lfrag[t];
  // Lane ids are recomputed from an opaque copy of threadIdx.x where they
  // are used inside the group loop, so the compiler does not hoist
  // per-lane 64-bit addresses out of it and hold them in registers.
  auto lane_ids = [&](int &tid, int &lane, int &el, int &q, int &i) {
    int tv = threadIdx.x;
    asm volatile("" : "+v"(tv));
    tid = tv;
    lane = tv & 63;
    el = lane & 15;
    q = lane >> 4;
    i = 4 * wave + q;
  };
  int tid, lane, el, q, i;
  lane_ids(tid, lane, el, q, i);

  double av[KS];
  {
    const double *ap = io.afrag7 + (size_t)wave * KS * 64 + lane;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) av[ks] = ap[ks * 64];
  }
  using HT = typename std::conditional<CIN, float2, double2>::type;
  HT sh[PPT];
  uint32_t sl[PPT];
  int nbn[FPT];
  int mnext[MPT];
  auto read_ahead = [&](int64_t g) {
    int tid, lane, el, q, i;
    lane_ids(tid, lane, el, q, i);
    const int64_t rg = io.n_elem - g * EPB;
    const int ng = rg < EPB ? (int)rg : EPB;
    const int64_t gb = g * EPB * NP * 6;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int t0 = tid + k * NT;
      const int t = t0 < ng * NP * 3 ? t0 : ng * NP * 3 - 1;
      if constexpr (CIN) {
        sh[k] = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + gb)[t];
        sl[k] = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + gb)[t];
      } else {
        sh[k] = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + gb)[t];
      }
    }
    const int64_t eg = g * EPB + (el < ng ? el : ng - 1);
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      const int fm = i + s * NP;
      const int v = io.nbase[eg * NF4 + (fm < NF4 ? fm : NF4 - 1)];
      nbn[s] = (i < NP && fm < NF4) ? v : kCodePec;
    }
#pragma unroll
    for (int k = 0; k < MPT; ++k) {
      const int t = tid + k * NT, te = t / 12;
      mnext[k] = io.mat[g * EPB + (te < ng ? te : ng - 1)];
    }
  };
  // The read-ahead state of a group with `ng` elements into sA, transposed
  // to component-major [c][node][element].
  auto write_state = [&](int ng) {
    int tid, lane, el, q, i;
    lane_ids(tid, lane, el, q, i);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int t = tid + k * NT;
      if (t < ng * NP * 3) {
        double vx, vy;
        if constexpr (CIN) {
          vx = (double)sh[k].x + (double)__uint_as_float(sl[k] << 16);
          vy = (double)sh[k].y + (double)__uint_as_float(sl[k] & 0xFFFF0000u);
        } else {
          vx = sh[k].x;
          vy = sh[k].y;
        }
        const int ee = t / (3 * NP), r2 = t - ee * (3 * NP);
        const int j = r2 / 3, c = 2 * (r2 - j * 3);
        sA[(c * NP + j) * E2 + ee] = vx;
        sA[((c + 1) * NP + j) * E2 + ee] = vy;
      }
    }
  };
  // Every load of the read-ahead is unconditional (the last group reads its
  // own data again), so the compiler can count outstanding loads exactly
  // and never has to drain them all at a merge point. The launch holds at
  // most ngroups blocks.
  const int64_t b0 = xcd_block();
  read_ahead(b0);
  {
    const int64_t r0 = io.n_elem - b0 * EPB;
    write_state(r0 < EPB ? (int)r0 : EPB);
  }

#pragma unroll 1
  for (int64_t grp = b0; grp < ngroups; grp += gridDim.x) {
    lane_ids(tid, lane, el, q, i);
    const int64_t e0 = grp * EPB;
    const int64_t rem = io.n_elem - e0;
    const int nel = rem < EPB ? (int)rem : EPB;
    const bool active = i < NP && el < nel;
    const int64_t e = e0 + (el < nel ? el : nel - 1);  // clamped: loads stay in range
    const int64_t base = (e * NP + (i < NP ? i : NP - 1)) * 6;

    // Group start: this group's reads, then the next group's read-ahead.
    double ji[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) ji[c] = io.jinv[e * 9 + c];
    TraceRaw<CIN> nr[FPT];
    int fcodes[FPT];
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      fcodes[s] = nbn[s] >= 0 ? 0 : nbn[s];
      load_trace_raw<CIN>(io, nbn[s] >= 0 ? nbn[s] : 0, nr[s]);
    }
    using PT = typename std::conditional<EPI == 2, float, double>::type;
    PT pold[6];
    if constexpr (EPI == 1) {
      const double2 *pp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.p) + base);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double2 v = pp[c];
        pold[2 * c] = v.x;
        pold[2 * c + 1] = v.y;
      }
    } else if constexpr (EPI == 2) {
      const float2 *pp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.p) + base);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float2 v = pp[c];
        pold[2 * c] = v.x;
        pold[2 * c + 1] = v.y;
      }
    }
    double mrow[MPT];
#pragma unroll
    for (int k = 0; k < MPT; ++k) mrow[k] = io.mat_tab[12 * mnext[k] + (tid + k * NT) % 12];
    read_ahead(grp + gridDim.x < ngroups ? grp + gridDim.x : grp);
    __syncthreads();

    // Face flux into LDS, then the volume products with their curls (MFMA,
    // no global reads).
    double own[6], acc[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) own[c] = sA[(c * NP + (i < NP ? i : 0)) * E2 + el];
    if (active) {
#pragma unroll
      for (int s = 0; s < FPT; ++s) {
        const int fm = i + s * NP;
        if (fm >= NF4) continue;
        const int f = fm / NFP;
        double n[3], cf;
        face_geom(ji, f, n, cf);
        const int vi = sfn[fm];
        double u[6], nv[6], fo[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = sA[(c * NP + vi) * E2 + el];
        trace_f64<CIN>(nr[s], nv);
        face_flux(fcodes[s], io.alpha, n, cf, u, nv, fo);
#pragma unroll
        for (int c = 0; c < 6; ++c) sF[(c * NF4 + fm) * E2 + el] = fo[c];
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int cbase = h == 0 ? 3 : 0;
      mf_d4 cacc[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) cacc[c] = mf_d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const bool jv = NP % 4 == 0 || ks < KS - 1 || ks * 4 + q < NP;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double b = jv ? sA[((cbase + c) * NP + ks * 4 + q) * E2 + el] : 0.0;
          cacc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], b, cacc[c], 0, 0, 0);
        }
      }
      auto dd = [&](int p, int c) { return ji[p] * cacc[c][0] + ji[3 + p] * cacc[c][1] + ji[6 + p] * cacc[c][2]; };
      const double sgn = h == 0 ? 1.0 : -1.0;
      acc[3 * h + 0] = sgn * (dd(1, 2) - dd(2, 1));
      acc[3 * h + 1] = sgn * (dd(2, 0) - dd(0, 2));
      acc[3 * h + 2] = sgn * (dd(0, 1) - dd(1, 0));
    }
#pragma unroll
    for (int k = 0; k < MPT; ++k)
      if (tid + k * NT < nel * 12) sm[tid + k * NT] = mrow[k];
    __syncthreads();  // state reads done, flux and materials complete

    // LIFT of all six flux components into sA as L[c][i][e].
    {
      const int lcol = lane & 15, lrow = lane >> 4;
      if constexpr (LFIX) {
        const int mt = wave % MTL;
        const int nwm = NW / MTL + (mt < NW % MTL ? 1 : 0);
        const double *ap = sL + mt * KSL * 64 + lane;
        for (int nt = wave / MTL; nt < 6; nt += nwm) {
          const double *bp = sF + (nt * NF4 + lrow) * E2 + lcol;
          mf_d4 c4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < KSL; ++ks)
            c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(ap[ks * 64], bp[ks * 4 * E2], c4, 0, 0, 0);
          double *lp = sA + (nt * NP + mt * 16 + lrow) * E2 + lcol;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (mt * 16 + r * 4 + lrow < NP) lp[r * 4 * E2] = c4[r];
        }
      } else {
        for (int t = wave; t < MTL * 6; t += NW) {
          const int mt = t / 6, nt = t - (t / 6) * 6;
          const double *ap = sL + mt * KSL * 64 + lane;
          const double *bp = sF + (nt * NF4 + lrow) * E2 + lcol;
          mf_d4 c4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
          for (int ks = 0; ks < KSL; ++ks)
            c4 = __builtin_amdgcn_mfma_f64_16x16x4f64(ap[ks * 64], bp[ks * 4 * E2], c4, 0, 0, 0);
          double *lp = sA + (nt * NP + mt * 16 + lrow) * E2 + lcol;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (mt * 16 + r * 4 + lrow < NP) lp[r * 4 * E2] = c4[r];
        }
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[c] += sA[(c * NP + i) * E2 + el];
    }
    __syncthreads();  // L is read; sA takes the next group's state
    // Written before this group's global stores, so waiting for the
    // read-ahead never waits for stores.
    if (grp + gridDim.x < ngroups) {
      const int64_t rn = io.n_elem - (grp + gridDim.x) * EPB;
      write_state(rn < EPB ? (int)rn : EPB);
    }
    if (active) {
      double po[6], srcv[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) po[c] = (double)pold[c];
      if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
        for (int c = 0; c < 6; ++c) srcv[c] = io.src_vec[base + c];
      }
      node_update<EPI>(io, base, sm + el * 12, acc, own, po, srcv);
    }
  }
}

// k_fused9: the k_fused3 layout (one thread per element node, D and LIFT in
// LDS, VALU contractions) with the group's memory traffic pipelined:
//   * the next group's state, neighbour-table entries and material indices
//     are read into registers while the current group computes (the state
//     goes to LDS after the group's last LDS reads);
//   * the group's J^-1, neighbour traces, old p and material rows are all
//     requested at its start, J^-1 first, then the read-ahead; the volume
//     contraction (which needs J^-1 only) runs while the traces arrive;
//   * every load is unconditional, with indices clamped to valid entries on
//     inactive lanes and boundary faces, so no conversion or wait is forced
//     where a load is issued and the compiler counts outstanding loads
//     exactly;
//   * face geometry is formed in the flux and the face-node table is in LDS.
template <int NP, int NFP, bool CIN, int EPI, int UNR>
__global__ void __launch_bounds__(kFusedWG) __attribute__((amdgpu_waves_per_eu(3))) k_fused9(StageArgs io) {
  constexpr int EPB = kFusedWG / NP;
  constexpr int NF4 = 4 * NFP;
  constexpr int FPT = (NF4 + NP - 1) / NP;
  // D_k as [k][i][j] (row stride DS) and LIFT as [i][fm] (row stride LS):
  // a 16-byte LDS read then returns two consecutive j (or fm) entries of a
  // thread's row; the padded strides spread the rows over the banks. At
  // Np = 35 LIFT's rows go unpadded: the block's LDS then fits twice in a
  // CU (81,664 of 163,840 bytes; 82,224 with the pad, one block per CU).
  constexpr int DS = NP + (NP & 1) + 2, LS = NF4 + (NP == 35 ? 0 : 2);
  constexpr int ND = 3 * NP * DS;
  constexpr int SD = ND + NP * LS;
  constexpr int SFS = NF4 * 6 + 2;
  constexpr int PPT = (EPB * NP * 3 + kFusedWG - 1) / kFusedWG;
  __shared__ __align__(16) double su[EPB * NP * 6];
  __shared__ __align__(16) double sf[EPB * SFS];
  __shared__ __align__(16) double sd[SD];
  __shared__ __align__(16) double sm[EPB * 12];
  __shared__ int sfn[NF4];
  for (int q = threadIdx.x; q < SD; q += kFusedWG) {
    double v = 0.0;
    if (q < ND) {
      const int k = q / (NP * DS), r = q - k * NP * DS, i = r / DS, j = r - i * DS;
      if (j < NP) v = io.dT[(k * NP + j) * NP + i];
    } else {
      const int i = (q - ND) / LS, fm = (q - ND) - i * LS;
      if (fm < NF4) v = io.lift[fm * NP + i];
    }
    sd[q] = v;
  }
  for (int q = threadIdx.x; q < NF4; q += kFusedWG) sfn[q] = io.face_nodes[q];
  const double *dT = sd;
  const double *lift = sd + ND;
  const int64_t ngroups = (io.n_elem + EPB - 1) / EPB;
  // Lane ids come from an opaque copy of threadIdx.x wherever the group loop
  // uses them, so per-lane addresses are not hoisted out of it and held in
  // registers.
  auto lane_ids = [&](int &tid, int &le, int &i) {
    int tv = threadIdx.x;
    asm volatile("" : "+v"(tv));
    tid = tv;
    le = tv / NP;
    i = tv - le * NP;
  };

  using HT = typename std::conditional<CIN, float2, double2>::type;
  HT sh[PPT];
  uint32_t sl[PPT];
  int nbn[FPT];
  int mnext;
  auto read_ahead = [&](int64_t g) {
    int tid, le, i;
    lane_ids(tid, le, i);
    const int64_t rg = io.n_elem - g * EPB;
    const int ng = rg < EPB ? (int)rg : EPB;
    const int64_t gb = g * EPB * NP * 6;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int q0 = tid + k * kFusedWG;
      const int q = q0 < ng * NP * 3 ? q0 : ng * NP * 3 - 1;
      if constexpr (CIN) {
        sh[k] = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + gb)[q];
        sl[k] = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + gb)[q];
      } else {
        sh[k] = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + gb)[q];
      }
    }
    const int64_t eg = g * EPB + (le < ng ? le : ng - 1);
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      const int fm = i + s * NP;
      const int v = io.nbase[eg * NF4 + (fm < NF4 ? fm : NF4 - 1)];
      nbn[s] = fm < NF4 ? v : kCodePec;
    }
    mnext = io.mat[eg];
  };
  auto write_state = [&](int ng) {
    int tid, le, i;
    lane_ids(tid, le, i);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int q = tid + k * kFusedWG;
      if (q < ng * NP * 3) {
        double2 v;
        if constexpr (CIN) {
          v.x = (double)sh[k].x + (double)__uint_as_float(sl[k] << 16);
          v.y = (double)sh[k].y + (double)__uint_as_float(sl[k] & 0xFFFF0000u);
        } else {
          v = sh[k];
        }
        reinterpret_cast<double2 *>(su)[q] = v;
      }
    }
  };
  // The launch holds at most ngroups blocks; the last group of a block reads
  // itself ahead again, which keeps every read-ahead load unconditional.
  const int64_t b0 = xcd_block();
  read_ahead(b0);
  {
    const int64_t r0 = io.n_elem - b0 * EPB;
    write_state(r0 < EPB ? (int)r0 : EPB);
  }

#pragma unroll 1
  for (int64_t grp = b0; grp < ngroups; grp += gridDim.x) {
    int tid, le, i;
    lane_ids(tid, le, i);
    const int64_t e0 = grp * EPB;
    const int64_t rem = io.n_elem - e0;
    const int nel = rem < EPB ? (int)rem : EPB;
    const bool active = tid < nel * NP;
    const int64_t e = e0 + (le < nel ? le : nel - 1);
    const int64_t base = (e * NP + i) * 6;

    double ji[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) ji[q] = io.jinv[e * 9 + q];
    TraceRaw<CIN> nr[FPT];
    int fcodes[FPT];
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      fcodes[s] = nbn[s] >= 0 ? 0 : nbn[s];
      load_trace_raw<CIN>(io, nbn[s] >= 0 ? nbn[s] : 0, nr[s]);
    }
    // material row entries i, i + Np, ... (< 12) of element le
    constexpr int MR = (12 + NP - 1) / NP;
    double mrow[MR];
#pragma unroll
    for (int k = 0; k < MR; ++k) {
      const int q = i + k * NP;
      mrow[k] = io.mat_tab[12 * mnext + (q < 12 ? q : 0)];
    }
    read_ahead(grp + gridDim.x < ngroups ? grp + gridDim.x : grp);
    __syncthreads();  // the group's state is in LDS
    // Needed only after the flux; issued here, behind the traces, to keep
    // them out of the registers of the volume contraction.
    using PT = typename std::conditional<EPI == 2, float, double>::type;
    PT pold[6];
    if constexpr (EPI == 1) {
      const double2 *pp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.p) + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double2 v = pp[q];
        pold[2 * q] = v.x;
        pold[2 * q + 1] = v.y;
      }
    } else if constexpr (EPI == 2) {
      const float2 *pp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.p) + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float2 v = pp[q];
        pold[2 * q] = v.x;
        pold[2 * q + 1] = v.y;
      }
    }

    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double *ub = su + le * NP * 6;
    if (active) {
      double r[3][6];
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) r[k][c] = 0.0;
      const double *dr = dT + i * DS;
      auto vol_step = [&](int j, double d0, double d1, double d2) {
        double u[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = ub[j * 6 + c];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          r[0][c] = fma(d0, u[c], r[0][c]);
          r[1][c] = fma(d1, u[c], r[1][c]);
          r[2][c] = fma(d2, u[c], r[2][c]);
        }
      };
#pragma unroll UNR
      for (int j = 0; j < NP - 1; j += 2) {
        const double2 a0 = *reinterpret_cast<const double2 *>(dr + j);
        const double2 a1 = *reinterpret_cast<const double2 *>(dr + NP * DS + j);
        const double2 a2 = *reinterpret_cast<const double2 *>(dr + 2 * NP * DS + j);
        vol_step(j, a0.x, a1.x, a2.x);
        vol_step(j + 1, a0.y, a1.y, a2.y);
      }
      if constexpr (NP & 1) vol_step(NP - 1, dr[NP - 1], dr[NP * DS + NP - 1], dr[2 * NP * DS + NP - 1]);
      auto pd = [&](int p, int c) { return ji[p] * r[0][c] + ji[3 + p] * r[1][c] + ji[6 + p] * r[2][c]; };
      acc[0] = pd(1, 5) - pd(2, 4);
      acc[1] = pd(2, 3) - pd(0, 5);
      acc[2] = pd(0, 4) - pd(1, 3);
      acc[3] = -(pd(1, 2) - pd(2, 1));
      acc[4] = -(pd(2, 0) - pd(0, 2));
      acc[5] = -(pd(0, 1) - pd(1, 0));
      // The curl is formed here, before the flux: left to itself the
      // compiler sinks it past the flux and keeps the 18 reference
      // derivatives live there.
      asm volatile("" ::"v"(acc[0]), "v"(acc[1]), "v"(acc[2]), "v"(acc[3]), "v"(acc[4]), "v"(acc[5]) : "memory");

#pragma unroll
      for (int s = 0; s < FPT; ++s) {
        const int fm = i + s * NP;
        if (fm >= NF4) continue;
        const int f = fm / NFP;
        double n[3], cf;
        face_geom(ji, f, n, cf);
        double u[6], nv[6], fo[6];
        const double *uv = ub + sfn[fm] * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = uv[c];
        trace_f64<CIN>(nr[s], nv);
        face_flux(fcodes[s], io.alpha, n, cf, u, nv, fo);
        double2 *out = reinterpret_cast<double2 *>(sf + le * SFS + fm * 6);
        out[0] = make_double2(fo[0], fo[1]);
        out[1] = make_double2(fo[2], fo[3]);
        out[2] = make_double2(fo[4], fo[5]);
      }
#pragma unroll
      for (int k = 0; k < MR; ++k)
        if (i + k * NP < 12) sm[le * 12 + i + k * NP] = mrow[k];
    }
    __syncthreads();  // flux and materials complete
    // Every register loaded for this group is taken here on all lanes: left
    // to lanes that are active only, a lane-predicated branch would leave the
    // loads pending on its skip path, and the next group start would wait for
    // all outstanding loads before reusing those registers.
#pragma unroll
    for (int q = 0; q < 9; ++q) asm volatile("" ::"v"(ji[q]));
#pragma unroll
    for (int s = 0; s < FPT; ++s) touch_trace<CIN>(nr[s]);
    if constexpr (EPI != 0) {
#pragma unroll
      for (int c = 0; c < 6; ++c) asm volatile("" ::"v"(pold[c]));
    }
#pragma unroll
    for (int k = 0; k < MR; ++k) asm volatile("" ::"v"(mrow[k]));
    double own[6];
    if (active) {
      const double *fb = sf + le * SFS;
      const double *lr = lift + i * LS;
#pragma unroll UNR
      for (int fm = 0; fm < NF4; fm += 2) {  // NF4 = 4 Nfp is even
        const double2 w = *reinterpret_cast<const double2 *>(lr + fm);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(w.x, fb[fm * 6 + c], acc[c]);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(w.y, fb[fm * 6 + 6 + c], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) own[c] = ub[i * 6 + c];
    }
    __syncthreads();  // su and sf are read; su takes the next group's state
    if (grp + gridDim.x < ngroups) {
      const int64_t rn = io.n_elem - (grp + gridDim.x) * EPB;
      write_state(rn < EPB ? (int)rn : EPB);
    }
    // The remaining read-ahead registers are taken here, before this
    // group's stores: a wait placed after the (lane-predicated) stores would
    // be counted as if they had not been issued and would wait on them.
#pragma unroll
    for (int s = 0; s < FPT; ++s) asm volatile("" : "+v"(nbn[s]));
    asm volatile("" : "+v"(mnext));
    if (active) {
      double po[6], srcv[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) po[c] = (double)pold[c];
      if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
        for (int c = 0; c < 6; ++c) srcv[c] = io.src_vec[base + c];
      }
      node_update<EPI>(io, base, sm + le * 12, acc, own, po, srcv);
    }
  }
}

// k_fused11: k_fused9 with whole elements per wave (EPB = 64/Np elements,
// one lane per element node; the block's 4 waves work independently on
// their own groups). Every LDS hand-off of a group (state, flux, materials,
// the LIFT reads) stays inside one wave, so after the initial table load
// there are no block-wide barriers: a wave stalled on memory or LDS does not
// hold the block's other waves.
// Waves per SIMD the compiler fits k_fused11's registers to: 3 for the
// compressed-register stages (EPI 2, 3) but at Np = 10, else 2 (the FP64
// registers and the apply, EPI 0 and 1; Np = 10; the volume on the matrix
// cores). Fitted to 3, the FP64-register stage spilled 30-38 VGPRs and ran
// 0.520 ms at N=3 against 0.418 at 2, and the compressed-register stage at
// N=2 0.70 / 0.47 ms (unrolled by 2 / 1) against 0.41; the compressed stage
// at N=3 runs the same at 2 and 3 (0.363 / 0.354 ms).
template <int NP, int EPI, int VMF>
struct K11Waves {
  static constexpr int v = VMF ? 2 : (NP != 10 && (EPI == 2 || EPI == 3)) ? 3 : 2;
};
template <int NP, int NFP, bool CIN, int EPI, int UNR, int VMF = 0>
__global__ void __launch_bounds__(kFusedWG) __attribute__((amdgpu_waves_per_eu(K11Waves<NP, EPI, VMF>::v))) k_fused11(StageArgs io) {
  constexpr int EPB = 64 / NP;  // elements per WAVE: a wave owns whole elements
  constexpr int NWB = kFusedWG / 64;
  constexpr int NF4 = 4 * NFP;
  constexpr int FPT = (NF4 + NP - 1) / NP;
  // D_k as [k][i][j] (row stride DS) and LIFT as [i][fm] (row stride LS):
  // a 16-byte LDS read then returns two consecutive j (or fm) entries of a
  // thread's row; the padded strides spread the rows over the banks.
  constexpr int DS = NP + (NP & 1) + 2, LS = NF4 + 2;
  constexpr int ND = VMF ? 0 : 3 * NP * DS;  // with the volume on the matrix cores D is not in LDS
  constexpr int SD = ND + NP * LS;
  constexpr int SFS = NF4 * 6 + 2;
  constexpr int PPT = (EPB * NP * 3 + 63) / 64;
  // VMF > 0: the volume contraction R = [Dr; Ds; Dt] U of the wave's group
  // (3Np x Np times Np x 6 EPB, columns element-major field components) on
  // v_mfma_f32_16x16x32_bf16, both operands split into VMF bf16 pieces
  // (D_k from the host, the state when it is written to LDS), the products
  // of pieces p, q with p + q <= VMF + 1 summed smallest first in FP32.
  // The pieces of the state, B[piece][col][node] with a node stride SBS,
  // share the wave's flux buffer (free between the LIFT of one group and
  // the flux of the next); the products go back through the same buffer as
  // R[e][i][k][c] in f32. J^-1, the curl, the flux, LIFT and the update stay
  // FP64.
  constexpr int KSV = (NP + 31) / 32;        // k-steps of 32 nodes
  constexpr int MTV = (3 * NP + 15) / 16;    // 16-row tiles of [Dr; Ds; Dt]
  constexpr int NCOL = 6 * EPB;              // columns of U
  constexpr int NTV = (NCOL + 15) / 16;      // 16-column tiles
  constexpr int SBS = (NP + 3) & ~3;         // node stride of a column (8-byte aligned reads)
  constexpr int SBE = NTV * 16 * SBS + 32 * KSV;  // entries per piece, with the read slack of the last tile
  constexpr int SFW_V = VMF ? ((VMF * SBE * 2 + 7) / 8 > (EPB * NP * 18 * 4 + 7) / 8 ? (VMF * SBE * 2 + 7) / 8
                                                                                       : (EPB * NP * 18 * 4 + 7) / 8)
                            : 0;
  constexpr int SFW = EPB * SFS > SFW_V ? EPB * SFS : SFW_V;
  __shared__ __align__(16) double su_b[NWB][EPB * NP * 6];
  __shared__ __align__(16) double sf_b[NWB][SFW];
  __shared__ __align__(16) double sd[SD];
  __shared__ __align__(16) double sm_b[NWB][EPB * 12];
  __shared__ int sfn[NF4];
  for (int q = threadIdx.x; q < SD; q += kFusedWG) {
    double v = 0.0;
    if (q < ND) {
      const int k = q / (NP * DS), r = q - k * NP * DS, i = r / DS, j = r - i * DS;
      if (j < NP) v = io.dT[(k * NP + j) * NP + i];
    } else {
      const int i = (q - ND) / LS, fm = (q - ND) - i * LS;
      if (fm < NF4) v = io.lift[fm * NP + i];
    }
    sd[q] = v;
  }
  for (int q = threadIdx.x; q < NF4; q += kFusedWG) sfn[q] = io.face_nodes[q];
  const double *dT = sd;
  const double *lift = sd + ND;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double *su = su_b[wave];
  double *sf = sf_b[wave];
  double *sm = sm_b[wave];
  const int64_t ngroups = (io.n_elem + EPB - 1) / EPB;  // groups of EPB elements, one wave each
  const int64_t gw0 = xcd_block() * NWB + wave, gws = (int64_t)gridDim.x * NWB;
  // The only block-wide barrier: D, LIFT and the face-node table are in LDS.
  __syncthreads();
  // Every other hand-off goes through the wave's own LDS slice; a wave's LDS
  // operations complete in order, so ordering the compiler's code is enough.
  auto wave_sync = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  // Lane ids come from an opaque copy of threadIdx.x wherever the group loop
  // uses them, so per-lane addresses are not hoisted out of it and held in
  // registers.
  auto lane_ids = [&](int &tid, int &le, int &i) {
    int tv = threadIdx.x;
    asm volatile("" : "+v"(tv));
    tid = tv & 63;  // lane
    le = tid / NP;  // element of the wave's group (EPB: spare lanes)
    i = tid - le * NP;
  };

  using HT = typename std::conditional<CIN, float2, double2>::type;
  HT sh[PPT];
  uint32_t sl[PPT];
  int nbn[FPT];
  int mnext;
  auto read_ahead = [&](int64_t g) {
    int tid, le, i;
    lane_ids(tid, le, i);
    const int64_t rg = io.n_elem - g * EPB;
    const int ng = rg < EPB ? (int)rg : EPB;
    const int64_t gb = g * EPB * NP * 6;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int q0 = tid + k * 64;
      const int q = q0 < ng * NP * 3 ? q0 : ng * NP * 3 - 1;
      if constexpr (CIN) {
        sh[k] = reinterpret_cast<const float2 *>(static_cast<const float *>(io.yin) + gb)[q];
        sl[k] = reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(io.yin_lo) + gb)[q];
      } else {
        sh[k] = reinterpret_cast<const double2 *>(static_cast<const double *>(io.yin) + gb)[q];
      }
    }
    const int64_t eg = g * EPB + (le < ng ? le : ng - 1);
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      const int fm = i + s * NP;
      const int v = io.nbase[eg * NF4 + (fm < NF4 ? fm : NF4 - 1)];
      nbn[s] = fm < NF4 ? v : kCodePec;
    }
    mnext = io.mat[eg];
  };
  auto write_state = [&](int ng) {
    int tid, le, i;
    lane_ids(tid, le, i);
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int q = tid + k * 64;
      if (q < ng * NP * 3) {
        double2 v;
        if constexpr (CIN) {
          v.x = (double)sh[k].x + (double)__uint_as_float(sl[k] << 16);
          v.y = (double)sh[k].y + (double)__uint_as_float(sl[k] & 0xFFFF0000u);
        } else {
          v = sh[k];
        }
        reinterpret_cast<double2 *>(su)[q] = v;
      }
    }
    if constexpr (VMF > 0) {
      // bf16 pieces of the state, transposed to B[piece][col][node]; zeros for
      // the node padding, the elements past ng and the read slack after the
      // last column, so every entry a valid column's fragment reads is finite.
      uint16_t *sb = reinterpret_cast<uint16_t *>(sf);
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        const int q = tid + k * 64;
        if (q < ng * NP * 3) {
          const int e = q / (NP * 3), r = q - e * (NP * 3), j = r / 3, c = 2 * (r - j * 3);
          uint16_t pc[2][VMF];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            if constexpr (CIN) {
              // f32 hi + bf16 lo: the hi part splits exactly, lo joins the last piece
              float t = h ? sh[k].y : sh[k].x;
              const float lo = __uint_as_float(h ? (sl[k] & 0xFFFF0000u) : (sl[k] << 16));
#pragma unroll
              for (int pz = 0; pz < VMF; ++pz) {
                const float x = pz == VMF - 1 ? t + lo : t;
                pc[h][pz] = bf16_rne(x);
                t -= bf16_to_f(pc[h][pz]);
              }
            } else {
              double t = h ? sh[k].y : sh[k].x;
#pragma unroll
              for (int pz = 0; pz < VMF; ++pz) {
                pc[h][pz] = bf16_rne((float)t);
                t -= (double)bf16_to_f(pc[h][pz]);
              }
            }
          }
          const int col = 6 * e + c;
#pragma unroll
          for (int pz = 0; pz < VMF; ++pz) {
            sb[pz * SBE + col * SBS + j] = pc[0][pz];
            sb[pz * SBE + (col + 1) * SBS + j] = pc[1][pz];
          }
        }
      }
      constexpr int NPAD = SBS - NP;  // node padding per column
      const int nvalid = ng * 6;
      const int zcount = nvalid * NPAD + (NCOL - nvalid) * SBS + 32 * KSV;
      for (int q = tid; q < VMF * zcount; q += 64) {
        const int pz = q / zcount, z = q - pz * zcount;
        int at;
        if (z < nvalid * NPAD) {
          const int col = NPAD ? z / (NPAD ? NPAD : 1) : 0;
          at = col * SBS + NP + (z - col * NPAD);
        } else {
          at = nvalid * SBS + (z - nvalid * NPAD);
        }
        sb[pz * SBE + at] = 0;
      }
    }
  };
  // KCL stage (EPI 3, 4): the wave's sum of the error-norm terms over its
  // nodes, left in io.kpart[wave of the grid] by the last stage.
  const int64_t wslot = (int64_t)blockIdx.x * NWB + wave;
  double wsum = 0.0;
  // Waves without a group at all (the last block) just leave (with a zero
  // partial); the last group of a wave reads itself ahead again, which
  // keeps every read-ahead load unconditional.
  if (gw0 >= ngroups) {
    if constexpr (EPI == 3 || EPI == 4)
      if (io.klast && (threadIdx.x & 63) == 0) io.kpart[wslot] = 0.0;
    return;
  }
  read_ahead(gw0);
  {
    const int64_t r0 = io.n_elem - gw0 * EPB;
    write_state(r0 < EPB ? (int)r0 : EPB);
  }

#pragma unroll 1
  for (int64_t grp = gw0; grp < ngroups; grp += gws) {
    int tid, le, i;
    lane_ids(tid, le, i);
    const int64_t e0 = grp * EPB;
    const int64_t rem = io.n_elem - e0;
    const int nel = rem < EPB ? (int)rem : EPB;
    const bool active = tid < nel * NP;  // tid is the lane
    const int64_t e = e0 + (le < nel ? le : nel - 1);
    const int64_t base = (e * NP + i) * 6;

    double ji[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) ji[q] = io.jinv[e * 9 + q];
    TraceRaw<CIN> nr[FPT];
    int fcodes[FPT];
#pragma unroll
    for (int s = 0; s < FPT; ++s) {
      fcodes[s] = nbn[s] >= 0 ? 0 : nbn[s];
      load_trace_raw<CIN>(io, nbn[s] >= 0 ? nbn[s] : 0, nr[s]);
    }
    // material row entries i, i + Np, ... (< 12) of element le
    constexpr int MR = (12 + NP - 1) / NP;
    double mrow[MR];
#pragma unroll
    for (int k = 0; k < MR; ++k) {
      const int q = i + k * NP;
      mrow[k] = io.mat_tab[12 * mnext + (q < 12 ? q : 0)];
    }
    read_ahead(grp + gws < ngroups ? grp + gws : grp);
    wave_sync();  // the group's state is in the wave's LDS
    // Needed only after the flux; issued here, behind the traces, to keep
    // them out of the registers of the volume contraction.
    using PT = typename std::conditional<EPI == 2, float, double>::type;
    PT pold[6];
    if constexpr (EPI == 1) {
      const double2 *pp = reinterpret_cast<const double2 *>(static_cast<const double *>(io.p) + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double2 v = pp[q];
        pold[2 * q] = v.x;
        pold[2 * q + 1] = v.y;
      }
    } else if constexpr (EPI == 2) {
      const float2 *pp = reinterpret_cast<const float2 *>(static_cast<const float *>(io.p) + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float2 v = pp[q];
        pold[2 * q] = v.x;
        pold[2 * q + 1] = v.y;
      }
    }

    double acc[6] = {0, 0, 0, 0, 0, 0};
    const double *ub = su + le * NP * 6;
    double r[3][6];
    if constexpr (VMF > 0) {
      // every lane takes part in the MFMAs, active or not
      const uint16_t *sb = reinterpret_cast<const uint16_t *>(sf);
      const uint4 *afr = static_cast<const uint4 *>(io.afragb);
      const int lc = tid & 15, lq = tid >> 4;
      // all B pieces of the group into registers first, so each product tile
      // can go back to the (shared) buffer as soon as it is complete
      bf16x8 bq[KSV][NTV][VMF];
#pragma unroll
      for (int ks = 0; ks < KSV; ++ks)
#pragma unroll
        for (int nt = 0; nt < NTV; ++nt)
#pragma unroll
          for (int pz = 0; pz < VMF; ++pz) {
            const uint2 *src = reinterpret_cast<const uint2 *>(sb + pz * SBE + (16 * nt + lc) * SBS + 32 * ks + 8 * lq);
            const uint2 w0 = src[0], w1 = src[1];
            bq[ks][nt][pz] = __builtin_bit_cast(bf16x8, make_uint4(w0.x, w0.y, w1.x, w1.y));
          }
      wave_sync();  // the pieces are read; the products take their place
      float *sR = reinterpret_cast<float *>(sf);
#pragma unroll
      for (int mt = 0; mt < MTV; ++mt) {
        f32x4 cacc[NTV];
#pragma unroll
        for (int nt = 0; nt < NTV; ++nt) cacc[nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < KSV; ++ks) {
          bf16x8 aq[VMF];
#pragma unroll
          for (int pz = 0; pz < VMF; ++pz)
            aq[pz] = __builtin_bit_cast(bf16x8, afr[((size_t)(pz * MTV + mt) * KSV + ks) * 64 + tid]);
#pragma unroll
          for (int nt = 0; nt < NTV; ++nt)
            // pieces pa, pb (0-based) with pa + pb <= VMF - 1, smallest products first
#pragma unroll
            for (int sum = VMF - 1; sum >= 0; --sum)
#pragma unroll
              for (int pa = 0; pa <= sum; ++pa)
                cacc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aq[pa], bq[ks][nt][sum - pa], cacc[nt], 0, 0, 0);
        }
#pragma unroll
        for (int nt = 0; nt < NTV; ++nt)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = 16 * mt + 4 * lq + q, col = 16 * nt + lc;
            if (row < 3 * NP && col < NCOL) {
              const int k = row / NP, ii = row - k * NP, ee = col / 6, c = col - ee * 6;
              sR[((ee * NP + ii) * 3 + k) * 6 + c] = cacc[nt][q];
            }
          }
      }
      wave_sync();
      if (active) {
        const float *rr = sR + (le * NP + i) * 18;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int c = 0; c < 6; ++c) r[k][c] = (double)rr[k * 6 + c];
      }
      wave_sync();  // the products are read before the flux overwrites them
    }
    if (active) {
      if constexpr (VMF == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int c = 0; c < 6; ++c) r[k][c] = 0.0;
        const double *dr = dT + i * DS;
        auto vol_step = [&](int j, double d0, double d1, double d2) {
          double u[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) u[c] = ub[j * 6 + c];
#pragma unroll
          for (int c = 0; c < 6; ++c) {
            r[0][c] = fma(d0, u[c], r[0][c]);
            r[1][c] = fma(d1, u[c], r[1][c]);
            r[2][c] = fma(d2, u[c], r[2][c]);
          }
        };
#pragma unroll UNR
        for (int j = 0; j < NP - 1; j += 2) {
          const double2 a0 = *reinterpret_cast<const double2 *>(dr + j);
          const double2 a1 = *reinterpret_cast<const double2 *>(dr + NP * DS + j);
          const double2 a2 = *reinterpret_cast<const double2 *>(dr + 2 * NP * DS + j);
          vol_step(j, a0.x, a1.x, a2.x);
          vol_step(j + 1, a0.y, a1.y, a2.y);
        }
        if constexpr (NP & 1) vol_step(NP - 1, dr[NP - 1], dr[NP * DS + NP - 1], dr[2 * NP * DS + NP - 1]);
      }
      auto pd = [&](int p, int c) { return ji[p] * r[0][c] + ji[3 + p] * r[1][c] + ji[6 + p] * r[2][c]; };
      acc[0] = pd(1, 5) - pd(2, 4);
      acc[1] = pd(2, 3) - pd(0, 5);
      acc[2] = pd(0, 4) - pd(1, 3);
      acc[3] = -(pd(1, 2) - pd(2, 1));
      acc[4] = -(pd(2, 0) - pd(0, 2));
      acc[5] = -(pd(0, 1) - pd(1, 0));
      // The curl is formed here, before the flux: left to itself the
      // compiler sinks it past the flux and keeps the 18 reference
      // derivatives live there.
      asm volatile("" ::"v"(acc[0]), "v"(acc[1]), "v"(acc[2]), "v"(acc[3]), "v"(acc[4]), "v"(acc[5]) : "memory");

#pragma unroll
      for (int s = 0; s < FPT; ++s) {
        const int fm = i + s * NP;
        if (fm >= NF4) continue;
        const int f = fm / NFP;
        double n[3], cf;
        face_geom(ji, f, n, cf);
        double u[6], nv[6], fo[6];
        const double *uv = ub + sfn[fm] * 6;
#pragma unroll
        for (int c = 0; c < 6; ++c) u[c] = uv[c];
        trace_f64<CIN>(nr[s], nv);
        face_flux(fcodes[s], io.alpha, n, cf, u, nv, fo);
        double2 *out = reinterpret_cast<double2 *>(sf + le * SFS + fm * 6);
        out[0] = make_double2(fo[0], fo[1]);
        out[1] = make_double2(fo[2], fo[3]);
        out[2] = make_double2(fo[4], fo[5]);
      }
#pragma unroll
      for (int k = 0; k < MR; ++k)
        if (i + k * NP < 12) sm[le * 12 + i + k * NP] = mrow[k];
    }
    wave_sync();  // flux and materials complete
    // Every register loaded for this group is taken here on all lanes: left
    // to lanes that are active only, a lane-predicated branch would leave the
    // loads pending on its skip path, and the next group start would wait for
    // all outstanding loads before reusing those registers.
#pragma unroll
    for (int q = 0; q < 9; ++q) asm volatile("" ::"v"(ji[q]));
#pragma unroll
    for (int s = 0; s < FPT; ++s) touch_trace<CIN>(nr[s]);
    if constexpr (EPI == 1 || EPI == 2) {
#pragma unroll
      for (int c = 0; c < 6; ++c) asm volatile("" ::"v"(pold[c]));
    }
#pragma unroll
    for (int k = 0; k < MR; ++k) asm volatile("" ::"v"(mrow[k]));
    // KCL registers of the lane's node (every lane, clamped like the other
    // group loads), requested here so the LIFT covers their latency
    float kyh[6], kev[6];
    uint32_t kyl[3];
    if constexpr (EPI == 3) {
      const float2 *yp = reinterpret_cast<const float2 *>(io.ky + base);
      const uint32_t *lp = reinterpret_cast<const uint32_t *>(io.ky_lo + base);
      const float2 *ep = reinterpret_cast<const float2 *>(io.kerr + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const float2 a = yp[q], b = ep[q];
        kyh[2 * q] = a.x;
        kyh[2 * q + 1] = a.y;
        kev[2 * q] = b.x;
        kev[2 * q + 1] = b.y;
        kyl[q] = lp[q];
      }
    }
    double kyd[6], ked[6];
    if constexpr (EPI == 4) {
      const double2 *yp = reinterpret_cast<const double2 *>(io.kyd + base);
      const double2 *ep = reinterpret_cast<const double2 *>(io.kerrd + base);
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double2 a = yp[q], b = ep[q];
        kyd[2 * q] = a.x;
        kyd[2 * q + 1] = a.y;
        ked[2 * q] = b.x;
        ked[2 * q + 1] = b.y;
      }
    }
    double own[6];
    if (active) {
      const double *fb = sf + le * SFS;
      const double *lr = lift + i * LS;
#pragma unroll UNR
      for (int fm = 0; fm < NF4; fm += 2) {  // NF4 = 4 Nfp is even
        const double2 w = *reinterpret_cast<const double2 *>(lr + fm);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(w.x, fb[fm * 6 + c], acc[c]);
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] = fma(w.y, fb[fm * 6 + 6 + c], acc[c]);
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) own[c] = ub[i * 6 + c];
    }
    wave_sync();  // su and sf are read; su takes the next group's state
    if (grp + gws < ngroups) {
      const int64_t rn = io.n_elem - (grp + gws) * EPB;
      write_state(rn < EPB ? (int)rn : EPB);
    }
    // The remaining read-ahead registers are taken here, before this
    // group's stores: a wait placed after the (lane-predicated) stores would
    // be counted as if they had not been issued and would wait on them.
#pragma unroll
    for (int s = 0; s < FPT; ++s) asm volatile("" : "+v"(nbn[s]));
    asm volatile("" : "+v"(mnext));
    if constexpr (EPI == 3) {
#pragma unroll
      for (int c = 0; c < 6; ++c) asm volatile("" : "+v"(kyh[c]), "+v"(kev[c]));
#pragma unroll
      for (int q = 0; q < 3; ++q) asm volatile("" : "+v"(kyl[q]));
    } else if constexpr (EPI == 4) {
#pragma unroll
      for (int c = 0; c < 6; ++c) asm volatile("" : "+v"(kyd[c]), "+v"(ked[c]));
    }
    if (active) {
      double srcv[6];
      if (io.src_kind == RFHIP_DRIVE_VECTOR) {
#pragma unroll
        for (int c = 0; c < 6; ++c) srcv[c] = io.src_vec[base + c];
      }
      if constexpr (EPI == 3) {
        node_kcl<CIN>(io, base, sm + le * 12, acc, own, srcv, kyh, kyl, kev, wsum);
      } else if constexpr (EPI == 4) {
        node_kcl64(io, base, sm + le * 12, acc, own, srcv, kyd, ked, wsum);
      } else {
        double po[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) po[c] = (double)pold[c];
        node_update<EPI>(io, base, sm + le * 12, acc, own, po, srcv);
      }
    }
  }
  if constexpr (EPI == 3 || EPI == 4) {
    if (io.klast) {
      // fixed butterfly order: the same sum for the same grid
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o, 64);
      if ((threadIdx.x & 63) == 0) io.kpart[wslot] = wsum;
    }
  }
}

// ---- component lanes ------------------------------------------------------
//
// k_clane maps lane 6q + c of a wave to field component c (Ex..Hz) of
// element q of the wave's group of kClEpw = 10 elements (lanes 60..63 idle).
// Every lane then runs the same node sequence, so the entries of D_k and
// LIFT are the same for all lanes of a wave: they come through the scalar
// cache as SGPR operands of the FMAs, read in the order the kernel uses them
// from a table the host packs once per operator (StageArgs::clane), and the
// zeros of the equispaced D_k (ElemTab::DM: 258 of 400 per D_k at N=3) are
// skipped when the kernel is compiled. A lane holds its component at every
// node of its element in registers (u) and accumulates its component of the
// right-hand side (acc); D and LIFT are never read from LDS. What one
// component needs of the others moves between the element's six lanes with
// ds_bpermute: two physical derivatives per node for the curl, and four
// face jumps per face node for the flux.
//
// The J^-1 entries are loaded with the lane's component axis rotated to the
// front (jr[k][a] = J^-1[k][(p + a) % 3] for field axis p), so the face
// normal each lane forms is (n_p, n_p+1, n_p+2) and the flux of component p
// is one formula for every lane:
//   F_p = cf (a ((n . j_same) n_p - j_p) + s (n_p+1 j'_p+2 - n_p+2 j'_p+1)),
// s = -1 for E, +1 for H, j' the other field's jumps (rhs.rs:1372-1440; the
// PEC and port rules enter through the lane's own jump, see face_flux).
constexpr int kClEpw = 10;

template <int NP>
struct ClaneWaves {
  static constexpr int v = NP >= 20 ? 2 : 3;
};

// Entries of the packed table before (i, j, k) in (i, j, k) order (k
// fastest): the nonzeros of D_r, D_s, D_t. clane_nd<NP>(NP, 0, 0) counts all
// of them, so LIFT's entries start there.
template <int NP>
constexpr int clane_nd(int i_end, int j_end, int k_end) {
  int s = 0;
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j)
      for (int k = 0; k < 3; ++k) {
        if (i > i_end || (i == i_end && (j > j_end || (j == j_end && k >= k_end)))) return s;
        s += (int)((ElemTab<NP>::DM[k][i] >> j) & 1ull);
      }
  return s;
}

// f(std::integral_constant<int, B>), ..., f(... E - 1): loop bodies whose
// index must be a compile-time constant (register arrays, table entries).
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

__device__ __forceinline__ double bperm_f64(int src_lane, double v) {
  const int2 w = __builtin_bit_cast(int2, v);
  const int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, w.x);
  const int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, w.y);
  return __builtin_bit_cast(double, make_int2(lo, hi));
}

template <int NP, bool CIN, int EPI>
__global__ void __launch_bounds__(kFusedWG) __attribute__((amdgpu_waves_per_eu(ClaneWaves<NP>::v)))
k_clane(StageArgs io) {
  using T = ElemTab<NP>;
  constexpr int NFP = T::NFP;
  constexpr int NWB = kFusedWG / 64;
  constexpr int NDV = clane_nd<NP>(NP, 0, 0);
  typedef const double __attribute__((address_space(4))) *cptr;
  // neighbour volume node per ((f*4 + nlf)*6 + o, m); rows 96 + f: the
  // element's own face nodes (boundary faces read their own trace)
  __shared__ int snb[100 * NFP];
  for (int q = threadIdx.x; q < 100 * NFP; q += kFusedWG)
    snb[q] = q < 96 * NFP ? io.nbr_node[q] : io.face_nodes[q - 96 * NFP];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int qg = lane / 6 < kClEpw ? lane / 6 : kClEpw - 1;  // idle lanes mirror element 9
  const int c = lane % 6;
  const int fld = c >= 3 ? 1 : 0;
  const int pc = c - 3 * fld;
  const int p1 = pc == 2 ? 0 : pc + 1, p2 = pc == 0 ? 2 : pc - 1;
  const int srcS1 = 6 * qg + 3 * fld + p1, srcS2 = 6 * qg + 3 * fld + p2;
  const int srcO1 = 6 * qg + 3 * (1 - fld) + p1, srcO2 = 6 * qg + 3 * (1 - fld) + p2;
  const double sgn = fld ? -1.0 : 1.0;  // E: +curl H, H: -curl E
  const int64_t ngroups = (io.n_elem + kClEpw - 1) / kClEpw;
  const float *yh = static_cast<const float *>(io.yin);
  const uint16_t *yl = static_cast<const uint16_t *>(io.yin_lo);
  const double *yd = static_cast<const double *>(io.yin);
  using RT = typename std::conditional<CIN, float, double>::type;
  struct Raw {
    RT h;
    uint32_t l;
  };
  auto load_raw = [&](int64_t idx, Raw &r) {
    if constexpr (CIN) {
      r.h = yh[idx];
      r.l = yl[idx];
    } else {
      r.h = yd[idx];
      r.l = 0;
    }
  };
  auto raw_f64 = [&](const Raw &r) -> double {
    if constexpr (CIN)
      return (double)r.h + (double)__uint_as_float(r.l << 16);
    else
      return r.h;
  };
#pragma unroll 1
  for (int64_t grp = xcd_block() * NWB + wave; grp < ngroups; grp += (int64_t)gridDim.x * NWB) {
    const int64_t e0 = grp * kClEpw;
    const int64_t rem = io.n_elem - e0;
    const int nel = rem < kClEpw ? (int)rem : kClEpw;
    const bool active = lane < 6 * nel;
    const int64_t e = e0 + (qg < nel ? qg : nel - 1);
    const int64_t b0 = e * NP * 6 + c;  // index of node 0 of this lane's component
    // The table's address is laundered per group: its entries are loop
    // invariant, and hoisted out of the group loop they would all have to
    // stay in registers.
    const double *csp = io.clane;
    asm volatile("" : "+s"(csp));
    // Loads from the constant address space are taken never to be written,
    // so the scheduler would issue every table load of the group at once
    // (more SGPRs than a wave has). The entries are used in batches (the D_k
    // rows of a node, a LIFT row); after each batch the table address is
    // laundered with the batch's result as an input, so the next batch's
    // entries load once it is done. (Letting them load one batch earlier
    // raised the register count and measured slower.)
    cptr cs = (cptr)csp;  // wave-uniform: scalar loads
    auto next_batch = [&](double dep) {
      asm volatile("" : "+s"(csp) : "v"(dep));
      cs = (cptr)csp;
    };
    double jr[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      jr[k][0] = io.jinv[e * 9 + 3 * k + pc];
      jr[k][1] = io.jinv[e * 9 + 3 * k + p1];
      jr[k][2] = io.jinv[e * 9 + 3 * k + p2];
    }
    int code[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) code[f] = io.fcode[e * 4 + f];
    const int mi = io.mat[e];
    const double m_inv = io.mat_tab[12 * mi + c];
    const double m_sig = io.mat_tab[12 * mi + 6 + c];
    double u[NP];
    {
      Raw ur[NP];
      static_for<0, NP>([&](auto j) { load_raw(b0 + 6 * j.value, ur[j.value]); });
      static_for<0, NP>([&](auto j) { u[j.value] = raw_f64(ur[j.value]); });
    }
    // neighbour trace of face f, this lane's component (own node on a
    // boundary face, so every load is valid and unconditional)
    Raw nb[NFP];
    auto issue_face = [&](auto fc) {
      constexpr int f = decltype(fc)::value;
      const int cd = code[f];
      const int64_t ne = cd >= 0 ? (int64_t)(cd >> 5) : e;
      const int tb = (cd >= 0 ? (f * 4 + ((cd >> 3) & 3)) * 6 + (cd & 7) : 96 + f) * NFP;
      static_for<0, NFP>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        load_raw((ne * NP + snb[tb + m]) * 6 + c, nb[m]);
      });
    };
    issue_face(std::integral_constant<int, 0>{});

    // volume: reference derivatives of u at node i (sparse D_k rows), the
    // two physical derivatives this component feeds into the curl, and the
    // curl component of the lane's own output from its element's lanes
    double acc[NP];
    static_for<0, NP>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      double d[3] = {0.0, 0.0, 0.0};
      static_for<0, NP>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr bool any = ((T::DM[0][i] | T::DM[1][i] | T::DM[2][i]) >> j) & 1ull;
        if constexpr (any) {
          const double uj = u[j];
          static_for<0, 3>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            if constexpr ((T::DM[k][i] >> j) & 1ull) {
              constexpr int at = clane_nd<NP>(i, j, k);
              d[k] = fma(cs[at], uj, d[k]);
            }
          });
        }
      });
      const double d1 = jr[0][1] * d[0] + jr[1][1] * d[1] + jr[2][1] * d[2];  // along axis p+1
      const double d2 = jr[0][2] * d[0] + jr[1][2] * d[1] + jr[2][2] * d[2];  // along axis p+2
      // (curl H)_p = d/dx_p+1 H_p+2 - d/dx_p+2 H_p+1: lane H_p+2's d2 and lane H_p+1's d1
      acc[i] = sgn * (bperm_f64(srcO2, d2) - bperm_f64(srcO1, d1));
      next_batch(acc[i]);
    });

    // faces: jumps, flux, LIFT (entries after the volume's in the table)
    static_for<0, 4>([&](auto fc) {
      constexpr int f = decltype(fc)::value;
      const int cd = code[f];
      // g_f = {-row2, -row0, -row1, row0 + row1 + row2} of J^-1, rotated
      double g[3];
#pragma unroll
      for (int x = 0; x < 3; ++x)
        g[x] = f == 0 ? -jr[2][x] : f == 1 ? -jr[0][x] : f == 2 ? -jr[1][x] : jr[0][x] + jr[1][x] + jr[2][x];
      const double s2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
      const double hs = 0.5 * s2;
      double rs = __builtin_amdgcn_rsq(s2);
      rs = rs * fma(-hs * rs, rs, 1.5);
      rs = rs * fma(-hs * rs, rs, 1.5);
      const double n0 = g[0] * rs, n1 = g[1] * rs, n2 = g[2] * rs;
      const double cf = hs * rs;
      // own jump = wo * u - wn * neighbour: interior (1, 1), PEC E (2, 0),
      // PEC H (0, 0), port (1, 0); a = 1 on ports
      const double wo = cd >= 0 ? 1.0 : cd == kCodePec ? (fld ? 0.0 : 2.0) : 1.0;
      const double wn = cd >= 0 ? 1.0 : 0.0;
      const double a = cd >= 0 || cd == kCodePec ? io.alpha : 1.0;
      double fl[NFP];
      static_for<0, NFP>([&](auto mc) {
        constexpr int m = decltype(mc)::value;
        const double jo = fma(-wn, raw_f64(nb[m]), wo * u[T::FN[f][m]]);
        const double js1 = bperm_f64(srcS1, jo), js2 = bperm_f64(srcS2, jo);
        const double jo1 = bperm_f64(srcO1, jo), jo2 = bperm_f64(srcO2, jo);
        const double sd = n0 * jo + n1 * js1 + n2 * js2;
        fl[m] = cf * (a * (sd * n0 - jo) - sgn * (n1 * jo2 - n2 * jo1));
      });
      // the next face's traces load while this face is lifted
      if constexpr (f < 3) issue_face(std::integral_constant<int, f + 1>{});
      // LIFT rows of this face: node i takes its Nfp entries in one go
      static_for<0, NP>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        static_for<0, NFP>([&](auto mc) {
          constexpr int m = decltype(mc)::value;
          acc[i] = fma(cs[NDV + (f * NP + i) * NFP + m], fl[m], acc[i]);
        });
        next_batch(acc[i]);
      });
    });

    // materials, held source, stage update
    using PT = typename std::conditional<EPI == 2, float, double>::type;
    if (active) {
      static_for<0, NP>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int64_t idx = b0 + 6 * i;
        const double ui = u[i];
        double kv = m_inv * acc[i] - m_sig * ui;
        if (io.src_kind == RFHIP_DRIVE_POINT) {
          if (io.src_dof == idx) kv += io.g;
        } else if (io.src_kind == RFHIP_DRIVE_VECTOR) {
          kv += io.src_vec[idx] * io.g;
        }
        if constexpr (EPI == 0) {
          io.dy[idx] = kv;
        } else if constexpr (EPI == 1) {
          const double pn = io.sa * static_cast<const double *>(io.p)[idx] + io.dt * kv;
          static_cast<double *>(io.p)[idx] = pn;
          static_cast<double *>(io.yout)[idx] = ui + io.sb * pn;
        } else {
          const float pn = (float)(io.sa * (double)static_cast<const float *>(io.p)[idx] + io.dt * kv);
          static_cast<float *>(io.p)[idx] = pn;
          const double yn = ui + io.sb * (double)pn;
          const float h = (float)yn;
          static_cast<float *>(io.yout)[idx] = h;
          static_cast<uint16_t *>(io.yout_lo)[idx] = bf16_rne((float)(yn - (double)h));
        }
      });
    }
  }
}

template <int NP>
struct Nfp {
  static constexpr int v = NP == 4 ? 3 : NP == 10 ? 6 : NP == 20 ? 10 : 15;
};

// Blocks of `kernel` (launched with `nt` threads) resident on the whole
// device at once, computed once per (kernel, device, block size) and kept
// (the lookups run on every launch, from any host thread).
inline int64_t resident_blocks_nt(const void *kernel, int nt) {
  int dev = 0;
  RF_HIP(hipGetDevice(&dev));
  static std::mutex mu;
  static std::map<std::tuple<const void *, int, int>, int64_t> cache;
  const auto key = std::make_tuple(kernel, dev, nt);
  {
    std::lock_guard<std::mutex> lock(mu);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int cus = 0, per_cu = 0;
  RF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  RF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, nt, 0));
  const int64_t r = (int64_t)(cus > 0 ? cus : 1) * (per_cu > 0 ? per_cu : 1);
  std::lock_guard<std::mutex> lock(mu);
  cache.emplace(key, r);
  return r;
}

// Blocks of `kernel` resident on the whole device at once (grid-stride
// launches use this many).
inline int64_t resident_blocks(const void *kernel) { return resident_blocks_nt(kernel, kFusedWG); }

// Kernel variant per launch, each measured fastest for some order and
// register form on MI355X:
//   0     k_fused3, D and LIFT in LDS, grid-stride over element groups;
//   1, 2  k_fused3 with the next group's state read ahead into registers,
//         loops unrolled by 4 (1) or by 2 with the read-ahead issued after
//         the group's own reads (2);
//   3     k_fused8: FP64 MFMA volume with the curl in registers, persistent,
//         read-ahead (Np <= 20);
//   4, 5  k_fused9: the k_fused3 layout with the group's memory traffic
//         pipelined, loops unrolled by 2 / 4 nodes;
//   6, 7  k_fused11: k_fused9 with whole elements per wave and no
//         block-wide barriers in the group loop, unrolled by 2 / 4;
//   8     k_clane: one lane per (element, field component), D_k and LIFT
//         as wave-uniform scalar operands, the zeros of D_k skipped
//         (operators whose reference element matches ElemTab).
// Variants 3 to 7 read the neighbour-trace table (built for conforming
// meshes); variant 3 needs Np <= 20 (variant_available).

template <int NP, bool CIN, int EPI>
void fused_np(int variant, const StageArgs &a, hipStream_t s) {
  constexpr int F = Nfp<NP>::v;
  constexpr int EPB1 = kFusedWG / NP;
  const int64_t groups1 = (a.n_elem + EPB1 - 1) / EPB1;
  auto persistent = [&](const void *kf, int64_t groups) {
    const int64_t res = resident_blocks(kf);
    return groups < res ? groups : res;
  };
  if (variant == 3 && a.nbase && NP <= 20) {
    const int64_t res = resident_blocks_nt(reinterpret_cast<const void *>(&k_fused8<NP, F, CIN, EPI>), Mfma7Cfg<NP>::NT);
    const int64_t groups = (a.n_elem + 15) / 16;
    const int64_t grid = groups < res ? groups : res;
    if (grid > 0)
      hipLaunchKernelGGL((k_fused8<NP, F, CIN, EPI>), dim3((unsigned)grid), dim3(Mfma7Cfg<NP>::NT), 0, s, a);
  } else if ((variant == 4 || variant == 5) && a.nbase) {
    const void *kf = variant == 4 ? reinterpret_cast<const void *>(&k_fused9<NP, F, CIN, EPI, 1>)
                                  : reinterpret_cast<const void *>(&k_fused9<NP, F, CIN, EPI, 2>);
    const int64_t grid = persistent(kf, groups1);
    if (grid > 0) {
      if (variant == 4)
        hipLaunchKernelGGL((k_fused9<NP, F, CIN, EPI, 1>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
      else
        hipLaunchKernelGGL((k_fused9<NP, F, CIN, EPI, 2>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
    }
  } else if ((variant == 6 || variant == 7) && a.nbase) {
    const void *kf = variant == 6 ? reinterpret_cast<const void *>(&k_fused11<NP, F, CIN, EPI, 1>)
                                  : reinterpret_cast<const void *>(&k_fused11<NP, F, CIN, EPI, 2>);
    const int64_t epw = 64 / NP;
    const int64_t wgroups = (a.n_elem + epw - 1) / epw;
    const int64_t grid = persistent(kf, (wgroups + 3) / 4);  // blocks of 4 waves
    if (grid > 0) {
      if (variant == 6)
        hipLaunchKernelGGL((k_fused11<NP, F, CIN, EPI, 1>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
      else
        hipLaunchKernelGGL((k_fused11<NP, F, CIN, EPI, 2>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
    }
  } else if (variant == 8 && a.clane) {
    const void *kf = reinterpret_cast<const void *>(&k_clane<NP, CIN, EPI>);
    const int64_t wgroups = (a.n_elem + kClEpw - 1) / kClEpw;
    const int64_t grid = persistent(kf, (wgroups + 3) / 4);  // blocks of 4 waves
    if (grid > 0) hipLaunchKernelGGL((k_clane<NP, CIN, EPI>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
  } else if (variant == 2) {
    const void *kf = reinterpret_cast<const void *>(&k_fused3<NP, F, CIN, EPI, true, false, false, true, 2, true>);
    hipLaunchKernelGGL((k_fused3<NP, F, CIN, EPI, true, false, false, true, 2, true>),
                       dim3((unsigned)persistent(kf, groups1)), dim3(kFusedWG), 0, s, a);
  } else if (variant == 1) {
    const void *kf = reinterpret_cast<const void *>(&k_fused3<NP, F, CIN, EPI, true, false, false, true, 4>);
    hipLaunchKernelGGL((k_fused3<NP, F, CIN, EPI, true, false, false, true, 4>),
                       dim3((unsigned)persistent(kf, groups1)), dim3(kFusedWG), 0, s, a);
  } else {
    const void *kf = reinterpret_cast<const void *>(&k_fused3<NP, F, CIN, EPI, true, false>);
    hipLaunchKernelGGL((k_fused3<NP, F, CIN, EPI, true, false>), dim3((unsigned)persistent(kf, groups1)),
                       dim3(kFusedWG), 0, s, a);
  }
}

// The precision forms: k_fused11 with the volume on the matrix cores in
// VMF bf16 pieces, variant 0 / 1 = loops unrolled by 1 / 2.
template <int NP, bool CIN, int EPI, int VMF>
void vmf_np(int variant, const StageArgs &a, hipStream_t s) {
  constexpr int F = Nfp<NP>::v;
  const void *kf = variant == 0 ? reinterpret_cast<const void *>(&k_fused11<NP, F, CIN, EPI, 1, VMF>)
                                : reinterpret_cast<const void *>(&k_fused11<NP, F, CIN, EPI, 2, VMF>);
  const int64_t res = resident_blocks(kf);
  const int64_t epw = 64 / NP;
  const int64_t wgroups = (a.n_elem + epw - 1) / epw;
  const int64_t groups = (wgroups + 3) / 4;  // blocks of 4 waves
  const int64_t grid = groups < res ? groups : res;
  if (grid <= 0) return;
  if (variant == 0)
    hipLaunchKernelGGL((k_fused11<NP, F, CIN, EPI, 1, VMF>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
  else
    hipLaunchKernelGGL((k_fused11<NP, F, CIN, EPI, 2, VMF>), dim3((unsigned)grid), dim3(kFusedWG), 0, s, a);
}

template <int NP, bool CIN, int EPI>
void vmf_pieces(int vmf, int variant, const StageArgs &a, hipStream_t s) {
  if (vmf == 2)
    vmf_np<NP, CIN, EPI, 2>(variant, a, s);
  else if (vmf == 3)
    vmf_np<NP, CIN, EPI, 3>(variant, a, s);
  else
    vmf_np<NP, CIN, EPI, 4>(variant, a, s);
}

template <bool CIN, int EPI>
void vmf_dispatch(int np, int vmf, int variant, const StageArgs &a, hipStream_t s) {
  switch (np) {
    case 10: vmf_pieces<10, CIN, EPI>(vmf, variant, a, s); break;
    case 20: vmf_pieces<20, CIN, EPI>(vmf, variant, a, s); break;
    case 35: vmf_pieces<35, CIN, EPI>(vmf, variant, a, s); break;
    default: throw Error{RFHIP_EUNSUPPORTED, "the precision forms run orders 2 to 4"};
  }
  RF_HIP(hipGetLastError());
}

template <bool CIN, int EPI>
void fused_dispatch(int np, int variant, const StageArgs &a, hipStream_t s) {
  switch (np) {
    case 4: fused_np<4, CIN, EPI>(variant, a, s); break;
    case 10: fused_np<10, CIN, EPI>(variant, a, s); break;
    case 20: fused_np<20, CIN, EPI>(variant, a, s); break;
    case 35: fused_np<35, CIN, EPI>(variant, a, s); break;
    default: throw Error{RFHIP_EUNSUPPORTED, "order not compiled (Np must be 4, 10, 20 or 35)"};
  }
  RF_HIP(hipGetLastError());
}

}  // namespace

// Port source (rhs.rs:1157-1199): for port face (e, f) and face node m with
// the mode's (e_t, h_t), the incident-field flux jumps [E] = -e_t,
// [H] = -h_t give, with the characteristic flux,
//   s_E = n x h_t - n x (n x e_t),  s_H = -n x e_t - n x (n x h_t),
// lifted to every node i of e with 0.5*fscale*LIFT[i][f*Nfp + m] and scaled
// by the element's 1/eps, 1/mu. One thread per (element with port faces,
// node) sums over that element's port faces: no two threads write the same
// entry. FUSED reads J^-1 (face geometry as in the stage kernels), the
// material table and the column-major LIFT of the fused forms; otherwise the
// plain forms' face normals, fscale, materials and row-major LIFT (type T).
template <bool FUSED, typename T>
__global__ void __launch_bounds__(256) k_port_source(StageArgs sa, DevOp dv, int64_t n_elems,
                                                     const int64_t *__restrict__ elem,
                                                     const int64_t *__restrict__ off,
                                                     const int32_t *__restrict__ lf,
                                                     const double *__restrict__ prof, double *__restrict__ b) {
  const int np = dv.np, nfp = dv.nfp;
  const int64_t n = n_elems * np;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int64_t u = t / np;
    const int i = (int)(t - u * np);
    const int64_t e = elem[u];
    double ae[3] = {0, 0, 0}, ah[3] = {0, 0, 0};
    for (int64_t q = off[u]; q < off[u + 1]; ++q) {
      const int f = lf[q];
      double nv[3], cf;
      if constexpr (FUSED) {
        face_geom(sa.jinv + 9 * e, f, nv, cf);
      } else {
        const T *nn = static_cast<const T *>(dv.normal) + (4 * e + f) * 3;
        nv[0] = (double)nn[0];
        nv[1] = (double)nn[1];
        nv[2] = (double)nn[2];
        cf = 0.5 * (double)static_cast<const T *>(dv.fscale)[4 * e + f];
      }
      for (int m = 0; m < nfp; ++m) {
        const double *pr = prof + (q * nfp + m) * 6;
        const double et[3] = {pr[0], pr[1], pr[2]}, ht[3] = {pr[3], pr[4], pr[5]};
        double nxe[3], nxh[3], nne[3], nnh[3];
        cross3(nv, et, nxe);
        cross3(nv, ht, nxh);
        cross3(nv, nxe, nne);
        cross3(nv, nxh, nnh);
        double w;
        if constexpr (FUSED)
          w = cf * sa.lift[(size_t)(f * nfp + m) * np + i];
        else
          w = cf * (double)static_cast<const T *>(dv.lift)[(size_t)i * 4 * nfp + f * nfp + m];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          ae[c] += w * (nxh[c] - nne[c]);
          ah[c] += w * (-nxe[c] - nnh[c]);
        }
      }
    }
    double ie[3], im[3];
    if constexpr (FUSED) {
      const double *row = sa.mat_tab + 12 * sa.mat[e];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ie[c] = row[c];
        im[c] = row[3 + c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ie[c] = (double)static_cast<const T *>(dv.inv_eps)[3 * e + c];
        im[c] = (double)static_cast<const T *>(dv.inv_mu)[3 * e + c];
      }
    }
    const int64_t base = (e * np + i) * 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      b[base + c] = ie[c] * ae[c];
      b[base + 3 + c] = im[c] * ah[c];
    }
  }
}

void launch_port_source(const rfhip_op *op, int64_t n_elems, const int64_t *elem, const int64_t *off,
                        const int32_t *lf, const double *prof, double *b, hipStream_t s) {
  if (n_elems == 0) return;
  int64_t blocks = (n_elems * op->np + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  const unsigned g = (unsigned)blocks;
  if (form_fused(op->form))
    hipLaunchKernelGGL((k_port_source<true, double>), dim3(g), dim3(256), 0, s, op->sa, op->dev, n_elems, elem, off,
                       lf, prof, b);
  else if (op->form == RFHIP_FORM_PLAIN_F64)
    hipLaunchKernelGGL((k_port_source<false, double>), dim3(g), dim3(256), 0, s, op->sa, op->dev, n_elems, elem,
                       off, lf, prof, b);
  else
    hipLaunchKernelGGL((k_port_source<false, float>), dim3(g), dim3(256), 0, s, op->sa, op->dev, n_elems, elem,
                       off, lf, prof, b);
  RF_HIP(hipGetLastError());
}

// The held port drive of one LSERK4 stage with its waveform sample on the
// device, g = gseq[*gidx], over the DOFs of the elements with port faces
// (the only nonzero entries of the port source b, rhs.rs:1157-1199):
//   mode 0 (plain forms, before the stage kernel):  k += b g;
//   mode 1 (fused FP64, after the stage kernel):    p += dt b g, y += sb dt b g;
//   mode 2 (fused compressed registers): the same on p (f32) and on y
//          (f32 hi + bf16 lo, recompressed as the stage kernel does).
// The stage kernel has run without the drive, so together they give
// p = a p + dt (A y + b g), y = y + sb p.
template <int MODE, class T>
__global__ void __launch_bounds__(256) k_port_stage(int64_t n_elems, const int64_t *__restrict__ elem, int ndof_e,
                                                    const double *__restrict__ b, const double *__restrict__ gseq,
                                                    const int64_t *__restrict__ gidx, T *__restrict__ kp,
                                                    void *__restrict__ yh, uint16_t *__restrict__ yl, double dt,
                                                    double sb) {
  const double g = gseq[gidx[0]];
  const int64_t total = n_elems * ndof_e;
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < total; q += (int64_t)gridDim.x * 256) {
    const int64_t u = q / ndof_e;
    const int64_t d = elem[u] * ndof_e + (q - u * ndof_e);
    const double bg = b[d] * g;
    if constexpr (MODE == 0) {
      kp[d] = (T)((double)kp[d] + bg);
    } else if constexpr (MODE == 1) {
      double *y = static_cast<double *>(yh);
      const double pn = (double)kp[d] + dt * bg;
      kp[d] = (T)pn;
      y[d] += sb * dt * bg;
    } else {
      float *y = static_cast<float *>(yh);
      kp[d] = (T)((double)kp[d] + dt * bg);
      const double yn = (double)y[d] + (double)bf16_to_f(yl[d]) + sb * dt * bg;
      const float h = (float)yn;
      y[d] = h;
      yl[d] = bf16_rne((float)(yn - (double)h));
    }
  }
}

__global__ void k_step_advance(int64_t *gidx) {
  if (threadIdx.x == 0) gidx[0] += 1;
}

void launch_port_stage(const rfhip_op *op, bool before, const StageArgs &a, void *k, double dt, double sb,
                       hipStream_t s) {
  const int64_t ne = op->ps_count;
  if (ne == 0) return;
  const int ndof_e = 6 * op->np;
  int64_t blocks = (ne * ndof_e + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  const dim3 g((unsigned)blocks), b(256);
  const int64_t *el = op->ps_elem.as<int64_t>();
  const double *src = op->src.as<double>(), *gs = op->gseq.as<double>();
  const int64_t *gi = op->gctr.as<int64_t>();
  if (before) {  // plain forms: k += b g before the stage kernel
    if (op->form == RFHIP_FORM_PLAIN_F32)
      hipLaunchKernelGGL((k_port_stage<0, float>), g, b, 0, s, ne, el, ndof_e, src, gs, gi, static_cast<float *>(k),
                         nullptr, nullptr, dt, sb);
    else
      hipLaunchKernelGGL((k_port_stage<0, double>), g, b, 0, s, ne, el, ndof_e, src, gs, gi,
                         static_cast<double *>(k), nullptr, nullptr, dt, sb);
  } else if (!form_compressed(op->form)) {
    hipLaunchKernelGGL((k_port_stage<1, double>), g, b, 0, s, ne, el, ndof_e, src, gs, gi, static_cast<double *>(a.p),
                       a.yout, nullptr, dt, sb);
  } else {
    hipLaunchKernelGGL((k_port_stage<2, float>), g, b, 0, s, ne, el, ndof_e, src, gs, gi, static_cast<float *>(a.p),
                       a.yout, static_cast<uint16_t *>(a.yout_lo), dt, sb);
  }
  RF_HIP(hipGetLastError());
}

void launch_step_advance(const rfhip_op *op, hipStream_t s) {
  hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, op->gctr.as<int64_t>());
  RF_HIP(hipGetLastError());
}

bool variant_available(int order, int form, int variant, const StageArgs &a) {
  const int np = (order + 1) * (order + 2) * (order + 3) / 6;
  if (form_vmf(form)) return variant >= 0 && variant < kStageVariantsBf16 && a.nbase != nullptr && order >= 2 && order <= 4;
  if (variant < 0 || variant >= kStageVariants) return false;
  if (variant == 8) return a.clane != nullptr;
  if (variant == 3) return a.nbase != nullptr && np <= 20;
  if (variant >= 4) return a.nbase != nullptr;
  return true;
}

namespace {
template <int NP>
const void *kcl_kernel(int variant, bool r64) {
  constexpr int F = Nfp<NP>::v;
  if (r64)
    return variant == 6 ? reinterpret_cast<const void *>(&k_fused11<NP, F, false, 4, 1>)
                        : reinterpret_cast<const void *>(&k_fused11<NP, F, false, 4, 2>);
  return variant == 6 ? reinterpret_cast<const void *>(&k_fused11<NP, F, true, 3, 1>)
                      : reinterpret_cast<const void *>(&k_fused11<NP, F, true, 3, 2>);
}

// Blocks of the KCL stage launch: one wave per group of 64/Np elements,
// four waves per block, at most the resident blocks (persistent grid).
template <int NP>
int64_t kcl_grid(const StageArgs &a, bool r64) {
  const int64_t epw = 64 / NP;
  const int64_t wgroups = (a.n_elem + epw - 1) / epw;
  const int64_t groups = (wgroups + 3) / 4;
  const int64_t res = resident_blocks(kcl_kernel<NP>(a.variant, r64));
  return groups < res ? groups : res;
}

template <int NP>
void kcl_np(const StageArgs &a, bool r64, hipStream_t s) {
  constexpr int F = Nfp<NP>::v;
  const int64_t grid = kcl_grid<NP>(a, r64);
  if (grid <= 0) return;
  const dim3 g((unsigned)grid), b(kFusedWG);
  if (r64) {
    if (a.variant == 6)
      hipLaunchKernelGGL((k_fused11<NP, F, false, 4, 1>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_fused11<NP, F, false, 4, 2>), g, b, 0, s, a);
  } else {
    if (a.variant == 6)
      hipLaunchKernelGGL((k_fused11<NP, F, true, 3, 1>), g, b, 0, s, a);
    else
      hipLaunchKernelGGL((k_fused11<NP, F, true, 3, 2>), g, b, 0, s, a);
  }
}

template <int NP>
bool clane_pack_np(const int *face_nodes, const double *const D[3], const double *lift, std::vector<double> &out) {
  using T = ElemTab<NP>;
  constexpr int NFP = T::NFP;
  for (int f = 0; f < 4; ++f)
    for (int m = 0; m < NFP; ++m)
      if (face_nodes[f * NFP + m] != T::FN[f][m]) return false;
  for (int k = 0; k < 3; ++k) {
    double mx = 0.0;
    for (int q = 0; q < NP * NP; ++q) mx = std::max(mx, std::fabs(D[k][q]));
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j)
        if ((std::fabs(D[k][i * NP + j]) > 1e-9 * mx) != (((T::DM[k][i] >> j) & 1ull) != 0)) return false;
  }
  out.clear();
  for (int i = 0; i < NP; ++i)
    for (int j = 0; j < NP; ++j)
      for (int k = 0; k < 3; ++k)
        if ((T::DM[k][i] >> j) & 1ull) out.push_back(D[k][i * NP + j]);
  for (int f = 0; f < 4; ++f)
    for (int i = 0; i < NP; ++i)
      for (int m = 0; m < NFP; ++m) out.push_back(lift[(size_t)i * 4 * NFP + f * NFP + m]);
  return true;
}
}  // namespace

bool kcl_stage_fusable(int form, const StageArgs &a) { return !form_vmf(form) && a.nbase != nullptr; }

int64_t kcl_stage_slots(int order, const StageArgs &a, bool r64) {
  switch ((order + 1) * (order + 2) * (order + 3) / 6) {
    case 4: return kcl_grid<4>(a, r64) * (kFusedWG / 64);
    case 10: return kcl_grid<10>(a, r64) * (kFusedWG / 64);
    case 20: return kcl_grid<20>(a, r64) * (kFusedWG / 64);
    case 35: return kcl_grid<35>(a, r64) * (kFusedWG / 64);
    default: throw Error{RFHIP_EUNSUPPORTED, "the fused stage kernels run orders 1 to 4"};
  }
}

void launch_kcl_stage(int order, const StageArgs &a, bool r64, hipStream_t s) {
  RF_REQUIRE(a.nbase != nullptr, "the fused KCL stage needs the neighbour-trace table");
  RF_REQUIRE(!a.klast || a.kpart != nullptr, "the last KCL stage needs the error partials");
  switch ((order + 1) * (order + 2) * (order + 3) / 6) {
    case 4: kcl_np<4>(a, r64, s); break;
    case 10: kcl_np<10>(a, r64, s); break;
    case 20: kcl_np<20>(a, r64, s); break;
    case 35: kcl_np<35>(a, r64, s); break;
    default: throw Error{RFHIP_EUNSUPPORTED, "the fused stage kernels run orders 1 to 4"};
  }
  RF_HIP(hipGetLastError());
}

bool clane_pack(int np, const int *face_nodes, const double *const D[3], const double *lift, std::vector<double> &out) {
  switch (np) {
    case 4: return clane_pack_np<4>(face_nodes, D, lift, out);
    case 10: return clane_pack_np<10>(face_nodes, D, lift, out);
    case 20: return clane_pack_np<20>(face_nodes, D, lift, out);
    case 35: return clane_pack_np<35>(face_nodes, D, lift, out);
    default: return false;
  }
}

void launch_stage(int order, int form, int epi, const StageArgs &a, hipStream_t s) {
  const int np = (order + 1) * (order + 2) * (order + 3) / 6;
  const int variant = a.variant;
  if (const int vmf = form_vmf(form)) {
    if (!a.nbase || !a.afragb) throw Error{RFHIP_EUNSUPPORTED, "the precision forms need a conforming mesh"};
    if (epi == 0)
      vmf_dispatch<false, 0>(np, vmf, variant, a, s);
    else
      vmf_dispatch<true, 2>(np, vmf, variant, a, s);
    return;
  }
  if (epi == 0)
    fused_dispatch<false, 0>(np, variant, a, s);
  else if (form_compressed(form))
    fused_dispatch<true, 2>(np, variant, a, s);
  else
    fused_dispatch<false, 1>(np, variant, a, s);
}

}  // namespace rfhip
