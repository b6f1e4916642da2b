// Sparse state-space export on the device: MaxwellOperator::assemble_sparse
// (rhs.rs:1621-1730) into a CsrMatrix (rhs.rs:1575-1591).
//
// The reference finds row block e of A by probing every column of e's
// stencil (e and its distinct face neighbours, ascending, rhs.rs:1597-1610,
// then e's own polarisation columns) with a unit vector through
// apply_element and keeping the outputs that are not 0.0 (rhs.rs:409-421).
// Here one thread owns one row (e, node i, component co) and walks the same
// columns in ascending order, evaluating what apply_element computes for
// each unit column with the same operations in the same order, so every
// stored value is the reference's bit for bit and a value is stored exactly
// when the reference stores it:
//  * volume (rhs.rs:39-85): a unit in component cc at node n gives the
//    reference derivatives rd_k = D_k[i][n] (one nonzero term of the sum),
//    pd_p = J^-1[0][p] rd_0 + J^-1[1][p] rd_1 + J^-1[2][p] rd_2, and a curl
//    component is +-pd_p or zero; dH = -curl E;
//  * flux (rhs.rs:1370-1440): a unit column enters the jump at face node
//    (f, m) only when it is the own trace there (face_nodes[f][m] == n) or
//    the neighbour trace behind it (face_nodes[nlf][perm[m]] == n on an
//    interior face). The jump vector, its cross products and
//    (-n x [H] + a n x n x [E]), (n x [E] + a n x n x [H]) depend on
//    (face, component, side) only, so a block tabulates them once per face
//    and a row adds w = (0.5 fscale) LIFT[i][f Nfp + m] times the table
//    entry, face by face as the reference's loop does;
//  * materials and the Debye current (rhs.rs:1443-1472) with the unit's
//    own value (1 on the diagonal, 0 elsewhere).
// Adding a zero term or multiplying by 1 changes no value, and no sum the
// reference forms is reassociated or fused, so the file is compiled without
// FMA contraction. The appended P rows carry g and a whatever their values
// (rhs.rs:1704-1727).
//
// Pass 1 counts each row's entries into row_ptr[row + 1], an inclusive scan
// turns the counts into offsets, pass 2 writes (column, value) pairs.
#include <algorithm>

#include "rfhip_impl.hpp"

#pragma clang fp contract(off)

namespace rfhip {
namespace {

struct SparseArgs {
  const double *dr, *ds, *dt;   // Np*Np row-major
  const double *lift;           // Np*4Nfp row-major
  const int *face_nodes;        // 4*Nfp
  const double *jinv;           // 9K: jacobian_inv[k][p] at 9e + 3k + p
  const double *inv_eps, *inv_mu, *sig_eps, *sig_mu;  // 3K
  const double *normal, *fscale;                      // 12K, 4K
  const int *nbr, *nbr_lf, *port, *perm;              // 4K, 4K, 4K, 4K*Nfp
  const int64_t *disp_slot;     // K: Debye slot of the element, -1 none
  const double *disp_coef;      // slot -> (a, g, 1/eps_inf)
  const int64_t *disp_elem;     // slot -> element
  int64_t n_disp;
  double alpha;
  int64_t eh;                   // 6 Np K, the first P row / column
  int64_t *row_ptr;             // n + 1
  int64_t *col_idx;             // nnz (pass 2)
  double *values;               // nnz (pass 2)
};

__device__ __forceinline__ void cross3(const double *a, const double *b, double *o) {
  // crates/rapidfem-core/src/geom.rs:32-34
  const double t0 = a[1] * b[2] - a[2] * b[1];
  const double t1 = a[2] * b[0] - a[0] * b[2];
  const double t2 = a[0] * b[1] - a[1] * b[0];
  o[0] = t0;
  o[1] = t1;
  o[2] = t2;
}

// (-n x [H] + a n x n x [E])_c for c < 3 and (n x [E] + a n x n x [H])_c-3
// for c >= 3 when the jump comes from a unit in component ci on the own
// side (side 0) or the neighbour side (side 1) of face q (rhs.rs:1380-1440).
__device__ void jump_row(const SparseArgs &a, int64_t q, int side, int ci, double *s6) {
  const double nrm[3] = {a.normal[3 * q], a.normal[3 * q + 1], a.normal[3 * q + 2]};
  const bool port = a.port[q] >= 0, pec = !port && a.nbr[q] < 0;
  double u[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  u[ci] = 1.0;
  double je[3], jh[3];
  if (side == 1) {  // E- - E+ with the unit behind the face: 0 - u
    for (int k = 0; k < 3; ++k) {
      je[k] = 0.0 - u[k];
      jh[k] = 0.0 - u[3 + k];
    }
  } else if (pec) {  // PEC ghost: [E] = 2 E_t, [H] = 2 H_n
    const double edn = nrm[0] * u[0] + nrm[1] * u[1] + nrm[2] * u[2];
    const double hdn = nrm[0] * u[3] + nrm[1] * u[4] + nrm[2] * u[5];
    for (int k = 0; k < 3; ++k) {
      je[k] = 2.0 * (u[k] - edn * nrm[k]);
      jh[k] = 2.0 * hdn * nrm[k];
    }
  } else {  // port (zero incident field) or interior with a zero neighbour
    for (int k = 0; k < 3; ++k) {
      je[k] = u[k];
      jh[k] = u[3 + k];
    }
  }
  double nxjh[3], nxje[3], t[3], pe[3], ph[3];
  cross3(nrm, jh, nxjh);
  cross3(nrm, je, nxje);
  const double al = port ? 1.0 : a.alpha;
  cross3(nrm, je, t);
  cross3(nrm, t, pe);
  cross3(nrm, jh, t);
  cross3(nrm, t, ph);
  for (int c = 0; c < 3; ++c) {
    s6[c] = -nxjh[c] + al * pe[c];
    s6[3 + c] = nxje[c] + al * ph[c];
  }
}

template <int NP, int NFP, bool FILL>
__global__ void __launch_bounds__(256) k_sparse_rows(SparseArgs a) {
  constexpr int ST = 6 * NP, NF4 = 4 * NFP;
  __shared__ double S[2][4][6][6];  // [side][face][co][ci]
  __shared__ double coef[4];
  __shared__ int64_t sten[5];
  __shared__ int nsten;
  __shared__ int fnb[4];             // neighbour of an interior face, else -1
  __shared__ signed char own_m[NP][4];  // m with face_nodes[f][m] == n, else -1
  __shared__ signed char nbr_m[4][NP];  // m whose neighbour trace is node n, else -1
  const int64_t e = blockIdx.x;
  const int t = threadIdx.x;

  for (int x = t; x < 4 * NP; x += blockDim.x) {
    own_m[x / 4][x % 4] = -1;
    nbr_m[x / NP][x % NP] = -1;
  }
  if (t < 4) {
    const int64_t q = 4 * e + t;
    coef[t] = 0.5 * a.fscale[q];
    fnb[t] = (a.port[q] < 0 && a.nbr[q] >= 0) ? a.nbr[q] : -1;
  }
  if (t == 0) {  // rhs.rs:1597-1610 element_stencil
    int n = 1;
    sten[0] = e;
    for (int f = 0; f < 4; ++f) {
      const int64_t nb = a.nbr[4 * e + f];
      bool seen = false;
      for (int k = 0; k < n; ++k) seen |= sten[k] == nb;
      if (nb >= 0 && !seen) sten[n++] = nb;
    }
    for (int x = 1; x < n; ++x)
      for (int y = x; y > 0 && sten[y - 1] > sten[y]; --y) {
        const int64_t tmp = sten[y];
        sten[y] = sten[y - 1];
        sten[y - 1] = tmp;
      }
    nsten = n;
  }
  if (t < 2 * 4 * 6) {
    const int side = t / 24, f = (t / 6) % 4, ci = t % 6;
    double s6[6];
    jump_row(a, 4 * e + f, side, ci, s6);
    for (int co = 0; co < 6; ++co) S[side][f][co][ci] = s6[co];
  }
  __syncthreads();
  for (int x = t; x < NF4; x += blockDim.x) {
    const int f = x / NFP, m = x % NFP;
    own_m[a.face_nodes[x]][f] = (signed char)m;
    const int64_t q = 4 * e + f;
    if (a.port[q] < 0 && a.nbr[q] >= 0) {
      const int vj = a.face_nodes[a.nbr_lf[q] * NFP + a.perm[q * NFP + m]];
      nbr_m[f][vj] = (signed char)m;
    }
  }
  __syncthreads();
  if (t >= ST) return;

  const int i = t / 6, co = t % 6, oc = co % 3;
  const int64_t row = e * ST + t;
  const double J[3][3] = {{a.jinv[9 * e + 0], a.jinv[9 * e + 1], a.jinv[9 * e + 2]},
                          {a.jinv[9 * e + 3], a.jinv[9 * e + 4], a.jinv[9 * e + 5]},
                          {a.jinv[9 * e + 6], a.jinv[9 * e + 7], a.jinv[9 * e + 8]}};
  const bool erow = co < 3;
  const double mat = erow ? a.inv_eps[3 * e + oc] : a.inv_mu[3 * e + oc];
  const double sig = erow ? a.sig_eps[3 * e + oc] : a.sig_mu[3 * e + oc];
  const int64_t slot = a.disp_slot ? a.disp_slot[e] : -1;
  const bool debye = erow && slot >= 0;
  const double pa = debye ? a.disp_coef[3 * slot] : 0.0;
  const double pg = debye ? a.disp_coef[3 * slot + 1] : 0.0;
  const double iei = debye ? a.disp_coef[3 * slot + 2] : 0.0;
  const double *lrow = a.lift + (size_t)i * NF4;

  int64_t o = FILL ? a.row_ptr[row] : 0;
  auto emit = [&](int64_t col, double v) {
    if (v != 0.0) {
      if constexpr (FILL) {
        a.col_idx[o] = col;
        a.values[o] = v;
      }
      ++o;
    }
  };
  // rhs.rs:1443-1472 for the row's component, own value dg of the unit
  auto finish = [&](double de, double dg, double pv) {
    double v = mat * de - sig * dg;
    if (debye) v = v - iei * (pa * pv + pg * dg);
    return v;
  };

  const int ns = nsten;
  for (int b = 0; b < ns; ++b) {
    const int64_t c = sten[b], cbase = c * ST;
    if (c == e) {
      for (int n = 0; n < NP; ++n) {
        const double r0 = a.dr[i * NP + n], r1 = a.ds[i * NP + n], r2 = a.dt[i * NP + n];
        for (int ci = 0; ci < 6; ++ci) {
          // volume: curl component oc of a unit in component cc
          double de = 0.0;
          if (erow != (ci < 3)) {
            const int cc = ci % 3;
            int ph = -1;
            bool neg = false;
            if (oc == 0) {
              if (cc == 2) ph = 1;
              else if (cc == 1) ph = 2, neg = true;
            } else if (oc == 1) {
              if (cc == 0) ph = 2;
              else if (cc == 2) ph = 0, neg = true;
            } else {
              if (cc == 1) ph = 0;
              else if (cc == 0) ph = 1, neg = true;
            }
            if (ph >= 0) {
              const double pd = J[0][ph] * r0 + J[1][ph] * r1 + J[2][ph] * r2;
              de = neg ? -pd : pd;
              if (!erow) de = -de;
            }
          }
          for (int f = 0; f < 4; ++f) {
            const int m = own_m[n][f];
            if (m >= 0) {
              const double w = coef[f] * lrow[f * NFP + m];
              de = de + w * S[0][f][co][ci];
            }
          }
          emit(cbase + n * 6 + ci, finish(de, (n == i && ci == co) ? 1.0 : 0.0, 0.0));
        }
      }
    } else {
      for (int n = 0; n < NP; ++n) {
        bool touch = false;
        for (int f = 0; f < 4; ++f) touch |= fnb[f] == c && nbr_m[f][n] >= 0;
        if (!touch) continue;  // every value of this node is exactly zero
        for (int ci = 0; ci < 6; ++ci) {
          double de = 0.0;
          for (int f = 0; f < 4; ++f) {
            const int m = nbr_m[f][n];
            if (fnb[f] == c && m >= 0) {
              const double w = coef[f] * lrow[f * NFP + m];
              de = de + w * S[1][f][co][ci];
            }
          }
          emit(cbase + n * 6 + ci, finish(de, 0.0, 0.0));
        }
      }
    }
  }
  // the element's own polarisation columns: only P(i, co) reaches this row
  if (debye) emit(a.eh + slot * 3 * NP + i * 3 + co, finish(0.0, 0.0, 1.0));
  if constexpr (!FILL) a.row_ptr[row + 1] = o;
}

// rhs.rs:1704-1727: P row (slot, node, c) holds g at E(node, c) of its
// element, then a on the diagonal.
template <bool FILL>
__global__ void __launch_bounds__(256) k_sparse_prows(SparseArgs a, int np) {
  const int64_t n = a.n_disp * 3 * np;
  for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
    const int64_t row = a.eh + r;
    if constexpr (!FILL) {
      a.row_ptr[row + 1] = 2;
    } else {
      const int64_t slot = r / (3 * np), jl = r - slot * 3 * np;
      const int64_t o = a.row_ptr[row];
      a.col_idx[o] = a.disp_elem[slot] * 6 * np + (jl / 3) * 6 + jl % 3;
      a.values[o] = a.disp_coef[3 * slot + 1];
      a.col_idx[o + 1] = row;
      a.values[o + 1] = a.disp_coef[3 * slot];
    }
  }
}

// ---- inclusive scan of int64 counts ---------------------------------------

constexpr int kScanT = 256, kScanI = 8, kScanTile = kScanT * kScanI;

__global__ void __launch_bounds__(kScanT) k_scan_tiles(int64_t *x, int64_t n, int64_t *tile_sum) {
  __shared__ int64_t sh[kScanT];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanI;
  int64_t v[kScanI];
  int64_t s = 0;
  for (int q = 0; q < kScanI; ++q) {
    v[q] = base + q < n ? x[base + q] : 0;
    s += v[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kScanT; off <<= 1) {
    const int64_t add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = sh[threadIdx.x] - s;
  for (int q = 0; q < kScanI; ++q) {
    run += v[q];
    if (base + q < n) x[base + q] = run;
  }
  if (threadIdx.x == kScanT - 1) tile_sum[blockIdx.x] = sh[kScanT - 1];
}

__global__ void __launch_bounds__(kScanT) k_scan_add(int64_t *x, int64_t n, const int64_t *tile_incl) {
  if (blockIdx.x == 0) return;
  const int64_t add = tile_incl[blockIdx.x - 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  for (int q = threadIdx.x; q < kScanTile; q += kScanT)
    if (base + q < n) x[base + q] += add;
}

void scan_inclusive(int64_t *x, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  DevBuf sums;
  sums.alloc(sizeof(int64_t) * (size_t)tiles);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(kScanT), 0, s, x, n, sums.as<int64_t>());
  RF_HIP(hipGetLastError());
  if (tiles > 1) {
    scan_inclusive(sums.as<int64_t>(), tiles, s);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(kScanT), 0, s, x, n, sums.as<int64_t>());
    RF_HIP(hipGetLastError());
  }
  RF_HIP(hipStreamSynchronize(s));  // sums is freed on return
}

template <int NP, int NFP>
void launch_rows(const SparseArgs &a, int64_t K, bool fill, hipStream_t s) {
  const unsigned threads = (unsigned)((6 * NP + 63) / 64 * 64);
  if (fill)
    hipLaunchKernelGGL((k_sparse_rows<NP, NFP, true>), dim3((unsigned)K), dim3(threads), 0, s, a);
  else
    hipLaunchKernelGGL((k_sparse_rows<NP, NFP, false>), dim3((unsigned)K), dim3(threads), 0, s, a);
  RF_HIP(hipGetLastError());
}

}  // namespace

void assemble_sparse(rfhip_csr *csr, const rfhip_op_desc &d, int64_t n_disp, const int64_t *disp_elem,
                     const double *disp_coef) {
  hipStream_t s = csr->stream;
  const int64_t K = d.n_elem;
  const int np = d.np, nfp = d.nfp;
  DevBuf dr, ds, dt, lift, fn, jinv, ie, im, se, sm, nrm, fsc, nbr, nlf, port, perm, dslot, dcoef, delem;
  upload(dr, d.diff_r, (size_t)np * np, s);
  upload(ds, d.diff_s, (size_t)np * np, s);
  upload(dt, d.diff_t, (size_t)np * np, s);
  upload(lift, d.lift, (size_t)np * 4 * nfp, s);
  upload(fn, d.face_nodes, 4 * (size_t)nfp, s);
  upload(jinv, d.jinv, 9 * (size_t)K, s);
  upload(ie, d.inv_eps, 3 * (size_t)K, s);
  upload(im, d.inv_mu, 3 * (size_t)K, s);
  upload(se, d.sigma_eps, 3 * (size_t)K, s);
  upload(sm, d.sigma_mu, 3 * (size_t)K, s);
  upload(nrm, d.face_normal, 12 * (size_t)K, s);
  upload(fsc, d.face_fscale, 4 * (size_t)K, s);
  upload(nbr, d.face_neighbor, 4 * (size_t)K, s);
  upload(nlf, d.face_nbr_local, 4 * (size_t)K, s);
  upload(port, d.face_port, 4 * (size_t)K, s);
  upload(perm, d.face_perm, 4 * (size_t)K * nfp, s);
  std::vector<int64_t> slot;
  if (n_disp > 0) {
    slot.assign((size_t)K, -1);
    for (int64_t q = 0; q < n_disp; ++q) slot[(size_t)disp_elem[q]] = q;
    upload(dslot, slot.data(), slot.size(), s);
    upload(dcoef, disp_coef, 3 * (size_t)n_disp, s);
    upload(delem, disp_elem, (size_t)n_disp, s);
  }
  SparseArgs a{};
  a.dr = dr.as<double>();
  a.ds = ds.as<double>();
  a.dt = dt.as<double>();
  a.lift = lift.as<double>();
  a.face_nodes = fn.as<int>();
  a.jinv = jinv.as<double>();
  a.inv_eps = ie.as<double>();
  a.inv_mu = im.as<double>();
  a.sig_eps = se.as<double>();
  a.sig_mu = sm.as<double>();
  a.normal = nrm.as<double>();
  a.fscale = fsc.as<double>();
  a.nbr = nbr.as<int>();
  a.nbr_lf = nlf.as<int>();
  a.port = port.as<int>();
  a.perm = perm.as<int>();
  a.disp_slot = n_disp > 0 ? dslot.as<int64_t>() : nullptr;
  a.disp_coef = n_disp > 0 ? dcoef.as<double>() : nullptr;
  a.disp_elem = n_disp > 0 ? delem.as<int64_t>() : nullptr;
  a.n_disp = n_disp;
  a.alpha = d.flux_alpha;
  a.eh = 6 * (int64_t)np * K;
  const int64_t n = a.eh + 3 * (int64_t)np * n_disp;
  csr->n = n;
  csr->row_ptr.alloc(sizeof(int64_t) * (size_t)(n + 1));
  a.row_ptr = csr->row_ptr.as<int64_t>();
  RF_HIP(hipMemsetAsync(a.row_ptr, 0, sizeof(int64_t), s));

  auto rows = [&](bool fill) {
    switch (np) {
      case 4: launch_rows<4, 3>(a, K, fill, s); break;
      case 10: launch_rows<10, 6>(a, K, fill, s); break;
      case 20: launch_rows<20, 10>(a, K, fill, s); break;
      case 35: launch_rows<35, 15>(a, K, fill, s); break;
      default: throw Error{RFHIP_EUNSUPPORTED, "orders 1..4 are compiled (Np 4, 10, 20, 35)"};
    }
    if (n_disp > 0) {
      const int64_t pr = n_disp * 3 * np;
      const unsigned grid = (unsigned)std::min<int64_t>((pr + 255) / 256, 8192);
      if (fill)
        hipLaunchKernelGGL(k_sparse_prows<true>, dim3(grid), dim3(256), 0, s, a, np);
      else
        hipLaunchKernelGGL(k_sparse_prows<false>, dim3(grid), dim3(256), 0, s, a, np);
      RF_HIP(hipGetLastError());
    }
  };
  hipEvent_t ev[4];
  for (auto &x : ev) RF_HIP(hipEventCreate(&x));
  try {
    RF_HIP(hipEventRecord(ev[0], s));
    rows(false);
    RF_HIP(hipEventRecord(ev[1], s));
    scan_inclusive(a.row_ptr + 1, n, s);
    int64_t nnz = 0;
    RF_HIP(hipMemcpyAsync(&nnz, a.row_ptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    csr->nnz = nnz;
    csr->col_idx.alloc(sizeof(int64_t) * (size_t)nnz);
    csr->values.alloc(sizeof(double) * (size_t)nnz);
    a.col_idx = csr->col_idx.as<int64_t>();
    a.values = csr->values.as<double>();
    RF_HIP(hipEventRecord(ev[2], s));
    rows(true);
    RF_HIP(hipEventRecord(ev[3], s));
    RF_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    RF_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    csr->count_ms = ms;
    RF_HIP(hipEventElapsedTime(&ms, ev[2], ev[3]));
    csr->fill_ms = ms;
  } catch (...) {
    (void)hipStreamSynchronize(s);
    for (auto &x : ev) (void)hipEventDestroy(x);
    throw;
  }
  for (auto &x : ev) RF_HIP(hipEventDestroy(x));
}

}  // namespace rfhip
