// Sparse state-space export on the device: MaxwellOperator::assemble_sparse
// (rhs.rs:1621-1730) into a CsrMatrix (rhs.rs:1575-1591).
//
// The reference finds row block e of A by probing every column of e's
// stencil (e and its distinct face neighbours, ascending, rhs.rs:1597-1610,
// then e's own polarisation columns) with a unit vector through
// apply_element and keeping the outputs that are not 0.0 (rhs.rs:409-421).
// Here a block takes element e and lists, in LDS, the columns of its rows that
// can be nonzero, ascending (all of e's own, and of each neighbour the nodes
// behind the shared face). A wave takes one row (e, node i, component co) at
// a time, each lane evaluating one candidate column: what apply_element
// computes for that unit column, with the same operations in the same
// order, so every stored value is the reference's bit for bit and a value is
// stored exactly when the reference stores it. A ballot compacts the
// nonzeros, so consecutive lanes write a row's entries to consecutive
// addresses:
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

// Persistent: each block loads D and LIFT into LDS once and then takes
// elements e = blockIdx.x, blockIdx.x + gridDim.x, ...
template <int NP, int NFP, bool FILL>
__global__ void __launch_bounds__(256) k_sparse_rows(SparseArgs a, int64_t K) {
  constexpr int ST = 6 * NP, NF4 = 4 * NFP, MAXC = ST + 4 * NFP * 6;
  __shared__ double Dl[3][NP][NP];  // D_k[i][n]
  __shared__ double Ll[NP][NF4];    // LIFT
  __shared__ double S[2][4][6][6];  // [side][face][co][ci]
  __shared__ double coef[4];
  __shared__ int64_t sten[5];
  __shared__ int nsten;
  __shared__ int fnb[4];                // neighbour of an interior face, else -1
  __shared__ signed char own_m[NP][4];  // m with face_nodes[f][m] == n, else -1
  __shared__ signed char nbr_m[4][NP];  // m whose neighbour trace is node n, else -1
  // candidate columns of every row of the element, ascending: all of the
  // element's own, and of each neighbour the nodes behind a shared face
  __shared__ int64_t cand_col[MAXC];
  __shared__ unsigned cand_key[MAXC];  // own << 16 | n << 3 | ci
  __shared__ int cand_fm[MAXC];        // m at face f in byte f (0xff: none)
  __shared__ int cand_off[6];
  const int t = threadIdx.x;
  for (int x = t; x < NP * NP; x += blockDim.x) {
    Dl[0][x / NP][x % NP] = a.dr[x];
    Dl[1][x / NP][x % NP] = a.ds[x];
    Dl[2][x / NP][x % NP] = a.dt[x];
  }
  for (int x = t; x < NP * NF4; x += blockDim.x) Ll[x / NF4][x % NF4] = a.lift[x];
  const int lane = t & 63, wave = t >> 6, nwave = blockDim.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;

  for (int64_t e = blockIdx.x; e < K; e += gridDim.x) {
    __syncthreads();  // the previous element's tables are no longer read
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
    // Of a neighbour c only the nodes behind a face shared with it can give
    // a nonzero (rhs.rs:1403-1418): the face-node bytes of such a node, 0xff
    // on the other faces.
    auto nbr_bytes = [&](int64_t c, int n) {
      unsigned v = 0;
      for (int f = 0; f < 4; ++f) {
        const int m = (fnb[f] == c) ? nbr_m[f][n] : -1;
        v |= (unsigned)(m >= 0 ? m : 0xff) << (8 * f);
      }
      return v;
    };
    if (t == 0) {  // offsets of the stencil slots in the candidate list
      int off = 0;
      for (int b = 0; b < nsten; ++b) {
        cand_off[b] = off;
        if (sten[b] == e) {
          off += ST;
        } else {
          for (int n = 0; n < NP; ++n) off += nbr_bytes(sten[b], n) != 0xffffffffu ? 6 : 0;
        }
      }
      cand_off[nsten] = off;
    }
    __syncthreads();
    for (int b = 0; b < nsten; ++b) {
      const int64_t c = sten[b];
      if (c == e) {
        for (int x = t; x < ST; x += blockDim.x) {
          const int n = x / 6;
          unsigned fm = 0;
          for (int f = 0; f < 4; ++f) fm |= (unsigned)(own_m[n][f] >= 0 ? own_m[n][f] : 0xff) << (8 * f);
          cand_col[cand_off[b] + x] = c * ST + x;
          cand_key[cand_off[b] + x] = 1u << 16 | (unsigned)n << 3 | (unsigned)(x % 6);
          cand_fm[cand_off[b] + x] = (int)fm;
        }
      } else if (t == 0) {
        int k = cand_off[b];
        for (int n = 0; n < NP; ++n) {
          const unsigned fm = nbr_bytes(c, n);
          if (fm == 0xffffffffu) continue;
          for (int ci = 0; ci < 6; ++ci, ++k) {
            cand_col[k] = c * ST + n * 6 + ci;
            cand_key[k] = (unsigned)n << 3 | (unsigned)ci;
            cand_fm[k] = (int)fm;
          }
        }
      }
    }
    __syncthreads();
    const int ncand = cand_off[nsten];

    const int64_t slot = a.disp_slot ? a.disp_slot[e] : -1;
    // One wave per row at a time: lane l evaluates candidate c0 + l, the
    // nonzeros are compacted with a ballot, so a row's entries are written
    // by consecutive lanes to consecutive addresses.
    for (int r = wave; r < ST; r += nwave) {
      const int i = r / 6, co = r % 6, oc = co % 3;
      const int64_t row = e * ST + r;
      const bool erow = co < 3;
      // curl component oc (rhs.rs:80-84) of a unit in component cc is
      // +d/dx_php for cc == phm and -d/dx_phm for cc == php, else zero
      const int php = oc == 2 ? 0 : oc + 1, phm = oc == 0 ? 2 : oc - 1;
      const double *jv = a.jinv + 9 * e;  // jacobian_inv[k][p] at 3k + p
      const double jp0 = jv[php], jp1 = jv[3 + php], jp2 = jv[6 + php];
      const double jm0 = jv[phm], jm1 = jv[3 + phm], jm2 = jv[6 + phm];
      const double mat = erow ? a.inv_eps[3 * e + oc] : a.inv_mu[3 * e + oc];
      const double sig = erow ? a.sig_eps[3 * e + oc] : a.sig_mu[3 * e + oc];
      const bool debye = erow && slot >= 0;
      const double pa = debye ? a.disp_coef[3 * slot] : 0.0;
      const double pg = debye ? a.disp_coef[3 * slot + 1] : 0.0;
      const double iei = debye ? a.disp_coef[3 * slot + 2] : 0.0;
      // rhs.rs:1443-1472 for the row's component, own value dg of the unit
      auto finish = [&](double de, double dg, double pv) {
        double v = mat * de - sig * dg;
        if (debye) v = v - iei * (pa * pv + pg * dg);
        return v;
      };
      int64_t o = FILL ? a.row_ptr[row] : 0;
      const int64_t end = FILL ? a.row_ptr[row + 1] : 0;
      for (int c0 = 0; c0 < ncand; c0 += 64) {
        const int k = c0 + lane, kk = k < ncand ? k : ncand - 1;
        const int64_t col = cand_col[kk];
        const unsigned key = cand_key[kk], fm = (unsigned)cand_fm[kk];
        const bool own = key >> 16;
        const int n = (key >> 3) & 511, ci = key & 7, cc = ci < 3 ? ci : ci - 3;
        // volume (own columns only); pd = sum_k J^-1[k][ph] D_k[i][n]
        const bool plus = cc == phm;
        const double pd = (plus ? jp0 : jm0) * Dl[0][i][n] + (plus ? jp1 : jm1) * Dl[1][i][n] +
                          (plus ? jp2 : jm2) * Dl[2][i][n];
        double de = plus ? pd : -pd;
        if (!erow) de = -de;
        if (!(own && erow != (ci < 3) && (plus || cc == php))) de = 0.0;
        // flux, face by face; a face the column does not reach adds w = 0
        // times its jump term, exactly zero, as the reference's loop does
        const int side = own ? 0 : 1;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
          const unsigned m = (fm >> (8 * f)) & 0xffu;
          const bool on = m != 0xffu;
          const double w = coef[f] * Ll[i][f * NFP + (on ? m : 0u)];
          de = de + (on ? w : 0.0) * S[side][f][co][ci];
        }
        const double v = finish(de, (own && n == i && ci == co) ? 1.0 : 0.0, 0.0);
      const bool nz = k < ncand && v != 0.0;
        const uint64_t mask = __ballot(nz);
        if constexpr (FILL) {
          const int64_t pos = o + __popcll(mask & below);
          if (nz && pos < end) {
            a.col_idx[pos] = col;
            a.values[pos] = v;
          }
        }
        o += __popcll(mask);
      }
      // the element's own polarisation columns: only P(i, co) reaches this row
      if (debye) {
        const double v = finish(0.0, 0.0, 1.0);
        if (v != 0.0) {
          if constexpr (FILL) {
            if (lane == 0 && o < end) {
              a.col_idx[o] = a.eh + slot * 3 * NP + i * 3 + co;
              a.values[o] = v;
            }
          }
          ++o;
        }
      }
      if constexpr (!FILL) {
        if (lane == 0) a.row_ptr[row + 1] = o;
      }
    }
  }
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
  const void *kern = fill ? reinterpret_cast<const void *>(&k_sparse_rows<NP, NFP, true>)
                          : reinterpret_cast<const void *>(&k_sparse_rows<NP, NFP, false>);
  int dev = 0, cus = 0, per_cu = 0;
  RF_HIP(hipGetDevice(&dev));
  RF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  RF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
  const int64_t grid = std::min<int64_t>(K, (int64_t)std::max(cus, 1) * std::max(per_cu, 1));
  if (fill)
    hipLaunchKernelGGL((k_sparse_rows<NP, NFP, true>), dim3((unsigned)grid), dim3(256), 0, s, a, K);
  else
    hipLaunchKernelGGL((k_sparse_rows<NP, NFP, false>), dim3((unsigned)grid), dim3(256), 0, s, a, K);
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
