// C ABI of the MI355X DGTD backend (include/rfhip.h): the drop-in for
// rapidfem_td::gpu::{GpuContext, GpuOperator} (crates/rapidfem-td/src/gpu/).
// Host orchestration only; kernels live in rfhip_plain.hip, rfhip_fused.hip
// and rfhip_vec.hip.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <type_traits>

#include "rfhip_impl.hpp"
#include "rfhip_krylov.hpp"

namespace rfhip {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

// LSERK4 (constants.rs:106-122), KCL RK4(3)5[2R+]C (constants.rs:151-178).
constexpr double kLsA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                            -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
constexpr double kLsB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                            1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                            2277821191437.0 / 14882151754819.0};
constexpr double kKclA[4] = {970286171893.0 / 4311952581923.0, 6584761158862.0 / 12103376702013.0,
                             2251764453980.0 / 15575788980749.0, 26877169314380.0 / 34165994151039.0};
constexpr double kKclB[5] = {1153189308089.0 / 22510343858157.0, 1772645290293.0 / 4653164025191.0,
                             -1672844663538.0 / 4480602732383.0, 2114624349019.0 / 3568978502595.0,
                             5198255086312.0 / 14908931495163.0};
constexpr double kKclBhat[5] = {1016888040809.0 / 7410784769900.0, 11231460423587.0 / 58533540763752.0,
                                -1563879915014.0 / 6823010717585.0, 606302364029.0 / 971179775848.0,
                                1097981568119.0 / 3980877426909.0};
constexpr double kKclErrFloor = 1e-12;   // constants.rs KCL_ERR_FLOOR
constexpr double kKclNonfinite = 10.0;   // constants.rs KCL_NONFINITE_ERR
constexpr double kArnoldiBreakdown = 1e-12;  // constants.rs:62
constexpr int kMaxKrylov = 64;

template <class F>
int guarded(F &&f) {
  try {
    f();
    return RFHIP_OK;
  } catch (const Error &e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return RFHIP_ENOMEM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return RFHIP_EINVAL;
  }
}

bool is_fused(int form) { return form_fused(form); }
int n_stage_variants(int form) { return form_vmf(form) ? kStageVariantsBf16 : kStageVariants; }
bool f64_regs(int form) { return form != RFHIP_FORM_PLAIN_F32; }

// ---- reference-element conventions used to check and encode faces -------

struct Ijk {
  int i, j, k;
};
std::vector<Ijk> node_ijk(int p) {  // dg_basis.rs:270-280
  std::vector<Ijk> v;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= p - i; ++j)
      for (int k = 0; k <= p - i - j; ++k) v.push_back({i, j, k});
  return v;
}
int order_of_np(int np) {
  for (int p = 1; p <= 8; ++p)
    if ((p + 1) * (p + 2) * (p + 3) / 6 == np) return p;
  return -1;
}
constexpr int kFaceVerts[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};  // core mesh.rs:45-50
int bary(const Ijk &q, int a, int p) {
  switch (a) {
    case 0: return p - q.i - q.j - q.k;
    case 1: return q.i;
    case 2: return q.j;
    default: return q.k;
  }
}
// The six slot correspondences s -> sigma[s] between the two copies of a
// shared triangle.
constexpr int kSigma[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

template <class T>
std::vector<T> cvt(const double *src, size_t n) {
  std::vector<T> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = (T)src[i];
  return v;
}

void sync(rfhip_op *op) { RF_HIP(hipStreamSynchronize(op->stream)); }

void check_drive(const rfhip_op *op, int kind, int64_t dof, const double *src, int64_t n_g, int64_t need, const void *g) {
  RF_REQUIRE(kind >= RFHIP_DRIVE_FREE && kind <= RFHIP_DRIVE_VECTOR, "unknown drive kind");
  if (kind == RFHIP_DRIVE_POINT) RF_REQUIRE(dof >= 0 && dof < op->n_dof, "source_dof out of range");
  if (kind == RFHIP_DRIVE_VECTOR)
    RF_REQUIRE(src != nullptr || op->src_resident,
               "a vector drive needs a source pattern (or a port source assembled on the device)");
  if (kind != RFHIP_DRIVE_FREE) {
    RF_REQUIRE(g != nullptr, "a driven run needs waveform samples");
    RF_REQUIRE(n_g == need, "wrong number of waveform samples");
  }
}

// ---- operator construction --------------------------------------------

void build_plain(rfhip_op *op, const rfhip_op_desc &d) {
  const bool f64 = op->form == RFHIP_FORM_PLAIN_F64;
  const size_t K = (size_t)d.n_elem, np = d.np, nfp = d.nfp;
  hipStream_t s = op->stream;
  auto up = [&](DevBuf &b, const double *src, size_t n) {
    if (f64) {
      upload(b, src, n, s);
    } else {
      auto v = cvt<float>(src, n);
      upload(b, v.data(), n, s);
      sync(op);  // keep v alive until the copy is done
    }
  };
  up(op->dr, d.diff_r, np * np);
  up(op->ds, d.diff_s, np * np);
  up(op->dt, d.diff_t, np * np);
  up(op->lift, d.lift, np * 4 * nfp);
  up(op->jinv, d.jinv, 9 * K);
  up(op->inv_eps, d.inv_eps, 3 * K);
  up(op->inv_mu, d.inv_mu, 3 * K);
  up(op->sig_eps, d.sigma_eps, 3 * K);
  up(op->sig_mu, d.sigma_mu, 3 * K);
  up(op->normal, d.face_normal, 12 * K);
  up(op->fscale, d.face_fscale, 4 * K);
  upload(op->face_nodes, d.face_nodes, 4 * nfp, s);
  upload(op->nbr, d.face_neighbor, 4 * K, s);
  upload(op->nbr_lf, d.face_nbr_local, 4 * K, s);
  upload(op->port, d.face_port, 4 * K, s);
  upload(op->perm, d.face_perm, 4 * K * nfp, s);
  const size_t w = f64 ? 8 : 4;
  op->y[0].alloc(w * op->n_dof);
  op->k.alloc(w * op->n_dof);
  op->p.alloc(w * op->n_dof);
  DevOp &dv = op->dev;
  dv.dr = op->dr.p;
  dv.ds = op->ds.p;
  dv.dt = op->dt.p;
  dv.lift = op->lift.p;
  dv.face_nodes = op->face_nodes.as<int>();
  dv.jinv = op->jinv.p;
  dv.inv_eps = op->inv_eps.p;
  dv.inv_mu = op->inv_mu.p;
  dv.sig_eps = op->sig_eps.p;
  dv.sig_mu = op->sig_mu.p;
  dv.normal = op->normal.p;
  dv.fscale = op->fscale.p;
  dv.nbr = op->nbr.as<int>();
  dv.nbr_lf = op->nbr_lf.as<int>();
  dv.port = op->port.as<int>();
  dv.perm = op->perm.as<int>();
}
// Stage-kernel variant an operator starts with: the fastest per order and
// register form in a sweep of every variant over LSERK4 steps on MI355X
// (pulse boxes of 1.3M / 393k / 324k / 96k tets at N = 1..4; N = 2 and the
// FP64 registers at N = 3 again on 768k / 324k tets once k_fused11 fitted
// those cases to 2 waves per SIMD, N = 4 on the 998k-tet cavity once
// k_fused9's LDS fitted two blocks per CU). autotune re-picks on the
// operator's own mesh.
int default_variant(int order, int form) {
  if (form_vmf(form)) return 1;
  const bool c = form == RFHIP_FORM_FUSED_F64C;
  switch (order) {
    case 1: return c ? 7 : 5;
    case 2: return c ? 7 : 6;
    case 3: return 7;
    case 4: return 4;
    default: return 0;
  }
}

void build_fused(rfhip_op *op, const rfhip_op_desc &d) {
  const int64_t K = d.n_elem;
  const int np = d.np, nfp = d.nfp, nf4 = 4 * nfp;
  const int p = order_of_np(np);
  hipStream_t s = op->stream;
  // Node convention check: the fused kernels encode neighbour nodes through
  // the equispaced node order of ReferenceElement::new.
  const auto ijk = node_ijk(p);
  {
    int c[4] = {0, 0, 0, 0};
    std::vector<int32_t> fn(4 * nfp, -1);
    for (int a = 0; a < np; ++a) {
      const Ijk &q = ijk[a];
      const bool on[4] = {q.k == 0, q.i == 0, q.j == 0, q.i + q.j + q.k == p};
      for (int f = 0; f < 4; ++f)
        if (on[f]) fn[f * nfp + c[f]++] = a;
    }
    for (int q = 0; q < 4 * nfp; ++q)
      if (fn[q] != d.face_nodes[q])
        throw Error{RFHIP_EUNSUPPORTED, "fused forms need the equispaced node order of ReferenceElement::new"};
  }
  // Face node lookup by barycentric weights on the local face slots.
  std::vector<int> lookup(4 * (p + 1) * (p + 1), -1);
  for (int f = 0; f < 4; ++f)
    for (int m = 0; m < nfp; ++m) {
      const Ijk &q = ijk[d.face_nodes[f * nfp + m]];
      lookup[(f * (p + 1) + bary(q, kFaceVerts[f][1], p)) * (p + 1) + bary(q, kFaceVerts[f][2], p)] = m;
    }
  // Candidate permutations perm[(f, nlf, sigma)][m] and the neighbour-node table.
  std::vector<int32_t> cand(4 * 4 * 6 * nfp), nbr_node(4 * 4 * 6 * nfp);
  for (int f = 0; f < 4; ++f)
    for (int g = 0; g < 4; ++g)
      for (int o = 0; o < 6; ++o)
        for (int m = 0; m < nfp; ++m) {
          const Ijk &q = ijk[d.face_nodes[f * nfp + m]];
          int w[3];
          for (int sl = 0; sl < 3; ++sl) w[kSigma[o][sl]] = bary(q, kFaceVerts[f][sl], p);
          const int m2 = lookup[(g * (p + 1) + w[1]) * (p + 1) + w[2]];
          const size_t at = ((size_t)(f * 4 + g) * 6 + o) * nfp + m;
          cand[at] = m2;
          nbr_node[at] = m2 < 0 ? 0 : d.face_nodes[g * nfp + m2];
        }
  // Per-face codes, and a check that normal*fscale is the J^-1 gradient.
  std::vector<int32_t> fcode(4 * K), mat(K);
  double worst = 0.0;
  std::vector<double> tab;
  std::map<std::array<double, 12>, int> seen;
  for (int64_t e = 0; e < K; ++e) {
    const double *ji = d.jinv + 9 * e;
    for (int f = 0; f < 4; ++f) {
      const int64_t q = 4 * e + f;
      double g[3];
      for (int x = 0; x < 3; ++x) {
        const double r0 = ji[0 * 3 + x], r1 = ji[1 * 3 + x], r2 = ji[2 * 3 + x];
        g[x] = f == 0 ? -r2 : f == 1 ? -r0 : f == 2 ? -r1 : (r0 + r1 + r2);
      }
      const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      for (int x = 0; x < 3; ++x)
        worst = std::max(worst, std::fabs(g[x] - d.face_fscale[q] * d.face_normal[3 * q + x]) / gn);
      const int nb = d.face_neighbor[q];
      if (nb < 0) {
        fcode[q] = d.face_port[q] >= 0 ? -2 - d.face_port[q] : kCodePec;
        continue;
      }
      RF_REQUIRE(nb < K && d.face_nbr_local[q] >= 0 && d.face_nbr_local[q] < 4, "bad neighbour data");
      if (nb >= (1 << 26)) throw Error{RFHIP_EUNSUPPORTED, "fused forms encode neighbours in 26 bits"};
      const int g2 = d.face_nbr_local[q];
      int found = -1;
      for (int o = 0; o < 6 && found < 0; ++o) {
        bool ok = true;
        for (int m = 0; m < nfp && ok; ++m)
          ok = cand[((size_t)(f * 4 + g2) * 6 + o) * nfp + m] == d.face_perm[q * nfp + m];
        if (ok) found = o;
      }
      if (found < 0)
        throw Error{RFHIP_EUNSUPPORTED, "face permutation is not a triangle symmetry (non-conforming face)"};
      fcode[q] = (int32_t)((nb << 5) | (g2 << 3) | found);
    }
    std::array<double, 12> row;
    for (int c = 0; c < 3; ++c) {
      row[c] = d.inv_eps[3 * e + c];
      row[3 + c] = d.inv_mu[3 * e + c];
      row[6 + c] = d.sigma_eps[3 * e + c];
      row[9 + c] = d.sigma_mu[3 * e + c];
    }
    auto it = seen.find(row);
    if (it == seen.end()) {
      it = seen.emplace(row, (int)(tab.size() / 12)).first;
      tab.insert(tab.end(), row.begin(), row.end());
    }
    mat[e] = it->second;
  }
  if (worst > 1e-9)
    throw Error{RFHIP_EUNSUPPORTED, "face normal*fscale is not the J^-1 face gradient (rel " + std::to_string(worst) + ")"};
  // Neighbour trace base per (element, face node) for k_fused5: the AoS
  // index of the neighbour node's first field value, or the face code on a
  // boundary face (negative). Needs n_dof < 2^31; otherwise left empty and
  // that kernel variant is not offered.
  std::vector<int32_t> nbase;
  if (6 * (int64_t)np * K < ((int64_t)1 << 31)) {
    nbase.resize((size_t)K * nf4);
    for (int64_t e = 0; e < K; ++e)
      for (int f = 0; f < 4; ++f) {
        const int32_t code = fcode[4 * e + f];
        for (int m = 0; m < nfp; ++m) {
          int32_t v = code;
          if (code >= 0) {
            const int64_t nb = code >> 5;
            const int nlf = (code >> 3) & 3, o = code & 7;
            const int vj = nbr_node[((size_t)(f * 4 + nlf) * 6 + o) * nfp + m];
            v = (int32_t)((nb * np + vj) * 6);
          }
          nbase[(size_t)e * nf4 + f * nfp + m] = v;
        }
      }
  }
  // Transposed reference matrices.
  std::vector<double> dT(3 * np * np), liftc((size_t)nf4 * np);
  const double *D[3] = {d.diff_r, d.diff_s, d.diff_t};
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i) dT[(k * np + j) * np + i] = D[k][i * np + j];
  for (int fm = 0; fm < nf4; ++fm)
    for (int i = 0; i < np; ++i) liftc[(size_t)fm * np + i] = d.lift[(size_t)i * nf4 + fm];
  upload(op->dT, dT.data(), dT.size(), s);
  upload(op->liftc, liftc.data(), liftc.size(), s);
  // A-operand fragments of LIFT (Np x 4Nfp) for the FP64 MFMA 16x16x4 lift
  // of k_fused8: lfrag[(mt*KSL + ks)*64 + lane] holds
  // LIFT[16 mt + lane%16][4 ks + lane/16], zero padded.
  const int ltiles = (np + 15) / 16, lsteps = (nf4 + 3) / 4;
  std::vector<double> lfrag((size_t)ltiles * lsteps * 64, 0.0);
  for (int mt = 0; mt < ltiles; ++mt)
    for (int ks = 0; ks < lsteps; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int i = 16 * mt + l % 16, fm = 4 * ks + l / 16;
        if (i < np && fm < nf4) lfrag[((size_t)mt * lsteps + ks) * 64 + l] = d.lift[(size_t)i * nf4 + fm];
      }
  // ... and the node-interleaved [Dr; Ds; Dt] of its volume: wave w owns nodes 4w..4w+3 and
  // its A rows are (k, node) = (row / 4, 4w + row % 4), row 4k + q, k < 3
  // (row tile 16 = 3 derivatives x 4 nodes + a zero row per node), so one
  // lane receives all three reference derivatives of its node.
  const int nwaves7 = (np + 3) / 4, ksteps = (np + 3) / 4;
  std::vector<double> afrag7((size_t)nwaves7 * ksteps * 64, 0.0);
  for (int w = 0; w < nwaves7; ++w)
    for (int ks = 0; ks < ksteps; ++ks)
      for (int l = 0; l < 64; ++l) {
        const int row = l % 16, k = row / 4, node = 4 * w + row % 4, j = 4 * ks + l / 16;
        if (k < 3 && node < np && j < np) afrag7[((size_t)w * ksteps + ks) * 64 + l] = D[k][node * np + j];
      }
  upload(op->lfrag, lfrag.data(), lfrag.size(), s);
  upload(op->afrag7, afrag7.data(), afrag7.size(), s);
  // Split-BF16 A fragments of [Dr; Ds; Dt] for the precision forms' volume on
  // v_mfma_f32_16x16x32_bf16: afragb[((p*MT + mt)*KS + ks)*64 + lane] holds
  // 8 bf16, piece p of D_k[i][j] for row 16 mt + lane%16 = k*Np + i and
  // j = 32 ks + 8 (lane/16) + t, zero outside; with t_0 = D_k[i][j], piece
  // q = bf16(t_q) rounded to nearest even and t_{q+1} = t_q - piece q in FP64.
  std::vector<uint16_t> fragb;
  if (const int vmf = form_vmf(op->form)) {
    const int mtv = (3 * np + 15) / 16, ksv = (np + 31) / 32;
    fragb.assign((size_t)vmf * mtv * ksv * 64 * 8, 0);
    auto bf16_bits = [](float x) {
      uint32_t u;
      std::memcpy(&u, &x, 4);
      u += 0x7FFFu + ((u >> 16) & 1u);
      return (uint16_t)(u >> 16);
    };
    auto bf16_value = [](uint16_t b) {
      const uint32_t u = (uint32_t)b << 16;
      float f;
      std::memcpy(&f, &u, 4);
      return f;
    };
    for (int mt = 0; mt < mtv; ++mt)
      for (int ks = 0; ks < ksv; ++ks)
        for (int l = 0; l < 64; ++l)
          for (int t = 0; t < 8; ++t) {
            const int row = 16 * mt + l % 16, j = 32 * ks + 8 * (l / 16) + t;
            if (row >= 3 * np || j >= np) continue;
            double v = D[row / np][(row % np) * np + j];
            for (int p = 0; p < vmf; ++p) {
              const uint16_t b = bf16_bits((float)v);
              fragb[(((size_t)(p * mtv + mt) * ksv + ks) * 64 + l) * 8 + t] = b;
              v -= (double)bf16_value(b);
            }
          }
    upload(op->afragb, fragb.data(), fragb.size(), s);
  }
  // Component-lane stage kernel table (variant 8), when the reference
  // element is the compiled equispaced one.
  std::vector<double> clane;
  const bool has_clane = clane_pack(np, d.face_nodes, D, d.lift, clane);
  if (has_clane) upload(op->clane, clane.data(), clane.size(), s);
  upload(op->jinv64, d.jinv, 9 * (size_t)K, s);
  upload(op->fcode, fcode.data(), fcode.size(), s);
  upload(op->mat, mat.data(), mat.size(), s);
  upload(op->mat_tab, tab.data(), tab.size(), s);
  upload(op->nbr_node, nbr_node.data(), nbr_node.size(), s);
  if (!nbase.empty()) upload(op->nbase, nbase.data(), nbase.size(), s);
  upload(op->face_nodes, d.face_nodes, 4 * (size_t)nfp, s);
  sync(op);
  const size_t n = op->n_vec;
  if (op->form == RFHIP_FORM_FUSED_F64) {
    op->y[0].alloc(8 * n);
    op->y[1].alloc(8 * n);
    op->p.alloc(8 * n);
  } else {
    op->y[0].alloc(4 * n);
    op->y[1].alloc(4 * n);
    op->ylo[0].alloc(2 * n);
    op->ylo[1].alloc(2 * n);
    op->p.alloc(4 * n);
  }
  op->k.alloc(8 * n);
  StageArgs &a = op->sa;
  a = StageArgs{};
  a.n_elem = K;
  a.alpha = d.flux_alpha;
  a.dT = op->dT.as<double>();
  a.lift = op->liftc.as<double>();
  a.lfrag = op->lfrag.as<double>();
  a.afragb = form_vmf(op->form) ? op->afragb.p : nullptr;
  a.afrag7 = op->afrag7.as<double>();
  a.jinv = op->jinv64.as<double>();
  a.fcode = op->fcode.as<int>();
  a.mat = op->mat.as<int>();
  a.mat_tab = op->mat_tab.as<double>();
  a.face_nodes = op->face_nodes.as<int>();
  a.nbr_node = op->nbr_node.as<int>();
  a.nbase = nbase.empty() ? nullptr : op->nbase.as<int>();
  a.clane = has_clane ? op->clane.as<double>() : nullptr;
  a.variant = default_variant(op->order, op->form);
}

// ---- vectors at the boundary ---------------------------------------------
//
// "Layout" vectors are f64 AoS vectors on the device. "Register" vectors
// are what the KCL registers and the apply kernels work on: f64 AoS, except
// f32 AoS for PLAIN_F32.

int64_t nvec(const rfhip_op *op) { return op->n_vec; }

void h2d_layout_f64(rfhip_op *op, const double *host, double *dev) {
  copy_h2d(dev, host, 8 * op->n_dof, op->stream);
}
void d2h_layout_f64(rfhip_op *op, const double *dev, double *host) {
  copy_d2h(host, dev, 8 * op->n_dof, op->stream);
}
void h2d_reg_f64(rfhip_op *op, const double *host, void *dev) {
  if (op->form != RFHIP_FORM_PLAIN_F32) {
    h2d_layout_f64(op, host, static_cast<double *>(dev));
    return;
  }
  op->aos.alloc(8 * op->n_dof);
  copy_h2d(op->aos.p, host, 8 * op->n_dof, op->stream);
  launch_cast_d2f(op->aos.as<double>(), static_cast<float *>(dev), op->n_dof, op->stream);
}
void d2h_reg_f64(rfhip_op *op, const void *dev, double *host) {
  if (op->form != RFHIP_FORM_PLAIN_F32) {
    d2h_layout_f64(op, static_cast<const double *>(dev), host);
    return;
  }
  op->aos.alloc(8 * op->n_dof);
  launch_cast_f2d(static_cast<const float *>(dev), op->aos.as<double>(), op->n_dof, op->stream);
  copy_d2h(host, op->aos.p, 8 * op->n_dof, op->stream);
}
std::vector<double> widen(const float *x, int64_t n) { return std::vector<double>(x, x + n); }
void narrow(const std::vector<double> &x, float *out) {
  for (size_t i = 0; i < x.size(); ++i) out[i] = (float)x[i];
}

// ---- stepping -----------------------------------------------------------

void fill_source(rfhip_op *op, StageArgs &a, int kind, int64_t dof, double g) {
  a.src_kind = kind;
  a.g = g;
  a.src_vec = op->src.as<double>();
  a.src_dof = dof;
}

// k = A*x for the plain forms: the element kernel for [E,H], then the Debye
// rows (rhs.rs:1262-1325).
void plain_apply_full(rfhip_op *op, const void *x, void *k, hipStream_t s) {
  launch_plain_apply(op, x, k, s);
  launch_debye_apply(op, x, k, op->form == RFHIP_FORM_PLAIN_F32, RFHIP_DRIVE_FREE, 0, nullptr, 0.0, s);
}

// k = A*x (+ b*g) on register vectors.
void rhs_into(rfhip_op *op, const void *x, void *k, int kind, int64_t dof, double g) {
  hipStream_t s = op->stream;
  if (is_fused(op->form)) {
    StageArgs a = op->sa;
    a.yin = x;
    a.dy = static_cast<double *>(k);
    fill_source(op, a, kind, dof, g);
    launch_stage(op->order, op->form, 0, a, s);
    // the fused kernel drives [E,H]; the Debye kernel drives the P rows
    launch_debye_apply(op, x, k, false, kind, dof, a.src_vec, g, s);
    return;
  }
  plain_apply_full(op, x, k, s);
  if (kind == RFHIP_DRIVE_POINT)
    launch_add_point(op, k, dof, g, s);
  else if (kind == RFHIP_DRIVE_VECTOR)
    launch_add_vec(op, k, op->src.p, g, s);
}

// One LSERK4 step on the resident state; stage_ev (optional) gets an event
// pair around each dominant-kernel launch. dev_wave: the drive is the
// device-assembled port source held at the waveform sample
// op->gseq[*op->gctr] (kind/dof/g unused), and the step advances *op->gctr.
void lserk_step(rfhip_op *op, double h, int kind, int64_t dof, double g, hipEvent_t *stage_ev,
                bool dev_wave = false) {
  hipStream_t s = op->stream;
  if (dev_wave) kind = RFHIP_DRIVE_FREE;
  for (int st = 0; st < 5; ++st) {
    if (stage_ev) RF_HIP(hipEventRecord(stage_ev[2 * st], s));
    if (is_fused(op->form)) {
      StageArgs a = op->sa;
      a.yin = op->y[op->cur].p;
      a.yin_lo = op->ylo[op->cur].p;
      a.yout = op->y[1 - op->cur].p;
      a.yout_lo = op->ylo[1 - op->cur].p;
      a.p = op->p.p;
      a.sa = kLsA[st];
      a.sb = kLsB[st];
      a.dt = h;
      fill_source(op, a, kind, dof, g);
      launch_stage(op->order, op->form, 1, a, s);
      if (stage_ev) RF_HIP(hipEventRecord(stage_ev[2 * st + 1], s));
      launch_debye_stage(op, a, s);
      if (dev_wave) launch_port_stage(op, false, a, nullptr, h, kLsB[st], s);
      op->cur = 1 - op->cur;
    } else {
      plain_apply_full(op, op->y[0].p, op->k.p, s);
      if (stage_ev) RF_HIP(hipEventRecord(stage_ev[2 * st + 1], s));
      if (kind == RFHIP_DRIVE_POINT)
        launch_add_point(op, op->k.p, dof, g, s);
      else if (kind == RFHIP_DRIVE_VECTOR)
        launch_add_vec(op, op->k.p, op->src.p, g, s);
      if (dev_wave) launch_port_stage(op, true, op->sa, op->k.p, h, kLsB[st], s);
      launch_plain_stage(op, op->p.p, op->k.p, op->y[0].p, kLsA[st], kLsB[st], h, s);
    }
  }
  if (dev_wave) launch_step_advance(op, s);
}

void zero_p(rfhip_op *op) {
  const size_t w = (op->form == RFHIP_FORM_PLAIN_F32 || form_compressed(op->form)) ? 4 : 8;
  launch_fill(op->p.p, (int64_t)(w * nvec(op)), op->stream);
}

// Load an f64 host state into the form's state registers.
void state_from_host_f64(rfhip_op *op, const double *y0) {
  op->cur = 0;
  if (form_compressed(op->form)) {
    h2d_layout_f64(op, y0, op->k.as<double>());
    launch_f64_to_c(op->k.as<double>(), op->y[0].as<float>(), op->ylo[0].as<uint16_t>(), nvec(op), op->stream);
  } else {
    h2d_reg_f64(op, y0, op->y[0].p);
  }
}
// Current state as an f64 layout vector on the device.
const double *state_layout_f64(rfhip_op *op) {
  if (form_compressed(op->form)) {
    launch_c_to_f64(op->y[op->cur].as<float>(), op->ylo[op->cur].as<uint16_t>(), op->k.as<double>(), nvec(op),
                    op->stream);
    return op->k.as<double>();
  }
  if (op->form == RFHIP_FORM_PLAIN_F32) {
    op->kst.alloc(8 * op->n_dof);
    launch_cast_f2d(op->y[0].as<float>(), op->kst.as<double>(), op->n_dof, op->stream);
    return op->kst.as<double>();
  }
  return op->y[is_fused(op->form) ? op->cur : 0].as<double>();
}
void state_to_host_f64(rfhip_op *op, double *out) { d2h_layout_f64(op, state_layout_f64(op), out); }
// Current state as an f64 AoS vector on the device.
const double *state_aos_f64(rfhip_op *op) { return state_layout_f64(op); }

void load_source(rfhip_op *op, int kind, const double *src) {
  if (kind != RFHIP_DRIVE_VECTOR) return;
  if (src == nullptr) return;  // the device-assembled port source (checked by check_drive)
  op->src.alloc(8 * nvec(op));
  h2d_layout_f64(op, src, op->src.as<double>());
  op->src_resident = false;
  sync(op);
}

template <class G>
void transient_impl(rfhip_op *op, int64_t steps, int64_t substeps, double dt, int kind, int64_t dof,
                    const G *g, bool record, const std::function<void(int64_t row)> &emit) {
  substeps = std::max<int64_t>(1, substeps);
  const double h = dt / (double)substeps;
  zero_p(op);
  if (record) emit(0);
  for (int64_t k = 0; k < steps; ++k) {
    for (int64_t j = 0; j < substeps; ++j) {
      const double gv = kind == RFHIP_DRIVE_FREE ? 0.0 : (double)g[k * substeps + j];
      lserk_step(op, h, kind, dof, gv, nullptr);
    }
    if (record) emit(k + 1);
  }
  if (!record) emit(-1);
}

// ---- KCL ------------------------------------------------------------------

double factor(const rfhip_controller &c, double err_norm, double prev, bool reject) {  // session.rs:174-186
  const double err = std::max(err_norm, kKclErrFloor);
  const double f = (reject || prev <= 0.0) ? c.safety * std::pow(err, -c.pi_alpha)
                                           : c.safety * std::pow(err, -c.pi_alpha) * std::pow(prev, c.pi_beta);
  if (reject) return std::max(f, c.shrink_limit);
  return std::min(std::max(f, c.shrink_limit), c.growth_limit);
}

// Bytes of a compressed register of n entries: f32 hi parts, then bf16 lo
// parts padded to an even count (a whole number of 4-byte words).
size_t creg_bytes(int64_t n) { return 4 * (size_t)n + 2 * (size_t)((n + 1) & ~int64_t(1)); }

// KCL RK4(3)5[2R+]C transient with the PI controller (gpu/operator.rs:823-893
// structure, session.rs:781-827 control logic) on the register vector ys.
// The controller runs on the host; each substep reads back one double.
template <class G>
void kcl_impl(rfhip_op *op, void *ys, double dt, int64_t steps, int kind, int64_t dof, const G *g,
              const rfhip_controller &ctl, rfhip_kcl_stats &st, const std::function<void(int64_t)> &emit) {
  const int64_t n = nvec(op);
  const bool f64 = f64_regs(op->form);
  // compressed registers: y and its backup as f32 hi + bf16 lo ([hi | lo]
  // in one buffer), err as f32; with the separate update pass every stage's
  // apply reads st in FP64, the last stage leaving the accepted y there for
  // the next substep
  const bool cmp = (op->kcl_reg == RFHIP_KCL_REG_COMPRESSED || op->kcl_reg == RFHIP_KCL_REG_COMPRESSED_SEPARATE) &&
                   form_compressed(op->form);
  // with them, the stage kernel does the register update in its epilogue
  // where it can (no Debye P block, which the stage kernel does not
  // update; conforming mesh): k is never written, and the stage input is
  // held like y (f32 hi + bf16 lo) in two buffers of kst that alternate
  // (neighbours still read a stage's input while its next one is written)
  const bool fusable = form_fused(op->form) && op->n_disp == 0 && n == 6 * (int64_t)op->np * op->n_elem &&
                       kcl_stage_fusable(op->form, op->sa);
  const bool fuse = cmp && op->kcl_reg == RFHIP_KCL_REG_COMPRESSED && fusable;
  // FP64 registers on a fused form: the same fusion, y / err / the backup
  // and the two alternating stage inputs in FP64
  const bool fuse64 = !cmp && f64 && op->kcl_reg == RFHIP_KCL_REG_F64 && fusable;
  const size_t w = f64 ? 8 : 4;
  const size_t creg = creg_bytes(n), cstride = (creg + 255) & ~size_t(255);
  const size_t dstride = (8 * (size_t)n + 255) & ~size_t(255);
  hipStream_t s = op->stream;
  op->kst.alloc(fuse ? 2 * cstride : fuse64 ? 2 * dstride : w * n);
  op->kerr.alloc((cmp ? 4 : w) * n);
  op->kbak.alloc(cmp ? creg_bytes(n) : w * n);
  for (DevBuf *b : {&op->kst, &op->kerr, &op->kbak}) launch_fill(b->p, (int64_t)b->bytes, s);
  float *yh = nullptr, *bh = nullptr;
  uint16_t *yl = nullptr, *bl = nullptr;
  if (cmp) {
    op->kyc.alloc(creg_bytes(n));
    yh = op->kyc.as<float>();
    yl = reinterpret_cast<uint16_t *>(yh + n);
    bh = op->kbak.as<float>();
    bl = reinterpret_cast<uint16_t *>(bh + n);
    launch_f64_to_c(static_cast<const double *>(ys), yh, yl, n, s);
    if (!fuse) launch_copy(op->kst.p, ys, 8 * n, s);
  }
  float *sth[2] = {nullptr, nullptr};  // fused: the stage inputs; sth[sc] is the next stage's
  uint16_t *stl[2] = {nullptr, nullptr};
  int sc = 0;
  if (fuse) {
    for (int i = 0; i < 2; ++i) {
      sth[i] = reinterpret_cast<float *>(static_cast<char *>(op->kst.p) + i * cstride);
      stl[i] = reinterpret_cast<uint16_t *>(sth[i] + n);
    }
    launch_copy(sth[0], op->kyc.p, (int64_t)creg, s);  // the first stage's input is y
  }
  double *std64[2] = {nullptr, nullptr};  // fused FP64: the stage inputs; std64[sc] is the next stage's
  if (fuse64) {
    for (int i = 0; i < 2; ++i) std64[i] = reinterpret_cast<double *>(static_cast<char *>(op->kst.p) + i * dstride);
    launch_copy(std64[0], ys, 8 * n, s);
  }
  const int nblk = weighted_err_blocks(n);
  const int64_t nslot = fuse || fuse64 ? kcl_stage_slots(op->order, op->sa, fuse64) : 0;
  const int64_t npart = std::max<int64_t>(nblk, nslot);
  op->red.alloc(8 * (npart + 1));
  double *part = op->red.as<double>();
  double *res = part + npart;
  double h = dt, prev = 0.0, hmin = INFINITY, hmax = 0.0;
  int64_t acc = 0, rej = 0;
  const double h_floor = ctl.min_step_factor * dt;
  // wall time of the substeps, the frame downloads left out
  double emit_ms = 0.0;
  auto timed_emit = [&](int64_t r) {
    const auto a = std::chrono::steady_clock::now();
    emit(r);
    emit_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
  };
  sync(op);
  const auto t_begin = std::chrono::steady_clock::now();
  auto finish_timing = [&] {
    sync(op);
    op->last_call_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count() - emit_ms;
  };
  timed_emit(0);
  // the frame of the current y in ys (FP64), where the caller reads it
  auto frame = [&](int64_t r) {
    if (cmp) launch_c_to_f64(yh, yl, static_cast<double *>(ys), n, s);
    timed_emit(r);
  };
  for (int64_t fr = 0; fr < steps; ++fr) {
    double t_rel = 0.0;
    const double gv = kind == RFHIP_DRIVE_FREE ? 0.0 : (double)g[fr];
    while (t_rel < dt) {
      const double ht = std::min(h, dt - t_rel);
      // explicit_adaptive.rs:107-149. After each stage's apply one pass
      // does f = h k, y += b f, err = (or +=) (bhat - b) f and, but for the
      // last stage, builds the next stage's input y + (a - b) f. The first
      // pass also saves the starting y (kbak); the last one reduces the
      // error norm (session.rs:160-171) against it.
      for (int sg = 0; (fuse || fuse64) && sg < 5; ++sg) {
        StageArgs a = op->sa;
        if (fuse64) {
          a.yin = std64[sc];
          a.yout = std64[1 - sc];
          a.kyd = static_cast<double *>(ys);
          a.kerrd = op->kerr.as<double>();
          a.kbakd = op->kbak.as<double>();
        } else {
          a.yin = sth[sc];
          a.yin_lo = stl[sc];
          a.yout = sth[1 - sc];
          a.yout_lo = stl[1 - sc];
          a.ky = yh;
          a.ky_lo = yl;
          a.kerr = op->kerr.as<float>();
          a.kbak = bh;
          a.kbak_lo = bl;
        }
        fill_source(op, a, kind, dof, gv);
        a.dt = ht;
        a.kb = kKclB[sg];
        a.kew = kKclBhat[sg] - kKclB[sg];
        a.kamb = sg < 4 ? kKclA[sg] - kKclB[sg] : 0.0;
        a.katol = ctl.atol;
        a.krtol = ctl.rtol;
        a.kfirst = sg == 0;
        a.klast = sg == 4;
        a.kpart = part;
        launch_kcl_stage(op->order, a, fuse64, s);
        sc = 1 - sc;
        if (sg == 4) launch_sum_partials(part, (int)nslot, res, s);
      }
      for (int sg = 0; !fuse && !fuse64 && sg < 5; ++sg) {
        rhs_into(op, sg == 0 && !cmp ? ys : op->kst.p, op->k.p, kind, dof, gv);
        KclStage a{};
        a.y = cmp ? (void *)yh : ys;
        a.y_lo = yl;
        a.err = op->kerr.p;
        a.st = op->kst.p;
        a.bak = op->kbak.p;
        a.bak_lo = bl;
        a.k = op->k.p;
        a.dt = ht;
        a.b = kKclB[sg];
        a.ew = kKclBhat[sg] - kKclB[sg];
        a.amb = sg < 4 ? kKclA[sg] - kKclB[sg] : 0.0;
        a.first = sg == 0;
        a.next = sg < 4;
        a.n = n;
        a.atol = ctl.atol;
        a.rtol = ctl.rtol;
        a.partials = part;
        a.nblk = nblk;
        a.err_sum = res;
        if (cmp)
          launch_kcl_update_c(a, s);
        else
          launch_kcl_update(f64, a, s);
      }
      double ssum = 0.0;
      RF_HIP(hipMemcpyAsync(&ssum, res, 8, hipMemcpyDeviceToHost, s));
      sync(op);
      const double e = std::sqrt(ssum / (double)std::max<int64_t>(op->n_dof, 1));
      if (std::isfinite(e) && e <= 1.0) {
        t_rel += ht;
        ++acc;
        h = ht * factor(ctl, e, prev, false);
        prev = std::max(e, kKclErrFloor);
        hmin = std::min(hmin, h);
        hmax = std::max(hmax, h);
      } else {
        ++rej;
        if (cmp) {
          launch_copy(op->kyc.p, op->kbak.p, (int64_t)creg, s);
          if (fuse)
            launch_copy(sth[sc], op->kbak.p, (int64_t)creg, s);
          else
            launch_c_to_f64(bh, bl, op->kst.as<double>(), n, s);
        } else {
          launch_copy(ys, op->kbak.p, (int64_t)(w * n), s);
          if (fuse64) launch_copy(std64[sc], op->kbak.p, 8 * n, s);
        }
        h = ht * factor(ctl, std::isfinite(e) ? e : kKclNonfinite, prev, true);
      }
      if (h < h_floor) {
        st = {acc, rej, hmin, hmax};
        finish_timing();
        throw Error{RFHIP_ECOLLAPSE, "GPU KCL: step size collapsed below " + std::to_string(ctl.min_step_factor) +
                                         "*dt after " + std::to_string(acc) + " accepted, " + std::to_string(rej) +
                                         " rejected substeps"};
      }
    }
    frame(fr + 1);
  }
  finish_timing();
  st = {acc, rej, hmin, hmax};
}

// ---- Krylov (propagator.rs:186-373) ---------------------------------------

// Arnoldi with CGS2 on the device, Hessenberg on the host. aug: the ETD
// augmentation [[A, b], [0, 0]] with the scalar component carried here and
// b = op->src (a layout vector).
int krylov_impl(rfhip_op *op, const double *v_host, double t, int max_dim, double tol, bool aug, double *out) {
  const int64_t n = nvec(op);
  RF_REQUIRE(max_dim >= 1 && max_dim <= kMaxKrylov, "Krylov dimension must be in [1, 64]");
  hipStream_t s = op->stream;
  if (op->kb_cap < max_dim) {
    op->kb_basis.alloc(8 * (size_t)(max_dim + 1) * n);
    op->kb_cap = max_dim;
  }
  op->kb_w.alloc(8 * n);
  op->kb_x.alloc(8 * n);
  op->kb_out.alloc(8 * n);
  op->kb_proj.alloc(8 * (kMaxKrylov + 8));
  op->kb_host.alloc(8 * (2 * kMaxKrylov + 8));
  launch_fill(op->kb_w.p, 8 * n, s);
  const int nblk = krylov_dot_blocks(n);
  op->kb_part.alloc(8 * (size_t)nblk * (kMaxKrylov + 1));
  double *basis = op->kb_basis.as<double>(), *w = op->kb_w.as<double>();
  double *proj = op->kb_proj.as<double>(), *part = op->kb_part.as<double>();
  double *norm_dev = proj + kMaxKrylov;
  const bool f32op = op->form == RFHIP_FORM_PLAIN_F32;
  if (f32op) {
    op->kst.alloc(4 * n);
    op->k.alloc(4 * n);
  }
  h2d_layout_f64(op, v_host, op->kb_x.as<double>());
  sync(op);
  const auto t_begin = std::chrono::steady_clock::now();
  // beta = |v| (with the augmented unit component), reduced on the device
  double nrm2 = 0.0;
  launch_multi_dot(op->kb_x.as<double>(), n, 1, op->kb_x.as<double>(), part, nblk, proj, s);
  RF_HIP(hipMemcpyAsync(&nrm2, proj, 8, hipMemcpyDeviceToHost, s));
  sync(op);
  if (aug) nrm2 += 1.0;
  const double beta = std::sqrt(nrm2);
  if (beta == 0.0) {
    std::fill(out, out + op->n_dof, 0.0);
    op->last_call_ms = 0.0;
    return 0;
  }
  launch_scale(op->kb_x.as<double>(), basis, n, 1.0 / beta, s);
  std::vector<double> H((size_t)max_dim * max_dim, 0.0), sc(max_dim + 1, 0.0), pj(kMaxKrylov);
  sc[0] = aug ? 1.0 / beta : 0.0;
  int dim = max_dim;
  std::vector<double> th, eth;
  for (int j = 0; j < max_dim; ++j) {
    const double *bj = basis + (int64_t)j * n;
    // matvec (the reference GPU path drops to f32 through cast kernels, gpu/operator.rs:1329-1353)
    if (f32op) {
      launch_cast_d2f(bj, op->kst.as<float>(), n, s);
      plain_apply_full(op, op->kst.p, op->k.p, s);
      launch_cast_f2d(op->k.as<float>(), w, n, s);
    } else {
      rhs_into(op, bj, w, RFHIP_DRIVE_FREE, 0, 0.0);
    }
    if (aug && sc[j] != 0.0) launch_axpy(w, op->src.as<double>(), sc[j], n, s);
    double ws = 0.0;  // augmented scalar of w (zero row of the augmented matrix)
    const int cols = j + 1;
    double norm2 = 0.0;
    // Classical Gram-Schmidt, two passes. The first pass's update and the
    // second pass's projections share one sweep over the basis.
    launch_multi_dot(basis, n, cols, w, part, nblk, proj, s);
    if (!aug) {
      // The host only accumulates H here: both passes' projections and the
      // norm come back asynchronously, one wait per step.
      double *hp = op->kb_host.as<double>();
      RF_HIP(hipMemcpyAsync(hp, proj, 8 * cols, hipMemcpyDeviceToHost, s));
      launch_axpy_dot(w, basis, n, cols, proj, part, nblk, proj, s);
      RF_HIP(hipMemcpyAsync(hp + kMaxKrylov, proj, 8 * cols, hipMemcpyDeviceToHost, s));
      launch_axpy_basis_norm(w, basis, n, cols, proj, part, nblk, norm_dev, s);
      RF_HIP(hipMemcpyAsync(hp + 2 * kMaxKrylov, norm_dev, 8, hipMemcpyDeviceToHost, s));
      sync(op);
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < cols; ++i) H[(size_t)i * max_dim + j] += hp[pass * kMaxKrylov + i];
      norm2 = hp[2 * kMaxKrylov];
    }
    if (aug) {
      // Augmented: each pass's projections take the augmented scalar's
      // share, sc[i] * ws, before the update uses them. ws is 0 in the first
      // pass, so its projections go to the update unchanged and only the
      // second pass waits for the host.
      double *hp = op->kb_host.as<double>();
      RF_HIP(hipMemcpyAsync(hp, proj, 8 * cols, hipMemcpyDeviceToHost, s));
      launch_axpy_dot(w, basis, n, cols, proj, part, nblk, proj, s);
      RF_HIP(hipMemcpyAsync(hp + kMaxKrylov, proj, 8 * cols, hipMemcpyDeviceToHost, s));
      sync(op);
      for (int pass = 0; pass < 2; ++pass) {
        double *p = hp + pass * kMaxKrylov;
        double dscal = 0.0;
        for (int i = 0; i < cols; ++i) {
          p[i] += sc[i] * ws;
          H[(size_t)i * max_dim + j] += p[i];
          dscal += p[i] * sc[i];
        }
        ws -= dscal;
      }
      RF_HIP(hipMemcpyAsync(proj, hp + kMaxKrylov, 8 * cols, hipMemcpyHostToDevice, s));
      launch_axpy_basis_norm(w, basis, n, cols, proj, part, nblk, norm_dev, s);
      RF_HIP(hipMemcpyAsync(&norm2, norm_dev, 8, hipMemcpyDeviceToHost, s));
      sync(op);
    }
    const double hnext = std::sqrt(norm2 + ws * ws);
    const int d = j + 1;
    if (hnext < kArnoldiBreakdown || d == max_dim) {
      dim = d;
      break;
    }
    if (tol > 0.0) {
      th.assign((size_t)d * d, 0.0);
      eth.assign((size_t)d * d, 0.0);
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) th[(size_t)a * d + b] = t * H[(size_t)a * max_dim + b];
      expm_host(th.data(), d, eth.data());
      if (hnext * std::fabs(eth[(size_t)(d - 1) * d]) < tol) {
        dim = d;
        break;
      }
    }
    H[(size_t)(j + 1) * max_dim + j] = hnext;
    launch_scale(w, basis + (int64_t)(j + 1) * n, n, 1.0 / hnext, s);
    sc[j + 1] = ws / hnext;
  }
  th.assign((size_t)dim * dim, 0.0);
  eth.assign((size_t)dim * dim, 0.0);
  for (int a = 0; a < dim; ++a)
    for (int b = 0; b < dim; ++b) th[(size_t)a * dim + b] = t * H[(size_t)a * max_dim + b];
  expm_host(th.data(), dim, eth.data());
  for (int i = 0; i < dim; ++i) pj[i] = beta * eth[(size_t)i * dim];
  RF_HIP(hipMemcpyAsync(proj, pj.data(), 8 * dim, hipMemcpyHostToDevice, s));
  launch_lincomb(basis, n, dim, proj, op->kb_out.as<double>(), s);
  sync(op);
  op->last_call_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count();
  d2h_layout_f64(op, op->kb_out.as<double>(), out);
  return dim;
}

// ---- adaptive Krylov exponential ------------------------------------------
//
// The adaptive exponential's Arnoldi basis, built one column at a time by
// classical Gram-Schmidt with delayed reorthogonalisation (DCGS2,
// Swirydowicz, Langou, Ananthan, Yang and Thomas, "Low synchronization
// Gram-Schmidt and generalized minimal residual algorithms", 2021): column
// j's second projection is taken in the same sweep as column j+1's first.
// Column j of the basis holds the provisional vector q_j (projected once,
// not normalised) until the next column finishes it. Per column:
//  * z = A q_j, written into column j+1;
//  * one Gram sweep: s = V_j^T q_j, |q_j|^2, p = V_j^T z, q_j . z;
//  * on the host: r = |q_j - V_j s| (Pythagoras), column j-1 of H corrected
//    by s and closed with r, and column j's first projections of
//    A v_j = (z - A V_j s) / r = (z - V_{j+1} H s) / r;
//  * one update sweep: v_j = (q_j - V_j s) / r in place, and the next
//    provisional vector q_{j+1} = (z - V_j a - b q_j) / r.
// Two passes over the basis per column where CGS2 takes three, and no
// normalisation pass (q_j is applied unnormalised; the scale folds into
// the coefficients). finalize() reorthogonalises the last q before H is
// used. The step's state vector is column 0 (q_0), and the step's result
// is written back there.
struct Arnoldi {
  rfhip_op *op = nullptr;
  int64_t n = 0;
  int cap = 0, cols = 0, nblk = 0;
  bool broke = false, pending = false;
  double beta = 0.0;
  std::vector<double> H;
  // compressed basis (RFHIP_KRYLOV_BASIS_COMPRESSED): the finished columns
  // as f32 hi [cap][n] + bf16 lo [cap][nl] in kb_cbasis; the provisional
  // vector (which starts as the state) and the product A q in kb_basis
  bool cb = false;
  int64_t nl = 0;

  Arnoldi(rfhip_op *o, int capacity) : op(o), n(nvec(o)), cap(capacity) {
    cb = op->kb_mode == RFHIP_KRYLOV_BASIS_COMPRESSED;
    if (cb) {
      nl = (n + 1) & ~int64_t(1);
      op->kb_cbasis.alloc((size_t)cap * (4 * (size_t)n + 2 * (size_t)nl));
      op->kb_basis.alloc(16 * (size_t)n);
    } else if (op->kb_cap < cap) {
      op->kb_basis.alloc(8 * (size_t)(cap + 1) * n);
      op->kb_cap = cap;
    }
    op->kb_proj.alloc(8 * (size_t)(6 * kMaxKrylov + 16));
    op->kb_host.alloc(8 * (size_t)(6 * kMaxKrylov + 16));
    nblk = krylov_dot_blocks(n);
    op->kb_part.alloc(8 * (size_t)nblk * (2 * kMaxKrylov + 4));
    if (op->form == RFHIP_FORM_PLAIN_F32) {
      op->kst.alloc(4 * n);
      op->k.alloc(4 * n);
    }
    H.assign((size_t)(cap + 1) * cap, 0.0);
  }
  double *basis() const { return op->kb_basis.as<double>(); }
  double *col(int c) const { return basis() + (int64_t)c * n; }
  float *chi() const { return op->kb_cbasis.as<float>(); }
  uint16_t *clo() const { return reinterpret_cast<uint16_t *>(chi() + (int64_t)cap * n); }
  // the FP64 state: each step's basis starts from it and its result goes
  // back there (column 0, or the provisional vector of the compressed basis)
  double *state() const { return col(0); }
  // the provisional vector q_j and the product A q_j of column j
  double *qv(int j) const { return cb ? col(0) : col(j); }
  double *zv(int j) const { return cb ? col(1) : col(j + 1); }
  double &h(int r, int c) { return H[(size_t)r * cap + c]; }
  // device / pinned-host regions: Gram results, update coefficients, |q|^2
  static constexpr int kGram = 0, kCoef = 2 * kMaxKrylov + 4, kNorm = 4 * kMaxKrylov + 8;
  // A new basis from the state vector in column 0.
  void begin() {
    std::fill(H.begin(), H.end(), 0.0);
    cols = 0;
    pending = true;
    broke = false;
    beta = 0.0;
  }
  void apply(const double *x, double *w) {
    hipStream_t s = op->stream;
    if (op->form == RFHIP_FORM_PLAIN_F32) {
      launch_cast_d2f(x, op->kst.as<float>(), n, s);
      plain_apply_full(op, op->kst.p, op->k.p, s);
      launch_cast_f2d(op->k.as<float>(), w, n, s);
    } else {
      rhs_into(op, x, w, RFHIP_DRIVE_FREE, 0, 0.0);
    }
  }
  // Finishes v_j (j = cols) and starts q_{j+1}; at j = 0 sets beta = |q_0|.
  // A breakdown (|q_j - V_j s| below the threshold of propagator.rs) leaves
  // cols = j, H[j][j-1] = 0 and broke set.
  void extend() {
    RF_REQUIRE(pending && !broke && cols < cap, "adaptive Arnoldi: no provisional column to extend");
    hipStream_t s = op->stream;
    const int j = cols;
    double *dev = op->kb_proj.as<double>(), *hp = op->kb_host.as<double>(), *part = op->kb_part.as<double>();
    apply(qv(j), zv(j));
    if (cb)
      launch_gram2c(chi(), clo(), n, nl, j, qv(j), zv(j), part, nblk, dev + kGram, s);
    else
      launch_gram2(basis(), n, j + 2, part, nblk, dev + kGram, s);
    RF_HIP(hipMemcpyAsync(hp + kGram, dev + kGram, 8 * (size_t)(2 * (j + 1)), hipMemcpyDeviceToHost, s));
    sync(op);
    const double *g = hp + kGram;  // g[2c] = b_c . q_j, g[2c + 1] = b_c . z
    const double alpha = g[2 * j], gamma = g[2 * j + 1];
    RF_REQUIRE(std::isfinite(alpha) && std::isfinite(gamma), "adaptive Arnoldi: non-finite basis vector");
    double r2 = alpha, sp = 0.0;
    for (int c = 0; c < j; ++c) {
      r2 -= g[2 * c] * g[2 * c];
      sp += g[2 * c] * g[2 * c + 1];
    }
    const double r = r2 > 0.0 ? std::sqrt(r2) : 0.0;
    if (j == 0) {
      beta = r;
      if (beta == 0.0) {
        broke = true;
        pending = false;
        return;
      }
    } else {
      for (int c = 0; c < j; ++c) h(c, j - 1) += g[2 * c];
      if (!(r >= kArnoldiBreakdown)) {
        h(j, j - 1) = 0.0;
        broke = true;
        pending = false;
        return;
      }
      h(j, j - 1) = r;
    }
    // t = H[0..j][0..j-1] s, the first projections of A v_j onto V_{j+1}
    // and the next provisional vector's coefficients
    std::vector<double> t((size_t)j + 1, 0.0), cv((size_t)j + 1);
    for (int k = 0; k <= j; ++k)
      for (int i = 0; i < j; ++i) t[k] += h(k, i) * g[2 * i];
    for (int c = 0; c < j; ++c) h(c, j) = (g[2 * c + 1] - t[c]) / r;
    h(j, j) = ((gamma - sp) / r - t[j]) / r;
    for (int k = 0; k <= j; ++k) cv[k] = t[k] + r * h(k, j);
    double *cf = hp + kCoef;
    for (int c = 0; c < j; ++c) {
      cf[c] = g[2 * c];
      cf[j + c] = cv[c] - g[2 * c] * cv[j] / r;
    }
    cf[2 * j] = cv[j] / r;
    cf[2 * j + 1] = 1.0 / r;
    RF_HIP(hipMemcpyAsync(dev + kCoef, cf, 8 * (size_t)(2 * j + 2), hipMemcpyHostToDevice, s));
    if (cb)
      launch_dcgs_updatec(chi(), clo(), n, nl, j, dev + kCoef, qv(j), zv(j), part, nblk, dev + kNorm, s);
    else
      launch_dcgs_update(basis(), n, j, dev + kCoef, part, nblk, dev + kNorm, s);
    RF_HIP(hipMemcpyAsync(hp + kNorm, dev + kNorm, 8, hipMemcpyDeviceToHost, s));
    cols = j + 1;
  }
  // H[cols][cols-1] from |q_cols| (the first-pass residual norm), for the
  // proposer's features before the column is finished.
  void provisional() {
    if (!pending || broke || cols == 0) return;
    sync(op);
    h(cols, cols - 1) = std::sqrt(std::max(op->kb_host.as<double>()[kNorm], 0.0));
  }
  // Reorthogonalises the last provisional vector: closes column cols-1 of H
  // (its correction and H[cols][cols-1]); a breakdown sets broke.
  void finalize() {
    if (!pending || broke || cols == 0) {
      pending = false;
      return;
    }
    hipStream_t s = op->stream;
    const int j = cols;
    double *dev = op->kb_proj.as<double>(), *hp = op->kb_host.as<double>();
    if (cb) {
      // v_c . q and |q|^2 as the q-column of the Gram sweep (z = q)
      launch_gram2c(chi(), clo(), n, nl, j, qv(j), qv(j), op->kb_part.as<double>(), nblk, dev + kGram, s);
      RF_HIP(hipMemcpyAsync(hp + kGram, dev + kGram, 8 * (size_t)(2 * (j + 1)), hipMemcpyDeviceToHost, s));
      sync(op);
      for (int c = 0; c <= j; ++c) hp[kGram + c] = hp[kGram + 2 * c];
    } else {
      launch_multi_dot(basis(), n, j + 1, col(j), op->kb_part.as<double>(), nblk, dev + kGram, s);
      RF_HIP(hipMemcpyAsync(hp + kGram, dev + kGram, 8 * (size_t)(j + 1), hipMemcpyDeviceToHost, s));
      sync(op);
    }
    double r2 = hp[kGram + j];
    for (int c = 0; c < j; ++c) {
      h(c, j - 1) += hp[kGram + c];
      r2 -= hp[kGram + c] * hp[kGram + c];
    }
    const double r = r2 > 0.0 ? std::sqrt(r2) : 0.0;
    pending = false;
    if (!(r >= kArnoldiBreakdown)) {
      h(j, j - 1) = 0.0;
      broke = true;
    } else {
      h(j, j - 1) = r;
    }
  }
  // out = beta V_d exp(h H_d) e_1 (f64 layout vector on the device; may be
  // column 0).
  void combine(int d, double hstep, double *out, KrylovScratch &ws) {
    ws.m.assign((size_t)d * d, 0.0);
    ws.e.assign((size_t)d * d, 0.0);
    for (int a = 0; a < d; ++a)
      for (int b = 0; b < d; ++b) ws.m[(size_t)a * d + b] = hstep * H[(size_t)a * cap + b];
    expm_pade13(ws.m.data(), d, ws.e.data(), ws.work);
    double *cf = op->kb_host.as<double>() + kCoef;
    for (int i = 0; i < d; ++i) cf[i] = beta * ws.e[(size_t)i * d];
    double *dev = op->kb_proj.as<double>() + kCoef;
    RF_HIP(hipMemcpyAsync(dev, cf, 8 * (size_t)d, hipMemcpyHostToDevice, op->stream));
    if (cb)
      launch_lincombc(chi(), clo(), n, nl, d, dev, out, op->stream);
    else
      launch_lincomb_alias(basis(), n, d, dev, out, op->stream);
    sync(op);
  }
};

// exp(t_end A) v with a step size and Krylov dimension chosen per step: the
// plain error-controlled controller, or the network's proposals: its m sets
// the basis and its h seeds a search for the largest step the same
// a-posteriori estimate admits on that basis.
void adaptive_impl(rfhip_op *op, const double *v_host, double t_end, const rfhip_expctl &ctl, rfhip_mlp *mlp,
                   rfhip_exp_step *log, int64_t log_cap, double *hess, int hess_dim, int64_t *n_steps, double *out) {
  const int cap = std::max(ctl.m, ctl.m_max);
  Arnoldi ar(op, cap);
  double *x = ar.state();
  h2d_layout_f64(op, v_host, x);
  sync(op);
  const auto t_begin = std::chrono::steady_clock::now();
  KrylovScratch ws;
  ProposalHistory hist;
  double t = 0.0, h_next = ctl.h0 > 0.0 ? ctl.h0 : 0.0, rate = 0.0, hs = 1.0;
  int64_t nst = 0;
  while (t_end - t > 1e-14 * t_end) {
    RF_REQUIRE(nst < ctl.max_steps, "adaptive exponential: more than max_steps steps");
    const auto t_step = std::chrono::steady_clock::now();
    ar.begin();
    ar.extend();
    if (ar.beta == 0.0) break;  // a zero state stays zero
    if (nst == 0) rate = ctl.tol * ar.beta / t_end;
    const double t_left = t_end - t;
    if (!ar.broke) ar.extend();
    ar.provisional();
    rfhip_exp_step rec{};
    rec.t = t;
    rec.beta = ar.beta;
    rec.t_left = t_left;
    const bool early = ar.broke;  // an invariant subspace of dimension 1 or 2
    if (nst == 0) {
      hs = ctl.h_scale > 0.0 ? ctl.h_scale : 1.0 / std::max(std::fabs(ar.H[(size_t)cap]), 1e-300);
      hist.t_end = t_end;
      hist.beta0 = ar.beta;
      hist.h10_0 = std::fabs(ar.H[(size_t)cap]);
    }
    if (!early) proposal_features(ar.H.data(), cap, hist, ar.beta, ctl.tol, t_left, hs, op->n_dof, rec.feat);
    double h_prop;
    int m_prop;
    if (early) {
      h_prop = t_left;
      m_prop = ar.cols;
    } else if (mlp) {
      float yo[2];
      int rc = rfhip_mlp_forward(mlp, rec.feat, 1, yo);
      if (rc != RFHIP_OK) throw Error{rc, "proposer forward pass failed"};
      const double hp = hs * std::exp(std::clamp((double)yo[0], -700.0, 700.0));
      h_prop = std::isfinite(hp) ? std::clamp(hp, 1e-9 * t_end, t_left) : t_left;
      m_prop = std::clamp((int)std::lround(std::isfinite(yo[1]) ? yo[1] : (float)ctl.m), std::max(ctl.m_min, 2), ctl.m_max);
      rec.proposer = 1;
    } else {
      h_prop = h_next > 0.0 ? std::min(h_next, t_left) : 0.0;
      m_prop = ctl.m;
    }
    while (!ar.broke && ar.cols < m_prop) ar.extend();
    ar.finalize();
    const int d = ar.cols;
    rec.m = d;
    rec.breakdown = ar.broke ? 1 : 0;
    if (ar.broke) h_prop = t_left;  // exact in an invariant subspace
    if (h_prop <= 0.0) h_prop = max_admissible_h(ar.H.data(), cap, d, ar.beta, rate, t_left * 1e-3, t_left, 12, ws);
    RF_REQUIRE(h_prop > 0.0, "adaptive exponential: no admissible step on the first basis");
    double est = ar.broke ? 0.0 : krylov_estimate(ar.H.data(), cap, d, ar.beta, h_prop, ws);
    double h = h_prop;
    rec.h_first = h_prop;
    rec.est_first = est;
    rec.bound_first = rate * h_prop;
    rec.n_judged = 1;
    if (mlp && !ar.broke) {
      // the network's step seeds a search for the largest step the estimate
      // admits on this basis (up or down from it); a proposal the estimate
      // rejects counts as a second judgement
      if (!(est <= rate * h)) ++rec.n_judged;
      h = max_admissible_h(ar.H.data(), cap, d, ar.beta, rate, h_prop, t_left, 10, ws);
      RF_REQUIRE(h > 0.0, "adaptive exponential: no admissible step on the proposed basis");
      est = krylov_estimate(ar.H.data(), cap, d, ar.beta, h, ws);
    } else if (!(est <= rate * h)) {  // shrink on the same basis until the estimate accepts
      for (;;) {
        const double r = est / (rate * h);
        h *= std::max(ctl.shrink_limit, std::min(ctl.safety * std::pow(r, -1.0 / d), 1.0 - 1e-3));
        est = krylov_estimate(ar.H.data(), cap, d, ar.beta, h, ws);
        ++rec.n_judged;
        if (est <= rate * h) break;
        RF_REQUIRE(h > 1e-12 * t_end && rec.n_judged < 200,
                   "adaptive exponential: step size collapsed (tolerance too tight for this dimension)");
      }
    }
    rec.h = h;
    rec.est = est;
    rec.bound = rate * h;
    ar.combine(d, h, x, ws);
    if (hess && hess_dim > 0 && nst < log_cap) {
      double *blk = hess + (size_t)nst * (hess_dim + 1) * hess_dim;
      std::fill(blk, blk + (size_t)(hess_dim + 1) * hess_dim, 0.0);
      for (int a = 0; a <= std::min(d, hess_dim); ++a)
        for (int b = 0; b < std::min(d, hess_dim); ++b) blk[(size_t)a * hess_dim + b] = ar.H[(size_t)a * cap + b];
    }
    rec.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_step).count();
    if (log && nst < log_cap) log[nst] = rec;
    ++nst;
    // next step: the controller's growth from this step's estimate
    const double r = est / (rate * h);
    const double grow = r > 0.0 ? ctl.safety * std::pow(r, -1.0 / std::max(d, 1)) : ctl.growth_limit;
    h_next = h * std::clamp(grow, ctl.shrink_limit, ctl.growth_limit);
    hist.h_prev = h;
    hist.beta_prev = ar.beta;
    t += h;
  }
  sync(op);
  op->last_call_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count();
  if (n_steps) *n_steps = nst;
  d2h_layout_f64(op, x, out);
}


// nsteps free (or, with waves, port-driven) LSERK4 steps of every operator's
// resident state, two steps of all of them captured into one graph that
// forks from ops[0]'s stream to every operator's stream and joins back.
// waves[i] (when waves is given and waves[i] is not null): nsteps waveform
// samples of ops[i]'s device-assembled port source, held per step; they are
// uploaded once and read on the device through a step index, so the same
// graph replays every pair of steps.
void batch_lserk_impl(rfhip_op *const *ops, int32_t n_ops, const double *dt, int64_t nsteps,
                      const double *const *waves) {
  RF_REQUIRE(ops && dt && n_ops >= 1 && nsteps >= 0, "bad arguments");
  std::vector<char> driven((size_t)n_ops, 0);
  for (int32_t i = 0; i < n_ops; ++i) {
    RF_REQUIRE(ops[i], "null operator in the batch");
    RF_REQUIRE(ops[i]->ctx == ops[0]->ctx, "batched operators must share one context");
    RF_REQUIRE(std::isfinite(dt[i]), "non-finite step");
    for (int32_t j = 0; j < i; ++j) RF_REQUIRE(ops[j] != ops[i], "operator listed twice in the batch");
    if (waves && waves[i]) {
      RF_REQUIRE(ops[i]->src_resident, "a port-driven operator needs its port source assembled on the device "
                                       "(rfhip_op_port_source_f64) and not replaced since");
      for (int64_t k = 0; k < nsteps; ++k) RF_REQUIRE(std::isfinite(waves[i][k]), "non-finite waveform sample");
      driven[(size_t)i] = 1;
    }
  }
  for (int32_t i = 0; i < n_ops; ++i) {
    rfhip_op *op = ops[i];
    zero_p(op);
    if (driven[(size_t)i]) {
      op->gseq.alloc(8 * (size_t)std::max<int64_t>(nsteps, 1));
      op->gctr.alloc(8);
      if (nsteps > 0)
        RF_HIP(hipMemcpyAsync(op->gseq.p, waves[i], 8 * (size_t)nsteps, hipMemcpyHostToDevice, op->stream));
      launch_fill(op->gctr.p, 8, op->stream);
    }
    sync(op);
  }
  hipStream_t s0 = ops[0]->stream;
  std::vector<hipEvent_t> ev((size_t)n_ops + 1, nullptr);
  hipGraphExec_t exec = nullptr;
  auto cleanup = [&] {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  auto step_all = [&] {
    for (int32_t i = 0; i < n_ops; ++i)
      lserk_step(ops[i], dt[i], RFHIP_DRIVE_FREE, 0, 0.0, nullptr, driven[(size_t)i] != 0);
  };
  try {
    for (auto &e : ev) RF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (nsteps >= 2) {
      // The capture runs nothing, so each operator's buffer parity is restored.
      std::vector<int> cur0((size_t)n_ops);
      for (int32_t i = 0; i < n_ops; ++i) cur0[(size_t)i] = ops[i]->cur;
      hipGraph_t graph = nullptr;
      RF_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
      try {
        RF_HIP(hipEventRecord(ev[0], s0));
        for (int32_t i = 1; i < n_ops; ++i) RF_HIP(hipStreamWaitEvent(ops[i]->stream, ev[0], 0));
        step_all();
        step_all();
        for (int32_t i = 1; i < n_ops; ++i) {
          RF_HIP(hipEventRecord(ev[(size_t)i], ops[i]->stream));
          RF_HIP(hipStreamWaitEvent(s0, ev[(size_t)i], 0));
        }
      } catch (...) {
        (void)hipStreamEndCapture(s0, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        for (int32_t i = 0; i < n_ops; ++i) ops[i]->cur = cur0[(size_t)i];
        throw;
      }
      RF_HIP(hipStreamEndCapture(s0, &graph));
      for (int32_t i = 0; i < n_ops; ++i) ops[i]->cur = cur0[(size_t)i];
      const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      RF_HIP(e);
      for (int64_t k = 0; k + 1 < nsteps; k += 2) RF_HIP(hipGraphLaunch(exec, s0));
      RF_HIP(hipStreamSynchronize(s0));
    }
    if (nsteps % 2) step_all();
    for (int32_t i = 0; i < n_ops; ++i) sync(ops[i]);
  } catch (...) {
    for (int32_t i = 0; i < n_ops; ++i) (void)hipStreamSynchronize(ops[i]->stream);
    cleanup();
    throw;
  }
  cleanup();
}
}  // namespace

// propagator.rs:69-114: infinity-norm scaling to <= 0.5, Taylor 18, squaring.
// Kept step for step (and with the reference's zero-skip products) so the
// Hessenberg exponential matches the CPU path it replaces.
void expm_host(const double *a, int n, double *out) {
  const size_t nn = (size_t)n * n;
  double norm = 0.0;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += std::fabs(a[(size_t)i * n + j]);
    norm = std::max(norm, row);
  }
  int sq = 0;
  if (norm > 0.5) sq = (int)std::max<long>(0, (long)std::ceil(std::log2(norm)) + 1);
  double scale = 1.0;
  for (int q = 0; q < sq; ++q) scale *= 2.0;
  std::vector<double> b(nn), res(nn, 0.0), term(nn, 0.0), t2(nn), tmp(nn);
  for (size_t k = 0; k < nn; ++k) b[k] = a[k] / scale;
  for (int i = 0; i < n; ++i) res[(size_t)i * n + i] = term[(size_t)i * n + i] = 1.0;
  auto mm = [n](const std::vector<double> &x, const std::vector<double> &y, std::vector<double> &o) {
    std::fill(o.begin(), o.end(), 0.0);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < n; ++k) {
        const double xik = x[(size_t)i * n + k];
        if (xik == 0.0) continue;
        for (int j = 0; j < n; ++j) o[(size_t)i * n + j] += xik * y[(size_t)k * n + j];
      }
  };
  for (int k = 1; k <= 18; ++k) {
    mm(term, b, t2);
    const double inv = 1.0 / (double)k;
    for (size_t x = 0; x < nn; ++x) t2[x] *= inv;
    for (size_t x = 0; x < nn; ++x) res[x] += t2[x];
    term.swap(t2);
  }
  for (int q = 0; q < sq; ++q) {
    mm(res, res, tmp);
    res.swap(tmp);
  }
  std::copy(res.begin(), res.end(), out);
}

}  // namespace rfhip

using namespace rfhip;

namespace {

template <class T, class G>
int kcl_entry(rfhip_op *op, const T *y0, double dt, int64_t steps, int kind, int64_t dof, const double *src,
              const G *g, int64_t n_g, const rfhip_controller *ctl, T *out, rfhip_kcl_stats *stats) {
  return guarded([&] {
    RF_REQUIRE(op && y0 && out && ctl, "null argument");
    RF_REQUIRE(steps >= 0, "negative step count");
    check_drive(op, kind, dof, src, n_g, steps, g);
    load_source(op, kind, src);
    const int64_t n = op->n_dof;
    op->kb_w.alloc(8 * nvec(op));  // the KCL state register
    launch_fill(op->kb_w.p, 8 * nvec(op), op->stream);
    void *ys = op->kb_w.p;
    std::vector<double> row(n);
    if constexpr (std::is_same<T, double>::value) {
      h2d_reg_f64(op, y0, ys);
    } else {
      row = widen(y0, n);
      h2d_reg_f64(op, row.data(), ys);
    }
    rfhip_kcl_stats st{};
    try {
      kcl_impl<G>(op, ys, dt, steps, kind, dof, g, *ctl, st, [&](int64_t r) {
        if constexpr (std::is_same<T, double>::value) {
          d2h_reg_f64(op, ys, out + r * n);
        } else {
          d2h_reg_f64(op, ys, row.data());
          narrow(row, out + r * n);
        }
      });
    } catch (const Error &) {
      if (stats) *stats = st;
      throw;
    }
    if (stats) *stats = st;
  });
}

}  // namespace


extern "C" {

size_t rfhip_last_error(char *buf, size_t len) {
  if (buf && len) {
    const size_t c = std::min(len - 1, g_last_error.size());
    std::memcpy(buf, g_last_error.data(), c);
    buf[c] = 0;
  }
  return g_last_error.size();
}

int rfhip_ctx_create(rfhip_ctx **out) {
  return guarded([&] {
    RF_REQUIRE(out, "null output handle");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1) throw Error{RFHIP_ENODEV, "no HIP GPU device found"};
    auto ctx = std::make_unique<rfhip_ctx>();
    ctx->device = 0;
    RF_HIP(hipSetDevice(0));
    noalloc_init(0);
    hipDeviceProp_t prop;
    RF_HIP(hipGetDeviceProperties(&prop, 0));
    ctx->name = std::string(prop.name) + " (" + prop.gcnArchName + ")";
    *out = ctx.release();
  });
}

void rfhip_ctx_destroy(rfhip_ctx *ctx) { delete ctx; }

int rfhip_ctx_device_name(const rfhip_ctx *ctx, char *buf, size_t len) {
  return guarded([&] {
    RF_REQUIRE(ctx && buf && len, "bad arguments");
    const size_t c = std::min(len - 1, ctx->name.size());
    std::memcpy(buf, ctx->name.data(), c);
    buf[c] = 0;
  });
}

int rfhip_ctx_has_fp64(const rfhip_ctx *ctx) { return ctx ? 1 : 0; }

int rfhip_op_create(rfhip_ctx *ctx, const rfhip_op_desc *d, int form, rfhip_op **out) {
  return rfhip_op_create_dispersive(ctx, d, form, 0, nullptr, nullptr, nullptr, nullptr, out);
}

int rfhip_op_create_dispersive(rfhip_ctx *ctx, const rfhip_op_desc *d, int form, int64_t n_disp,
                               const int64_t *disp_elem, const double *disp_a, const double *disp_g,
                               const double *disp_inv_eps_inf, rfhip_op **out) {
  return guarded([&] {
    RF_REQUIRE(ctx && d && out, "null argument");
    RF_REQUIRE(n_disp >= 0, "negative number of dispersive elements");
    RF_REQUIRE(n_disp == 0 || (disp_elem && disp_a && disp_g && disp_inv_eps_inf), "missing dispersive array");
    RF_REQUIRE(form >= RFHIP_FORM_PLAIN_F32 && form <= RFHIP_FORM_FUSED_BF16X4, "unknown form");
    RF_REQUIRE(d->n_elem >= 1, "empty operator");
    const int p = order_of_np(d->np);
    RF_REQUIRE(p >= 1 && d->nfp == (p + 1) * (p + 2) / 2, "inconsistent Np / Nfp");
    if (p > 4) throw Error{RFHIP_EUNSUPPORTED, "orders 1..4 are compiled (Np 4, 10, 20, 35)"};
    RF_REQUIRE(d->diff_r && d->diff_s && d->diff_t && d->lift && d->face_nodes && d->jinv && d->inv_eps && d->inv_mu &&
                   d->sigma_eps && d->sigma_mu && d->face_normal && d->face_fscale && d->face_neighbor &&
                   d->face_nbr_local && d->face_port && d->face_perm,
               "missing operator array");
    RF_REQUIRE(6 * (int64_t)d->np * d->n_elem < ((int64_t)1 << 40), "operator too large");
    RF_HIP(hipSetDevice(ctx->device));
    auto op = std::make_unique<rfhip_op>();
    op->ctx = ctx;
    op->form = form;
    op->order = p;
    op->np = d->np;
    op->nfp = d->nfp;
    op->n_elem = d->n_elem;
    // rhs.rs:963-985: one P-block slot per listed element, in list order;
    // an element listed twice or out of range is refused (rhs.rs:970-974)
    std::vector<double> coef(3 * (size_t)n_disp);  // uploaded below; lives until the final sync
    if (n_disp > 0) {
      std::vector<char> seen((size_t)d->n_elem, 0);
      for (int64_t q = 0; q < n_disp; ++q) {
        const int64_t e = disp_elem[q];
        RF_REQUIRE(e >= 0 && e < d->n_elem, "dispersive element index out of range");
        RF_REQUIRE(!seen[(size_t)e], "element listed twice as dispersive");
        seen[(size_t)e] = 1;
        RF_REQUIRE(std::isfinite(disp_a[q]) && std::isfinite(disp_g[q]) && std::isfinite(disp_inv_eps_inf[q]),
                   "non-finite dispersive coefficient");
        coef[3 * q] = disp_a[q];
        coef[3 * q + 1] = disp_g[q];
        coef[3 * q + 2] = disp_inv_eps_inf[q];
      }
      op->n_disp = n_disp;
    }
    op->n_dof = 6 * (int64_t)d->np * d->n_elem + 3 * (int64_t)d->np * n_disp;
    op->n_vec = op->n_dof;
    op->alpha = d->flux_alpha;
    RF_HIP(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    if (n_disp > 0) {
      upload(op->disp_elem, disp_elem, (size_t)n_disp, op->stream);
      upload(op->disp_coef, coef.data(), coef.size(), op->stream);
    }
    op->port_host.assign(d->face_port, d->face_port + 4 * d->n_elem);
    op->face_nodes_host.assign(d->face_nodes, d->face_nodes + 4 * d->nfp);
    op->dev = DevOp{};
    op->dev.np = d->np;
    op->dev.nfp = d->nfp;
    op->dev.n_elem = d->n_elem;
    op->dev.alpha = d->flux_alpha;
    if (is_fused(form))
      build_fused(op.get(), *d);
    else
      build_plain(op.get(), *d);
    sync(op.get());
    *out = op.release();
  });
}

void rfhip_op_destroy(rfhip_op *op) {
  if (!op) return;
  if (op->stream) {
    (void)hipStreamSynchronize(op->stream);
    for (auto &g : op->step_graph)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    (void)hipStreamDestroy(op->stream);
  }
  delete op;
}

int64_t rfhip_op_n_dof(const rfhip_op *op) { return op ? op->n_dof : -1; }
int64_t rfhip_op_n_dispersive(const rfhip_op *op) { return op ? op->n_disp : -1; }
int rfhip_op_form(const rfhip_op *op) { return op ? op->form : -1; }

int rfhip_op_apply(rfhip_op *op, const float *y, float *dy) {
  return guarded([&] {
    RF_REQUIRE(op && y && dy, "null argument");
    const int64_t n = op->n_dof;
    hipStream_t s = op->stream;
    if (op->form == RFHIP_FORM_PLAIN_F32) {
      // staged in the scratch register: the resident state in y[0] stays
      op->kst.alloc(4 * nvec(op));
      RF_HIP(hipMemcpyAsync(op->kst.p, y, 4 * n, hipMemcpyHostToDevice, s));
      plain_apply_full(op, op->kst.p, op->k.p, s);
      RF_HIP(hipMemcpyAsync(dy, op->k.p, 4 * n, hipMemcpyDeviceToHost, s));
      sync(op);
      return;
    }
    const auto yd = widen(y, n);
    std::vector<double> out(n);
    op->kst.alloc(8 * nvec(op));
    h2d_reg_f64(op, yd.data(), op->kst.p);
    rhs_into(op, op->kst.p, op->k.p, RFHIP_DRIVE_FREE, 0, 0.0);
    d2h_reg_f64(op, op->k.p, out.data());
    narrow(out, dy);
  });
}

int rfhip_op_apply_f64(rfhip_op *op, const double *y, double *dy) {
  return guarded([&] {
    RF_REQUIRE(op && y && dy, "null argument");
    const size_t w = op->form == RFHIP_FORM_PLAIN_F32 ? 4 : 8;
    op->kst.alloc(w * nvec(op));
    h2d_reg_f64(op, y, op->kst.p);
    rhs_into(op, op->kst.p, op->k.p, RFHIP_DRIVE_FREE, 0, 0.0);
    d2h_reg_f64(op, op->k.p, dy);
  });
}

int rfhip_op_transient(rfhip_op *op, const float *y0, float dt, int64_t steps, int64_t substeps, int kind,
                       int64_t dof, const double *src, const float *g, int64_t n_g, int record, float *out) {
  return guarded([&] {
    RF_REQUIRE(op && y0 && out, "null argument");
    RF_REQUIRE(steps >= 0, "negative step count");
    substeps = std::max<int64_t>(1, substeps);
    check_drive(op, kind, dof, src, n_g, steps * substeps, g);
    load_source(op, kind, src);
    const int64_t n = op->n_dof;
    const auto y0d = widen(y0, n);
    state_from_host_f64(op, y0d.data());
    std::vector<double> row(n);
    transient_impl<float>(op, steps, substeps, (double)dt, kind, dof, g, record != 0, [&](int64_t r) {
      if (r == 0) {
        std::memcpy(out, y0, sizeof(float) * n);
        return;
      }
      state_to_host_f64(op, row.data());
      narrow(row, r < 0 ? out : out + r * n);
    });
  });
}

int rfhip_op_transient_f64(rfhip_op *op, const double *y0, double dt, int64_t steps, int64_t substeps, int kind,
                           int64_t dof, const double *src, const double *g, int64_t n_g, int record, double *out) {
  return guarded([&] {
    RF_REQUIRE(op && y0 && out, "null argument");
    RF_REQUIRE(steps >= 0, "negative step count");
    substeps = std::max<int64_t>(1, substeps);
    check_drive(op, kind, dof, src, n_g, steps * substeps, g);
    load_source(op, kind, src);
    const int64_t n = op->n_dof;
    state_from_host_f64(op, y0);
    transient_impl<double>(op, steps, substeps, dt, kind, dof, g, record != 0, [&](int64_t r) {
      if (r == 0)
        std::memcpy(out, y0, sizeof(double) * n);
      else
        state_to_host_f64(op, r < 0 ? out : out + r * n);
    });
  });
}

int rfhip_op_transient_probes_f64(rfhip_op *op, const double *y0, double dt, int64_t steps, int64_t substeps,
                                  int kind, int64_t dof, const double *src, const double *g, int64_t n_g,
                                  const int64_t *probes, int64_t n_probes, double *rows, double *y_final) {
  return guarded([&] {
    RF_REQUIRE(op && y0 && rows, "null argument");
    RF_REQUIRE(steps >= 0, "negative step count");
    RF_REQUIRE(n_probes >= 0 && (n_probes == 0 || probes), "bad probe list");
    for (int64_t q = 0; q < n_probes; ++q)
      RF_REQUIRE(probes[q] >= 0 && probes[q] < op->n_dof, "probe dof out of range");
    substeps = std::max<int64_t>(1, substeps);
    check_drive(op, kind, dof, src, n_g, steps * substeps, g);
    load_source(op, kind, src);
    hipStream_t s = op->stream;
    for (int64_t q = 0; q < n_probes; ++q) rows[q] = y0[probes[q]];
    if (n_probes > 0) {
      upload(op->probe_idx, probes, (size_t)n_probes, s);
      op->probe_rows.alloc(8 * (size_t)std::max<int64_t>(1, steps) * (size_t)n_probes);
    }
    state_from_host_f64(op, y0);
    transient_impl<double>(op, steps, substeps, dt, kind, dof, g, true, [&](int64_t r) {
      if (r <= 0 || n_probes == 0) return;
      const int cur = is_fused(op->form) ? op->cur : 0;
      const int kind_y = form_compressed(op->form) ? 2 : op->form == RFHIP_FORM_PLAIN_F32 ? 1 : 0;
      launch_gather(op->y[cur].p, op->ylo[cur].as<uint16_t>(), kind_y, op->probe_idx.as<int64_t>(), n_probes,
                    op->probe_rows.as<double>() + (r - 1) * n_probes, s);
    });
    if (n_probes > 0 && steps > 0)
      RF_HIP(hipMemcpyAsync(rows + n_probes, op->probe_rows.p, 8 * (size_t)steps * (size_t)n_probes,
                            hipMemcpyDeviceToHost, s));
    sync(op);
    if (y_final) state_to_host_f64(op, y_final);
  });
}

int rfhip_op_set_port(rfhip_op *op, int32_t port, int64_t n_faces, const int32_t *elem, const int32_t *local_face,
                      const double *profiles, const double *weights) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    RF_REQUIRE(port >= 0 && port < 65536, "port index out of range");
    RF_REQUIRE(n_faces >= 0 && (n_faces == 0 || (elem && local_face && profiles && weights)), "bad port arrays");
    const int nfp = op->nfp, np = op->np;
    rfhip_op::PortEntries pe;
    pe.dof.resize((size_t)n_faces * nfp);
    pe.prof.assign(profiles, profiles + (size_t)n_faces * nfp * 6);
    pe.w.assign(weights, weights + (size_t)n_faces * nfp);
    pe.face_elem.assign(elem, elem + n_faces);
    pe.face_lf.assign(local_face, local_face + n_faces);
    for (int64_t q = 0; q < n_faces; ++q) {
      RF_REQUIRE(elem[q] >= 0 && elem[q] < op->n_elem && local_face[q] >= 0 && local_face[q] < 4,
                 "port face out of range");
      for (int m = 0; m < nfp; ++m)
        pe.dof[(size_t)q * nfp + m] = ((int64_t)elem[q] * np + op->face_nodes_host[local_face[q] * nfp + m]) * 6;
    }
    for (size_t q = 0; q < pe.w.size(); ++q) {
      const double *pr = &pe.prof[6 * q];
      pe.e_norm += pe.w[q] * (pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2]);
      pe.h_norm += pe.w[q] * (pr[3] * pr[3] + pr[4] * pr[4] + pr[5] * pr[5]);
    }
    if ((int)op->ports.size() <= port) op->ports.resize(port + 1);
    op->ports[port] = std::move(pe);
    op->ports_dirty = true;
  });
}

int rfhip_op_n_ports(const rfhip_op *op) { return op ? (int)op->ports.size() : -1; }

int rfhip_op_port_source_f64(rfhip_op *op, int32_t port, double *b) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    RF_REQUIRE(port >= 0 && port < (int)op->ports.size(), "port index out of range");
    const auto &pe = op->ports[port];
    const int nfp = op->nfp;
    // port faces grouped by element, so one thread per (element, node) sums
    // every port face of its element
    std::vector<int64_t> order(pe.face_elem.size());
    for (size_t q = 0; q < order.size(); ++q) order[q] = (int64_t)q;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return pe.face_elem[x] < pe.face_elem[y]; });
    std::vector<int64_t> elems, offs(1, 0);
    std::vector<int32_t> lf;
    std::vector<double> prof;
    for (int64_t q : order) {
      if (elems.empty() || elems.back() != pe.face_elem[q]) {
        if (!elems.empty()) offs.push_back((int64_t)lf.size());
        elems.push_back(pe.face_elem[q]);
      }
      lf.push_back(pe.face_lf[q]);
      prof.insert(prof.end(), pe.prof.begin() + q * nfp * 6, pe.prof.begin() + (q + 1) * nfp * 6);
    }
    offs.push_back((int64_t)lf.size());
    hipStream_t s = op->stream;
    op->src.alloc(8 * nvec(op));
    launch_fill(op->src.p, 8 * nvec(op), s);
    if (!elems.empty()) {
      upload(op->ps_elem, elems.data(), elems.size(), s);
      upload(op->ps_off, offs.data(), offs.size(), s);
      upload(op->ps_lf, lf.data(), lf.size(), s);
      upload(op->ps_prof, prof.data(), prof.size(), s);
      launch_port_source(op, (int64_t)elems.size(), op->ps_elem.as<int64_t>(), op->ps_off.as<int64_t>(),
                         op->ps_lf.as<int32_t>(), op->ps_prof.as<double>(), op->src.as<double>(), s);
    }
    op->src_resident = true;
    op->ps_count = (int64_t)elems.size();
    if (b) d2h_layout_f64(op, op->src.as<double>(), b);
    sync(op);  // the host staging vectors above live until here
  });
}

int rfhip_op_transient_ports_f64(rfhip_op *op, const double *y0, double dt, int64_t steps, int64_t substeps, int kind,
                                 int64_t dof, const double *src, const double *g, int64_t n_g, double *signals,
                                 double *y_final) {
  return guarded([&] {
    RF_REQUIRE(op && y0 && signals, "null argument");
    RF_REQUIRE(steps >= 0, "negative step count");
    substeps = std::max<int64_t>(1, substeps);
    check_drive(op, kind, dof, src, n_g, steps * substeps, g);
    load_source(op, kind, src);
    hipStream_t s = op->stream;
    const int np_ = (int)op->ports.size();
    if (np_ > 0 && op->ports_dirty) {
      std::vector<int64_t> dofs, off(np_ + 1, 0);
      std::vector<double> prof, w, norm(2 * np_);
      for (int p = 0; p < np_; ++p) {
        const auto &pe = op->ports[p];
        dofs.insert(dofs.end(), pe.dof.begin(), pe.dof.end());
        prof.insert(prof.end(), pe.prof.begin(), pe.prof.end());
        w.insert(w.end(), pe.w.begin(), pe.w.end());
        off[p + 1] = (int64_t)dofs.size();
        norm[2 * p] = pe.e_norm;
        norm[2 * p + 1] = pe.h_norm;
      }
      if (dofs.empty()) {  // keep the device pointers valid
        dofs.push_back(0);
        prof.assign(6, 0.0);
        w.push_back(0.0);
      }
      upload(op->port_dof, dofs.data(), dofs.size(), s);
      upload(op->port_prof, prof.data(), prof.size(), s);
      upload(op->port_w, w.data(), w.size(), s);
      upload(op->port_off, off.data(), off.size(), s);
      upload(op->port_norm, norm.data(), norm.size(), s);
      sync(op);
      op->ports_dirty = false;
    }
    const size_t row = 2 * (size_t)np_;
    if (np_ > 0) op->port_rows.alloc(8 * row * (size_t)(steps + 1));
    state_from_host_f64(op, y0);
    transient_impl<double>(op, steps, substeps, dt, kind, dof, g, true, [&](int64_t r) {
      if (r < 0 || np_ == 0) return;
      const int cur = is_fused(op->form) ? op->cur : 0;
      const int kind_y = form_compressed(op->form) ? 2 : op->form == RFHIP_FORM_PLAIN_F32 ? 1 : 0;
      launch_port_proj(op->y[cur].p, op->ylo[cur].as<uint16_t>(), kind_y, op->port_dof.as<int64_t>(),
                       op->port_prof.as<double>(), op->port_w.as<double>(), op->port_off.as<int64_t>(),
                       op->port_norm.as<double>(), np_, op->port_rows.as<double>() + r * row, s);
    });
    if (np_ > 0)
      RF_HIP(hipMemcpyAsync(signals, op->port_rows.p, 8 * row * (size_t)(steps + 1), hipMemcpyDeviceToHost, s));
    sync(op);
    if (y_final) state_to_host_f64(op, y_final);
  });
}

int rfhip_op_transient_kcl(rfhip_op *op, const float *y0, float dt, int64_t steps, int kind, int64_t dof,
                           const double *src, const float *g, int64_t n_g, const rfhip_controller *ctl, float *out,
                           rfhip_kcl_stats *stats) {
  return kcl_entry<float, float>(op, y0, (double)dt, steps, kind, dof, src, g, n_g, ctl, out, stats);
}

int rfhip_op_set_kcl_registers(rfhip_op *op, int32_t mode) {
  return guarded([&] {
    RF_REQUIRE(op, "null operator");
    RF_REQUIRE(mode == RFHIP_KCL_REG_F64 || mode == RFHIP_KCL_REG_COMPRESSED ||
                   mode == RFHIP_KCL_REG_COMPRESSED_SEPARATE || mode == RFHIP_KCL_REG_F64_SEPARATE,
               "unknown KCL register mode");
    if ((mode == RFHIP_KCL_REG_COMPRESSED || mode == RFHIP_KCL_REG_COMPRESSED_SEPARATE) && !form_compressed(op->form))
      throw Error{RFHIP_EUNSUPPORTED, "compressed KCL registers need a form with compressed state registers"};
    op->kcl_reg = mode;
  });
}

int rfhip_op_set_krylov_basis(rfhip_op *op, int32_t mode) {
  return guarded([&] {
    RF_REQUIRE(op, "null operator");
    RF_REQUIRE(mode == RFHIP_KRYLOV_BASIS_F64 || mode == RFHIP_KRYLOV_BASIS_COMPRESSED, "unknown Krylov basis mode");
    op->kb_mode = mode;
  });
}

int rfhip_op_transient_kcl_f64(rfhip_op *op, const double *y0, double dt, int64_t steps, int kind, int64_t dof,
                               const double *src, const double *g, int64_t n_g, const rfhip_controller *ctl,
                               double *out, rfhip_kcl_stats *stats) {
  return kcl_entry<double, double>(op, y0, dt, steps, kind, dof, src, g, n_g, ctl, out, stats);
}

int rfhip_op_expmv(rfhip_op *op, const double *v, double t, int32_t max_dim, double tol, double *out,
                   int32_t *dim_used) {
  return guarded([&] {
    RF_REQUIRE(op && v && out, "null argument");
    const int dim = krylov_impl(op, v, t, max_dim, tol, false, out);
    if (dim_used) *dim_used = dim;
  });
}

int rfhip_op_etd_step(rfhip_op *op, const double *y, const double *b, double h, int32_t max_dim, double tol,
                      double *out, int32_t *dim_used) {
  return guarded([&] {
    RF_REQUIRE(op && y && b && out, "null argument");
    op->src.alloc(8 * nvec(op));
    h2d_layout_f64(op, b, op->src.as<double>());
    op->src_resident = false;
    const int dim = krylov_impl(op, y, h, max_dim, tol, true, out);
    if (dim_used) *dim_used = dim;
  });
}

int rfhip_op_expmv_adaptive(rfhip_op *op, const double *v, double t_end, const rfhip_expctl *ctl, rfhip_mlp *proposer,
                            rfhip_exp_step *steps, int64_t steps_cap, double *hess, int32_t hess_dim, int64_t *n_steps,
                            double *out) {
  return guarded([&] {
    RF_REQUIRE(op && v && ctl && out, "null argument");
    RF_REQUIRE(std::isfinite(t_end) && t_end > 0.0, "t_end must be positive and finite");
    RF_REQUIRE(std::isfinite(ctl->tol) && ctl->tol > 0.0, "tol must be positive");
    RF_REQUIRE(ctl->m >= 2 && ctl->m <= kMaxKrylov, "m must be in [2, 64]");
    RF_REQUIRE(ctl->m_min >= 2 && ctl->m_min <= ctl->m_max && ctl->m_max <= kMaxKrylov,
               "need 2 <= m_min <= m_max <= 64");
    RF_REQUIRE(ctl->safety > 0.0 && ctl->safety <= 1.0 && ctl->shrink_limit > 0.0 && ctl->shrink_limit < 1.0 &&
                   ctl->growth_limit > 1.0 && ctl->max_steps >= 1,
               "controller needs 0 < safety <= 1, 0 < shrink_limit < 1 < growth_limit, max_steps >= 1");
    RF_REQUIRE(steps_cap >= 0 && (steps || !hess), "a Hessenberg log needs the step log");
    if (hess) RF_REQUIRE(hess_dim >= std::max(ctl->m, proposer ? ctl->m_max : ctl->m),
                         "hess_dim must cover every Krylov dimension the run can build");
    adaptive_impl(op, v, t_end, *ctl, proposer, steps, steps ? steps_cap : 0, hess, hess ? hess_dim : 0, n_steps, out);
  });
}

int rfhip_op_last_call_ms(const rfhip_op *op, double *ms) {
  return guarded([&] {
    RF_REQUIRE(op && ms, "null argument");
    *ms = op->last_call_ms;
  });
}

int rfhip_op_n_variants(const rfhip_op *op) { return op && is_fused(op->form) ? n_stage_variants(op->form) : 1; }

int rfhip_op_variant_available(const rfhip_op *op, int variant) {
  if (!op) return 0;
  if (!is_fused(op->form)) return variant == 0 ? 1 : 0;
  return variant_available(op->order, op->form, variant, op->sa) ? 1 : 0;
}

int rfhip_op_variant(const rfhip_op *op) { return op && is_fused(op->form) ? op->sa.variant : 0; }

int rfhip_op_set_variant(rfhip_op *op, int variant) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    RF_REQUIRE(variant >= 0 && variant < rfhip_op_n_variants(op), "variant out of range");
    if (!is_fused(op->form)) return;
    if (!variant_available(op->order, op->form, variant, op->sa))
      throw Error{RFHIP_EUNSUPPORTED, "stage-kernel variant " + std::to_string(variant) +
                                          " cannot run on this operator (order or neighbour table)"};
    op->sa.variant = variant;
  });
}

int rfhip_op_poison_lds(rfhip_op *op) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    launch_poison_lds(op->stream);
    sync(op);
  });
}

int rfhip_op_autotune(rfhip_op *op, int *chosen, double *ms_best) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    if (!is_fused(op->form)) {
      if (chosen) *chosen = 0;
      if (ms_best) *ms_best = 0.0;
      return;
    }
    hipStream_t s = op->stream;
    const int64_t n = nvec(op);
    const bool c = form_compressed(op->form);
    DevBuf yo, ylo, pp;
    yo.alloc((c ? 4 : 8) * n);
    ylo.alloc(2 * n);
    pp.alloc((c ? 4 : 8) * n);
    launch_fill(pp.p, (int64_t)pp.bytes, s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RF_HIP(hipEventCreate(&e0));
    RF_HIP(hipEventCreate(&e1));
    // Three round-robin passes over the variants, 8 timed launches each,
    // keeping each variant's best pass: variants that differ by a few
    // percent are then not ordered by timing noise or clock drift.
    double v_ms[kStageVariants];
    const int nv = n_stage_variants(op->form);
    for (int v = 0; v < kStageVariants; ++v) v_ms[v] = 1e300;
    try {
      for (int pass = 0; pass < 3; ++pass)
      for (int v = 0; v < nv; ++v) {
        if (!variant_available(op->order, op->form, v, op->sa)) continue;
        StageArgs a = op->sa;
        a.variant = v;
        a.yin = op->y[op->cur].p;
        a.yin_lo = op->ylo[op->cur].p;
        a.yout = yo.p;
        a.yout_lo = ylo.p;
        a.p = pp.p;
        a.sa = kLsA[1];
        a.sb = kLsB[1];
        a.dt = 1e-3;
        a.src_kind = RFHIP_DRIVE_FREE;
        for (int w = 0; w < 2; ++w) launch_stage(op->order, op->form, 1, a, s);
        RF_HIP(hipEventRecord(e0, s));
        for (int w = 0; w < 8; ++w) launch_stage(op->order, op->form, 1, a, s);
        RF_HIP(hipEventRecord(e1, s));
        sync(op);
        float ms = 0.0f;
        RF_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms / 8.0 < v_ms[v]) v_ms[v] = ms / 8.0;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int best = op->sa.variant;
    for (int v = 0; v < nv; ++v)
      if (v_ms[v] < v_ms[best]) best = v;
    op->sa.variant = best;
    if (chosen) *chosen = best;
    if (ms_best) *ms_best = v_ms[best];
  });
}

int rfhip_op_state_load(rfhip_op *op, const double *y0) {
  return guarded([&] {
    RF_REQUIRE(op && y0, "null argument");
    state_from_host_f64(op, y0);
    sync(op);
  });
}

int rfhip_op_state_store(rfhip_op *op, double *y) {
  return guarded([&] {
    RF_REQUIRE(op && y, "null argument");
    state_to_host_f64(op, y);
  });
}

int rfhip_op_state_lserk(rfhip_op *op, double dt, int64_t nsteps) {
  return guarded([&] {
    RF_REQUIRE(op, "null argument");
    zero_p(op);
    for (int64_t k = 0; k < nsteps; ++k) lserk_step(op, dt, RFHIP_DRIVE_FREE, 0, 0.0, nullptr);
    sync(op);
  });
}

// Two free LSERK4 steps (ten stage launches, and the Debye kernels if any)
// captured into a graph: an even number of stages returns the fused forms
// to the state buffer they started from, so the graph for a starting buffer
// can be replayed back to back. Nothing in it is a memset node.
hipGraphExec_t step_graph(rfhip_op *op, double dt) {
  auto &g = op->step_graph[op->cur];
  if (g.exec && g.dt == dt && g.variant == op->sa.variant) return g.exec;
  if (g.exec) {
    RF_HIP(hipGraphExecDestroy(g.exec));
    g.exec = nullptr;
  }
  hipStream_t s = op->stream;
  const int cur0 = op->cur;
  hipGraph_t graph = nullptr;
  RF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    lserk_step(op, dt, RFHIP_DRIVE_FREE, 0, 0.0, nullptr);
    lserk_step(op, dt, RFHIP_DRIVE_FREE, 0, 0.0, nullptr);
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    op->cur = cur0;
    throw;
  }
  RF_HIP(hipStreamEndCapture(s, &graph));
  const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  RF_HIP(e);
  op->cur = cur0;  // capture did not run anything
  g.dt = dt;
  g.variant = op->sa.variant;
  return g.exec;
}

int rfhip_op_state_lserk_graph(rfhip_op *op, double dt, int64_t nsteps) {
  return guarded([&] {
    RF_REQUIRE(op && nsteps >= 0, "bad arguments");
    zero_p(op);
    hipStream_t s = op->stream;
    for (int64_t k = 0; k + 1 < nsteps; k += 2) RF_HIP(hipGraphLaunch(step_graph(op, dt), s));
    if (nsteps % 2) lserk_step(op, dt, RFHIP_DRIVE_FREE, 0, 0.0, nullptr);
    sync(op);
  });
}

int rfhip_batch_lserk(rfhip_op *const *ops, int32_t n_ops, const double *dt, int64_t nsteps) {
  return guarded([&] { batch_lserk_impl(ops, n_ops, dt, nsteps, nullptr); });
}

int rfhip_batch_lserk_ports(rfhip_op *const *ops, int32_t n_ops, const double *dt, int64_t nsteps,
                            const double *const *g) {
  return guarded([&] {
    RF_REQUIRE(g, "null waveform list");
    batch_lserk_impl(ops, n_ops, dt, nsteps, g);
  });
}

int rfhip_op_state_lserk_timed(rfhip_op *op, double dt, int64_t nsteps, double *total_ms, double *stage_ms_avg) {
  return guarded([&] {
    RF_REQUIRE(op && nsteps >= 1, "bad arguments");
    hipStream_t s = op->stream;
    std::vector<hipEvent_t> ev((size_t)10 * nsteps + 2, nullptr);
    struct Cleanup {
      std::vector<hipEvent_t> &v;
      ~Cleanup() {
        for (auto e : v)
          if (e) (void)hipEventDestroy(e);
      }
    } cleanup{ev};
    for (auto &e : ev) RF_HIP(hipEventCreate(&e));
    zero_p(op);
    RF_HIP(hipEventRecord(ev[10 * nsteps], s));
    for (int64_t k = 0; k < nsteps; ++k) lserk_step(op, dt, RFHIP_DRIVE_FREE, 0, 0.0, &ev[10 * k]);
    RF_HIP(hipEventRecord(ev[10 * nsteps + 1], s));
    sync(op);
    float ms = 0.0f;
    double kern = 0.0;
    for (int64_t q = 0; q < 5 * nsteps; ++q) {
      RF_HIP(hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
      kern += ms;
    }
    RF_HIP(hipEventElapsedTime(&ms, ev[10 * nsteps], ev[10 * nsteps + 1]));
    if (total_ms) *total_ms = ms;
    if (stage_ms_avg) *stage_ms_avg = kern / (double)(5 * nsteps);
  });
}

int rfhip_op_state_energy(rfhip_op *op, const double *mass_ref, const double *det, double *energy) {
  return guarded([&] {
    RF_REQUIRE(op && mass_ref && det && energy, "null argument");
    hipStream_t s = op->stream;
    const int np = op->np;
    const int64_t K = op->n_elem;
    const double *y = state_aos_f64(op);
    // per-element 1/eps and 1/mu in f64
    std::vector<double> hie(3 * K), him(3 * K);
    if (is_fused(op->form)) {
      std::vector<int32_t> mat(K);
      std::vector<double> tab(op->mat_tab.bytes / 8);
      RF_HIP(hipMemcpyAsync(mat.data(), op->mat.p, 4 * mat.size(), hipMemcpyDeviceToHost, s));
      RF_HIP(hipMemcpyAsync(tab.data(), op->mat_tab.p, op->mat_tab.bytes, hipMemcpyDeviceToHost, s));
      sync(op);
      for (int64_t e = 0; e < K; ++e)
        for (int c = 0; c < 3; ++c) {
          hie[3 * e + c] = tab[12 * mat[e] + c];
          him[3 * e + c] = tab[12 * mat[e] + 3 + c];
        }
    } else if (op->form == RFHIP_FORM_PLAIN_F64) {
      RF_HIP(hipMemcpyAsync(hie.data(), op->inv_eps.p, 24 * K, hipMemcpyDeviceToHost, s));
      RF_HIP(hipMemcpyAsync(him.data(), op->inv_mu.p, 24 * K, hipMemcpyDeviceToHost, s));
      sync(op);
    } else {
      std::vector<float> fe(3 * K), fmu(3 * K);
      RF_HIP(hipMemcpyAsync(fe.data(), op->inv_eps.p, 12 * K, hipMemcpyDeviceToHost, s));
      RF_HIP(hipMemcpyAsync(fmu.data(), op->inv_mu.p, 12 * K, hipMemcpyDeviceToHost, s));
      sync(op);
      for (int64_t q = 0; q < 3 * K; ++q) {
        hie[q] = fe[q];
        him[q] = fmu[q];
      }
    }
    DevBuf m, dd, ie, im, part;
    upload(m, mass_ref, (size_t)np * np, s);
    upload(dd, det, (size_t)K, s);
    upload(ie, hie.data(), hie.size(), s);
    upload(im, him.data(), him.size(), s);
    const int nblk = 1024;
    part.alloc(8 * (nblk + 1));
    launch_energy(y, m.as<double>(), dd.as<double>(), ie.as<double>(), im.as<double>(), np, K, part.as<double>(),
                  nblk, part.as<double>() + nblk, s);
    double half = 0.0;
    RF_HIP(hipMemcpyAsync(&half, part.as<double>() + nblk, 8, hipMemcpyDeviceToHost, s));
    sync(op);
    *energy = 0.5 * half;
  });
}

}  // extern "C"

// ---- sparse state-space export ----------------------------------------------

int rfhip_assemble_sparse(rfhip_ctx *ctx, const rfhip_op_desc *d, int64_t n_disp, const int64_t *disp_elem,
                          const double *disp_a, const double *disp_g, const double *disp_inv_eps_inf,
                          rfhip_csr **out) {
  return guarded([&] {
    RF_REQUIRE(ctx && d && out, "null argument");
    *out = nullptr;
    RF_REQUIRE(n_disp >= 0, "negative number of dispersive elements");
    RF_REQUIRE(n_disp == 0 || (disp_elem && disp_a && disp_g && disp_inv_eps_inf), "missing dispersive array");
    RF_REQUIRE(d->n_elem >= 1 && d->n_elem < ((int64_t)1 << 31), "element count out of range");
    const int p = order_of_np(d->np);
    RF_REQUIRE(p >= 1 && d->nfp == (p + 1) * (p + 2) / 2, "inconsistent Np / Nfp");
    if (p > 4) throw Error{RFHIP_EUNSUPPORTED, "orders 1..4 are compiled (Np 4, 10, 20, 35)"};
    RF_REQUIRE(d->diff_r && d->diff_s && d->diff_t && d->lift && d->face_nodes && d->jinv && d->inv_eps && d->inv_mu &&
                   d->sigma_eps && d->sigma_mu && d->face_normal && d->face_fscale && d->face_neighbor &&
                   d->face_nbr_local && d->face_port && d->face_perm,
               "missing operator array");
    const int64_t K = d->n_elem;
    const int np = d->np, nfp = d->nfp;
    // every index the row kernel follows (rhs.rs:1403-1418 on interior faces)
    for (int x = 0; x < 4 * nfp; ++x)
      RF_REQUIRE(d->face_nodes[x] >= 0 && d->face_nodes[x] < np, "face node out of range");
    for (int64_t q = 0; q < 4 * K; ++q) {
      const int nb = d->face_neighbor[q];
      RF_REQUIRE(nb >= -1 && nb < K, "face neighbour out of range");
      if (nb >= 0 && d->face_port[q] < 0) {
        RF_REQUIRE(d->face_nbr_local[q] >= 0 && d->face_nbr_local[q] < 4, "neighbour local face out of range");
        for (int m = 0; m < nfp; ++m)
          RF_REQUIRE(d->face_perm[q * nfp + m] >= 0 && d->face_perm[q * nfp + m] < nfp, "face perm out of range");
      }
    }
    std::vector<double> coef(3 * (size_t)n_disp);
    if (n_disp > 0) {  // rhs.rs:963-985
      std::vector<char> seen((size_t)K, 0);
      for (int64_t q = 0; q < n_disp; ++q) {
        const int64_t e = disp_elem[q];
        RF_REQUIRE(e >= 0 && e < K, "dispersive element index out of range");
        RF_REQUIRE(!seen[(size_t)e], "element listed twice as dispersive");
        seen[(size_t)e] = 1;
        coef[3 * q] = disp_a[q];
        coef[3 * q + 1] = disp_g[q];
        coef[3 * q + 2] = disp_inv_eps_inf[q];
      }
    }
    RF_HIP(hipSetDevice(ctx->device));
    auto csr = std::make_unique<rfhip_csr>();
    csr->ctx = ctx;
    RF_HIP(hipStreamCreateWithFlags(&csr->stream, hipStreamNonBlocking));
    assemble_sparse(csr.get(), *d, n_disp, disp_elem, coef.data());
    *out = csr.release();
  });
}

void rfhip_csr_destroy(rfhip_csr *csr) { delete csr; }

int rfhip_csr_shape(const rfhip_csr *csr, int64_t *n, int64_t *nnz) {
  return guarded([&] {
    RF_REQUIRE(csr, "null argument");
    if (n) *n = csr->n;
    if (nnz) *nnz = csr->nnz;
  });
}

int rfhip_csr_download(const rfhip_csr *csr, int64_t *row_ptr, int64_t *col_idx, double *values) {
  return guarded([&] {
    RF_REQUIRE(csr, "null argument");
    hipStream_t s = csr->stream;
    if (row_ptr) copy_d2h(row_ptr, csr->row_ptr.p, 8 * (size_t)(csr->n + 1), s);
    if (col_idx && csr->nnz) copy_d2h(col_idx, csr->col_idx.p, 8 * (size_t)csr->nnz, s);
    if (values && csr->nnz) copy_d2h(values, csr->values.p, 8 * (size_t)csr->nnz, s);
    RF_HIP(hipStreamSynchronize(s));
  });
}

int rfhip_csr_device(const rfhip_csr *csr, const int64_t **row_ptr, const int64_t **col_idx, const double **values) {
  return guarded([&] {
    RF_REQUIRE(csr, "null argument");
    if (row_ptr) *row_ptr = csr->row_ptr.as<int64_t>();
    if (col_idx) *col_idx = csr->col_idx.as<int64_t>();
    if (values) *values = csr->values.as<double>();
  });
}

int rfhip_csr_timing(const rfhip_csr *csr, double *count_ms, double *fill_ms) {
  return guarded([&] {
    RF_REQUIRE(csr, "null argument");
    if (count_ms) *count_ms = csr->count_ms;
    if (fill_ms) *fill_ms = csr->fill_ms;
  });
}
