// Host-side operator builder: the arrays GpuOperator::new reads from
// MaxwellOperator (crates/rapidfem-td/src/gpu/operator.rs:216-292), built
// from a tet mesh. Mirrors rhs.rs:805-1000 (MaxwellOperator::new),
// dg_basis.rs:61-385 (ReferenceElement::new), geom_factors.rs:33-88 and
// mesh_gen.rs:16-139, with its own construction:
//  * the reference element follows the reference's f64 operation sequence,
//    so D/M/LIFT equal the reference's tables bit for bit;
//  * faces are paired through a hash of their sorted vertex triple;
//  * the face-node permutation is found exactly from integer barycentric
//    coordinates on the shared triangle, O(Nfp) per face, instead of the
//    reference's O(Nfp^2) nearest-coordinate search (rhs.rs:546-568). For a
//    conforming mesh both give the same permutation.
#include "rfhost.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace {

// ---- reference element ------------------------------------------------

struct Ijk {
  int i, j, k;
};

// Node order of dg_basis.rs:270-280 (i outer, k inner); the state layout of
// every element follows it.
std::vector<Ijk> node_ijk(int p) {
  std::vector<Ijk> v;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= p - i; ++j)
      for (int k = 0; k <= p - i - j; ++k) v.push_back({i, j, k});
  return v;
}

// Monomial exponents (a, b, deg - a - b) for deg = 0..p, a ascending, then
// b ascending (dg_basis.rs:125-135). The order is part of the operation
// sequence that fixes the tables' round-off.
std::vector<std::array<int, 3>> monos3(int p) {
  std::vector<std::array<int, 3>> v;
  for (int d = 0; d <= p; ++d)
    for (int a = 0; a <= d; ++a)
      for (int b = 0; b <= d - a; ++b) v.push_back({a, b, d - a - b});
  return v;
}

// n! as a running product 1*1*2*...*n (dg_basis.rs:145-148).
double factorial(int n) {
  double r = 1.0;
  for (int k = 1; k <= n; ++k) r *= (double)k;
  return r;
}

// x^e the way f64::powi evaluates it for a run-time exponent (LLVM's powi:
// square-and-multiply over the exponent's bits), which is not always the
// same double as x*x*...*x.
double powi(double x, int e) {
  double r = 1.0;
  for (;;) {
    if (e & 1) r *= x;
    e /= 2;
    if (e == 0) break;
    x *= x;
  }
  return r;
}

// Gauss-Jordan inverse with partial pivoting (dg_basis.rs:171-214): the
// first row of largest |a| below the diagonal is the pivot, the pivot row is
// divided through, every other row with a nonzero factor is eliminated.
std::vector<double> gj_inverse(std::vector<double> a, int n) {
  std::vector<double> inv((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double best = std::fabs(a[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(a[(size_t)r * n + c]) > best) {
        best = std::fabs(a[(size_t)r * n + c]);
        piv = r;
      }
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)c * n + k], a[(size_t)piv * n + k]);
        std::swap(inv[(size_t)c * n + k], inv[(size_t)piv * n + k]);
      }
    const double d = a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] /= d;
      inv[(size_t)c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = a[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return inv;
}

// out = A B, rows of A accumulated left to right, zero entries of A skipped
// (dg_basis.rs:220-243).
std::vector<double> product(const std::vector<double> &a, const std::vector<double> &b, int rows, int inner,
                            int cols) {
  std::vector<double> out((size_t)rows * cols, 0.0);
  for (int i = 0; i < rows; ++i)
    for (int k = 0; k < inner; ++k) {
      const double aik = a[(size_t)i * inner + k];
      if (aik == 0.0) continue;
      for (int j = 0; j < cols; ++j) out[(size_t)i * cols + j] += aik * b[(size_t)k * cols + j];
    }
  return out;
}

// C^T G C as (C^T G) first, then the product above (dg_basis.rs:253-266).
std::vector<double> congruence(const std::vector<double> &c, const std::vector<double> &g, int n) {
  std::vector<double> ctg((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double acc = 0.0;
      for (int m = 0; m < n; ++m) acc += c[(size_t)m * n + i] * g[(size_t)m * n + k];
      ctg[(size_t)i * n + k] = acc;
    }
  return product(ctg, c, n, n, n);
}

struct RefElem {
  int p, np, nfp;
  std::vector<double> nodes, mass, dr, ds, dt, lift, fw;
  std::vector<int32_t> face_nodes;
};

// Face membership and in-face coordinates, aligned with TET_FACE_LOCAL
// (crates/rapidfem-core/src/mesh.rs:45-50): face 0 t=0, 1 r=0, 2 s=0,
// 3 r+s+t=1 (dg_basis.rs:295-312).
bool on_face(int f, const Ijk &q, int p) {
  switch (f) {
    case 0: return q.k == 0;
    case 1: return q.i == 0;
    case 2: return q.j == 0;
    default: return q.i + q.j + q.k == p;
  }
}
void face_uv(int f, const Ijk &q, int &u, int &v) {
  switch (f) {
    case 0: u = q.i; v = q.j; break;
    case 1: u = q.j; v = q.k; break;
    case 2: u = q.i; v = q.k; break;
    default: u = q.i; v = q.j; break;
  }
}

// ReferenceElement::new (dg_basis.rs:61-119) and build_lift (:288-385) in
// the reference's own f64 operation sequence, so that the tables carry the
// reference's round-off, residues of exact zeros included: which entries of
// A are stored by the sparse state-space export depends on them.
RefElem build_refelem(int p) {
  RefElem re;
  re.p = p;
  const auto ijk = node_ijk(p);
  const auto mon = monos3(p);
  const int n = (int)ijk.size();
  re.np = n;
  re.nfp = (p + 1) * (p + 2) / 2;
  const int nfp = re.nfp;
  const double pp = (double)p;
  re.nodes.resize((size_t)n * 3);
  for (int a = 0; a < n; ++a) {
    re.nodes[a * 3] = (double)ijk[a].i / pp;
    re.nodes[a * 3 + 1] = (double)ijk[a].j / pp;
    re.nodes[a * 3 + 2] = (double)ijk[a].k / pp;
  }
  auto mono = [&](const std::array<int, 3> &m, const double *x) {
    return powi(x[0], m[0]) * powi(x[1], m[1]) * powi(x[2], m[2]);
  };
  // V[node][mono]; the nodal basis is phi_i = sum_m C[m][i] mono_m, C = V^-1.
  std::vector<double> V((size_t)n * n);
  for (int a = 0; a < n; ++a)
    for (int m = 0; m < n; ++m) V[(size_t)a * n + m] = mono(mon[m], &re.nodes[a * 3]);
  const std::vector<double> C = gj_inverse(V, n);
  // M = C^T G C with the exact monomial integrals over the reference tet.
  std::vector<double> G((size_t)n * n);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) {
      const int x = mon[a][0] + mon[b][0], y = mon[a][1] + mon[b][1], z = mon[a][2] + mon[b][2];
      G[(size_t)a * n + b] = factorial(x) * factorial(y) * factorial(z) / factorial(x + y + z + 3);
    }
  re.mass = congruence(C, G, n);
  const std::vector<double> Minv = gj_inverse(re.mass, n);
  // D_axis = V_axis C, V_axis holding the monomials' derivatives.
  std::vector<double> *D[3] = {&re.dr, &re.ds, &re.dt};
  for (int ax = 0; ax < 3; ++ax) {
    std::vector<double> Vd((size_t)n * n);
    for (int a = 0; a < n; ++a)
      for (int m = 0; m < n; ++m) {
        const int e = mon[m][ax];
        if (e == 0) {
          Vd[(size_t)a * n + m] = 0.0;
          continue;
        }
        std::array<int, 3> lower = mon[m];
        lower[ax] -= 1;
        Vd[(size_t)a * n + m] = (double)e * mono(lower, &re.nodes[a * 3]);
      }
    *D[ax] = product(Vd, C, n, n, n);
  }
  // Face nodes, face mass matrices scattered into E (Np x 4Nfp), LIFT = M^-1 E.
  re.face_nodes.assign(4 * nfp, 0);
  const int cols = 4 * nfp;
  std::vector<double> E((size_t)n * cols, 0.0);
  re.fw.assign(4 * nfp, 0.0);
  std::vector<std::array<int, 2>> mon2;
  for (int d = 0; d <= p; ++d)
    for (int a = 0; a <= d; ++a) mon2.push_back({a, d - a});
  for (int f = 0; f < 4; ++f) {
    int c = 0;
    for (int a = 0; a < n; ++a)
      if (on_face(f, ijk[a], p)) re.face_nodes[f * nfp + c++] = a;
    std::vector<double> V2((size_t)nfp * nfp);
    for (int a = 0; a < nfp; ++a) {
      int u, v;
      face_uv(f, ijk[re.face_nodes[f * nfp + a]], u, v);
      const double x = (double)u / pp, y = (double)v / pp;
      for (int m = 0; m < nfp; ++m) V2[(size_t)a * nfp + m] = powi(x, mon2[m][0]) * powi(y, mon2[m][1]);
    }
    const std::vector<double> C2 = gj_inverse(V2, nfp);
    std::vector<double> G2((size_t)nfp * nfp);
    for (int a = 0; a < nfp; ++a)
      for (int b = 0; b < nfp; ++b) {
        const int x = mon2[a][0] + mon2[b][0], y = mon2[a][1] + mon2[b][1];
        G2[(size_t)a * nfp + b] = factorial(x) * factorial(y) / factorial(x + y + 2);
      }
    const std::vector<double> Mf = congruence(C2, G2, nfp);
    for (int a = 0; a < nfp; ++a) {
      double rs = 0.0;
      for (int m = 0; m < nfp; ++m) {
        E[(size_t)re.face_nodes[f * nfp + a] * cols + f * nfp + m] = Mf[(size_t)a * nfp + m];
        rs += Mf[(size_t)a * nfp + m];
      }
      re.fw[f * nfp + a] = rs;
    }
  }
  re.lift = product(Minv, E, n, n, cols);
  return re;
}

// ---- geometry helpers (crates/rapidfem-core/src/geom.rs:17-39) ----------

inline void sub(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
inline double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void cross(const double *a, const double *b, double *o) {
  double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  o[0] = x;
  o[1] = y;
  o[2] = z;
}

// geom_factors.rs:33-48 + :81-88: J columns v_k - v0, J^-1 = adj / det.
void geom(const double *v, double *jac, double *jinv, double *det) {
  double c0[3], c1[3], c2[3];
  sub(v + 3, v, c0);
  sub(v + 6, v, c1);
  sub(v + 9, v, c2);
  const double m[3][3] = {{c0[0], c1[0], c2[0]}, {c0[1], c1[1], c2[1]}, {c0[2], c1[2], c2[2]}};
  double adj[3][3];
  cross(m[1], m[2], adj[0]);
  cross(m[2], m[0], adj[1]);
  cross(m[0], m[1], adj[2]);
  const double d = dot(m[0], adj[0]);
  const double id = 1.0 / d;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (jac) jac[i * 3 + j] = m[i][j];
      jinv[i * 3 + j] = adj[j][i] * id;
    }
  *det = d;
}

// TET_FACE_LOCAL (crates/rapidfem-core/src/mesh.rs:45-50) and the vertex
// opposite each local face (crates/rapidfem-core/src/topology.rs:19).
constexpr int kFaceVerts[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};
constexpr int kFaceOpp[4] = {3, 1, 2, 0};

// Integer barycentric weight of reference node (i,j,k) on tet vertex a:
// vertex 0 = (0,0,0), 1 = (1,0,0), 2 = (0,1,0), 3 = (0,0,1).
inline int bary(const Ijk &q, int a, int p) {
  switch (a) {
    case 0: return p - q.i - q.j - q.k;
    case 1: return q.i;
    case 2: return q.j;
    default: return q.k;
  }
}

}  // namespace

extern "C" {

int32_t rfhost_np(int32_t p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
int32_t rfhost_nfp(int32_t p) { return (p + 1) * (p + 2) / 2; }

int rfhost_refelem(int32_t p, double *nodes, double *mass, double *dr, double *ds, double *dt,
                   int32_t *face_nodes, double *lift, double *face_weights) {
  if (p < 1) return 1;
  RefElem re = build_refelem(p);
  const size_t n = re.np, nfp = re.nfp;
  if (nodes) std::memcpy(nodes, re.nodes.data(), sizeof(double) * n * 3);
  if (mass) std::memcpy(mass, re.mass.data(), sizeof(double) * n * n);
  if (dr) std::memcpy(dr, re.dr.data(), sizeof(double) * n * n);
  if (ds) std::memcpy(ds, re.ds.data(), sizeof(double) * n * n);
  if (dt) std::memcpy(dt, re.dt.data(), sizeof(double) * n * n);
  if (face_nodes) std::memcpy(face_nodes, re.face_nodes.data(), sizeof(int32_t) * 4 * nfp);
  if (lift) std::memcpy(lift, re.lift.data(), sizeof(double) * n * 4 * nfp);
  if (face_weights) std::memcpy(face_weights, re.fw.data(), sizeof(double) * 4 * nfp);
  return 0;
}

// mesh_gen.rs:124-139: splitmix-style hash of (i, j, k, seed) -> [-1, 1]^3.
// The constants and mixing steps define the jittered test meshes, so they
// are reproduced exactly.
static double jitter_mix(uint64_t x) {
  x *= 0x9E3779B97F4A7C15ULL;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBULL;
  x ^= x >> 31;
  return ((double)x / (double)UINT64_MAX) * 2.0 - 1.0;
}

void rfhost_box_mesh(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                     double amplitude, uint64_t seed, int32_t jittered, double *xyz, int64_t *tets) {
  const double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  int64_t q = 0;
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i <= nx; ++i, ++q) {
        double *p = xyz + 3 * q;
        if (!jittered) {  // mesh_gen.rs:64-72
          p[0] = lx * (double)i / (double)nx;
          p[1] = ly * (double)j / (double)ny;
          p[2] = lz * (double)k / (double)nz;
          continue;
        }
        p[0] = hx * (double)i;  // mesh_gen.rs:100-117
        p[1] = hy * (double)j;
        p[2] = hz * (double)k;
        const uint64_t base = seed + (uint64_t)i * 0x01000193ULL + (uint64_t)j * 0x00000100000001B3ULL +
                              (uint64_t)k * 0xC2B2AE3D27D4EB4FULL;
        if (i != 0 && i != nx) p[0] += amplitude * hx * jitter_mix(base);
        if (j != 0 && j != ny) p[1] += amplitude * hy * jitter_mix(base ^ 0xAAAAULL);
        if (k != 0 && k != nz) p[2] += amplitude * hz * jitter_mix(base ^ 0x5555ULL);
      }
  // Kuhn split through the 0-7 diagonal, corner c = x + 2y + 4z
  // (mesh_gen.rs:16-23), each tet's vertex ids sorted (mesh_gen.rs:44).
  static const int kuhn[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
  const int64_t nxp = nx + 1, nyp = ny + 1;
  int64_t t = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        for (int s = 0; s < 6; ++s, ++t) {
          int64_t v[4];
          for (int a = 0; a < 4; ++a) {
            const int c = kuhn[s][a];
            v[a] = ((k + ((c >> 2) & 1)) * nyp + (j + ((c >> 1) & 1))) * nxp + (i + (c & 1));
          }
          std::sort(v, v + 4);
          std::memcpy(tets + 4 * t, v, sizeof v);
        }
}

int rfhost_operator_arrays(int64_t n_nodes, const double *xyz, int64_t K, const int64_t *tets, int32_t order,
                           const double *eps, const double *mu, const double *sigma, const double *sigma_m,
                           const int32_t *face_port, double *jinv, double *det, double *normal, double *fscale,
                           int32_t *nbr, int32_t *nbr_lf, int32_t *perm, int32_t *port, double *inv_eps,
                           double *inv_mu, double *sigma_eps, double *sigma_mu) {
  (void)n_nodes;
  const int p = order;
  const int nfp = (p + 1) * (p + 2) / 2;
  const auto ijk = node_ijk(p);
  std::vector<int> fnodes(4 * nfp);
  {
    int c[4] = {0, 0, 0, 0};
    for (int a = 0; a < (int)ijk.size(); ++a)
      for (int f = 0; f < 4; ++f)
        if (on_face(f, ijk[a], p)) fnodes[f * nfp + c[f]++] = a;
  }
  // Geometric factors.
  for (int64_t e = 0; e < K; ++e) {
    double v[12];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) v[3 * a + d] = xyz[3 * tets[4 * e + a] + d];
    geom(v, nullptr, jinv + 9 * e, det + e);
  }
  // Face pairing through a hash of the sorted vertex triple.
  std::unordered_map<uint64_t, int64_t> open;
  open.reserve((size_t)(2 * K + 16));
  auto sorted_face = [&](int64_t e, int f, int64_t *s) {
    s[0] = tets[4 * e + kFaceVerts[f][0]];
    s[1] = tets[4 * e + kFaceVerts[f][1]];
    s[2] = tets[4 * e + kFaceVerts[f][2]];
    std::sort(s, s + 3);
  };
  auto key_of = [&](int64_t e, int f) -> uint64_t {
    int64_t s[3];
    sorted_face(e, f, s);
    // 21 bits per vertex id (meshes up to 2M nodes); larger meshes fall
    // back to a mixing hash of the full ids checked on match below.
    if (s[2] < (1 << 21)) return (uint64_t)s[0] << 42 | (uint64_t)s[1] << 21 | (uint64_t)s[2];
    uint64_t h = (uint64_t)s[0] * 0x9E3779B97F4A7C15ULL;
    h ^= (uint64_t)s[1] + 0x7F4A7C159E3779B9ULL + (h << 6) + (h >> 2);
    h ^= (uint64_t)s[2] + 0x94D049BB133111EBULL + (h << 6) + (h >> 2);
    return h | (1ULL << 63);
  };
  for (int64_t q = 0; q < 4 * K; ++q) {
    nbr[q] = -1;
    nbr_lf[q] = -1;
  }
  for (int64_t e = 0; e < K; ++e)
    for (int f = 0; f < 4; ++f) {
      const uint64_t key = key_of(e, f);
      auto it = open.find(key);
      if (it == open.end()) {
        open.emplace(key, 4 * e + f);
        continue;
      }
      const int64_t other = it->second;
      if (other < 0) return 3;  // a third tet on this face
      const int64_t eo = other / 4;
      const int fo = (int)(other % 4);
      int64_t sa[3], sb[3];
      sorted_face(e, f, sa);
      sorted_face(eo, fo, sb);
      if (sa[0] != sb[0] || sa[1] != sb[1] || sa[2] != sb[2]) return 3;  // hash collision: refuse
      nbr[4 * e + f] = (int32_t)eo;
      nbr_lf[4 * e + f] = fo;
      nbr[other] = (int32_t)e;
      nbr_lf[other] = f;
      it->second = -1;
    }
  // Per-face normal, fscale, port and perm.
  // Face-node lookup by barycentric weights on the face's own vertex slots:
  // index[f][(w_slot1) * (p+1) + w_slot2] -> face-node position.
  std::vector<int> lookup(4 * (p + 1) * (p + 1), -1);
  for (int f = 0; f < 4; ++f)
    for (int m = 0; m < nfp; ++m) {
      const Ijk &q = ijk[fnodes[f * nfp + m]];
      lookup[f * (p + 1) * (p + 1) + bary(q, kFaceVerts[f][1], p) * (p + 1) + bary(q, kFaceVerts[f][2], p)] = m;
    }
  for (int64_t e = 0; e < K; ++e)
    for (int f = 0; f < 4; ++f) {
      const int64_t q = 4 * e + f;
      // topology.rs:56-73: normal from the face's first corner, flipped away
      // from the opposite vertex; area = |cross| / 2.
      const double *v0 = xyz + 3 * tets[4 * e + kFaceVerts[f][0]];
      const double *v1 = xyz + 3 * tets[4 * e + kFaceVerts[f][1]];
      const double *v2 = xyz + 3 * tets[4 * e + kFaceVerts[f][2]];
      const double *vo = xyz + 3 * tets[4 * e + kFaceOpp[f]];
      double d1[3], d2[3], n[3], dv[3];
      sub(v1, v0, d1);
      sub(v2, v0, d2);
      cross(d1, d2, n);
      const double len = std::sqrt(dot(n, n));
      for (int k = 0; k < 3; ++k) n[k] /= len;
      sub(vo, v0, dv);
      if (dot(dv, n) > 0.0)
        for (int k = 0; k < 3; ++k) n[k] = -n[k];
      std::memcpy(normal + 3 * q, n, sizeof n);
      fscale[q] = 2.0 * (0.5 * len) / std::fabs(det[e]);  // rhs.rs:829
      int32_t pt = face_port ? face_port[q] : -1;
      if (nbr[q] >= 0) pt = -1;
      port[q] = pt;
      for (int m = 0; m < nfp; ++m) perm[q * nfp + m] = 0;
      if (nbr[q] < 0) continue;
      const int64_t en = nbr[q];
      const int fn = nbr_lf[q];
      // Global vertex ids on each slot of both faces.
      int64_t g[3], h[3];
      for (int s = 0; s < 3; ++s) {
        g[s] = tets[4 * e + kFaceVerts[f][s]];
        h[s] = tets[4 * en + kFaceVerts[fn][s]];
      }
      int slot_in_nbr[3];
      for (int s = 0; s < 3; ++s) {
        slot_in_nbr[s] = -1;
        for (int t = 0; t < 3; ++t)
          if (h[t] == g[s]) slot_in_nbr[s] = t;
        if (slot_in_nbr[s] < 0) return 2;
      }
      for (int m = 0; m < nfp; ++m) {
        const Ijk &qn = ijk[fnodes[f * nfp + m]];
        int w[3];  // weight on the neighbour's slots
        for (int s = 0; s < 3; ++s) w[slot_in_nbr[s]] = bary(qn, kFaceVerts[f][s], p);
        const int m2 = lookup[fn * (p + 1) * (p + 1) + w[1] * (p + 1) + w[2]];
        if (m2 < 0) return 2;
        perm[q * nfp + m] = m2;
      }
    }
  // rhs.rs:908-926 per-element diagonal materials.
  for (int64_t e = 0; e < K; ++e)
    for (int c = 0; c < 3; ++c) {
      const double ep = eps ? eps[3 * e + c] : 1.0, m = mu ? mu[3 * e + c] : 1.0;
      const double s = sigma ? sigma[e] : 0.0, sm = sigma_m ? sigma_m[e] : 0.0;
      inv_eps[3 * e + c] = 1.0 / ep;
      inv_mu[3 * e + c] = 1.0 / m;
      sigma_eps[3 * e + c] = s / ep;
      sigma_mu[3 * e + c] = sm / m;
    }
  return 0;
}

void rfhost_node_coords(int64_t K, const double *xyz, const int64_t *tets, int32_t p, double *out) {
  const auto ijk = node_ijk(p);
  const int np = (int)ijk.size();
  for (int64_t e = 0; e < K; ++e) {
    const double *v0 = xyz + 3 * tets[4 * e];
    double jc[3][3];
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < 3; ++i) jc[i][k] = xyz[3 * tets[4 * e + k + 1] + i] - v0[i];
    for (int a = 0; a < np; ++a) {
      const double xi[3] = {(double)ijk[a].i / p, (double)ijk[a].j / p, (double)ijk[a].k / p};
      double *o = out + 3 * (e * np + a);
      for (int i = 0; i < 3; ++i) {
        double x = v0[i];
        for (int k = 0; k < 3; ++k) x += jc[i][k] * xi[k];
        o[i] = x;
      }
    }
  }
}

}  // extern "C"
