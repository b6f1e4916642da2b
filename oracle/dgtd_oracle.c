/*
 * dgtd_oracle.c - CPU FP64 restatement of rapidfem-td's DGTD Maxwell path.
 *
 * TEST INFRASTRUCTURE ONLY (see dgtd_oracle.h). Written in plain C from the
 * reference's Rust sources; every function cites the file:line it follows,
 * relative to the reference root. Loop orders follow the reference so the
 * rounding matches it as closely as C allows (build with
 * -ffp-contract=off: the reference's Rust never fuses multiply-adds).
 */
#include "dgtd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

/* llvm.powi semantics (compiler-rt __powidf2): binary exponentiation, the
 * routine Rust's f64::powi lowers to for a run-time exponent. */
static double powi_(double a, int b) {
  const int recip = b < 0;
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

static void sub3(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0];
  o[1] = a[1] - b[1];
  o[2] = a[2] - b[2];
}
/* crates/rapidfem-core/src/geom.rs:27-34 */
static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static void cross3(const double *a, const double *b, double *o) {
  double t0 = a[1] * b[2] - a[2] * b[1];
  double t1 = a[2] * b[0] - a[0] * b[2];
  double t2 = a[0] * b[1] - a[1] * b[0];
  o[0] = t0;
  o[1] = t1;
  o[2] = t2;
}

static int g_threads = 0;
void orc_op_set_threads(int nthreads) { g_threads = nthreads; }
static int nthreads_(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* reference element: crates/rapidfem-td/src/dg_basis.rs                    */
/* ------------------------------------------------------------------------ */

int orc_np(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
int orc_nfp(int p) { return (p + 1) * (p + 2) / 2; }

/* dg_basis.rs:123-133 monomials(p) */
static int monomials_(int p, int (*out)[3]) {
  int n = 0;
  for (int deg = 0; deg <= p; ++deg)
    for (int a = 0; a <= deg; ++a)
      for (int b = 0; b <= deg - a; ++b) {
        out[n][0] = a;
        out[n][1] = b;
        out[n][2] = deg - a - b;
        ++n;
      }
  return n;
}

/* dg_basis.rs:270-280 equispaced_ijk(p) */
static int equispaced_ijk_(int p, int (*out)[3]) {
  int n = 0;
  for (int i = 0; i <= p; ++i)
    for (int j = 0; j <= p - i; ++j)
      for (int k = 0; k <= p - i - j; ++k) {
        out[n][0] = i;
        out[n][1] = j;
        out[n][2] = k;
        ++n;
      }
  return n;
}

/* dg_basis.rs:145-148 factorial */
static double factorial_(int n) {
  double r = 1.0;
  for (int k = 1; k <= n; ++k) r *= (double)k;
  return r;
}
/* dg_basis.rs:151-153 */
static double mono_integral_(int a, int b, int c) {
  return factorial_(a) * factorial_(b) * factorial_(c) / factorial_(a + b + c + 3);
}
/* dg_basis.rs:283-285 */
static double tri_integral_(int a, int b) {
  return factorial_(a) * factorial_(b) / factorial_(a + b + 2);
}
/* dg_basis.rs:156-158 */
static double eval_mono_(const int *m, const double *x) {
  return powi_(x[0], m[0]) * powi_(x[1], m[1]) * powi_(x[2], m[2]);
}
/* dg_basis.rs:161-169 */
static double eval_mono_d_(const int *m, const double *x, int axis) {
  int e = m[axis];
  if (e == 0) return 0.0;
  int d[3] = {m[0], m[1], m[2]};
  d[axis] -= 1;
  return (double)e * eval_mono_(d, x);
}

/* dg_basis.rs:171-214 Gauss-Jordan inverse with partial pivoting */
static void invert_(const double *src, int n, double *inv) {
  double *a = (double *)malloc(sizeof(double) * n * n);
  memcpy(a, src, sizeof(double) * n * n);
  memset(inv, 0, sizeof(double) * n * n);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = fabs(a[col * n + col]);
    for (int r = col + 1; r < n; ++r) {
      double v = fabs(a[r * n + col]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (piv != col) {
      for (int k = 0; k < n; ++k) {
        double t = a[col * n + k];
        a[col * n + k] = a[piv * n + k];
        a[piv * n + k] = t;
        t = inv[col * n + k];
        inv[col * n + k] = inv[piv * n + k];
        inv[piv * n + k] = t;
      }
    }
    double d = a[col * n + col];
    for (int k = 0; k < n; ++k) {
      a[col * n + k] /= d;
      inv[col * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == col) continue;
      double f = a[r * n + col];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        a[r * n + k] -= f * a[col * n + k];
        inv[r * n + k] -= f * inv[col * n + k];
      }
    }
  }
  free(a);
}

/* dg_basis.rs:220-243 matmul_into, skipping zero entries of A */
static void matmul_(const double *a, const double *b, int rows, int inner,
                    int cols, double *out) {
  memset(out, 0, sizeof(double) * rows * cols);
  for (int i = 0; i < rows; ++i)
    for (int k = 0; k < inner; ++k) {
      double aik = a[i * inner + k];
      if (aik == 0.0) continue;
      for (int j = 0; j < cols; ++j) out[i * cols + j] += aik * b[k * cols + j];
    }
}

/* dg_basis.rs:253-266 C^T G C */
static void triple_product_(const double *c, const double *g, int n,
                            double *out) {
  double *ctg = (double *)malloc(sizeof(double) * n * n);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double acc = 0.0;
      for (int m = 0; m < n; ++m) acc += c[m * n + i] * g[m * n + k];
      ctg[i * n + k] = acc;
    }
  matmul_(ctg, c, n, n, n, out);
  free(ctg);
}

/* dg_basis.rs:61-119 ReferenceElement::new and :288-385 build_lift */
int orc_refelem(int p, double *nodes, double *mass, double *dr, double *ds,
                double *dt, int *face_nodes, double *lift, double *fweights) {
  if (p < 1) return -1;
  const int n = orc_np(p);
  const int nfp = orc_nfp(p);
  int(*monos)[3] = malloc(sizeof(int[3]) * n);
  int(*ijk)[3] = malloc(sizeof(int[3]) * n);
  monomials_(p, monos);
  equispaced_ijk_(p, ijk);
  double *xyz = malloc(sizeof(double) * n * 3);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) xyz[i * 3 + d] = (double)ijk[i][d] / (double)p;

  double *vander = malloc(sizeof(double) * n * n);
  for (int ni = 0; ni < n; ++ni)
    for (int mi = 0; mi < n; ++mi) vander[ni * n + mi] = eval_mono_(monos[mi], &xyz[ni * 3]);
  double *c = malloc(sizeof(double) * n * n);
  invert_(vander, n, c);

  double *gram = malloc(sizeof(double) * n * n);
  for (int mi = 0; mi < n; ++mi)
    for (int qi = 0; qi < n; ++qi)
      gram[mi * n + qi] = mono_integral_(monos[mi][0] + monos[qi][0],
                                         monos[mi][1] + monos[qi][1],
                                         monos[mi][2] + monos[qi][2]);
  double *m = malloc(sizeof(double) * n * n);
  triple_product_(c, gram, n, m);
  double *minv = malloc(sizeof(double) * n * n);
  invert_(m, n, minv);

  double *vd[3];
  for (int a = 0; a < 3; ++a) vd[a] = malloc(sizeof(double) * n * n);
  for (int ni = 0; ni < n; ++ni)
    for (int mi = 0; mi < n; ++mi)
      for (int a = 0; a < 3; ++a) vd[a][ni * n + mi] = eval_mono_d_(monos[mi], &xyz[ni * 3], a);
  double *dd[3] = {dr, ds, dt};
  for (int a = 0; a < 3; ++a) matmul_(vd[a], c, n, n, n, dd[a]);

  /* build_lift, dg_basis.rs:288-385 */
  for (int f = 0; f < 4; ++f) {
    int cnt = 0;
    for (int idx = 0; idx < n; ++idx) {
      const int *q = ijk[idx];
      int on = 0;
      switch (f) {
        case 0: on = q[2] == 0; break;
        case 1: on = q[0] == 0; break;
        case 2: on = q[1] == 0; break;
        default: on = q[0] + q[1] + q[2] == p; break;
      }
      if (on) face_nodes[f * nfp + cnt++] = idx;
    }
    if (cnt != nfp) return -2;
  }
  int(*monos2)[2] = malloc(sizeof(int[2]) * nfp);
  {
    int k = 0;
    for (int deg = 0; deg <= p; ++deg)
      for (int a = 0; a <= deg; ++a) {
        monos2[k][0] = a;
        monos2[k][1] = deg - a;
        ++k;
      }
  }
  const int cols = 4 * nfp;
  double *emat = calloc((size_t)n * cols, sizeof(double));
  double *v2 = malloc(sizeof(double) * nfp * nfp);
  double *c2 = malloc(sizeof(double) * nfp * nfp);
  double *g2 = malloc(sizeof(double) * nfp * nfp);
  double *mf = malloc(sizeof(double) * nfp * nfp);
  for (int f = 0; f < 4; ++f) {
    for (int i = 0; i < nfp; ++i) {
      const int *q = ijk[face_nodes[f * nfp + i]];
      int u, v;
      switch (f) {
        case 0: u = q[0]; v = q[1]; break;
        case 1: u = q[1]; v = q[2]; break;
        case 2: u = q[0]; v = q[2]; break;
        default: u = q[0]; v = q[1]; break;
      }
      double x = (double)u / (double)p, y = (double)v / (double)p;
      for (int j = 0; j < nfp; ++j)
        v2[i * nfp + j] = powi_(x, monos2[j][0]) * powi_(y, monos2[j][1]);
    }
    invert_(v2, nfp, c2);
    for (int i = 0; i < nfp; ++i)
      for (int j = 0; j < nfp; ++j)
        g2[i * nfp + j] = tri_integral_(monos2[i][0] + monos2[j][0], monos2[i][1] + monos2[j][1]);
    triple_product_(c2, g2, nfp, mf);
    for (int a = 0; a < nfp; ++a) {
      int vi = face_nodes[f * nfp + a];
      for (int mm = 0; mm < nfp; ++mm) emat[vi * cols + f * nfp + mm] = mf[a * nfp + mm];
    }
    for (int a = 0; a < nfp; ++a) {
      double s = 0.0;
      for (int mm = 0; mm < nfp; ++mm) s += mf[a * nfp + mm];
      fweights[f * nfp + a] = s;
    }
  }
  matmul_(minv, emat, n, n, cols, lift);
  memcpy(nodes, xyz, sizeof(double) * n * 3);
  memcpy(mass, m, sizeof(double) * n * n);

  free(monos); free(ijk); free(xyz); free(vander); free(c); free(gram);
  free(m); free(minv);
  for (int a = 0; a < 3; ++a) free(vd[a]);
  free(monos2); free(emat); free(v2); free(c2); free(g2); free(mf);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* mesher: crates/rapidfem-td/src/mesh_gen.rs                               */
/* ------------------------------------------------------------------------ */

/* mesh_gen.rs:124-139 jitter_offset */
static double jitter_mix_(uint64_t x) {
  x *= 0x9E3779B97F4A7C15ULL;
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ULL;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBULL;
  x ^= x >> 31;
  return ((double)x / (double)UINT64_MAX) * 2.0 - 1.0;
}
void orc_jitter_offset(int64_t i, int64_t j, int64_t k, uint64_t seed,
                       double *off) {
  uint64_t base = seed + (uint64_t)i * 0x01000193ULL +
                  (uint64_t)j * 0x00000100000001B3ULL +
                  (uint64_t)k * 0xC2B2AE3D27D4EB4FULL;
  off[0] = jitter_mix_(base);
  off[1] = jitter_mix_(base ^ 0xAAAAULL);
  off[2] = jitter_mix_(base ^ 0x5555ULL);
}

/* mesh_gen.rs:54-77 structured_box and :86-121 structured_box_jittered */
void orc_box_nodes(int nx, int ny, int nz, double lx, double ly, double lz,
                   double amplitude, uint64_t seed, int jittered, double *xyz) {
  int64_t q = 0;
  double hx = lx / nx, hy = ly / ny, hz = lz / nz;
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double *p = &xyz[3 * q++];
        if (!jittered) {
          p[0] = lx * (double)i / (double)nx;
          p[1] = ly * (double)j / (double)ny;
          p[2] = lz * (double)k / (double)nz;
        } else {
          p[0] = hx * (double)i;
          p[1] = hy * (double)j;
          p[2] = hz * (double)k;
          double off[3];
          orc_jitter_offset(i, j, k, seed, off);
          if (i != 0 && i != nx) p[0] += amplitude * hx * off[0];
          if (j != 0 && j != ny) p[1] += amplitude * hy * off[1];
          if (k != 0 && k != nz) p[2] += amplitude * hz * off[2];
        }
      }
}

/* mesh_gen.rs:16-23 KUHN, :26-51 kuhn_tets (vertices sorted per tet) */
static const int KUHN_[6][4] = {{0, 1, 3, 7}, {0, 3, 2, 7}, {0, 2, 6, 7},
                                {0, 6, 4, 7}, {0, 4, 5, 7}, {0, 5, 1, 7}};
void orc_kuhn_tets(int nx, int ny, int nz, int64_t *tets) {
  const int64_t nxp = nx + 1, nyp = ny + 1;
  int64_t t = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int s = 0; s < 6; ++s) {
          int64_t v[4];
          for (int a = 0; a < 4; ++a) {
            int c = KUHN_[s][a];
            v[a] = ((int64_t)(k + ((c >> 2) & 1)) * nyp + (j + ((c >> 1) & 1))) * nxp +
                   (i + (c & 1));
          }
          for (int a = 1; a < 4; ++a) /* sort_unstable on 4 distinct ids */
            for (int b = a; b > 0 && v[b - 1] > v[b]; --b) {
              int64_t tmp = v[b];
              v[b] = v[b - 1];
              v[b - 1] = tmp;
            }
          for (int a = 0; a < 4; ++a) tets[4 * t + a] = v[a];
          ++t;
        }
}

/* ------------------------------------------------------------------------ */
/* geometric factors: crates/rapidfem-td/src/geom_factors.rs                */
/* ------------------------------------------------------------------------ */

/* geom_factors.rs:33-48 for_tet and :81-88 inv3 */
void orc_geom_factors(const double *v, double *jac, double *jinv, double *det) {
  double c0[3], c1[3], c2[3];
  sub3(&v[3], &v[0], c0);
  sub3(&v[6], &v[0], c1);
  sub3(&v[9], &v[0], c2);
  double m[3][3] = {{c0[0], c1[0], c2[0]}, {c0[1], c1[1], c2[1]}, {c0[2], c1[2], c2[2]}};
  double adj[3][3];
  cross3(m[1], m[2], adj[0]);
  cross3(m[2], m[0], adj[1]);
  cross3(m[0], m[1], adj[2]);
  double d = dot3(m[0], adj[0]);
  double id = 1.0 / d;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      if (jac) jac[i * 3 + j] = m[i][j];
      jinv[i * 3 + j] = adj[j][i] * id;
    }
  *det = d;
}

/* geom_factors.rs:51-60 GeometricFactors::map */
static void geom_map_(const double *origin, const double *jac, const double *xi,
                      double *x) {
  x[0] = origin[0];
  x[1] = origin[1];
  x[2] = origin[2];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) x[i] += jac[i * 3 + k] * xi[k];
}

/* ------------------------------------------------------------------------ */
/* operator: crates/rapidfem-td/src/rhs.rs                                  */
/* ------------------------------------------------------------------------ */

struct orc_op {
  int p, np, nfp;
  int64_t K;
  double *nodes, *mass, *dr, *ds, *dt, *lift, *fweights;
  int *face_nodes;
  double *origin, *jac, *jinv, *det; /* K*3, K*9, K*9, K */
  double *normal, *fscale;           /* 4K*3, 4K */
  int64_t *nbr;                      /* 4K, -1 boundary */
  int *nbr_lf, *port, *perm;         /* 4K, 4K, 4K*nfp */
  double *inv_eps, *inv_mu, *sig_eps, *sig_mu; /* K*3 */
  double alpha;
  /* rhs.rs:469-487 DispersiveElem, one per P-block slot, and the element ->
   * slot map (-1: not dispersive); empty when the operator is not dispersive */
  int64_t n_disp;
  int64_t *disp_elem, *disp_slot;
  double *disp_a, *disp_g, *disp_iei;
};

/* rhs.rs:39-85 element_curl */
static void element_curl_(const orc_op *op, const double *jinv,
                          const double *field, double *out, double *rd,
                          double *pd) {
  const int n = op->np;
  const double *dref[3] = {op->dr, op->ds, op->dt};
  for (int comp = 0; comp < 3; ++comp) {
    for (int k = 0; k < 3; ++k)
      for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) acc += dref[k][i * n + j] * field[j * 3 + comp];
        rd[k * n + i] = acc;
      }
    for (int phys = 0; phys < 3; ++phys) {
      double j0 = jinv[0 * 3 + phys], j1 = jinv[1 * 3 + phys], j2 = jinv[2 * 3 + phys];
      int base = (phys * 3 + comp) * n;
      for (int i = 0; i < n; ++i) pd[base + i] = j0 * rd[i] + j1 * rd[n + i] + j2 * rd[2 * n + i];
    }
  }
#define PD(ph, c, i) pd[((ph) * 3 + (c)) * n + (i)]
  for (int i = 0; i < n; ++i) {
    out[i * 3] = PD(1, 2, i) - PD(2, 1, i);
    out[i * 3 + 1] = PD(2, 0, i) - PD(0, 2, i);
    out[i * 3 + 2] = PD(0, 1, i) - PD(1, 0, i);
  }
#undef PD
}

void orc_element_curl(int p, const double *tet_xyz, const double *field,
                      double *out) {
  orc_op op;
  memset(&op, 0, sizeof op);
  op.p = p;
  op.np = orc_np(p);
  op.nfp = orc_nfp(p);
  int n = op.np, nfp = op.nfp;
  double *nodes = malloc(sizeof(double) * n * 3), *mass = malloc(sizeof(double) * n * n);
  op.dr = malloc(sizeof(double) * n * n);
  op.ds = malloc(sizeof(double) * n * n);
  op.dt = malloc(sizeof(double) * n * n);
  int *fn = malloc(sizeof(int) * 4 * nfp);
  double *lift = malloc(sizeof(double) * n * 4 * nfp), *fw = malloc(sizeof(double) * 4 * nfp);
  orc_refelem(p, nodes, mass, op.dr, op.ds, op.dt, fn, lift, fw);
  double jac[9], jinv[9], det;
  orc_geom_factors(tet_xyz, jac, jinv, &det);
  double *rd = malloc(sizeof(double) * 3 * n), *pd = malloc(sizeof(double) * 9 * n);
  element_curl_(&op, jinv, field, out, rd, pd);
  free(nodes); free(mass); free(op.dr); free(op.ds); free(op.dt); free(fn);
  free(lift); free(fw); free(rd); free(pd);
}

/* Face record used to pair tets sharing a triangle (Mesh::from_tets face
 * dedup by sorted vertex triple, crates/rapidfem-core/src/mesh.rs:94-142). */
typedef struct {
  int64_t v[3];
  int64_t tet;
  int lf;
} face_rec_;
static int face_cmp_(const void *pa, const void *pb) {
  const face_rec_ *a = pa, *b = pb;
  for (int d = 0; d < 3; ++d) {
    if (a->v[d] < b->v[d]) return -1;
    if (a->v[d] > b->v[d]) return 1;
  }
  if (a->tet < b->tet) return -1;
  if (a->tet > b->tet) return 1;
  return 0;
}

/* crates/rapidfem-core/src/mesh.rs:45-50 TET_FACE_LOCAL and
 * crates/rapidfem-core/src/topology.rs:19 TET_FACE_OPPOSITE */
static const int TET_FACE_LOCAL_[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};
static const int TET_FACE_OPPOSITE_[4] = {3, 1, 2, 0};

/* rhs.rs:805-1000 MaxwellOperator::new (non-dispersive, no plates/periodic),
 * with crates/rapidfem-core/src/topology.rs:50-90 FaceTopology::build and
 * rhs.rs:546-568 match_face_nodes. */
orc_op *orc_op_new(int64_t n_nodes, const double *xyz, int64_t K,
                   const int64_t *tets, int order, double flux_alpha,
                   const double *eps, const double *mu, const double *sigma,
                   const double *sigma_m, const int *face_port) {
  (void)n_nodes;
  orc_op *op = calloc(1, sizeof(orc_op));
  op->p = order;
  op->np = orc_np(order);
  op->nfp = orc_nfp(order);
  op->K = K;
  op->alpha = flux_alpha;
  const int n = op->np, nfp = op->nfp;
  op->nodes = malloc(sizeof(double) * n * 3);
  op->mass = malloc(sizeof(double) * n * n);
  op->dr = malloc(sizeof(double) * n * n);
  op->ds = malloc(sizeof(double) * n * n);
  op->dt = malloc(sizeof(double) * n * n);
  op->face_nodes = malloc(sizeof(int) * 4 * nfp);
  op->lift = malloc(sizeof(double) * n * 4 * nfp);
  op->fweights = malloc(sizeof(double) * 4 * nfp);
  orc_refelem(order, op->nodes, op->mass, op->dr, op->ds, op->dt, op->face_nodes,
              op->lift, op->fweights);

  op->origin = malloc(sizeof(double) * K * 3);
  op->jac = malloc(sizeof(double) * K * 9);
  op->jinv = malloc(sizeof(double) * K * 9);
  op->det = malloc(sizeof(double) * K);
  for (int64_t e = 0; e < K; ++e) {
    double v[12];
    for (int a = 0; a < 4; ++a)
      for (int d = 0; d < 3; ++d) v[a * 3 + d] = xyz[tets[4 * e + a] * 3 + d];
    orc_geom_factors(v, &op->jac[9 * e], &op->jinv[9 * e], &op->det[e]);
    memcpy(&op->origin[3 * e], v, sizeof(double) * 3);
  }

  /* Face pairing. */
  face_rec_ *recs = malloc(sizeof(face_rec_) * 4 * K);
  for (int64_t e = 0; e < K; ++e)
    for (int lf = 0; lf < 4; ++lf) {
      face_rec_ *r = &recs[4 * e + lf];
      int64_t a = tets[4 * e + TET_FACE_LOCAL_[lf][0]], b = tets[4 * e + TET_FACE_LOCAL_[lf][1]],
              c = tets[4 * e + TET_FACE_LOCAL_[lf][2]], t;
      if (a > b) { t = a; a = b; b = t; }
      if (b > c) { t = b; b = c; c = t; }
      if (a > b) { t = a; a = b; b = t; }
      r->v[0] = a; r->v[1] = b; r->v[2] = c;
      r->tet = e;
      r->lf = lf;
    }
  qsort(recs, 4 * K, sizeof(face_rec_), face_cmp_);
  op->nbr = malloc(sizeof(int64_t) * 4 * K);
  op->nbr_lf = malloc(sizeof(int) * 4 * K);
  for (int64_t q = 0; q < 4 * K; ++q) {
    op->nbr[q] = -1;
    op->nbr_lf[q] = -1;
  }
  for (int64_t q = 0; q < 4 * K;) {
    int64_t r = q + 1;
    while (r < 4 * K && memcmp(recs[r].v, recs[q].v, sizeof(int64_t) * 3) == 0) ++r;
    if (r - q == 2) {
      face_rec_ *A = &recs[q], *B = &recs[q + 1];
      op->nbr[4 * A->tet + A->lf] = B->tet;
      op->nbr_lf[4 * A->tet + A->lf] = B->lf;
      op->nbr[4 * B->tet + B->lf] = A->tet;
      op->nbr_lf[4 * B->tet + B->lf] = A->lf;
    }
    q = r;
  }
  free(recs);

  op->normal = malloc(sizeof(double) * 4 * K * 3);
  op->fscale = malloc(sizeof(double) * 4 * K);
  op->perm = calloc((size_t)4 * K * nfp, sizeof(int));
  op->port = malloc(sizeof(int) * 4 * K);
  double *here = malloc(sizeof(double) * nfp * 3), *there = malloc(sizeof(double) * nfp * 3);
  for (int64_t e = 0; e < K; ++e) {
    for (int lf = 0; lf < 4; ++lf) {
      const int64_t q = 4 * e + lf;
      /* topology.rs:56-73 normal, area */
      const double *v0 = &xyz[tets[4 * e + TET_FACE_LOCAL_[lf][0]] * 3];
      const double *v1 = &xyz[tets[4 * e + TET_FACE_LOCAL_[lf][1]] * 3];
      const double *v2 = &xyz[tets[4 * e + TET_FACE_LOCAL_[lf][2]] * 3];
      const double *vo = &xyz[tets[4 * e + TET_FACE_OPPOSITE_[lf]] * 3];
      double d1[3], d2[3], nn[3], dv[3];
      sub3(v1, v0, d1);
      sub3(v2, v0, d2);
      cross3(d1, d2, nn);
      double len = sqrt(dot3(nn, nn));
      double area = 0.5 * len;
      for (int k = 0; k < 3; ++k) nn[k] /= len;
      sub3(vo, v0, dv);
      if (dot3(dv, nn) > 0.0)
        for (int k = 0; k < 3; ++k) nn[k] = -nn[k];
      memcpy(&op->normal[3 * q], nn, sizeof nn);
      /* rhs.rs:829 */
      op->fscale[q] = 2.0 * area / fabs(op->det[e]);
      op->port[q] = face_port ? face_port[q] : -1;
      if (op->port[q] >= 0 && op->nbr[q] >= 0) op->port[q] = -1; /* ports are boundary faces */
      if (op->nbr[q] >= 0) {
        /* rhs.rs:831-844 face_node_coords + match_face_nodes(tol2 1e-18) */
        int64_t en = op->nbr[q];
        int lfn = op->nbr_lf[q];
        for (int m = 0; m < nfp; ++m) {
          geom_map_(&op->origin[3 * e], &op->jac[9 * e], &op->nodes[3 * op->face_nodes[lf * nfp + m]],
                    &here[3 * m]);
          geom_map_(&op->origin[3 * en], &op->jac[9 * en], &op->nodes[3 * op->face_nodes[lfn * nfp + m]],
                    &there[3 * m]);
        }
        for (int m = 0; m < nfp; ++m) {
          double best = INFINITY;
          int bm = 0;
          for (int m2 = 0; m2 < nfp; ++m2) {
            double dx = here[3 * m] - there[3 * m2], dy = here[3 * m + 1] - there[3 * m2 + 1],
                   dz = here[3 * m + 2] - there[3 * m2 + 2];
            double d = dx * dx + dy * dy + dz * dz;
            if (d < best) {
              best = d;
              bm = m2;
            }
          }
          if (!(best < 1e-18)) {
            free(here); free(there);
            orc_op_free(op);
            return NULL; /* the reference panics: unmatched face node */
          }
          op->perm[q * nfp + m] = bm;
        }
      }
    }
  }
  free(here);
  free(there);

  /* rhs.rs:908-926 per-element diagonal materials */
  op->inv_eps = malloc(sizeof(double) * K * 3);
  op->inv_mu = malloc(sizeof(double) * K * 3);
  op->sig_eps = malloc(sizeof(double) * K * 3);
  op->sig_mu = malloc(sizeof(double) * K * 3);
  for (int64_t e = 0; e < K; ++e)
    for (int c = 0; c < 3; ++c) {
      double ep = eps ? eps[3 * e + c] : 1.0, mu_ = mu ? mu[3 * e + c] : 1.0;
      double s = sigma ? sigma[e] : 0.0, sm = sigma_m ? sigma_m[e] : 0.0;
      op->inv_eps[3 * e + c] = 1.0 / ep;
      op->inv_mu[3 * e + c] = 1.0 / mu_;
      op->sig_eps[3 * e + c] = s / ep;
      op->sig_mu[3 * e + c] = sm / mu_;
    }
  return op;
}

void orc_op_free(orc_op *op) {
  if (!op) return;
  free(op->nodes); free(op->mass); free(op->dr); free(op->ds); free(op->dt);
  free(op->lift); free(op->fweights); free(op->face_nodes);
  free(op->origin); free(op->jac); free(op->jinv); free(op->det);
  free(op->normal); free(op->fscale); free(op->nbr); free(op->nbr_lf);
  free(op->port); free(op->perm);
  free(op->inv_eps); free(op->inv_mu); free(op->sig_eps); free(op->sig_mu);
  free(op->disp_elem); free(op->disp_slot); free(op->disp_a); free(op->disp_g);
  free(op->disp_iei);
  free(op);
}

/* rhs.rs:1003-1010 n_dof: the [E,H] block plus 3*Np per Debye element. */
int64_t orc_op_n_dof(const orc_op *op) {
  return 6 * (int64_t)op->np * op->K + 3 * (int64_t)op->np * op->n_disp;
}

/* rhs.rs:963-985 resolve the Debye elements: slot s in list order, with
 * dispersive.rs:50-53 relaxation_coeffs a = -1/tau, g = (eps_s - eps_inf)/tau
 * and 1/eps_inf. Returns -1 for an element out of range or listed twice. The
 * element's own eps is left as given: the reference expects it to be eps_inf. */
int orc_op_set_dispersive(orc_op *op, int64_t n, const int64_t *elem,
                          const double *eps_inf, const double *eps_static,
                          const double *tau) {
  int64_t *slot = malloc(sizeof(int64_t) * op->K);
  for (int64_t e = 0; e < op->K; ++e) slot[e] = -1;
  for (int64_t s = 0; s < n; ++s) {
    if (elem[s] < 0 || elem[s] >= op->K || slot[elem[s]] >= 0) {
      free(slot);
      return -1;
    }
    slot[elem[s]] = s;
  }
  free(op->disp_elem); free(op->disp_slot); free(op->disp_a); free(op->disp_g);
  free(op->disp_iei);
  op->n_disp = n;
  op->disp_slot = slot;
  op->disp_elem = malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  op->disp_a = malloc(sizeof(double) * (n > 0 ? n : 1));
  op->disp_g = malloc(sizeof(double) * (n > 0 ? n : 1));
  op->disp_iei = malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int64_t s = 0; s < n; ++s) {
    op->disp_elem[s] = elem[s];
    op->disp_a[s] = -1.0 / tau[s];
    op->disp_g[s] = (eps_static[s] - eps_inf[s]) / tau[s];
    op->disp_iei[s] = 1.0 / eps_inf[s];
  }
  return 0;
}

int64_t orc_op_n_dispersive(const orc_op *op) { return op->n_disp; }
int64_t orc_op_n_elem(const orc_op *op) { return op->K; }
int orc_op_order(const orc_op *op) { return op->p; }

void orc_op_export(const orc_op *op, double *jinv, double *det, double *normal,
                   double *fscale, int *nbr, int *nbr_lf, int *port, int *perm,
                   double *inv_eps, double *inv_mu, double *sig_eps,
                   double *sig_mu) {
  const int64_t K = op->K, nfp = op->nfp;
  if (jinv) memcpy(jinv, op->jinv, sizeof(double) * 9 * K);
  if (det) memcpy(det, op->det, sizeof(double) * K);
  if (normal) memcpy(normal, op->normal, sizeof(double) * 12 * K);
  if (fscale) memcpy(fscale, op->fscale, sizeof(double) * 4 * K);
  for (int64_t q = 0; q < 4 * K; ++q) {
    if (nbr) nbr[q] = (int)op->nbr[q];
    if (nbr_lf) nbr_lf[q] = op->nbr_lf[q];
    if (port) port[q] = op->port[q];
  }
  if (perm) memcpy(perm, op->perm, sizeof(int) * 4 * K * nfp);
  if (inv_eps) memcpy(inv_eps, op->inv_eps, sizeof(double) * 3 * K);
  if (inv_mu) memcpy(inv_mu, op->inv_mu, sizeof(double) * 3 * K);
  if (sig_eps) memcpy(sig_eps, op->sig_eps, sizeof(double) * 3 * K);
  if (sig_mu) memcpy(sig_mu, op->sig_mu, sizeof(double) * 3 * K);
}

/* rhs.rs:1341-1454 apply_element (non-dispersive block). scratch holds
 * ee,hh,de,dh (3np each), rd (3np), pd (9np). */
/* ghost: NULL, or the incident field (E_inc, H_inc) per face node of every
 * face, [(4e + f)*Nfp + m][6], used as the ghost state on port faces
 * (rhs.rs:1150-1156: the absorbing operator is the zero-ghost case) */
static void apply_element_(const orc_op *op, int64_t e, const double *y,
                           double *out, double *s, const double *ghost) {
  const int np = op->np, nfp = op->nfp, cols = 4 * nfp;
  double *ee = s, *hh = s + 3 * np, *de = s + 6 * np, *dh = s + 9 * np;
  double *rd = s + 12 * np, *pd = s + 15 * np;
  const int64_t base = e * np * 6;
  for (int node = 0; node < np; ++node)
    for (int c = 0; c < 3; ++c) {
      ee[node * 3 + c] = y[base + node * 6 + c];
      hh[node * 3 + c] = y[base + node * 6 + 3 + c];
    }
  element_curl_(op, &op->jinv[9 * e], hh, de, rd, pd);
  element_curl_(op, &op->jinv[9 * e], ee, dh, rd, pd);
  for (int i = 0; i < 3 * np; ++i) dh[i] = -dh[i];

  for (int f = 0; f < 4; ++f) {
    const int64_t q = 4 * e + f;
    const double *nrm = &op->normal[3 * q];
    const double coef = 0.5 * op->fscale[q];
    const int port = op->port[q];
    const int64_t nb = op->nbr[q];
    for (int m = 0; m < nfp; ++m) {
      const int vi = op->face_nodes[f * nfp + m];
      double em[3] = {ee[vi * 3], ee[vi * 3 + 1], ee[vi * 3 + 2]};
      double hm[3] = {hh[vi * 3], hh[vi * 3 + 1], hh[vi * 3 + 2]};
      double je[3], jh[3];
      if (port >= 0) {
        /* rhs.rs:1380-1386 port: characteristic flux against the incident
         * field (zero unless a ghost is given) */
        const double *gh = ghost ? &ghost[(q * nfp + m) * 6] : NULL;
        for (int k = 0; k < 3; ++k) {
          je[k] = em[k] - (gh ? gh[k] : 0.0);
          jh[k] = hm[k] - (gh ? gh[3 + k] : 0.0);
        }
      } else if (nb < 0) {
        /* rhs.rs:1387-1402 PEC ghost */
        double edn = dot3(nrm, em), hdn = dot3(nrm, hm);
        for (int k = 0; k < 3; ++k) {
          je[k] = 2.0 * (em[k] - edn * nrm[k]);
          jh[k] = 2.0 * hdn * nrm[k];
        }
      } else {
        /* rhs.rs:1403-1418 interior: neighbour trace through perm */
        const int vj = op->face_nodes[op->nbr_lf[q] * nfp + op->perm[q * nfp + m]];
        const int64_t nbb = nb * np * 6;
        for (int k = 0; k < 3; ++k) {
          je[k] = em[k] - y[nbb + vj * 6 + k];
          jh[k] = hm[k] - y[nbb + vj * 6 + 3 + k];
        }
      }
      double nxjh[3], nxje[3], t[3], pe[3], ph[3];
      cross3(nrm, jh, nxjh);
      cross3(nrm, je, nxje);
      const double a = port >= 0 ? 1.0 : op->alpha;
      cross3(nrm, je, t);
      cross3(nrm, t, pe);
      cross3(nrm, jh, t);
      cross3(nrm, t, ph);
      for (int i = 0; i < np; ++i) {
        const double w = coef * op->lift[i * cols + f * nfp + m];
        for (int c = 0; c < 3; ++c) {
          de[i * 3 + c] += w * (-nxjh[c] + a * pe[c]);
          dh[i * 3 + c] += w * (nxje[c] + a * ph[c]);
        }
      }
    }
  }
  /* rhs.rs:1443-1454 materials epilogue */
  const double *ie = &op->inv_eps[3 * e], *im = &op->inv_mu[3 * e];
  const double *se = &op->sig_eps[3 * e], *sm = &op->sig_mu[3 * e];
  for (int node = 0; node < np; ++node)
    for (int c = 0; c < 3; ++c) {
      out[node * 6 + c] = ie[c] * de[node * 3 + c] - se[c] * ee[node * 3 + c];
      out[node * 6 + 3 + c] = im[c] * dh[node * 3 + c] - sm[c] * hh[node * 3 + c];
    }
  /* rhs.rs:1456-1472 Debye element: the polarisation current -Pdot/eps_inf
   * on the E rows, Pdot = a*P + g*E with P from the appended block */
  const int64_t slot = op->n_disp ? op->disp_slot[e] : -1;
  if (slot >= 0) {
    const double a = op->disp_a[slot], g = op->disp_g[slot], iei = op->disp_iei[slot];
    const double *P = y + 6 * (int64_t)np * op->K + slot * 3 * np;
    for (int node = 0; node < np; ++node)
      for (int c = 0; c < 3; ++c)
        out[node * 6 + c] -= iei * (a * P[node * 3 + c] + g * ee[node * 3 + c]);
  }
}

/* rhs.rs:1262-1294 apply_into: elements dealt out in contiguous runs of
 * K/(4*threads) (constants.rs:203 APPLY_TASKS_PER_THREAD = 4). */
static void apply_all_(const orc_op *op, const double *y, double *dy, const double *ghost);
void orc_apply(const orc_op *op, const double *y, double *dy) { apply_all_(op, y, dy, NULL); }
/* The operator with an incident field on the port faces: A y + b for the
 * port source b of that field (rhs.rs:1157-1199). */
void orc_apply_ghost(const orc_op *op, const double *y, const double *ghost, double *dy) {
  apply_all_(op, y, dy, ghost);
}
static void apply_all_(const orc_op *op, const double *y, double *dy, const double *ghost) {
  const int64_t K = op->K;
  const int np = op->np;
  const int nt = nthreads_();
  int64_t chunk = K / (4 * (int64_t)nt);
  if (chunk < 1) chunk = 1;
  const int64_t nruns = (K + chunk - 1) / chunk;
#pragma omp parallel num_threads(nt)
  {
    double *s = malloc(sizeof(double) * 24 * np);
#pragma omp for schedule(dynamic, 1)
    for (int64_t r = 0; r < nruns; ++r) {
      int64_t e1 = (r + 1) * chunk;
      if (e1 > K) e1 = K;
      for (int64_t e = r * chunk; e < e1; ++e) apply_element_(op, e, y, &dy[e * np * 6], s, ghost);
    }
    free(s);
  }
  /* rhs.rs:1295-1324 the appended polarisation block: dP = a*P + g*E per
   * Debye element, no spatial coupling */
  const int64_t eh = 6 * (int64_t)np * K;
#pragma omp parallel for num_threads(nt)
  for (int64_t slot = 0; slot < op->n_disp; ++slot) {
    const double a = op->disp_a[slot], g = op->disp_g[slot];
    const int64_t eb = op->disp_elem[slot] * np * 6, pb = eh + slot * 3 * np;
    for (int node = 0; node < np; ++node)
      for (int c = 0; c < 3; ++c)
        dy[pb + node * 3 + c] = a * y[pb + node * 3 + c] + g * y[eb + node * 6 + c];
  }
}

/* rhs.rs:1597-1610 element_stencil: the element and its distinct face
 * neighbours, ascending. Returns the count (1..5). */
static int element_stencil_(const orc_op *op, int64_t e, int64_t *s) {
  int n = 1;
  s[0] = e;
  for (int f = 0; f < 4; ++f) {
    const int64_t nb = op->nbr[4 * e + f];
    int seen = 0;
    for (int q = 0; q < n; ++q) seen |= s[q] == nb;
    if (nb >= 0 && !seen) s[n++] = nb;
  }
  for (int a = 1; a < n; ++a)
    for (int b = a; b > 0 && s[b - 1] > s[b]; --b) {
      const int64_t t = s[b];
      s[b] = s[b - 1];
      s[b - 1] = t;
    }
  return n;
}

/* rhs.rs:1621-1730 assemble_sparse: row block e of A found by probing the
 * columns of its stencil (the [E,H] columns of every stencil element, then
 * the element's own P columns when it is dispersive) with unit vectors
 * through apply_element (rhs.rs:409-421 probe_column keeps v != 0.0), each
 * row's columns ascending; then the appended P rows, two entries each (g at
 * the E column, a on the diagonal) whatever their values. Called with
 * col_idx == NULL it fills row_ptr (n_dof + 1) and returns nnz; called again
 * with col_idx/values of nnz entries it fills them. */
int64_t orc_assemble_sparse(const orc_op *op, int64_t *row_ptr, int64_t *col_idx, double *values) {
  const int np = op->np, stride = 6 * np, ps = 3 * np;
  const int64_t K = op->K, n = orc_op_n_dof(op), eh = 6 * (int64_t)np * K;
  const int fill = col_idx != NULL;
  if (!fill) row_ptr[0] = 0;
#pragma omp parallel num_threads(nthreads_())
  {
    double *probe = calloc((size_t)n, sizeof(double));
    double *out = malloc(sizeof(double) * stride);
    double *s = malloc(sizeof(double) * 24 * np);
    int64_t *cnt = malloc(sizeof(int64_t) * stride);
    /* (column, value) per local row, a row holding at most 5 stride + ps */
    const int cap = 5 * stride + ps;
    int64_t *rc = malloc(sizeof(int64_t) * (size_t)stride * cap);
    double *rv = malloc(sizeof(double) * (size_t)stride * cap);
#pragma omp for schedule(dynamic, 16)
    for (int64_t e = 0; e < K; ++e) {
      int64_t st[5];
      const int ns = element_stencil_(op, e, st);
      for (int il = 0; il < stride; ++il) cnt[il] = 0;
      const int64_t slot = op->n_disp ? op->disp_slot[e] : -1;
      const int ncol_blocks = ns + (slot >= 0);
      /* probing in ascending column order leaves every row sorted */
      for (int b = 0; b < ncol_blocks; ++b) {
        const int64_t c0 = b < ns ? st[b] * stride : eh + slot * ps;
        const int w = b < ns ? stride : ps;
        for (int jl = 0; jl < w; ++jl) {
          const int64_t j = c0 + jl;
          probe[j] = 1.0;
          for (int il = 0; il < stride; ++il) out[il] = 0.0;
          apply_element_(op, e, probe, out, s, NULL);
          probe[j] = 0.0;
          for (int il = 0; il < stride; ++il)
            if (out[il] != 0.0) {
              rc[(size_t)il * cap + cnt[il]] = j;
              rv[(size_t)il * cap + cnt[il]] = out[il];
              ++cnt[il];
            }
        }
      }
      for (int il = 0; il < stride; ++il) {
        const int64_t row = e * stride + il;
        if (!fill) {
          row_ptr[row + 1] = cnt[il];
        } else {
          const int64_t o = row_ptr[row];
          for (int64_t q = 0; q < cnt[il]; ++q) {
            col_idx[o + q] = rc[(size_t)il * cap + q];
            values[o + q] = rv[(size_t)il * cap + q];
          }
        }
      }
    }
    free(probe); free(out); free(s); free(cnt); free(rc); free(rv);
  }
  if (!fill) {
    for (int64_t r = eh; r < n; ++r) row_ptr[r + 1] = 2;
    for (int64_t r = 0; r < n; ++r) row_ptr[r + 1] += row_ptr[r];
    return row_ptr[n];
  }
  for (int64_t q = 0; q < op->n_disp; ++q) {
    const int64_t e = op->disp_elem[q];
    for (int node = 0; node < np; ++node)
      for (int c = 0; c < 3; ++c) {
        const int64_t row = eh + q * ps + node * 3 + c, o = row_ptr[row];
        col_idx[o] = e * stride + node * 6 + c;
        values[o] = op->disp_g[q];
        col_idx[o + 1] = row;
        values[o + 1] = op->disp_a[q];
      }
  }
  return row_ptr[n];
}

/* rhs.rs:1509-1544 field_energy */
double orc_field_energy(const orc_op *op, const double *y) {
  const int np = op->np;
  double half = 0.0;
#pragma omp parallel for reduction(+ : half) num_threads(nthreads_())
  for (int64_t e = 0; e < op->K; ++e) {
    double scale = fabs(op->det[e]);
    double w[6];
    for (int c = 0; c < 3; ++c) {
      w[c] = 1.0 / op->inv_eps[3 * e + c];
      w[3 + c] = 1.0 / op->inv_mu[3 * e + c];
    }
    const int64_t base = e * np * 6;
    double acc = 0.0;
    for (int ni = 0; ni < np; ++ni)
      for (int nj = 0; nj < np; ++nj) {
        double mij = scale * op->mass[ni * np + nj];
        if (mij == 0.0) continue;
        const int64_t bi = base + ni * 6, bj = base + nj * 6;
        for (int c = 0; c < 6; ++c) acc += w[c] * mij * y[bi + c] * y[bj + c];
      }
    half += acc;
  }
  return 0.5 * half;
}

/* ------------------------------------------------------------------------ */
/* matvec plumbing for the integrators                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  const orc_op *op;
  const double *a; /* dense row-major n x n when op == NULL */
  int64_t n;
} mv_;
static void mv_apply_(const mv_ *m, const double *x, double *ax) {
  if (m->op) {
    orc_apply(m->op, x, ax);
    return;
  }
  const int64_t n = m->n;
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t j = 0; j < n; ++j) s += m->a[i * n + j] * x[j];
    ax[i] = s;
  }
}

/* constants.rs:106-122 LSERK4 */
static const double LSERK4_A_[5] = {
    0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
    -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double LSERK4_B_[5] = {
    1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
    1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
    2277821191437.0 / 14882151754819.0};
/* constants.rs:151-178 KCL RK4(3)5[2R+]C */
static const double KCL_A_[4] = {970286171893.0 / 4311952581923.0, 6584761158862.0 / 12103376702013.0,
                                 2251764453980.0 / 15575788980749.0,
                                 26877169314380.0 / 34165994151039.0};
static const double KCL_B_[5] = {1153189308089.0 / 22510343858157.0, 1772645290293.0 / 4653164025191.0,
                                 -1672844663538.0 / 4480602732383.0, 2114624349019.0 / 3568978502595.0,
                                 5198255086312.0 / 14908931495163.0};
static const double KCL_BHAT_[5] = {1016888040809.0 / 7410784769900.0, 11231460423587.0 / 58533540763752.0,
                                    -1563879915014.0 / 6823010717585.0, 606302364029.0 / 971179775848.0,
                                    1097981568119.0 / 3980877426909.0};

/* explicit.rs:44-56 Source::add_to */
static void add_source_(double *k, int64_t n, int kind, int64_t dof,
                        const double *vec, double g) {
  if (kind == 1) {
    if (dof >= 0 && dof < n) k[dof] += g;
  } else if (kind == 2) {
#pragma omp parallel for num_threads(nthreads_())
    for (int64_t i = 0; i < n; ++i) k[i] += vec[i] * g;
  }
}

/* explicit.rs:100-135 LserkWorkspace::step_into */
static void lserk_step_(const mv_ *m, double *y, double *p, double *k, int64_t n,
                        double dt, int kind, int64_t dof, const double *vec,
                        double g) {
  memset(p, 0, sizeof(double) * n);
  for (int s = 0; s < 5; ++s) {
    mv_apply_(m, y, k);
    add_source_(k, n, kind, dof, vec, g);
    const double a = LSERK4_A_[s], b = LSERK4_B_[s];
#pragma omp parallel for num_threads(nthreads_())
    for (int64_t i = 0; i < n; ++i) {
      p[i] = a * p[i] + dt * k[i];
      y[i] += b * p[i];
    }
  }
}

void orc_lserk_run(const orc_op *op, double *y, double dt, int64_t nsteps,
                   int src_kind, int64_t src_dof, const double *src_vec,
                   const double *g) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  const int64_t n = m.n;
  double *p = malloc(sizeof(double) * n), *k = malloc(sizeof(double) * n);
  for (int64_t s = 0; s < nsteps; ++s)
    lserk_step_(&m, y, p, k, n, dt, src_kind, src_dof, src_vec, (src_kind && g) ? g[s] : 0.0);
  free(p);
  free(k);
}

void orc_lserk_run_dense(const double *a, int n, double *y, double dt,
                         int64_t nsteps) {
  mv_ m = {NULL, a, n};
  double *p = malloc(sizeof(double) * n), *k = malloc(sizeof(double) * n);
  for (int64_t s = 0; s < nsteps; ++s) lserk_step_(&m, y, p, k, n, dt, 0, 0, NULL, 0.0);
  free(p);
  free(k);
}

/* explicit_adaptive.rs:84-152 KclWorkspace::step_into */
static void kcl_step_(const mv_ *m, double *y, double *err, double *stage,
                      double *k, int64_t n, double dt, int kind, int64_t dof,
                      const double *vec, double g) {
  mv_apply_(m, y, k);
  add_source_(k, n, kind, dof, vec, g);
  const double b0 = KCL_B_[0], e0 = KCL_BHAT_[0] - b0;
  for (int64_t i = 0; i < n; ++i) {
    double f = dt * k[i];
    k[i] = f;
    y[i] += b0 * f;
    err[i] = e0 * f;
  }
  for (int s = 1; s < 5; ++s) {
    const double amb = KCL_A_[s - 1] - KCL_B_[s - 1];
    for (int64_t i = 0; i < n; ++i) stage[i] = y[i] + amb * k[i];
    mv_apply_(m, stage, k);
    add_source_(k, n, kind, dof, vec, g);
    const double b = KCL_B_[s], ew = KCL_BHAT_[s] - b;
    for (int64_t i = 0; i < n; ++i) {
      double f = dt * k[i];
      k[i] = f;
      y[i] += b * f;
      err[i] += ew * f;
    }
  }
}

void orc_kcl_step(const orc_op *op, double *y, double *err, double dt,
                  int src_kind, int64_t src_dof, const double *src_vec,
                  double g) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  double *st = malloc(sizeof(double) * m.n), *k = malloc(sizeof(double) * m.n);
  kcl_step_(&m, y, err, st, k, m.n, dt, src_kind, src_dof, src_vec, g);
  free(st);
  free(k);
}

void orc_kcl_step_dense(const double *a, int n, double *y, double *err,
                        double dt) {
  mv_ m = {NULL, a, n};
  double *st = malloc(sizeof(double) * n), *k = malloc(sizeof(double) * n);
  kcl_step_(&m, y, err, st, k, n, dt, 0, 0, NULL, 0.0);
  free(st);
  free(k);
}

/* session.rs:142-155 Controller::default, constants.rs:268-299 */
void orc_controller_default(orc_controller *c) {
  c->atol = 1e-8;
  c->rtol = 1e-4;
  c->safety = 0.9;
  c->growth_limit = 5.0;
  c->shrink_limit = 0.2;
  c->pi_alpha = 0.7 / 4.0;
  c->pi_beta = 0.4 / 4.0;
  c->min_step_factor = 1e-10;
}

/* session.rs:160-171 Controller::err_norm */
double orc_err_norm(const orc_controller *c, const double *y_old,
                    const double *y_new, const double *err, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double a = y_old[i], b = y_new[i];
    if (!isfinite(b)) return INFINITY;
    double r = err[i] / (c->atol + c->rtol * fmax(fabs(a), fabs(b)));
    s += r * r;
  }
  return sqrt(s / (double)(n > 1 ? n : 1));
}

#define KCL_ERR_FLOOR_ 1e-12   /* constants.rs KCL_ERR_FLOOR */
#define KCL_NONFINITE_ERR_ 10.0 /* constants.rs KCL_NONFINITE_ERR */

/* session.rs:174-186 Controller::factor */
double orc_factor(const orc_controller *c, double err_norm, double prev,
                  int reject) {
  double err = fmax(err_norm, KCL_ERR_FLOOR_);
  double f = (reject || prev <= 0.0)
                 ? c->safety * pow(err, -c->pi_alpha)
                 : c->safety * pow(err, -c->pi_alpha) * pow(prev, c->pi_beta);
  if (reject) return fmax(f, c->shrink_limit);
  return fmin(fmax(f, c->shrink_limit), c->growth_limit);
}

/* session.rs:781-827 adaptive_frame over steps frames (session.rs:754-767).
 * The waveform g[k] (when driven) is sampled at the substep start; here the
 * caller passes one sample per frame, the hold the GPU path uses
 * (gpu/operator.rs:845). */
int orc_adaptive_run(const orc_op *op, double *y, double dt, int64_t steps,
                     const orc_controller *c, int src_kind, int64_t src_dof,
                     const double *src_vec, const double *g, int64_t *stats,
                     double *hminmax) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  const int64_t n = m.n;
  double *ytry = malloc(sizeof(double) * n), *err = malloc(sizeof(double) * n);
  double *st = malloc(sizeof(double) * n), *k = malloc(sizeof(double) * n);
  double h = dt, prev = 0.0, hmin = INFINITY, hmax = 0.0;
  int64_t acc = 0, rej = 0;
  int rc = 0;
  const double h_floor = c->min_step_factor * dt;
  for (int64_t f = 0; f < steps && rc == 0; ++f) {
    double t_rel = 0.0;
    const double gv = (src_kind && g) ? g[f] : 0.0;
    while (t_rel < dt) {
      double h_try = fmin(h, dt - t_rel);
      memcpy(ytry, y, sizeof(double) * n);
      kcl_step_(&m, ytry, err, st, k, n, h_try, src_kind, src_dof, src_vec, gv);
      double e = orc_err_norm(c, y, ytry, err, n);
      if (isfinite(e) && e <= 1.0) {
        memcpy(y, ytry, sizeof(double) * n);
        t_rel += h_try;
        ++acc;
        h = h_try * orc_factor(c, e, prev, 0);
        prev = fmax(e, KCL_ERR_FLOOR_);
        hmin = fmin(hmin, h);
        hmax = fmax(hmax, h);
      } else {
        ++rej;
        double probe = isfinite(e) ? e : KCL_NONFINITE_ERR_;
        h = h_try * orc_factor(c, probe, prev, 1);
      }
      if (h < h_floor) {
        rc = -1;
        break;
      }
    }
  }
  if (stats) {
    stats[0] = acc;
    stats[1] = rej;
  }
  if (hminmax) {
    hminmax[0] = hmin;
    hminmax[1] = hmax;
  }
  free(ytry); free(err); free(st); free(k);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* CFL: session.rs:415-431 normal_vector, :514-569 cfl_dt                   */
/* ------------------------------------------------------------------------ */

void orc_normal_vector(int64_t n, uint64_t seed, double *out) {
  uint64_t state = seed;
  for (int64_t i = 0; i < n; ++i) {
    double uv[2];
    for (int t = 0; t < 2; ++t) {
      state += 0x9e3779b97f4a7c15ULL;
      uint64_t z = state;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      z ^= z >> 31;
      uv[t] = ((double)(z >> 11) + 0.5) / (double)(1ULL << 53);
    }
    out[i] = sqrt(-2.0 * log(uv[0])) * cos(2.0 * M_PI * uv[1]);
  }
}

static double norm_(const double *v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
  return sqrt(s);
}

/* constants.rs:239-262 CFL constants */
#define CFL_POWER_ITERS_ 40
#define CFL_PROBE_STEPS_ 64
#define CFL_BISECT_ITERS_ 7
#define CFL_Z_STABLE_ 3.0
#define CFL_Z_UNSTABLE_ 15.0
#define CFL_GROWTH_FACTOR_ 10.0
#define CFL_GROWTH_RATE_TOL_ 1e-3
#define CFL_SAFETY_ 0.8

static int cfl_stable_(const mv_ *m, const double *probe, double n0, double rho,
                       double z, double *y, double *p, double *k) {
  const int64_t n = m->n;
  const double h = z / rho;
  memcpy(y, probe, sizeof(double) * n);
  for (int s = 0; s < CFL_PROBE_STEPS_; ++s) lserk_step_(m, y, p, k, n, h, 0, 0, NULL, 0.0);
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(y[i])) return 0;
  double nf = norm_(y, n);
  if (nf > CFL_GROWTH_FACTOR_ * n0) return 0;
  double rate = nf > 0.0 ? pow(nf / n0, 1.0 / (double)CFL_PROBE_STEPS_) : 0.0;
  return rate <= 1.0 + CFL_GROWTH_RATE_TOL_;
}

double orc_cfl_dt(const orc_op *op) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  const int64_t n = m.n;
  double *v = malloc(sizeof(double) * n), *av = malloc(sizeof(double) * n);
  orc_normal_vector(n, 0, v);
  double s = 1.0 / norm_(v, n);
  for (int64_t i = 0; i < n; ++i) v[i] *= s;
  double rho = 1.0;
  for (int it = 0; it < CFL_POWER_ITERS_; ++it) {
    orc_apply(op, v, av);
    rho = norm_(av, n);
    for (int64_t i = 0; i < n; ++i) v[i] = av[i] / rho;
  }
  double n0 = norm_(v, n);
  double *y = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n), *k = av;
  double lo = CFL_Z_STABLE_, hi = CFL_Z_UNSTABLE_;
  while (!cfl_stable_(&m, v, n0, rho, lo, y, p, k) && lo > 0.1) lo *= 0.5;
  while (cfl_stable_(&m, v, n0, rho, hi, y, p, k) && hi < 1e3) hi *= 1.5;
  for (int it = 0; it < CFL_BISECT_ITERS_; ++it) {
    double mid = 0.5 * (lo + hi);
    if (cfl_stable_(&m, v, n0, rho, mid, y, p, k))
      lo = mid;
    else
      hi = mid;
  }
  free(v); free(av); free(y); free(p);
  return CFL_SAFETY_ * lo / rho;
}

/* ------------------------------------------------------------------------ */
/* Krylov exponential: crates/rapidfem-td/src/propagator.rs                 */
/* ------------------------------------------------------------------------ */

#define EXPM_SCALE_THRESHOLD_ 0.5 /* constants.rs:91 */
#define EXPM_TAYLOR_TERMS_ 18      /* constants.rs:96 */
#define ARNOLDI_BREAKDOWN_ 1e-12   /* constants.rs:62 */

/* propagator.rs:69-114 expm_into */
void orc_expm(const double *a, int n, double *out) {
  const int nn = n * n;
  double norm = 0.0;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += fabs(a[i * n + j]);
    norm = fmax(norm, row);
  }
  unsigned sq = 0;
  if (norm > EXPM_SCALE_THRESHOLD_) {
    long v = (long)ceil(log2(norm)) + 1;
    sq = v > 0 ? (unsigned)v : 0;
  }
  double scale = powi_(2.0, (int)sq);
  double *b = malloc(sizeof(double) * nn), *res = malloc(sizeof(double) * nn);
  double *term = malloc(sizeof(double) * nn), *term2 = malloc(sizeof(double) * nn);
  double *tmp = malloc(sizeof(double) * nn);
  for (int k = 0; k < nn; ++k) b[k] = a[k] / scale;
  memset(res, 0, sizeof(double) * nn);
  memset(term, 0, sizeof(double) * nn);
  for (int i = 0; i < n; ++i) res[i * n + i] = term[i * n + i] = 1.0;
  for (int k = 1; k <= EXPM_TAYLOR_TERMS_; ++k) {
    matmul_(term, b, n, n, n, term2);
    double inv = 1.0 / (double)k;
    for (int x = 0; x < nn; ++x) term2[x] *= inv;
    for (int x = 0; x < nn; ++x) res[x] += term2[x];
    memcpy(term, term2, sizeof(double) * nn);
  }
  for (unsigned s = 0; s < sq; ++s) {
    matmul_(res, res, n, n, n, tmp);
    memcpy(res, tmp, sizeof(double) * nn);
  }
  memcpy(out, res, sizeof(double) * nn);
  free(b); free(res); free(term); free(term2); free(tmp);
}

/* Augmented matvec for the ETD step: [[A,b],[0,0]] on [x; xi]
 * (propagator.rs:357-365). */
typedef struct {
  const mv_ *inner;
  const double *b;
} aug_;

static void kry_mv_(const mv_ *m, const aug_ *aug, const double *x, double *ax) {
  if (!aug) {
    mv_apply_(m, x, ax);
    return;
  }
  const int64_t n = aug->inner->n;
  mv_apply_(aug->inner, x, ax);
  const double xi = x[n];
  for (int64_t k = 0; k < n; ++k) ax[k] += xi * aug->b[k];
  ax[n] = 0.0;
}

/* propagator.rs:186-326 KrylovWorkspace::expmv_into (CGS2 + a-posteriori
 * stop). n is the vector length (n+1 for the augmented system). */
static int expmv_core_(const mv_ *m, const aug_ *aug, int64_t n,
                       const double *v, double t, int max_dim, double tol,
                       double *out) {
  double beta = norm_(v, n);
  if (beta == 0.0 || max_dim < 1) {
    memset(out, 0, sizeof(double) * n);
    return 0;
  }
  double *basis = malloc(sizeof(double) * (size_t)(max_dim + 1) * n);
  double *w = malloc(sizeof(double) * n);
  double *proj = malloc(sizeof(double) * (max_dim + 1));
  double *h = calloc((size_t)max_dim * max_dim, sizeof(double));
  double *th = malloc(sizeof(double) * max_dim * max_dim);
  double *eth = malloc(sizeof(double) * max_dim * max_dim);
  double inv_beta = 1.0 / beta;
  for (int64_t k = 0; k < n; ++k) basis[k] = v[k] * inv_beta;
  int dim = max_dim;
  for (int j = 0; j < max_dim; ++j) {
    kry_mv_(m, aug, &basis[(int64_t)j * n], w);
    const int cols = j + 1;
    for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for num_threads(nthreads_())
      for (int i = 0; i < cols; ++i) {
        const double *bi = &basis[(int64_t)i * n];
        double s = 0.0;
        for (int64_t k = 0; k < n; ++k) s += bi[k] * w[k];
        proj[i] = s;
      }
#pragma omp parallel for num_threads(nthreads_())
      for (int64_t k = 0; k < n; ++k) {
        double wk = w[k];
        for (int i = 0; i < cols; ++i) wk -= proj[i] * basis[(int64_t)i * n + k];
        w[k] = wk;
      }
      for (int i = 0; i < cols; ++i) h[i * max_dim + j] += proj[i];
    }
    double hnext = norm_(w, n);
    int d = j + 1;
    if (hnext < ARNOLDI_BREAKDOWN_ || d == max_dim) {
      dim = d;
      break;
    }
    if (tol > 0.0) {
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) th[a * d + b] = t * h[a * max_dim + b];
      orc_expm(th, d, eth);
      if (hnext * fabs(eth[(d - 1) * d]) < tol) {
        dim = d;
        break;
      }
    }
    h[(j + 1) * max_dim + j] = hnext;
    double inv = 1.0 / hnext;
    double *dst = &basis[(int64_t)(j + 1) * n];
    for (int64_t k = 0; k < n; ++k) dst[k] = w[k] * inv;
  }
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) th[i * dim + j] = t * h[i * max_dim + j];
  orc_expm(th, dim, eth);
  memset(out, 0, sizeof(double) * n);
  for (int i = 0; i < dim; ++i) {
    double c = beta * eth[i * dim];
    const double *bi = &basis[(int64_t)i * n];
    for (int64_t k = 0; k < n; ++k) out[k] += c * bi[k];
  }
  free(basis); free(w); free(proj); free(h); free(th); free(eth);
  return dim;
}

int orc_expmv_dense(const double *a, int n, const double *v, double t,
                    int max_dim, double tol, double *out) {
  mv_ m = {NULL, a, n};
  return expmv_core_(&m, NULL, n, v, t, max_dim, tol, out);
}

int orc_expmv_op(const orc_op *op, const double *v, double t, int max_dim,
                 double tol, double *out) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  return expmv_core_(&m, NULL, m.n, v, t, max_dim, tol, out);
}

/* propagator.rs:338-373 etd_step_into */
static void etd_core_(const mv_ *m, const double *y, const double *b, double h,
                      int max_dim, double tol, double *out) {
  const int64_t n = m->n;
  double *z = malloc(sizeof(double) * (n + 1)), *r = malloc(sizeof(double) * (n + 1));
  memcpy(z, y, sizeof(double) * n);
  z[n] = 1.0;
  aug_ aug = {m, b};
  expmv_core_(m, &aug, n + 1, z, h, max_dim, tol, r);
  memcpy(out, r, sizeof(double) * n);
  free(z);
  free(r);
}

void orc_etd_dense(const double *a, int n, const double *y, const double *b,
                   double h, int max_dim, double tol, double *out) {
  mv_ m = {NULL, a, n};
  etd_core_(&m, y, b, h, max_dim, tol, out);
}

void orc_etd_op(const orc_op *op, const double *y, const double *b, double h,
                int max_dim, double tol, double *out) {
  mv_ m = {op, NULL, orc_op_n_dof(op)};
  etd_core_(&m, y, b, h, max_dim, tol, out);
}
