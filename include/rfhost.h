/*
 * rfhost.h - host-side operator builder of the MI355X DGTD backend.
 *
 * The reference builds the operator data on the CPU in Rust
 * (MaxwellOperator::new, rhs.rs:805-1000) and GpuOperator::new uploads it.
 * The drop-in boundary (rfhip.h) therefore takes plain arrays; this header is
 * the C++ builder that produces those arrays for hosts that have no
 * MaxwellOperator of their own (the Python package, tests, benchmarks).
 * Each function names the reference function it mirrors.
 */
#ifndef RFHOST_H
#define RFHOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Np = (p+1)(p+2)(p+3)/6, Nfp = (p+1)(p+2)/2. */
int32_t rfhost_np(int32_t p);
int32_t rfhost_nfp(int32_t p);

/* ReferenceElement::new(p) (dg_basis.rs:61-119, build_lift :288-385):
 * equispaced nodes (np*3), mass, diff_r/s/t (np*np), face_nodes (4*nfp),
 * lift (np*4nfp), face_weights (4*nfp). Returns 0, or 1 for p < 1. */
int rfhost_refelem(int32_t p, double *nodes, double *mass, double *dr,
                   double *ds, double *dt, int32_t *face_nodes, double *lift,
                   double *face_weights);

/* structured_box / structured_box_jittered (mesh_gen.rs:54-121): node
 * coordinates ((nx+1)(ny+1)(nz+1)*3) and Kuhn tets (6*nx*ny*nz*4, each
 * tet's vertex ids ascending). jittered != 0 applies jitter_offset
 * (mesh_gen.rs:124-139) to interior coordinates. */
void rfhost_box_mesh(int32_t nx, int32_t ny, int32_t nz, double lx, double ly,
                     double lz, double amplitude, uint64_t seed,
                     int32_t jittered, double *xyz, int64_t *tets);

/* Operator arrays of MaxwellOperator::new (rhs.rs:805-1000) for a tet mesh,
 * non-dispersive, no plates or periodic pairs. Outputs (caller-allocated):
 *   jinv 9K, det K, normal 12K, fscale 4K, nbr 4K, nbr_lf 4K, perm 4K*nfp,
 *   port 4K, inv_eps/inv_mu/sigma_eps/sigma_mu 3K each.
 * Inputs: eps/mu K*3 (NULL = vacuum), sigma/sigma_m K (NULL = 0),
 * face_port 4K (NULL = no ports; a port tag on an interior face is dropped).
 * Returns 0; 2 when a face node has no partner (the reference panics,
 * rhs.rs:840-842); 3 when a face has more than two tets. */
int rfhost_operator_arrays(int64_t n_nodes, const double *xyz, int64_t n_tets,
                           const int64_t *tets, int32_t order,
                           const double *eps, const double *mu,
                           const double *sigma, const double *sigma_m,
                           const int32_t *face_port, double *jinv, double *det,
                           double *normal, double *fscale, int32_t *nbr,
                           int32_t *nbr_lf, int32_t *perm, int32_t *port,
                           double *inv_eps, double *inv_mu, double *sigma_eps,
                           double *sigma_mu);

/* Physical coordinates of every DG node (K*np*3), GeometricFactors::map
 * (geom_factors.rs:51-60) applied to the reference nodes. */
void rfhost_node_coords(int64_t n_tets, const double *xyz, const int64_t *tets,
                        int32_t p, double *out);

#ifdef __cplusplus
}
#endif
#endif
