/*
 * a2c_lattice.h -- seeded FCC Pd + octahedral H specimens in the host layout
 * of a2c_fe.h (host-only, no GPU).  Stands in for the reference's specimen
 * setup (SpaceInitializer::SetupLatticeVariables, SpaceInitializer.cpp:41-115)
 * for synthetic workloads; the reference has no bridge from LatticeVariables
 * to the legacy q_Pd/q_H arrays, so this writes that layout directly.
 *
 * Geometry: conventional cubic cell of edge a; Pd basis (0,0,0), (1/2,1/2,0),
 * (1/2,0,1/2), (0,1/2,1/2); octahedral H basis (1/2,0,0), (0,1/2,0),
 * (0,0,1/2), (1/2,1/2,1/2).  Site order: cell index (ix fastest, then iy, iz)
 * times 4 plus basis index; all Pd first, then all H.
 *
 * Random numbers: Philox4x32-10 keyed by the 64-bit seed, counter
 * (site, stream, lattice_id, 0); normals by Box-Muller on 53-bit uniforms.
 */
#ifndef A2C_LATTICE_H
#define A2C_LATTICE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct a2c_lattice_desc {
  int32_t nx, ny, nz;      /* cells per axis */
  double a;                /* lattice constant (3.89) */
  uint64_t seed;
  uint32_t lattice_id;
  double disp_std;         /* N(0, disp_std^2) per position component */
  double sigma_lo, sigma_hi;  /* sigma ~ U(lo, hi) (lo == hi: constant) */
  double x_lo, x_hi;       /* H occupancy ~ U(lo, hi) (lo == hi: constant) */
  double x_clip_lo, x_clip_hi; /* occupancy clipped to [clip_lo, clip_hi] */
  int32_t slab_cells;      /* > 0: x = slab_x for cells ix < slab_cells, x_lo elsewhere */
  double slab_x;
} a2c_lattice_desc;

void a2c_lattice_desc_default(a2c_lattice_desc* d);
/* n_pd = n_h = 4*nx*ny*nz. X: 4*(n_pd+n_h) doubles, x: n_h doubles,
 * box: 3 doubles (n*a per axis).  Returns 0 or A2C_EINVAL (-1). */
int a2c_lattice_fcc_pdh(const a2c_lattice_desc* d, double* X, double* x, double* box);
/* Philox4x32-10 block (for known-answer tests). */
void a2c_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
