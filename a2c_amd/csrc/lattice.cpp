// Seeded FCC Pd + octahedral-H specimens (host).  See include/a2c_lattice.h.
#include "a2c_lattice.h"

#include <cmath>
#include <cstdint>

namespace {

inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox4x32 round constants (Salmon et al., SC'11, "Parallel random numbers:
// as easy as 1, 2, 3").
constexpr uint32_t PH_M0 = 0xD2511F53u, PH_M1 = 0xCD9E8D57u;
constexpr uint32_t PH_W0 = 0x9E3779B9u, PH_W1 = 0xBB67AE85u;

void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(PH_M0, c0, hi0, lo0);
    mulhilo(PH_M1, c2, hi1, lo1);
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += PH_W0;
    k1 += PH_W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

struct Rng {
  uint32_t key[2];
  uint32_t lattice;
  // two 53-bit uniforms in [0,1) from one block
  void uniforms(uint32_t site, uint32_t stream, double u[2]) const {
    uint32_t ctr[4] = {site, stream, lattice, 0u}, o[4];
    philox(ctr, key, o);
    u[0] = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * (1.0 / 9007199254740992.0);
    u[1] = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * (1.0 / 9007199254740992.0);
  }
  // two standard normals (Box-Muller)
  void normals(uint32_t site, uint32_t stream, double z[2]) const {
    double u[2];
    uniforms(site, stream, u);
    const double r = std::sqrt(-2.0 * std::log(1.0 - u[0]));  // 1-u in (0,1]
    const double t = 6.283185307179586 * u[1];
    z[0] = r * std::cos(t);
    z[1] = r * std::sin(t);
  }
};

const double PD_BASIS[4][3] = {{0, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}, {0, 0.5, 0.5}};
const double H_BASIS[4][3] = {{0.5, 0, 0}, {0, 0.5, 0}, {0, 0, 0.5}, {0.5, 0.5, 0.5}};

}  // namespace

extern "C" void a2c_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}

extern "C" void a2c_lattice_desc_default(a2c_lattice_desc* d) {
  d->nx = d->ny = d->nz = 6;
  d->a = 3.89;
  d->seed = 1;
  d->lattice_id = 0;
  d->disp_std = 0.02;
  d->sigma_lo = d->sigma_hi = 0.10;
  d->x_lo = d->x_hi = 0.10;
  d->x_clip_lo = 1e-3;
  d->x_clip_hi = 1.0 - 1e-16;
  d->slab_cells = 0;
  d->slab_x = 0.0;
}

extern "C" int a2c_lattice_fcc_pdh(const a2c_lattice_desc* d, double* X, double* x, double* box) {
  if (!d || !X || !x || d->nx <= 0 || d->ny <= 0 || d->nz <= 0 || !(d->a > 0)) return -1;
  const int64_t ncell = (int64_t)d->nx * d->ny * d->nz;
  const int64_t nper = 4 * ncell;  // sites per species
  if (2 * nper > 0x7fffffff) return -1;
  Rng rng{{(uint32_t)d->seed, (uint32_t)(d->seed >> 32)}, d->lattice_id};
  if (box) { box[0] = d->nx * d->a; box[1] = d->ny * d->a; box[2] = d->nz * d->a; }
  for (int sp = 0; sp < 2; ++sp) {
    const double(*basis)[3] = sp == 0 ? PD_BASIS : H_BASIS;
    for (int64_t c = 0; c < ncell; ++c) {
      const int ix = (int)(c % d->nx), iy = (int)((c / d->nx) % d->ny), iz = (int)(c / ((int64_t)d->nx * d->ny));
      for (int b = 0; b < 4; ++b) {
        const int64_t s = sp * nper + 4 * c + b;  // global site index
        double z01[2], z23[2], u[2];
        rng.normals((uint32_t)s, 0u, z01);
        rng.normals((uint32_t)s, 1u, z23);
        rng.uniforms((uint32_t)s, 2u, u);
        X[4 * s + 0] = (ix + basis[b][0]) * d->a + d->disp_std * z01[0];
        X[4 * s + 1] = (iy + basis[b][1]) * d->a + d->disp_std * z01[1];
        X[4 * s + 2] = (iz + basis[b][2]) * d->a + d->disp_std * z23[0];
        X[4 * s + 3] = d->sigma_lo + (d->sigma_hi - d->sigma_lo) * u[0];
        if (sp == 1) {
          double v = d->x_lo + (d->x_hi - d->x_lo) * u[1];
          if (d->slab_cells > 0) v = (ix < d->slab_cells) ? d->slab_x : d->x_lo;
          if (v < d->x_clip_lo) v = d->x_clip_lo;
          if (v > d->x_clip_hi) v = d->x_clip_hi;
          x[s - nper] = v;
        }
      }
    }
  }
  return 0;
}
