// Internal declarations shared by the HIP sources of liba2c_fe.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "a2c_fe.h"

namespace a2c {

// Device buffers of the library (devmem.cpp): a caching allocator with
// hipMalloc/hipFree semantics (dfree synchronises the block's device first).
hipError_t dmalloc(void** ptr, size_t bytes);
template <class T>
inline hipError_t dmalloc(T** ptr, size_t bytes) {
  return dmalloc(reinterpret_cast<void**>(ptr), bytes);
}
hipError_t dfree(const void* ptr);
hipError_t hmalloc(void** ptr, size_t bytes);  // pinned host staging, cached alike
template <class T>
inline hipError_t hmalloc(T** ptr, size_t bytes) {
  return hmalloc(reinterpret_cast<void**>(ptr), bytes);
}
hipError_t hfree(const void* ptr);

constexpr double KB = 8.6173324e-5;  // Boltzmann constant, eV/K (minimizer.cpp:92)
constexpr int NBR_CAP = 256;          // per-site neighbor cap of the build kernel

// One neighbor list in ELL layout: entry t of row r is idx[t * rows + r], so
// lane r of a wave reading slot t touches consecutive words (coalesced).
// Indices are species-local (Pd index or H index), rows sorted ascending.
struct Ell {
  int32_t rows = 0;
  int32_t width = 0;
  int32_t* idx = nullptr;  // width * rows
  int32_t* cnt = nullptr;  // rows
  // Pair-once evaluation (each unordered pair evaluated once, by its owner).
  // Every row is split, keeping its order, into the entries this row owns
  // and the entries it receives; both halves are compact (slot-major, no
  // gaps), so a wave's lanes step through their halves without masking each
  // other off.  Widths are the longest half over all rows.
  int32_t owidth = 0, rwidth = 0;
  int32_t* ocnt = nullptr;  // rows: owned entries
  int32_t* oj = nullptr;    // owidth * rows: neighbor of owned entry t
  int32_t* orv = nullptr;   // owidth * rows: slot of the pair in the partner's received half
  int32_t* rcnt = nullptr;  // rows: received entries
  // Received shares, written by the partner's owner pass (already weighted
  // by the partner's occupancy) and summed by this row's receive pass of the
  // same evaluation: rwidth * rows entries of a float (density) or float4
  // (gradient) in the mixed passes, a double or double4 in the FP64 passes --
  // one buffer serves all four.  During the split it holds the map full-row
  // slot -> received slot.
  void* rbuf = nullptr;
  int64_t ocap = 0;            // allocated entries of oj/orv
  int64_t bcap = 0;            // allocated bytes of rbuf
  int64_t cap = 0;             // allocated entries of idx (grow-only across rebuilds)
};

// Device view of one lattice: what every kernel needs.  The eval kernels
// read it from device memory through a block -> lattice map, so one launch
// covers any set of lattices (a batch) and a single lattice is a batch of one.
// Each lattice's sites are split into blocks of 256 starting at its site 0 in
// every launch, so its results do not depend on the other lattices present.
struct LatDev {
  int32_t npd = 0, nh = 0;
  int32_t periodic[3] = {0, 0, 0};
  int32_t blk0 = 0;          // first block of this lattice in the current launch
  int32_t kl32 = 1, kl64 = 1;    // lanes per site in the mixed / FP64 pair-owner passes
  int32_t owg0_32 = 0, owg0_64 = 0;  // first workgroup of this lattice in those launches
  double box[3] = {0, 0, 0};
  double ibox[3] = {0, 0, 0};  // 1/box: the image shift of a listed pair is L*rint(d/L) with d/L as d*(1/L)
  double T = 0;
  int32_t rule = 0;          // A2C_QUAD_R12 or A2C_QUAD_GH3
  Ell pdpd, pdh, hpd, hh, hh1;
  const double* X = nullptr;  // 4N state (32-byte aligned)
  double* G = nullptr;        // 4N gradient
  const double* x = nullptr;  // nh occupancy
  double* rho = nullptr;      // N densities
  double* Fp = nullptr;       // N F'(rho) (FP64 path)
  float2* aux = nullptr;      // N {F'(rho), x} (mixed path)
  double* emb = nullptr;      // N embedding energy (x-weighted for H)
  double* epair = nullptr;    // N pair-energy partial (mixed path)
  double* part = nullptr;     // ceil(N/256) block partial sums of f
  double* f = nullptr;        // objective
};
using DevSys = LatDev;
constexpr int SITE_BLOCK = 256;  // sites per block of every site-parallel kernel
__host__ __device__ inline int site_blocks(int n) { return (n + SITE_BLOCK - 1) / SITE_BLOCK; }

// Global-memory views.  Pointers read from a descriptor in memory are
// generic to the compiler, which then issues flat_* instructions: those count
// against the LDS counter as well as the vector-memory one, so every wait for
// a table lookup also waits out outstanding gathers and stores.  Loads and
// stores through these views are global_* instructions.
#if defined(__HIP_DEVICE_COMPILE__)
#define A2C_GLOBAL __attribute__((address_space(1)))
#else
#define A2C_GLOBAL  // the host pass of a .hip file compiles device code too, without address spaces
#endif
template <class T>
__device__ __forceinline__ A2C_GLOBAL T* gmem(T* p) {
  return (A2C_GLOBAL T*)p;
}

// The n received shares of one row of a slot-major buffer (share t at
// base[t * rows + row]) passed to add in slot order.  The loads of RU
// consecutive shares are issued before any of them is added, so a row costs
// one memory latency per RU shares rather than one per share; the order of
// the additions, and so the result, is that of a plain loop.
template <int RU, class T, class Add>
__device__ __forceinline__ void for_row_shares(const A2C_GLOBAL T* __restrict__ base, size_t rows, int row, int n,
                                               Add add) {
  for (int t = 0; t < n; t += RU) {
    T v[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u)
      if (t + u < n) v[u] = base[(size_t)(t + u) * rows + row];
#pragma unroll
    for (int u = 0; u < RU; ++u)
      if (t + u < n) add(v[u]);
  }
}
// Loads in flight per lane: 8 for the density pass's scalar shares (faster
// at every size measured), 2 for the gradient pass's shares of 16 or 32
// bytes.  Measured on
// the mixed gradient receive pass, RU = 1/2/4/8: a 6^3-cell lattice (latency-
// bound) 19.6/15.0/12.2/10.9 us, a 40^3-cell one (streaming at HBM speed,
// where more rows in flight and the registers they hold cost waves)
// 74.6/76.3/83.0/87.9 us.
constexpr int RECV_RU_SCALAR = 8, RECV_RU_WIDE = 2;

// Lanes per site of the pair-owner passes.  A small lattice gives too few
// waves to fill the GPU with one lane per site, and each lane would walk its
// ~40 owned pairs x 12 points serially; splitting a site's pairs over K lanes
// (combined in a fixed shuffle order) cuts that latency K-fold.  Large
// lattices already fill the GPU in a batch, where one lane per site keeps the
// neighbor-list loads coalesced.  K comes from the process-wide table of
// a2c_tune.h (lattice size class and precision), so it depends only on the
// lattice: its results are the same in any batch.
int lanes_per_site(int n_sites, bool fp64);
constexpr int MAX_LANES_PER_SITE = 16;

// Fixed-order sum over the K lanes of a site (K a power of two <= 64, lanes
// contiguous within a wave); the total lands in the group's first lane.
template <typename T>
__device__ inline T site_sum(T v, int K) {
  for (int off = K >> 1; off > 0; off >>= 1) v += __shfl_down(v, off, K);
  return v;
}

// 3-point Gauss-Hermite product rule on the relative coordinate: per axis
// nodes -sqrt(3), 0, sqrt(3) (standard normal), weights 1/6, 2/3, 1/6.
// Node k in 0..26 has per-axis offsets a_c = (k / 3^c) % 3 - 1, so
// z = sqrt(3) a and weight w = prod_c gh_w(a_c).  The nodes of one pair are
// r_k = d + s z_k with s = sqrt(sigma_i^2 + sigma_j^2), and
// dr_k/dsigma_i = z_k sigma_i / s.
constexpr double GH_Z = 1.7320508075688772935;
__host__ __device__ constexpr double gh_w(int a) { return a == 0 ? 2.0 / 3.0 : 1.0 / 6.0; }
__device__ inline void gh_node(int k, double z[3], double& w) {
  const int a0 = k % 3 - 1, a1 = (k / 3) % 3 - 1, a2 = k / 9 - 1;
  z[0] = GH_Z * a0;
  z[1] = GH_Z * a1;
  z[2] = GH_Z * a2;
  w = gh_w(a0) * gh_w(a1) * gh_w(a2);
}


// General model of a context (a2c_fe_set_potential_*, a2c_fe_set_quadrature):
// the functions of a tabulated potential as cubic pieces on uniform grids,
// and the rule as a node table.  Function order: F_Pd, F_H (embedding),
// f_Pd, f_H (density), r phi_PdPd, r phi_PdH, r phi_HH (pair).
enum { GEN_F_PD = 0, GEN_F_H, GEN_RHO_PD, GEN_RHO_H, GEN_RPHI_PD, GEN_RPHI_PDH, GEN_RPHI_H, GEN_NFN };
struct GenFn {
  int32_t n = 0;        // nodes (n - 1 pieces), node k at k * h
  int32_t off = 0;      // first piece in the coefficient arrays
  double ih = 0.0;      // 1 / h
  double yend = 0.0;    // value and slope at the last node (embedding
  double dend = 0.0;    // functions continue along it)
};
struct GenDev {
  int32_t pot = 0;      // 0: compiled-in Pd-H closed forms; 1: tabulated
  int32_t K = 0;        // nodes of the rule
  int32_t rel = 0;      // 1: nodes z on the relative coordinate, r = d + s z,
                        //    s^2 = sigma_i^2 + sigma_j^2 (the Gauss-Hermite rule);
                        // 0: 6-D nodes, r = d + sqrt(2) (sigma_i xi1 - sigma_j xi2)
  const double* nodes = nullptr;  // K x 8: xi1[3] (or z), xi2[3], w, 0
  const double4* tab = nullptr;   // FP64 piece coefficients (c0..c3 in u = x/h - k)
  const float4* tab32 = nullptr;  // FP32 pieces of the five r-functions (mixed)
  GenFn fn[GEN_NFN];
  GenFn fn32[GEN_NFN];  // the FP32 pieces' grids (f_Pd .. r phi_HH)
};

// Minimum-image geometry of one lattice, copied out of its descriptor into
// registers before a pair loop (the loop's stores could otherwise alias the
// descriptor, and its fields would be reloaded for every pair).
struct Image {
  double box[3], ibox[3];
  bool per[3];
};
__device__ __forceinline__ Image image_of(const LatDev& S) {
  Image g;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    g.box[c] = S.box[c];
    g.ibox[c] = S.ibox[c];
    g.per[c] = S.periodic[c] != 0;
  }
  return g;
}

// Logical site block of a workgroup.  Workgroups are dealt to the 8 XCDs
// round-robin by index; this bijection gives XCD k one contiguous range of
// site blocks, so the neighbor data its waves gather stays within its own L2
// instead of every XCD touching the whole lattice.
constexpr int NUM_XCD = 8;
__device__ inline int xcd_block() {
  const int b = blockIdx.x, n = gridDim.x;
  const int x = b % NUM_XCD, q = n / NUM_XCD, r = n % NUM_XCD;
  return x * q + (x < r ? x : r) + b / NUM_XCD;
}

struct Timing;

// a2c_fe_create on a stream of the caller's (a batch's own lattices share
// its stream: creating and destroying one per lattice cost ~1.3 ms), or on a
// stream of its own when shared is null
int fe_create(const a2c_fe_desc* desc, hipStream_t shared, a2c_fe_ctx** out);

}  // namespace a2c

struct a2c_fe_ctx {
  a2c_fe_desc desc;
  int32_t N = 0;
  bool own_stream = true;
  hipStream_t stream = nullptr;
  std::string err;

  // state
  double* dX = nullptr;      // 4N
  double* dG = nullptr;      // 4N
  double* dx = nullptr;      // nH occupancy
  double* drho = nullptr;    // N
  double* dFp = nullptr;     // N  F'(rho)
  double* demb = nullptr;    // N  embedding energy per site (x-weighted for H)
  double* depair = nullptr;  // N  pair-energy partial of the mixed path
  double* dpart = nullptr;   // block partial sums
  double* dF = nullptr;      // 1 objective (device)
  int32_t nblk_red = 0;
  bool have_x = false;
  bool have_nbr = false;

  // neighbor lists
  a2c::Ell ell[5];           // PdPd, PdH, HPd, HH, HH1
  // neighbor-build scratch (grow-only)
  int* nb_cid = nullptr;     // N: cell of each site
  void* nb_order = nullptr;  // N double4: positions in cell order, site index in .w
  int* nb_cc = nullptr;      // cells + 1: cell starts
  int* nb_cursor = nullptr;  // cells + 1: histogram, then placement cursors
  int* nb_misc = nullptr;    // build counters (list widths, totals; copies per workgroup group), displacement, bounding box
  int64_t nb_cap_cells = 0;
  // skin lists: the pair lists at rc + skin and the positions they were built
  // at; while every site is within skin/2 of those positions the lists at rc
  // are filtered from them (nbr.hip)
  a2c::Ell nb_sup[5];
  double* nb_sup_q = nullptr;  // 4N: positions at the skin build
  bool nb_sup_ok = false;      // skin lists valid for the current positions' neighbors
  bool nb_sup_off = false;     // skin lists refused for this context (a row over the cap)
  int32_t nb_builds = 0;       // builds by walking the cells
  int32_t* dpdpdnum = nullptr;  // nPd*3
  a2c_nbr_stats stats{};

  // mixed-precision tables + FP32 mirrors
  float4* dtab = nullptr;    // cubic coefficient tables (5 x 500 intervals)
  double* dtab64 = nullptr;  // FP64 polynomial tables of the exponential functions
  float2* daux = nullptr;    // per site {F'(rho), x} in FP32

  // launch descriptors: this context is a batch of one lattice
  a2c::LatDev* hlat = nullptr;   // pinned host copy
  a2c::LatDev* dlat = nullptr;   // device copy
  int32_t* db2l = nullptr;       // block -> lattice (all zero)
  bool lat_dirty = true;

  // host staging (pinned)
  double* hX = nullptr;
  double* hG = nullptr;

  // occupancy scratch of the chemical potential and the diffusion step
  // (allocated at their first call, kept for the context's lifetime)
  double* occ_g = nullptr;     // nH gamma
  int32_t* occ_full = nullptr; // nH full_H
  double* occ_corr = nullptr;  // N correction sums
  double* occ_uv = nullptr;    // 2 nH edge-term factors of the diffusion step
  double* occ_xn = nullptr;    // nH new occupancy
  double* occ_part = nullptr;  // 4 per block of the balance
  double* occ_bal = nullptr;   // 4: the balance of the last step (a2c_diffusion_balance)
  a2c_diffusion_balance last_balance{};

  // general model (null gen_nodes: the descriptor's rule and the closed forms)
  bool gen_on = false;          // evaluate with the general kernels
  bool gen_table_rule = false;  // a node table was set (else the descriptor's rule)
  a2c::GenDev gen{};
  double* gen_nodes = nullptr;  // device: K x 8
  double4* gen_tab = nullptr;   // device: FP64 pieces
  float4* gen_tab32 = nullptr;  // device: FP32 pieces
  size_t gen_tab_n = 0, gen_tab32_n = 0;

  a2c::DevSys sys() const {
    a2c::DevSys s;
    s.npd = desc.n_pd;
    s.nh = desc.n_h;
    for (int c = 0; c < 3; ++c) {
      s.box[c] = desc.box[c];
      s.ibox[c] = desc.box[c] > 0.0 ? 1.0 / desc.box[c] : 0.0;
      s.periodic[c] = desc.periodic[c];
    }
    s.T = desc.temperature;
    s.rule = desc.quad_rule;
    s.pdpd = ell[0]; s.pdh = ell[1]; s.hpd = ell[2]; s.hh = ell[3]; s.hh1 = ell[4];
    s.x = dx; s.rho = drho; s.Fp = dFp; s.aux = daux; s.emb = demb; s.part = dpart; s.epair = depair;
    return s;
  }
};

namespace a2c {

// error helpers -------------------------------------------------------------
int set_err(a2c_fe_ctx* c, int code, const std::string& msg);
#define A2C_HIP(ctx, call)                                                       \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return ::a2c::set_err((ctx), A2C_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// kernels (launchers) ---------------------------------------------------------
int nbr_build(a2c_fe_ctx* c);                           // nbr.hip
void free_pair_buffers(Ell& L);                         // nbr.hip
// Multi-lattice launches: lats/b2l on the device, nblk = total site blocks,
// nlat = lattices.  The FP64 path fills rho, Fp, emb, G, part, f; the mixed
// path rho, aux, emb, G, part, f.
struct Launch {
  const LatDev* lats;
  const int32_t* b2l;
  int nblk, nlat;
  hipStream_t stream;
  int deriv_mode;
  double fd_eps;
  const float4* tab;  // mixed path
  const double* tab64 = nullptr;  // FP64 pair-owner passes
  const int32_t* ob2l = nullptr;  // pair-owner workgroup -> lattice
  int onblk = 0;                  // pair-owner workgroups
  bool mixed_full = false;    // mixed path: every pair from both ends (A/B reference)
  bool full_pairs64 = false;  // FP64 analytic path: every pair from both ends (A/B reference)
  int rules = 1;              // bit r set: some lattice of the launch uses quadrature rule r
};
bool mixed_full_pairs();  // A2C_MIXED_PAIRS=full selects the both-ends kernels
bool fp64_full_pairs();   // A2C_FP64_PAIRS=full likewise for the FP64 kernels
int run_density64(const Launch& L);
int run_eval64(const Launch& L, Timing* tm);
int run_eval32(const Launch& L, Timing* tm);
int run_final_sum(const Launch& L);
int sync_lattice(a2c_fe_ctx* c, const double* dX, double* dG, double* dF);  // refresh the one-lattice descriptor
int launch_eval64(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm);
int launch_density64(a2c_fe_ctx* c, const double* dX);  // fills drho, dFp, demb
int launch_eval32(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm);
int build_tables(a2c_fe_ctx* c);                        // eval32.hip
int tables_download(a2c_fe_ctx* c, float* out);         // 5*500 float4
void host_tables(std::vector<float4>& t);              // FP32 cubic tables (host)
// FP64 tables of the EAM functions: per function T64_N intervals of 0.05 A
// over [1, 6) A, each a polynomial in the local coordinate u in [0, 1).  Part
// one: the exponential functions (f_H, phi_PdH, phi_H), degree T64_D, T64_P
// doubles per row (c0..c8, one pad); part two, after it: the Pd polynomials
// (f_Pd, phi_Pd), degree T64Q_D, T64Q_P doubles per row (c0..c11).
constexpr int T64_N = 100, T64_D = 8, T64_P = 10, T64_NF = 3;
enum { T64_FH = 0, T64_PHPDH = 1, T64_PHH = 2 };
constexpr int T64Q_D = 11, T64Q_P = 12, T64Q_NF = 2;
enum { T64Q_FPD = 0, T64Q_PHPD = 1 };
constexpr int T64_PART1 = T64_NF * T64_N * T64_P;                // doubles
constexpr int T64_SIZE = T64_PART1 + T64Q_NF * T64_N * T64Q_P;  // doubles
void host_tables64(std::vector<double>& t);
// The FP32 and FP64 EAM tables on one device, uploaded at the first call for
// that device and shared by every context and batch of the process until it
// ends (they depend on nothing but the compiled-in potential); never freed.
struct DeviceTables {
  float4* t32 = nullptr;
  double* t64 = nullptr;
};
hipError_t device_tables(int device, DeviceTables* out);
int host_staging(a2c_fe_ctx* c);  // pinned hX, hG of the host-array entry points (capi.cpp)
// General model (evalgen.hip, general.cpp)
int run_density_gen(const LatDev* dlat, int N, const GenDev& g, int deriv_mode, double eps, hipStream_t st);
int run_eval_gen(const LatDev* dlat, int N, const GenDev& g, int deriv_mode, double eps, bool fp32, hipStream_t st,
                 Timing* tm);
int launch_density_gen(a2c_fe_ctx* c, const double* dX);  // fills drho, dFp, demb
int launch_eval_gen(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm);
int launch_chem_gen(a2c_fe_ctx* c, const double* dX, const double* coe_h, double gamma_bd, double* dgam,
                    int32_t* dfull, const double* dcorr);
void gen_free(a2c_fe_ctx* c);
int lmvm_minimize(a2c_fe_ctx* c, const double* X0, const a2c_lmvm_opts* o, double* Xout, a2c_lmvm_report* rep);
int launch_chem(a2c_fe_ctx* c, const double* dX, const double* coe_pd, const double* coe_h,
                double gamma_bd, double* dgam, int32_t* dfull, double* dcorr);  // occupancy.hip
// Scratch of one diffusion step: f (n_h doubles), block partials of the
// balance (4 per block of 256 H sites) and the balance itself (4 doubles).
struct DiffuseScratch {
  double* uv;   // 2 nH: (x e^{f/2}, (1-x) e^{-f/2}) per site, f = gamma - ln(x/(1-x)); NaN where |f| >= 700
  double* part;
  double* bal;
};
int diffuse_blocks(int nh);
int launch_diffuse(a2c_fe_ctx* c, double dt, double prefactor, double xLb, double xUb,
                   const double* dgam, const int32_t* dfull, const DiffuseScratch& w, double* dxnew,
                   Timing* tm = nullptr);

// Event timing of individual kernels on the context stream: mark() records
// the next preallocated event; durations are differences of consecutive marks.
struct Timing {
  hipEvent_t* ev = nullptr;
  int cap = 0;
  int n = 0;
  const char* name[8] = {};
  bool on = false;
  void mark(hipStream_t s, const char* nm) {
    if (!on || n >= cap) return;
    if (n < 8) name[n] = nm;
    hipEventRecord(ev[n], s);
    ++n;
  }
};

}  // namespace a2c
