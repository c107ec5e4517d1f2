// Cell-list neighbor build: replaces findNeighbors (neighbors.cpp:8-113) and
// Minimizer::findLocalNeighbors (minimizer.cpp:833-1144).
//
// Set semantics (bit-exact with the oracle): j != i and
// ((dx*dx + dy*dy) + dz*dz) <= rc*rc, d = q_j - q_i in FP64, evaluated with
// contraction off so no FMA changes a borderline pair (neighbors.cpp:79-82).
// Periodic axes use d -= L*rint(d/L).  Shells: PdPdnum counts d2 < 0.75a^2,
// 0.75a^2 < d2 < 1.25a^2, else (minimizer.cpp:935-947); HH1 = H neighbors with
// d2 < 0.75a^2 (:999-1004).
//
// Passes, all on the device (two host synchronisations: the list widths
// after the count pass, the totals at the end):
//   k_bin     cell of each site and the cell histogram
//   k_scan    exclusive scan of the histogram (one workgroup)
//   k_place   positions copied into cell order (with the site index), so a
//             candidate is one load; three x-adjacent cells of a stencil row
//             are one contiguous range
//   k_site<false>  one lane per site: the candidates of its 27 cells (four
//             loads in flight), row lengths, the owned/received split of each
//             pair list, and their maxima
//   k_site<true>   the same walk keeps each hit (neighbor index and shell
//             bits) in the lane's LDS column; the wave sorts each row by
//             ranking, then writes the rows slot by slot, 64 consecutive
//             rows per store
// Rows come out sorted whatever order the cell placement used.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fe_internal.hpp"

namespace a2c {
namespace {

struct Grid {
  int nc[3];
  double lo[3], w[3];  // origin and cell width per axis
  int periodic[3];
  double box[3];
};

__device__ __forceinline__ int cell_axis(const Grid& g, double x, int c) {
  if (g.periodic[c]) x -= g.box[c] * floor(x / g.box[c]);
  int k = (int)floor((x - g.lo[c]) / g.w[c]);
  return min(max(k, 0), g.nc[c] - 1);
}

__global__ void k_bin(const double* __restrict__ X, int N, Grid g, int* __restrict__ cid,
                      int* __restrict__ ccount) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int cx = cell_axis(g, X[4 * s], 0), cy = cell_axis(g, X[4 * s + 1], 1), cz = cell_axis(g, X[4 * s + 2], 2);
  int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
  cid[s] = c;
  atomicAdd(&ccount[c], 1);
}

// start[i] = sum of cnt[0..i), start[n] = the total; cursor = start[0..n).
// One workgroup: each thread scans a contiguous chunk, the chunk sums are
// scanned across the workgroup.
constexpr int SCAN_T = 1024;
__global__ __launch_bounds__(SCAN_T) void k_scan(const int* cnt, int n, int* __restrict__ start, int* cursor) {
  __shared__ int sm[SCAN_T];
  const int per = (n + SCAN_T - 1) / SCAN_T;
  const int b = min(n, (int)threadIdx.x * per), e = min(n, b + per);
  int acc = 0;
  for (int i = b; i < e; ++i) acc += cnt[i];
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < SCAN_T; off <<= 1) {  // inclusive Hillis-Steele scan of the chunk sums
    const int v = threadIdx.x >= off ? sm[threadIdx.x - off] : 0;
    __syncthreads();
    sm[threadIdx.x] += v;
    __syncthreads();
  }
  int run = sm[threadIdx.x] - acc;
  for (int i = b; i < e; ++i) {  // cnt and cursor may be one array: read before writing
    const int ci = cnt[i];
    start[i] = run;
    cursor[i] = run;
    run += ci;
  }
  if (threadIdx.x == SCAN_T - 1) start[n] = sm[SCAN_T - 1];
}

// positions in cell order, the site index in .w (exact in a double)
__global__ void k_place(const double* __restrict__ X, int N, const int* __restrict__ cid, int* __restrict__ cursor,
                        double4* __restrict__ qs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  const int pos = atomicAdd(&cursor[cid[s]], 1);
  qs[pos] = make_double4(X[4 * s], X[4 * s + 1], X[4 * s + 2], (double)s);
}

struct NbrArgs {
  const double* X;
  const double4* qs;   // cell-ordered positions
  int npd, N;
  Grid g;
  double ibox[3];
  const int* cid;
  const int* cstart;
  double rc2, sh1, sh2;
  int cap;
  int width;           // fill pass: LDS slots per lane (the longest row)
  int32_t* cnt[5];     // row lengths: PdPd, PdH (Pd rows); HPd, HH, HH1 (H rows)
  int32_t* ocnt[4];    // owned entries of the four pair lists
  int32_t* rcnt[4];    // received entries
  int32_t* idx[5];     // ELL arrays (fill pass)
  int32_t* pdpdnum;    // nPd*3 (fill pass)
  int32_t* maxw;       // 14 ints: width of the 5 lists, owned widths of the 4 pair lists, received
                       // widths, max total (count pass)
  int32_t* overflow;   // 1 if a site exceeds cap
  unsigned long long* totals;  // entries of the 5 lists (fill pass)
};

// Ownership of the unordered pair {s, j} (global site indices) in the
// pair-once evaluation: (s < j) XOR ((s ^ j) & 1), which hands each site
// about half of every list.
__device__ __forceinline__ bool pair_owner(int s, int j) { return (s < j) != (((s ^ j) & 1) != 0); }

// d - L rint(d / L) on a periodic axis.  The quotient is formed as d (1/L):
// the two agree except within an ulp of a half-integer, where the division
// itself decides (a candidate about L/2 away, so the chosen image and the set
// stay exactly those of the division).
__device__ __forceinline__ double wrap(double d, double L, double iL) {
  const double q = d * iL;
  const double f = fabs(q - trunc(q));
  return d - L * rint(fabs(f - 0.5) < 1e-9 ? d / L : q);
}

// Candidates of one stencil row (dy, dz) of cell (cx, cy, cz): the x cells
// cx-1..cx+1 are consecutive cell indices, one range [b0, e0), unless the
// row wraps a periodic x boundary (a second range [b1, e1)).  Axes with one
// or two cells contribute each cell once.
__device__ __forceinline__ void row_ranges(const NbrArgs& a, int cx, int yy, int zz, int& b0, int& e0, int& b1,
                                           int& e1) {
  const int n0 = a.g.nc[0];
  const int base = (zz * a.g.nc[1] + yy) * n0;
  int x0, x1, y0 = 0, y1 = -1;  // [x0, x1], then [y0, y1] (inclusive cell ranges)
  if (n0 < 3) {
    x0 = 0; x1 = n0 - 1;
  } else if (a.g.periodic[0] && cx == 0) {
    x0 = 0; x1 = 1; y0 = n0 - 1; y1 = n0 - 1;
  } else if (a.g.periodic[0] && cx == n0 - 1) {
    x0 = n0 - 2; x1 = n0 - 1; y0 = 0; y1 = 0;
  } else {
    x0 = max(cx - 1, 0); x1 = min(cx + 1, n0 - 1);
  }
  b0 = a.cstart[base + x0];
  e0 = a.cstart[base + x1 + 1];
  if (y1 >= y0) {
    b1 = a.cstart[base + y0];
    e1 = a.cstart[base + y1 + 1];
  } else {
    b1 = e1 = 0;
  }
}

// stencil coordinate k in 0..2 of axis ax around c: the wrapped (periodic)
// or clipped (-1: outside) cell, each distinct cell once
__device__ __forceinline__ int stencil_axis(const NbrArgs& a, int ax, int c, int k) {
  const int n = a.g.nc[ax];
  if (a.g.periodic[ax]) {
    if (n == 1) return k == 1 ? 0 : -1;  // periodic axes have 1 or >= 3 cells
    return (c + k - 1 + n) % n;
  }
  const int v = c + k - 1;
  return (v >= 0 && v < n) ? v : -1;
}

constexpr int NT = 64;  // sites (rows) per workgroup of the site passes: one lane each
constexpr int CS = NT + 1;  // LDS: slot t of lane l at t * CS + l (a column and a slot row are both
                            // conflict-free)
constexpr int MAXK = NBR_CAP / NT;  // keys of one row per lane in the ranking sort

// Build counters (list widths and totals) in SPREAD copies, a workgroup
// adding to copy blockIdx % SPREAD: tens of thousands of atomics on one
// address serialise in its L2 channel (measured: 64-thread workgroups
// spent ~0.6 ms of a 0.7 ms pass there).  The host reduces the copies.
// Layout of the context's misc buffer (bytes): SPREAD x 16 int32 (the 14
// maxima, the overflow flag in copy 0, one spare), SPREAD x 5 uint64
// (totals), the squared largest displacement, 6 bounding-box keys.
constexpr int SPREAD = 64;
constexpr size_t MISC_TOT = SPREAD * 16 * sizeof(int32_t);
constexpr size_t MISC_DMAX = MISC_TOT + SPREAD * 5 * sizeof(unsigned long long);
constexpr size_t MISC_BBOX = MISC_DMAX + sizeof(unsigned long long);
constexpr size_t MISC_BYTES = MISC_BBOX + 6 * sizeof(unsigned long long);
__device__ __forceinline__ int32_t* maxw_copy(int32_t* m) { return m + (blockIdx.x % SPREAD) * 16; }
__device__ __forceinline__ unsigned long long* tot_copy(unsigned long long* t) { return t + (blockIdx.x % SPREAD) * 5; }

// Sort each row of the wave (lane l's row in LDS column l) by key, one row at
// a time: every lane ranks its (at most K) keys of the row against all of the
// row's keys, read by broadcast, then writes them to their ranks (keys are
// distinct, so the ranks are a permutation).
template <int K>
__device__ __forceinline__ void rank_rows(uint32_t* buf, int n, int lane) {
  for (int rl = 0; rl < NT; ++rl) {
    const int nr = __shfl(n, rl);
    if (nr < 2) continue;
    uint32_t mine[K];
    int rk[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
      const int t = lane + NT * q;
      mine[q] = t < nr ? buf[t * CS + rl] : 0u;
      rk[q] = 0;
    }
    for (int t = 0; t < nr; ++t) {
      const uint32_t o = buf[t * CS + rl];
#pragma unroll
      for (int q = 0; q < K; ++q) rk[q] += o < mine[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < K; ++q)
      if (lane + NT * q < nr) buf[rk[q] * CS + rl] = mine[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// One lane per site (consecutive rows of one species per workgroup; Pd-row
// workgroups first).  Each lane walks the candidates of its 27 cells from
// the cell-ordered copy, four loads in flight.  Count pass (FILL = false):
// row lengths, owned/received counts and their maxima.  Fill pass: each hit's
// key (j << 2 | shell bits) goes into the lane's LDS column
// (slot t of lane l at t * CS + l) in the order found, the wave ranks each
// row's keys (the rows sorted by neighbor index), then the rows are written
// slot by slot: 64 consecutive rows per store, one 256-byte line.
template <bool FILL>
__global__ __launch_bounds__(NT) void k_site(NbrArgs a, int nblk_pd) {
  extern __shared__ uint32_t buf[];
  const int lane = threadIdx.x;
  const bool pdb = (int)blockIdx.x < nblk_pd;
  const int nrows = pdb ? a.npd : a.N - a.npd;
  const int row = (pdb ? (int)blockIdx.x : (int)blockIdx.x - nblk_pd) * NT + lane;
  const bool live = row < nrows;
  const int s = (pdb ? 0 : a.npd) + (live ? row : 0);
  const A2C_GLOBAL double4* __restrict__ qs = gmem(a.qs);
  int n = 0, npn = 0, nh1 = 0, op = 0, oh = 0, sh0 = 0, sh1n = 0;
  if (live) {
    const double qi0 = a.X[4 * (size_t)s], qi1 = a.X[4 * (size_t)s + 1], qi2 = a.X[4 * (size_t)s + 2];
    const int c = a.cid[s];
    const int cx = c % a.g.nc[0], cy = (c / a.g.nc[0]) % a.g.nc[1], cz = c / (a.g.nc[0] * a.g.nc[1]);
    for (int kz = 0; kz < 3; ++kz) {
      const int zz = stencil_axis(a, 2, cz, kz);
      if (zz < 0) continue;
      for (int ky = 0; ky < 3; ++ky) {
        const int yy = stencil_axis(a, 1, cy, ky);
        if (yy < 0) continue;
        int b0, e0, b1, e1;
        row_ranges(a, cx, yy, zz, b0, e0, b1, e1);
        const int l0 = e0 - b0, tot = l0 + (e1 - b1);
        for (int v0 = 0; v0 < tot; v0 += 4) {
          double4 cq[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int v = v0 + u;
            cq[u] = make_double4(0.0, 0.0, 0.0, -1.0);
            if (v < tot) cq[u] = qs[v < l0 ? b0 + v : b1 + (v - l0)];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = (int)cq[u].w;
            double d0 = cq[u].x - qi0, d1 = cq[u].y - qi1, dz = cq[u].z - qi2;
            if (a.g.periodic[0]) d0 = wrap(d0, a.g.box[0], a.ibox[0]);
            if (a.g.periodic[1]) d1 = wrap(d1, a.g.box[1], a.ibox[1]);
            if (a.g.periodic[2]) dz = wrap(dz, a.g.box[2], a.ibox[2]);
            const double d2 = d0 * d0 + d1 * d1 + dz * dz;
            if (!(j >= 0 && j != s && d2 <= a.rc2)) continue;
            const bool jpd = j < a.npd;
            ++n;
            if (jpd) ++npn;
            if (!FILL) {
              const bool own = pair_owner(s, j);
              if (!jpd && d2 < a.sh1) ++nh1;
              if (own) (jpd ? op : oh)++;
            } else if (n <= a.width) {
              if (jpd) {
                if (d2 < a.sh1) ++sh0;
                else if (d2 > a.sh1 && d2 < a.sh2) ++sh1n;
              }
              buf[(n - 1) * CS + lane] = ((uint32_t)j << 2) |
                                         (jpd ? (d2 < a.sh1 ? 1u : ((d2 > a.sh1 && d2 < a.sh2) ? 2u : 0u))
                                              : (d2 < a.sh1 ? 1u : 0u));
            }
          }
        }
      }
    }
    if (n > a.cap) atomicOr(a.overflow, 1);
  }
  const int nhn = n - npn;
  const int lp = pdb ? 0 : 2, lh = pdb ? 1 : 3;
  if (!FILL) {
    if (live) {
      a.cnt[lp][row] = npn;
      a.cnt[lh][row] = nhn;
      a.ocnt[lp][row] = op;
      a.rcnt[lp][row] = npn - op;
      a.ocnt[lh][row] = oh;
      a.rcnt[lh][row] = nhn - oh;
      if (!pdb) a.cnt[4][row] = nh1;
    }
    // maxima over the wave, then one atomic per quantity
    int v[14] = {0};
    v[lp] = npn; v[lh] = nhn; v[4] = pdb ? 0 : nh1;
    v[5 + lp] = op; v[5 + lh] = oh; v[9 + lp] = npn - op; v[9 + lh] = nhn - oh; v[13] = n;
#pragma unroll
    for (int k = 0; k < 14; ++k) {
      int m = v[k];
      for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
      v[k] = m;
    }
    if (lane == 0) {
      int32_t* mw = maxw_copy(a.maxw);
#pragma unroll
      for (int k = 0; k < 14; ++k)
        if (v[k]) atomicMax(&mw[k], v[k]);
    }
    return;
  }
  // each row of the wave sorted by neighbor index
  if (a.width <= 2 * NT) rank_rows<2>(buf, n, lane);
  else rank_rows<MAXK>(buf, n, lane);
  if (live && pdb) {
    a.pdpdnum[3 * (size_t)row] = sh0;
    a.pdpdnum[3 * (size_t)row + 1] = sh1n;
    a.pdpdnum[3 * (size_t)row + 2] = npn - sh0 - sh1n;
  }
  int mp = npn, mh = nhn;  // the wave's longest rows
  for (int off = 32; off > 0; off >>= 1) {
    mp = max(mp, __shfl_xor(mp, off));
    mh = max(mh, __shfl_xor(mh, off));
  }
  int32_t* __restrict__ oP = a.idx[lp];
  int32_t* __restrict__ oH = a.idx[lh];
  for (int t = 0; t < mp; ++t)
    if (t < npn) oP[(size_t)t * nrows + row] = (int)(buf[t * CS + lane] >> 2);
  int k1 = 0;
  for (int t = 0; t < mh; ++t)
    if (t < nhn) {
      const uint32_t key = buf[(npn + t) * CS + lane];
      const int h = (int)(key >> 2) - a.npd;
      oH[(size_t)t * nrows + row] = h;
      if (!pdb && (key & 1u)) a.idx[4][(size_t)(k1++) * nrows + row] = h;  // HH1, in order
    }
  // totals of the five lists: wave sums, one atomic each
  unsigned long long tp = npn, th = nhn, t1 = k1;
  for (int off = 32; off > 0; off >>= 1) {
    tp += __shfl_xor(tp, off);
    th += __shfl_xor(th, off);
    t1 += __shfl_xor(t1, off);
  }
  if (lane == 0) {
    unsigned long long* tt = tot_copy(a.totals);
    atomicAdd(&tt[lp], tp);
    atomicAdd(&tt[lh], th);
    if (!pdb) atomicAdd(&tt[4], t1);
  }
}

// Split of each pair-list row into owned and received halves (pair-once
// evaluation).  Ownership of {s, j} (global site indices):
// (s < j) XOR ((s ^ j) & 1), which hands each site about half of every list.
// Both halves keep the row's (sorted) order.

// Both halves keep the row's index order: consecutive lanes (sites) then
// gather consecutive neighbor sites at the same slot, so the state and F'
// loads of a pair slot coalesce.  (Ordering the owned half by distance, to
// align table rows across lanes, measured 40% slower: it scatters those
// gathers.)
__global__ void k_split(Ell L, int s_off, int j_off) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L.rows) return;
  int32_t* rpos = reinterpret_cast<int32_t*>(L.rbuf);  // full-row slot -> received slot (-1 if owned)
  const int n = L.cnt[r];
  const int s = s_off + r;
  int no = 0, nr = 0;
  for (int t = 0; t < n; ++t) {
    const size_t e = (size_t)t * L.rows + r;
    const int j = L.idx[e];
    if (pair_owner(s, j_off + j)) {
      L.oj[(size_t)no * L.rows + r] = j;
      rpos[e] = -1;
      ++no;
    } else {
      rpos[e] = nr;
      ++nr;
    }
  }
}

// For every owned entry (r, j): the received slot of r in the partner row j,
// found by binary search of r in that sorted row (-1 if absent, which the
// symmetric lists never produce).
__global__ void k_orv(Ell L, Ell P) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L.rows) return;
  const int32_t* prpos = reinterpret_cast<const int32_t*>(P.rbuf);
  const int n = L.ocnt[r];
  for (int t = 0; t < n; ++t) {
    const size_t e = (size_t)t * L.rows + r;
    const int j = L.oj[e];
    int lo = 0, hi = P.cnt[j] - 1, k = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = P.idx[(size_t)mid * P.rows + j];
      if (v == r) { k = mid; break; }
      if (v < r) lo = mid + 1;
      else hi = mid - 1;
    }
    L.orv[e] = k >= 0 ? prpos[(size_t)k * P.rows + j] : -1;
  }
}

// Bounding box of the positions (non-periodic axes): per block, then one
// atomic per block and quantity on the ordered-integer image of the double.
__device__ __forceinline__ unsigned long long dkey(double x) {  // monotone map double -> uint64
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__global__ __launch_bounds__(256) void k_bbox(const double* __restrict__ X, int N, unsigned long long* out) {
  __shared__ unsigned long long red[6][256];
  unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {0ull, 0ull, 0ull};
  for (int s = blockIdx.x * 256 + threadIdx.x; s < N; s += gridDim.x * 256)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const unsigned long long k = dkey(X[4 * (size_t)s + c]);
      lo[c] = min(lo[c], k);
      hi[c] = max(hi[c], k);
    }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    red[c][threadIdx.x] = lo[c];
    red[3 + c][threadIdx.x] = hi[c];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        red[c][threadIdx.x] = min(red[c][threadIdx.x], red[c][threadIdx.x + w]);
        red[3 + c][threadIdx.x] = max(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) atomicMin(&out[threadIdx.x], red[threadIdx.x][0]);
  else if (threadIdx.x < 6) atomicMax(&out[threadIdx.x], red[threadIdx.x][0]);
}
double dkey_inv(unsigned long long k) {
  const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  std::memcpy(&x, &u, 8);
  return x;
}

// ---------------------------------------------------------------- skin lists
// Largest squared displacement of any site from its position at the skin
// build (minimum image on periodic axes), as the bit pattern of a
// non-negative double (ordered like the value).
__global__ __launch_bounds__(256) void k_maxdisp(const double* __restrict__ X, const double* __restrict__ X0, int N,
                                                 Grid g, unsigned long long* out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int s = blockIdx.x * 256 + threadIdx.x; s < N; s += gridDim.x * 256) {
    double d2 = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double d = X[4 * (size_t)s + c] - X0[4 * (size_t)s + c];
      if (g.periodic[c]) d -= g.box[c] * rint(d / g.box[c]);
      d2 += d * d;
    }
    m = fmax(m, d2);
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

struct RefreshArgs {
  const double* X;
  int npd, N;
  Grid g;
  double ibox[3];
  double rc2, sh1, sh2;
  Ell sup[4];          // skin lists: PdPd, PdH (Pd rows), HPd, HH (H rows); rbuf: per entry its fate (below)
  Ell out[5];          // lists at rc (idx, cnt; pair lists also ocnt, rcnt, oj)
  int32_t* pdpdnum;
  int32_t* maxw;       // 14 ints as in the count pass
  unsigned long long* totals;
  const unsigned long long* dmax;  // squared largest displacement (a double's bits), 0 after a skin build
};

// One lane per row of one list (one wave per workgroup; the workgroups of
// PdPd, PdH, HPd and HH rows in that order): the row's skin-list entries
// still within rc, in the skin list's (sorted) order, into the list at rc
// with its owned/received split, and the Pd shell counts (PdPd) or HH1
// (HH) -- the count and fill passes' results for the same positions,
// without the cell walk.  An entry whose distance at the skin build (kept as
// a float beside it) is further than 2 delta + 1e-6 A from rc and from the
// shell radii, delta the largest displacement since then, is decided by that
// distance alone: no site pair moved closer or apart by more than 2 delta,
// and the float carries 3e-7 A of rounding.  The others are tested with the
// current FP64 positions exactly as the cell walk tests them.  The fate of
// every skin entry goes to the skin list's rbuf: -2 dropped, -3 - k kept as
// the row's k-th owned entry, else its slot in the received half.  Kept
// entries are staged in the lane's LDS column (neighbor << 2 | owned << 1 |
// first shell), then each output array is written slot by slot, 64
// consecutive rows per store.
__global__ __launch_bounds__(NT) void k_refresh(RefreshArgs a, int4 blk0) {
  extern __shared__ uint32_t col[];
  const int b = blockIdx.x;
  const int w = b < blk0.y ? 0 : (b < blk0.z ? 1 : (b < blk0.w ? 2 : 3));  // the list of this workgroup
  const int b0 = w == 0 ? 0 : (w == 1 ? blk0.y : (w == 2 ? blk0.z : blk0.w));
  const bool pdr = w < 2;                                 // Pd rows
  const int nrows = pdr ? a.npd : a.N - a.npd;
  const int lane = (int)threadIdx.x;
  const int row = (b - b0) * NT + lane;
  const bool live = row < nrows;
  const int s = (pdr ? 0 : a.npd) + (live ? row : 0);
  const int joff = (w == 0 || w == 2) ? 0 : a.npd;       // neighbors: Pd for PdPd and HPd
  const bool shells = w == 0, first = w == 3;             // PdPdnum / HH1
  const A2C_GLOBAL double4* __restrict__ X = gmem(reinterpret_cast<const double4*>(a.X));
  const double band = 2.0 * sqrt(__longlong_as_double((long long)*a.dmax)) + 1e-6;
  const double r1 = sqrt(a.sh1), r2 = sqrt(a.sh2), rc = sqrt(a.rc2);
  const double4 qi = X[s];
  const Ell S = a.sup[w];
  const Ell L = a.out[w];
  const A2C_GLOBAL int32_t* __restrict__ sidx = gmem(S.idx);
  const A2C_GLOBAL float* __restrict__ sdist = gmem(reinterpret_cast<const float*>(S.oj));
  A2C_GLOBAL int32_t* __restrict__ fate = gmem(reinterpret_cast<int32_t*>(S.rbuf));
  const size_t srows = S.rows, orows = L.rows;
  const int n = live ? gmem(S.cnt)[row] : 0;
  int nk = 0, no = 0, nr = 0, nf = 0, sh0 = 0, sh1n = 0;
  constexpr int B = 8;  // entries whose index and distance are loaded before any is tested
  int jb[B];
  float db[B];
  for (int t = 0; t < n; ++t) {
    const size_t se = (size_t)t * srows + row;
    if (t % B == 0) {
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const bool in = t + u < n;
        jb[u] = in ? sidx[se + (size_t)u * srows] : 0;
        db[u] = in ? sdist[se + (size_t)u * srows] : 0.0f;
      }
    }
    const int jl = jb[t % B];
    const int j = joff + jl;
    const double df = (double)db[t % B];
    const bool near_rc = fabs(df - rc) <= band;
    const bool near_sh = (shells || first) && (fabs(df - r1) <= band || (shells && fabs(df - r2) <= band));
    bool keep;
    int cls;  // shell class: 0 below r1, 1 between, 2 beyond
    if (near_rc || near_sh) {
      const double4 qj = X[j];
      double d0 = qj.x - qi.x, d1 = qj.y - qi.y, dz = qj.z - qi.z;
      if (a.g.periodic[0]) d0 = wrap(d0, a.g.box[0], a.ibox[0]);
      if (a.g.periodic[1]) d1 = wrap(d1, a.g.box[1], a.ibox[1]);
      if (a.g.periodic[2]) dz = wrap(dz, a.g.box[2], a.ibox[2]);
      const double d2 = d0 * d0 + d1 * d1 + dz * dz;
      keep = d2 <= a.rc2;
      cls = d2 < a.sh1 ? 0 : ((d2 > a.sh1 && d2 < a.sh2) ? 1 : 2);
    } else {
      keep = df < rc;
      cls = df < r1 ? 0 : (df < r2 ? 1 : 2);
    }
    if (!keep) {
      fate[se] = -2;
      continue;
    }
    const bool owned = pair_owner(s, j);
    fate[se] = owned ? -3 - no : nr;
    owned ? ++no : ++nr;
    if (shells) {
      if (cls == 0) ++sh0;
      else if (cls == 1) ++sh1n;
    }
    const bool f1 = first && cls == 0;
    nf += f1;
    col[nk * CS + lane] = ((uint32_t)jl << 2) | (owned ? 2u : 0u) | (f1 ? 1u : 0u);
    ++nk;
  }
  // the wave's longest row, owned half and HH1 row: slot loops every lane joins
  int mk = nk, mo = no, mf = nf;
  for (int off = 32; off > 0; off >>= 1) {
    mk = max(mk, __shfl_xor(mk, off));
    mo = max(mo, __shfl_xor(mo, off));
    mf = max(mf, __shfl_xor(mf, off));
  }
  A2C_GLOBAL int32_t* __restrict__ oidx = gmem(L.idx);
  for (int u = 0; u < mk; ++u)
    if (u < nk) oidx[(size_t)u * orows + row] = (int)(col[u * CS + lane] >> 2);
  A2C_GLOBAL int32_t* __restrict__ ooj = gmem(L.oj);
  for (int u = 0, q = 0; u < mo; ++u) {
    if (u >= no) continue;
    while (!(col[q * CS + lane] & 2u)) ++q;
    ooj[(size_t)u * orows + row] = (int)(col[q * CS + lane] >> 2);
    ++q;
  }
  if (first) {
    A2C_GLOBAL int32_t* __restrict__ h1 = gmem(a.out[4].idx);
    const size_t rows4 = a.out[4].rows;
    for (int u = 0, q = 0; u < mf; ++u) {
      if (u >= nf) continue;
      while (!(col[q * CS + lane] & 1u)) ++q;
      h1[(size_t)u * rows4 + row] = (int)(col[q * CS + lane] >> 2);
      ++q;
    }
  }
  if (live) {
    gmem(L.cnt)[row] = nk;
    gmem(L.ocnt)[row] = no;
    gmem(L.rcnt)[row] = nr;
    if (first) gmem(a.out[4].cnt)[row] = nf;
    if (shells) {
      A2C_GLOBAL int32_t* pn = gmem(a.pdpdnum);
      pn[3 * (size_t)row] = sh0;
      pn[3 * (size_t)row + 1] = sh1n;
      pn[3 * (size_t)row + 2] = nk - sh0 - sh1n;
    }
  }
  // maxima (row length, owned and received halves, HH1) and totals of the list
  int v[4] = {nk, no, nr, nf};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int m = v[k];
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    v[k] = m;
  }
  unsigned long long tk = nk, tf = nf;
  for (int off = 32; off > 0; off >>= 1) {
    tk += __shfl_xor(tk, off);
    tf += __shfl_xor(tf, off);
  }
  if (lane == 0) {
    int32_t* mw = maxw_copy(a.maxw);
    unsigned long long* tt = tot_copy(a.totals);
    if (v[0]) atomicMax(&mw[w], v[0]);
    if (v[1]) atomicMax(&mw[5 + w], v[1]);
    if (v[2]) atomicMax(&mw[9 + w], v[2]);
    if (first && v[3]) atomicMax(&mw[4], v[3]);
    atomicAdd(&tt[w], tk);
    if (first) atomicAdd(&tt[4], tf);
  }
}

// The longest total row (Pd rows: PdPd + PdH, H rows: HPd + HH) after a
// refresh, into maxw[13].
__global__ void k_rowmax(const int32_t* __restrict__ c0, const int32_t* __restrict__ c1, int n, int32_t* maxw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  int m = r < n ? c0[r] + c1[r] : 0;
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(&maxw_copy(maxw)[13], m);
}

// Skin build: the FP64 distance of every skin entry (as the cell walk
// computes it), rounded to a float, into the skin list's oj.
__global__ void k_sdist(Ell L, const double* __restrict__ X, int s_off, int j_off, Grid g, double ix, double iy,
                        double iz) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (r >= L.rows || t >= L.cnt[r]) return;
  const size_t e = (size_t)t * L.rows + r;
  const A2C_GLOBAL double4* __restrict__ Q = gmem(reinterpret_cast<const double4*>(X));
  const double4 qi = Q[s_off + r], qj = Q[j_off + L.idx[e]];
  double d0 = qj.x - qi.x, d1 = qj.y - qi.y, dz = qj.z - qi.z;
  if (g.periodic[0]) d0 = wrap(d0, g.box[0], ix);
  if (g.periodic[1]) d1 = wrap(d1, g.box[1], iy);
  if (g.periodic[2]) dz = wrap(dz, g.box[2], iz);
  reinterpret_cast<float*>(L.oj)[e] = (float)sqrt(d0 * d0 + d1 * d1 + dz * dz);
}

// Skin build: for every skin entry (row r, slot t = blockIdx.y, neighbor j)
// the slot of r in j's row of the partner skin list (binary search of the
// sorted row; -1 if absent, which the symmetric lists never produce), into
// the skin list's orv.
__global__ void k_srev(Ell L, Ell P) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (r >= L.rows || t >= L.cnt[r]) return;
  const size_t e = (size_t)t * L.rows + r;
  const int j = L.idx[e];
  int lo = 0, hi = P.cnt[j] - 1, k = -1;
  while (lo <= hi) {
    const int mid = (lo + hi) >> 1;
    const int v = P.idx[(size_t)mid * P.rows + j];
    if (v == r) { k = mid; break; }
    if (v < r) lo = mid + 1;
    else hi = mid - 1;
  }
  L.orv[e] = k;
}

// Reverse slots after a refresh, one thread per skin entry (slot t =
// blockIdx.y): a kept, owned entry (fate -3 - k) takes the fate of its twin
// in the partner's skin row -- the slot of this pair in the partner's
// received half -- as the reverse slot of the row's k-th owned entry.
__global__ void k_orv_skin(Ell S, Ell SP, Ell L) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = blockIdx.y;
  if (r >= S.rows || t >= gmem(S.cnt)[r]) return;
  const size_t e = (size_t)t * S.rows + r;
  const int f = gmem(reinterpret_cast<const int32_t*>(S.rbuf))[e];
  if (f > -3) return;
  const int j = gmem(S.idx)[e];
  const int u = gmem(S.orv)[e];
  gmem(L.orv)[(size_t)(-3 - f) * L.rows + r] =
      u >= 0 ? gmem(reinterpret_cast<const int32_t*>(SP.rbuf))[(size_t)u * SP.rows + j] : -1;
}

}  // namespace

namespace {

// The counters of a build read back from the misc buffer (the first bytes
// of it): maxima over the copies, the overflow flag, totals summed.
void reduce_counters(const unsigned char* hb, size_t bytes, int hm[16], unsigned long long ht[5]) {
  const int32_t* m = reinterpret_cast<const int32_t*>(hb);
  for (int k = 0; k < 16; ++k) hm[k] = 0;
  for (int c = 0; c < SPREAD; ++c)
    for (int k = 0; k < 14; ++k) hm[k] = std::max(hm[k], m[c * 16 + k]);
  hm[14] = m[14];
  if (!ht || bytes < MISC_DMAX) return;
  const unsigned long long* t = reinterpret_cast<const unsigned long long*>(hb + MISC_TOT);
  for (int k = 0; k < 5; ++k) {
    ht[k] = 0;
    for (int c = 0; c < SPREAD; ++c) ht[k] += t[c * 5 + k];
  }
}

// Device buffers of one set of five lists for rows of npd / nh sites: row
// counts always; ELL arrays, owned halves and received-share buffers for
// list widths w[0..4], owned widths ow[0..3] and received widths rw[0..3]
// (grow-only, with slack so small width changes reuse the arrays).
int size_lists(a2c_fe_ctx* c, Ell* L5, const int* w, const int* ow, const int* rw) {
  const int npd = c->desc.n_pd, nh = c->desc.n_h;
  for (int k = 0; k < 5; ++k) {
    Ell& L = L5[k];
    L.rows = (k < 2) ? npd : nh;
    const int64_t rows = std::max(L.rows, 1);
    if (!L.cnt) A2C_HIP(c, dmalloc(&L.cnt, sizeof(int32_t) * rows));
    if (k < 4 && !L.ocnt) {
      A2C_HIP(c, dmalloc(&L.ocnt, sizeof(int32_t) * rows));
      A2C_HIP(c, dmalloc(&L.rcnt, sizeof(int32_t) * rows));
    }
    if (!w) continue;
    const int64_t n = (int64_t)std::max(w[k], 1) * rows;
    if (n > L.cap) {
      (void)dfree(L.idx);
      L.idx = nullptr;
      L.cap = (int64_t)(std::max(w[k], 1) + 4) * rows;
      A2C_HIP(c, dmalloc(&L.idx, sizeof(int32_t) * L.cap));
    }
    if (k == 4 || !ow) continue;
    const int64_t on = (int64_t)std::max(ow[k], 1) * rows, rn = (int64_t)std::max(rw[k], 1) * rows;
    const int64_t bn = std::max<int64_t>(32 * rn, 4 * (int64_t)std::max(w[k], 1) * rows);
    if (on > L.ocap) {
      (void)dfree(L.oj); (void)dfree(L.orv);
      L.oj = L.orv = nullptr;
      L.ocap = on + 4 * rows;
      A2C_HIP(c, dmalloc(&L.oj, sizeof(int32_t) * L.ocap));
      A2C_HIP(c, dmalloc(&L.orv, sizeof(int32_t) * L.ocap));
    }
    if (bn > L.bcap) {
      (void)dfree(L.rbuf);
      L.rbuf = nullptr;
      L.bcap = bn + 32 * 4 * rows;
      A2C_HIP(c, dmalloc(&L.rbuf, L.bcap));
    }
  }
  return A2C_OK;
}

// Reverse slots of the four pair lists (after their split).
int reverse_slots(a2c_fe_ctx* c, Ell* L5) {
  static const int partner[4] = {0, 2, 1, 3};  // PdPd, PdH<->HPd, HH
  for (int w = 0; w < 4; ++w) {
    const Ell& L = L5[w];
    if (L.rows == 0) continue;
    hipLaunchKernelGGL(k_orv, dim3((L.rows + 255) / 256), dim3(256), 0, c->stream, L, L5[partner[w]]);
    A2C_HIP(c, hipGetLastError());
  }
  return A2C_OK;
}

Grid box_grid(a2c_fe_ctx* c) {
  Grid g{};
  for (int ax = 0; ax < 3; ++ax) {
    g.periodic[ax] = c->desc.periodic[ax];
    g.box[ax] = c->desc.box[ax];
  }
  return g;
}

// The cell walk: lists of radius rc from the positions c->dX into L5 (the
// context's lists, or its skin lists).  halves: also the owned/received
// split and the reverse slots.  hm: widths (as the count pass's maxima),
// ht: totals.
int walk_cells(a2c_fe_ctx* c, double rc, Ell* L5, bool halves, int hm[16], unsigned long long ht[5]) {
  const int N = c->N, npd = c->desc.n_pd, nh = c->desc.n_h;
  const int TB = 256;
  // scratch (grow-only): misc = 16 ints (widths, overflow) + 5 totals + 6 bbox keys
  if (!c->nb_cid) {
    A2C_HIP(c, dmalloc(&c->nb_cid, sizeof(int) * std::max(N, 1)));
    A2C_HIP(c, dmalloc(&c->nb_order, sizeof(double4) * std::max(N, 1)));  // cell-ordered positions
  }
  if (!c->nb_misc) A2C_HIP(c, dmalloc(&c->nb_misc, MISC_BYTES));
  int* misc = c->nb_misc;
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(misc) + MISC_TOT);
  unsigned long long* bbox = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(misc) + MISC_BBOX);
  Grid g = box_grid(c);
  bool need_bbox = false;
  for (int ax = 0; ax < 3; ++ax) need_bbox |= !c->desc.periodic[ax];
  double lo3[3] = {0, 0, 0}, hi3[3] = {1, 1, 1};
  if (need_bbox && N > 0) {
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    A2C_HIP(c, hipMemcpyAsync(bbox, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_bbox, dim3(std::min((N + TB - 1) / TB, 256)), dim3(TB), 0, c->stream, c->dX, N, bbox);
    A2C_HIP(c, hipGetLastError());
    unsigned long long hb[6];
    A2C_HIP(c, hipMemcpyAsync(hb, bbox, sizeof(hb), hipMemcpyDeviceToHost, c->stream));
    A2C_HIP(c, hipStreamSynchronize(c->stream));
    for (int ax = 0; ax < 3; ++ax) {
      lo3[ax] = dkey_inv(hb[ax]);
      hi3[ax] = dkey_inv(hb[3 + ax]);
    }
  }
  for (int ax = 0; ax < 3; ++ax) {
    double lo = 0.0, hi = 1.0;
    if (g.periodic[ax]) {
      lo = 0.0; hi = g.box[ax];
    } else if (N > 0) {
      lo = lo3[ax]; hi = hi3[ax];
    }
    int nc = std::max(1, (int)std::floor((hi - lo) / rc));
    if (g.periodic[ax] && nc < 3) nc = 1;
    g.nc[ax] = nc;
    g.lo[ax] = lo;
    g.w[ax] = (hi - lo) / nc;
    if (!(g.w[ax] > 0)) g.w[ax] = 1.0;
  }
  const long long ncell = (long long)g.nc[0] * g.nc[1] * g.nc[2];
  if (ncell > (1LL << 28)) return set_err(c, A2C_EINVAL, "neighbor grid too large");
  if (ncell + 1 > c->nb_cap_cells) {
    (void)dfree(c->nb_cc);
    (void)dfree(c->nb_cursor);
    c->nb_cc = c->nb_cursor = nullptr;
    A2C_HIP(c, dmalloc(&c->nb_cc, sizeof(int) * (ncell + 1)));
    A2C_HIP(c, dmalloc(&c->nb_cursor, sizeof(int) * (ncell + 1)));
    c->nb_cap_cells = ncell + 1;
  }
  int *dcid = c->nb_cid, *dcc = c->nb_cc, *dcursor = c->nb_cursor;
  double4* qs = reinterpret_cast<double4*>(c->nb_order);
  A2C_HIP(c, hipMemsetAsync(dcursor, 0, sizeof(int) * (ncell + 1), c->stream));  // the histogram
  A2C_HIP(c, hipMemsetAsync(misc, 0, MISC_DMAX, c->stream));
  if (N > 0) {
    hipLaunchKernelGGL(k_bin, dim3((N + TB - 1) / TB), dim3(TB), 0, c->stream, c->dX, N, g, dcid, dcursor);
    A2C_HIP(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(SCAN_T), 0, c->stream, dcursor, (int)ncell, dcc, dcursor);
  A2C_HIP(c, hipGetLastError());
  if (N > 0) {
    hipLaunchKernelGGL(k_place, dim3((N + TB - 1) / TB), dim3(TB), 0, c->stream, c->dX, N, dcid, dcursor, qs);
    A2C_HIP(c, hipGetLastError());
  }
  if (int rc2 = size_lists(c, L5, nullptr, nullptr, nullptr)) return rc2;  // row counts
  NbrArgs a{};
  a.X = c->dX; a.qs = qs; a.npd = npd; a.N = N; a.g = g; a.cid = dcid; a.cstart = dcc;
  for (int ax = 0; ax < 3; ++ax) a.ibox[ax] = g.box[ax] > 0.0 ? 1.0 / g.box[ax] : 0.0;
  a.rc2 = rc * rc;
  a.sh1 = c->desc.a_pd * c->desc.a_pd * 0.75;
  a.sh2 = c->desc.a_pd * c->desc.a_pd * 1.25;
  a.cap = std::min(c->desc.max_neighbors > 0 ? c->desc.max_neighbors : NBR_CAP, NBR_CAP);
  for (int w = 0; w < 5; ++w) a.cnt[w] = L5[w].cnt;
  for (int w = 0; w < 4; ++w) { a.ocnt[w] = L5[w].ocnt; a.rcnt[w] = L5[w].rcnt; }
  a.maxw = misc;
  a.overflow = misc + 14;
  a.totals = totals;
  const int nblk_pd = (npd + NT - 1) / NT, nblk_h = (nh + NT - 1) / NT;
  if (nblk_pd + nblk_h > 0) {
    hipLaunchKernelGGL(k_site<false>, dim3(nblk_pd + nblk_h), dim3(NT), 0, c->stream, a, nblk_pd);
    A2C_HIP(c, hipGetLastError());
  }
  {
    std::vector<unsigned char> hb(MISC_TOT);
    A2C_HIP(c, hipMemcpyAsync(hb.data(), misc, MISC_TOT, hipMemcpyDeviceToHost, c->stream));
    A2C_HIP(c, hipStreamSynchronize(c->stream));
    reduce_counters(hb.data(), MISC_TOT, hm, nullptr);
  }
  if (hm[14]) return A2C_ENBR_OVERFLOW;
  if (int rc2 = size_lists(c, L5, hm, halves ? hm + 5 : nullptr, hm + 9)) return rc2;
  for (int w = 0; w < 5; ++w) {
    a.idx[w] = L5[w].idx;
    L5[w].width = hm[w];
    if (w < 4) {
      L5[w].owidth = hm[5 + w];
      L5[w].rwidth = hm[9 + w];
    }
  }
  if (!c->dpdpdnum) A2C_HIP(c, dmalloc(&c->dpdpdnum, sizeof(int32_t) * 3 * std::max(npd, 1)));
  a.pdpdnum = c->dpdpdnum;
  a.width = std::max(hm[13], 1);  // LDS slots per lane: the longest row
  if (nblk_pd + nblk_h > 0) {
    hipLaunchKernelGGL(k_site<true>, dim3(nblk_pd + nblk_h), dim3(NT), sizeof(uint32_t) * CS * a.width, c->stream, a,
                       nblk_pd);
    A2C_HIP(c, hipGetLastError());
  }
  if (halves) {
    // owned/received split of the four pair lists, then the partner slots
    static const int s_sp[4] = {0, 0, 1, 1}, j_sp[4] = {0, 1, 0, 1};  // species of rows / of neighbors
    for (int w = 0; w < 4; ++w) {
      const Ell& L = L5[w];
      if (L.rows == 0) continue;
      hipLaunchKernelGGL(k_split, dim3((L.rows + TB - 1) / TB), dim3(TB), 0, c->stream, L, s_sp[w] * npd,
                         j_sp[w] * npd);
      A2C_HIP(c, hipGetLastError());
    }
    if (int rc2 = reverse_slots(c, L5)) return rc2;
  }
  {
    std::vector<unsigned long long> tb(SPREAD * 5);
    A2C_HIP(c, hipMemcpyAsync(tb.data(), totals, sizeof(unsigned long long) * tb.size(), hipMemcpyDeviceToHost,
                              c->stream));
    A2C_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 5; ++k) {
      ht[k] = 0;
      for (int cp = 0; cp < SPREAD; ++cp) ht[k] += tb[cp * 5 + k];
    }
  }
  return A2C_OK;
}

// Per-entry arrays of the skin lists, filled at the skin build: the
// distance at the build as a float (oj), the partner slot (orv); and room
// for the fate of every entry in a refresh (rbuf, one int32 each).
int size_skin_extras(a2c_fe_ctx* c) {
  for (int k = 0; k < 4; ++k) {
    Ell& S = c->nb_sup[k];
    const int64_t n = (int64_t)std::max(S.width, 1) * std::max(S.rows, 1);
    if (n > S.ocap) {
      (void)dfree(S.orv);
      (void)dfree(S.oj);
      S.orv = S.oj = nullptr;
      S.ocap = n + 4 * (int64_t)std::max(S.rows, 1);
      A2C_HIP(c, dmalloc(&S.orv, sizeof(int32_t) * S.ocap));
      A2C_HIP(c, dmalloc(&S.oj, sizeof(float) * S.ocap));
    }
    if ((int64_t)sizeof(int32_t) * n > S.bcap) {
      (void)dfree(S.rbuf);
      S.rbuf = nullptr;
      S.bcap = (int64_t)sizeof(int32_t) * (n + 4 * (int64_t)std::max(S.rows, 1));
      A2C_HIP(c, dmalloc(&S.rbuf, S.bcap));
    }
  }
  const Grid g = box_grid(c);
  double ib[3];
  for (int ax = 0; ax < 3; ++ax) ib[ax] = g.box[ax] > 0.0 ? 1.0 / g.box[ax] : 0.0;
  static const int s_sp[4] = {0, 0, 1, 1}, j_sp[4] = {0, 1, 0, 1};  // species of rows / of neighbors
  for (int k = 0; k < 4; ++k) {
    const Ell& S = c->nb_sup[k];
    if (S.rows == 0 || S.width == 0) continue;
    hipLaunchKernelGGL(k_sdist, dim3((S.rows + 255) / 256, S.width), dim3(256), 0, c->stream, S, c->dX,
                       s_sp[k] * c->desc.n_pd, j_sp[k] * c->desc.n_pd, g, ib[0], ib[1], ib[2]);
    A2C_HIP(c, hipGetLastError());
  }
  static const int partner[4] = {0, 2, 1, 3};
  for (int k = 0; k < 4; ++k) {
    const Ell& S = c->nb_sup[k];
    if (S.rows == 0) continue;
    if (S.width == 0) continue;
    hipLaunchKernelGGL(k_srev, dim3((S.rows + 255) / 256, S.width), dim3(256), 0, c->stream, S, c->nb_sup[partner[k]]);
    A2C_HIP(c, hipGetLastError());
  }
  return A2C_OK;
}

// The context's lists at rc filtered from its skin lists, and (check) the
// largest displacement since the skin build, in one round trip: the caller
// keeps the result only if that displacement is under skin/2.
int refresh(a2c_fe_ctx* c, bool check, int hm[16], unsigned long long ht[5], double* maxd) {
  const int N = c->N, npd = c->desc.n_pd, nh = c->desc.n_h;
  // every list at rc fits in the capacity of its skin list
  int w[5], ow[4], rw[4];
  for (int k = 0; k < 5; ++k) w[k] = c->nb_sup[k < 4 ? k : 3].width;
  for (int k = 0; k < 4; ++k) ow[k] = rw[k] = c->nb_sup[k].width;
  if (int rc = size_lists(c, c->ell, w, ow, rw)) return rc;
  int* misc = c->nb_misc;
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(misc) + MISC_TOT);
  unsigned long long* dmax = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(misc) + MISC_DMAX);
  A2C_HIP(c, hipMemsetAsync(misc, 0, MISC_BBOX, c->stream));
  if (!c->dpdpdnum) A2C_HIP(c, dmalloc(&c->dpdpdnum, sizeof(int32_t) * 3 * std::max(npd, 1)));
  const Grid g = box_grid(c);
  if (check && N > 0) {
    hipLaunchKernelGGL(k_maxdisp, dim3(std::min((N + 255) / 256, 256)), dim3(256), 0, c->stream, c->dX,
                       c->nb_sup_q, N, g, dmax);
    A2C_HIP(c, hipGetLastError());
  }
  RefreshArgs a{};
  a.X = c->dX;
  a.npd = npd;
  a.N = N;
  a.g = g;
  for (int ax = 0; ax < 3; ++ax) a.ibox[ax] = g.box[ax] > 0.0 ? 1.0 / g.box[ax] : 0.0;
  a.rc2 = c->desc.rc * c->desc.rc;
  a.sh1 = c->desc.a_pd * c->desc.a_pd * 0.75;
  a.sh2 = c->desc.a_pd * c->desc.a_pd * 1.25;
  for (int k = 0; k < 4; ++k) a.sup[k] = c->nb_sup[k];
  for (int k = 0; k < 5; ++k) a.out[k] = c->ell[k];
  a.pdpdnum = c->dpdpdnum;
  a.maxw = misc;
  a.totals = totals;
  a.dmax = dmax;
  const int bp = (npd + NT - 1) / NT, bh = (nh + NT - 1) / NT;  // workgroups per list of Pd / H rows
  const int4 blk0 = make_int4(0, bp, 2 * bp, 2 * bp + bh);     // first workgroup of each list
  int wmax = 1;  // LDS column of a lane: the longest skin row of any list
  for (int k = 0; k < 4; ++k) wmax = std::max(wmax, c->nb_sup[k].width);
  if (bp + bh > 0) {
    hipLaunchKernelGGL(k_refresh, dim3(2 * (bp + bh)), dim3(NT), sizeof(uint32_t) * CS * wmax, c->stream, a, blk0);
    A2C_HIP(c, hipGetLastError());
  }
  if (npd > 0) {
    hipLaunchKernelGGL(k_rowmax, dim3((npd + 255) / 256), dim3(256), 0, c->stream, c->ell[0].cnt, c->ell[1].cnt, npd,
                       misc);
    A2C_HIP(c, hipGetLastError());
  }
  if (nh > 0) {
    hipLaunchKernelGGL(k_rowmax, dim3((nh + 255) / 256), dim3(256), 0, c->stream, c->ell[2].cnt, c->ell[3].cnt, nh,
                       misc);
    A2C_HIP(c, hipGetLastError());
  }
  static const int partner[4] = {0, 2, 1, 3};
  for (int k = 0; k < 4; ++k) {
    const Ell& S = c->nb_sup[k];
    if (S.rows == 0) continue;
    if (S.width == 0) continue;
    hipLaunchKernelGGL(k_orv_skin, dim3((S.rows + 255) / 256, S.width), dim3(256), 0, c->stream, S,
                       c->nb_sup[partner[k]], c->ell[k]);
    A2C_HIP(c, hipGetLastError());
  }
  std::vector<unsigned char> hb(MISC_BBOX);
  A2C_HIP(c, hipMemcpyAsync(hb.data(), misc, MISC_BBOX, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  reduce_counters(hb.data(), MISC_BBOX, hm, ht);
  double d2;
  std::memcpy(&d2, hb.data() + MISC_DMAX, sizeof(d2));
  *maxd = std::sqrt(d2);
  for (int k = 0; k < 5; ++k) {
    c->ell[k].width = hm[k];
    if (k < 4) {
      c->ell[k].owidth = hm[5 + k];
      c->ell[k].rwidth = hm[9 + k];
    }
  }
  return A2C_OK;
}

double skin_length() {
  static const double skin = [] {
    const char* e = std::getenv("A2C_NBR_SKIN");
    const double v = e ? std::atof(e) : 0.2;
    return v > 0.0 ? v : 0.0;
  }();
  return skin;
}

}  // namespace

// Host driver.  The positions are the context state c->dX (on the device).
// First build: the cell walk at rc.  Later builds: lists at rc + skin by
// the cell walk, kept with the positions they were built at, and the lists
// at rc filtered from them -- again at the next builds while no site has
// moved by skin/2 (a pair within rc now was within rc + skin then).
int nbr_build(a2c_fe_ctx* c) {
  const int N = c->N;
  const double rc = c->desc.rc, skin = skin_length();
  int hm[16] = {0};
  unsigned long long ht[5] = {0, 0, 0, 0, 0};
  bool refreshed = false;
  double maxd = 0.0;
  if (!c->nb_misc) A2C_HIP(c, dmalloc(&c->nb_misc, MISC_BYTES));
  if (c->nb_sup_ok) {
    if (int rc2 = refresh(c, true, hm, ht, &maxd)) return rc2;
    refreshed = 2.0 * maxd < skin * (1.0 - 1e-9);  // else the filtered lists may miss a pair
    c->nb_sup_ok = refreshed;
  }
  bool fits = skin > 0.0 && !c->nb_sup_off && c->nb_builds > 0;
  for (int ax = 0; ax < 3; ++ax)
    if (c->desc.periodic[ax] && c->desc.box[ax] < 2.0 * (rc + skin)) fits = false;
  if (!refreshed && fits) {  // new skin lists, then the lists at rc filtered from them
    unsigned long long hs[5];
    const int rc2 = walk_cells(c, rc + skin, c->nb_sup, false, hm, hs);
    if (rc2 == A2C_OK) {
      if (int rc3 = size_skin_extras(c)) return rc3;
      if (!c->nb_sup_q) A2C_HIP(c, dmalloc(&c->nb_sup_q, 32 * (size_t)std::max(N, 1)));
      A2C_HIP(c, hipMemcpyAsync(c->nb_sup_q, c->dX, 32 * (size_t)N, hipMemcpyDeviceToDevice, c->stream));
      if (int rc3 = refresh(c, false, hm, ht, &maxd)) return rc3;
      c->nb_sup_ok = refreshed = true;
      maxd = 0.0;
    } else if (rc2 == A2C_ENBR_OVERFLOW) {
      c->nb_sup_off = true;  // rows over the cap at rc + skin: walk the cells at rc from now on
    } else {
      return rc2;
    }
  }
  if (!refreshed) {
    const int rc2 = walk_cells(c, rc, c->ell, true, hm, ht);
    if (rc2 == A2C_ENBR_OVERFLOW) {
      c->have_nbr = false;
      return set_err(c, A2C_ENBR_OVERFLOW, "a site has more neighbors than max_neighbors");
    }
    if (rc2) return rc2;
    ++c->nb_builds;
  }
  if (!c->depair) A2C_HIP(c, dmalloc(&c->depair, sizeof(double) * std::max(N, 1)));
  c->stats.max_neib = hm[13];
  for (int w = 0; w < 5; ++w) {
    c->stats.width[w] = hm[w];
    c->stats.total[w] = (int64_t)ht[w];
  }
  c->stats.refreshed = refreshed ? 1 : 0;
  c->stats.max_displacement = maxd;
  c->have_nbr = true;
  c->lat_dirty = true;
  return A2C_OK;
}

void free_pair_buffers(Ell& L) {
  (void)dfree(L.ocnt); (void)dfree(L.rcnt); (void)dfree(L.oj); (void)dfree(L.orv);
  (void)dfree(L.rbuf);
  L.ocnt = L.rcnt = L.oj = L.orv = nullptr;
  L.rbuf = nullptr;
  L.ocap = L.bcap = 0;
  L.owidth = L.rwidth = 0;
}

}  // namespace a2c
