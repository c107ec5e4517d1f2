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
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 14; ++k)
        if (v[k]) atomicMax(&a.maxw[k], v[k]);
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
    atomicAdd(&a.totals[lp], tp);
    atomicAdd(&a.totals[lh], th);
    if (!pdb) atomicAdd(&a.totals[4], t1);
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

}  // namespace

// Host driver: grid, bin, scan, place, count, allocate, fill, split.  The
// positions are the context state c->dX (on the device).
int nbr_build(a2c_fe_ctx* c) {
  const int N = c->N, npd = c->desc.n_pd, nh = c->desc.n_h;
  const double rc = c->desc.rc;
  const int TB = 256;
  // scratch (grow-only): misc = 16 ints (widths, overflow) + 5 totals + 6 bbox keys
  if (!c->nb_cid) {
    A2C_HIP(c, dmalloc(&c->nb_cid, sizeof(int) * std::max(N, 1)));
    A2C_HIP(c, dmalloc(&c->nb_order, sizeof(double4) * std::max(N, 1)));  // cell-ordered positions
    A2C_HIP(c, dmalloc(&c->nb_misc, 256));
  }
  int* misc = c->nb_misc;
  unsigned long long* totals = reinterpret_cast<unsigned long long*>(misc + 16);
  unsigned long long* bbox = totals + 5;
  Grid g;
  bool need_bbox = false;
  for (int ax = 0; ax < 3; ++ax) need_bbox |= !c->desc.periodic[ax];
  double lo3[3] = {0, 0, 0}, hi3[3] = {1, 1, 1};
  if (need_bbox && N > 0) {
    unsigned long long init[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    A2C_HIP(c, hipMemcpyAsync(bbox, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_bbox, dim3(std::min((N + TB - 1) / TB, 1024)), dim3(TB), 0, c->stream, c->dX, N, bbox);
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
    g.periodic[ax] = c->desc.periodic[ax];
    g.box[ax] = c->desc.box[ax];
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
  A2C_HIP(c, hipMemsetAsync(misc, 0, 16 * sizeof(int) + 5 * sizeof(unsigned long long), c->stream));
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
  // per-row counts (rows are fixed for a context)
  for (int w = 0; w < 5; ++w) {
    Ell& L = c->ell[w];
    L.rows = (w < 2) ? npd : nh;
    if (!L.cnt) A2C_HIP(c, dmalloc(&L.cnt, sizeof(int32_t) * std::max(L.rows, 1)));
    if (w < 4 && !L.ocnt) {
      A2C_HIP(c, dmalloc(&L.ocnt, sizeof(int32_t) * std::max(L.rows, 1)));
      A2C_HIP(c, dmalloc(&L.rcnt, sizeof(int32_t) * std::max(L.rows, 1)));
    }
  }
  NbrArgs a{};
  a.X = c->dX; a.qs = qs; a.npd = npd; a.N = N; a.g = g; a.cid = dcid; a.cstart = dcc;
  for (int ax = 0; ax < 3; ++ax) a.ibox[ax] = g.box[ax] > 0.0 ? 1.0 / g.box[ax] : 0.0;
  a.rc2 = rc * rc;
  a.sh1 = c->desc.a_pd * c->desc.a_pd * 0.75;
  a.sh2 = c->desc.a_pd * c->desc.a_pd * 1.25;
  a.cap = std::min(c->desc.max_neighbors > 0 ? c->desc.max_neighbors : NBR_CAP, NBR_CAP);
  for (int w = 0; w < 5; ++w) a.cnt[w] = c->ell[w].cnt;
  for (int w = 0; w < 4; ++w) { a.ocnt[w] = c->ell[w].ocnt; a.rcnt[w] = c->ell[w].rcnt; }
  a.maxw = misc;
  a.overflow = misc + 14;
  a.totals = totals;
  const int nblk_pd = (npd + NT - 1) / NT, nblk_h = (nh + NT - 1) / NT;
  if (nblk_pd + nblk_h > 0) {
    hipLaunchKernelGGL(k_site<false>, dim3(nblk_pd + nblk_h), dim3(NT), 0, c->stream, a, nblk_pd);
    A2C_HIP(c, hipGetLastError());
  }
  int hm[16];
  A2C_HIP(c, hipMemcpyAsync(hm, misc, sizeof(hm), hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  if (hm[14]) {
    c->have_nbr = false;
    return set_err(c, A2C_ENBR_OVERFLOW, "a site has more neighbors than max_neighbors");
  }
  // storage: ELL lists, owned halves, received-share buffers (grow-only, with slack)
  for (int w = 0; w < 5; ++w) {
    Ell& L = c->ell[w];
    const int64_t rows = std::max(L.rows, 1);
    L.width = hm[w];
    const int64_t n = (int64_t)std::max(hm[w], 1) * rows;
    if (n > L.cap) {
      (void)dfree(L.idx);
      L.idx = nullptr;
      L.cap = (int64_t)(std::max(hm[w], 1) + 4) * rows;
      A2C_HIP(c, dmalloc(&L.idx, sizeof(int32_t) * L.cap));
    }
    a.idx[w] = L.idx;
    if (w == 4) continue;
    L.owidth = hm[5 + w];
    L.rwidth = hm[9 + w];
    const int64_t on = (int64_t)std::max(L.owidth, 1) * rows, rn = (int64_t)std::max(L.rwidth, 1) * rows;
    const int64_t bn = std::max<int64_t>(32 * rn, 4 * (int64_t)std::max(L.width, 1) * rows);
    if (on > L.ocap) {
      (void)dfree(L.oj); (void)dfree(L.orv);
      L.oj = L.orv = nullptr;
      L.ocap = on + 4 * rows;  // slack: small changes of the widest half reuse the arrays
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
  if (!c->dpdpdnum) A2C_HIP(c, dmalloc(&c->dpdpdnum, sizeof(int32_t) * 3 * std::max(npd, 1)));
  a.pdpdnum = c->dpdpdnum;
  a.width = std::max(hm[13], 1);  // LDS slots per lane: the longest row
  if (nblk_pd + nblk_h > 0) {
    hipLaunchKernelGGL(k_site<true>, dim3(nblk_pd + nblk_h), dim3(NT), sizeof(uint32_t) * CS * a.width, c->stream, a,
                       nblk_pd);
    A2C_HIP(c, hipGetLastError());
  }
  // owned/received split of the four pair lists, then the partner slots
  static const int partner[4] = {0, 2, 1, 3};  // PdPd, PdH<->HPd, HH
  static const int s_sp[4] = {0, 0, 1, 1}, j_sp[4] = {0, 1, 0, 1};  // species of rows / of neighbors
  for (int w = 0; w < 4; ++w) {
    const Ell& L = c->ell[w];
    if (L.rows == 0) continue;
    hipLaunchKernelGGL(k_split, dim3((L.rows + TB - 1) / TB), dim3(TB), 0, c->stream, L, s_sp[w] * npd,
                       j_sp[w] * npd);
    A2C_HIP(c, hipGetLastError());
  }
  for (int w = 0; w < 4; ++w) {
    const Ell& L = c->ell[w];
    if (L.rows == 0) continue;
    hipLaunchKernelGGL(k_orv, dim3((L.rows + TB - 1) / TB), dim3(TB), 0, c->stream, L, c->ell[partner[w]]);
    A2C_HIP(c, hipGetLastError());
  }
  if (!c->depair) A2C_HIP(c, dmalloc(&c->depair, sizeof(double) * std::max(N, 1)));
  unsigned long long ht[5];
  A2C_HIP(c, hipMemcpyAsync(ht, totals, sizeof(ht), hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  c->stats.max_neib = hm[13];
  for (int w = 0; w < 5; ++w) {
    c->stats.width[w] = hm[w];
    c->stats.total[w] = (int64_t)ht[w];
  }
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
