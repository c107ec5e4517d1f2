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
// Kernels: k_bin (cell id + histogram), host scan, k_fill (scatter into cell
// order), k_nbr<false> (row lengths), k_width (ELL widths), k_nbr<true>
// (sorted ELL rows).  One 64-lane wave per site: lanes test consecutive members of a cell,
// a ballot compacts hits into LDS, and each hit's slot is its rank among the
// hits (an O(n) count per lane over LDS broadcasts), so rows come out sorted
// whatever order the atomics binned the sites in.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
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

__global__ void k_fill(int N, const int* __restrict__ cid, int* __restrict__ cursor, int* __restrict__ order) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= N) return;
  int pos = atomicAdd(&cursor[cid[s]], 1);
  order[pos] = s;
}

struct NbrArgs {
  const double* X;
  int npd, N;
  Grid g;
  const int* cid;
  const int* cstart;
  const int* order;
  double rc2, sh1, sh2;
  int cap;
  int32_t* cnt[5];    // row lengths: PdPd, PdH (Pd rows); HPd, HH, HH1 (H rows)
  int32_t* idx[5];    // ELL arrays (fill pass)
  int32_t width[5];   // ELL widths (fill pass)
  int32_t* pdpdnum;   // nPd*3 (fill pass)
  int32_t* maxw;      // 6 ints: max row length per list, max total (count pass)
  int32_t* overflow;  // 1 if a site exceeds cap
};

template <bool FILL>
__global__ __launch_bounds__(64) void k_nbr(NbrArgs a) {
  __shared__ int32_t sj[NBR_CAP];
  __shared__ double sd[NBR_CAP];
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  if (s >= a.N) return;  // uniform over the block (one site per block)
  const double qi0 = a.X[4 * s], qi1 = a.X[4 * s + 1], qi2 = a.X[4 * s + 2];
  const int c = a.cid[s];
  const int ci[3] = {c % a.g.nc[0], (c / a.g.nc[0]) % a.g.nc[1], c / (a.g.nc[0] * a.g.nc[1])};
  int lst[3][3], cnt[3];
  for (int ax = 0; ax < 3; ++ax) {  // distinct stencil cells per axis
    cnt[ax] = 0;
    for (int o = -1; o <= 1; ++o) {
      int k = ci[ax] + o;
      if (a.g.periodic[ax]) k = (k + a.g.nc[ax]) % a.g.nc[ax];
      else if (k < 0 || k >= a.g.nc[ax]) continue;
      bool dup = false;
      for (int t = 0; t < cnt[ax]; ++t) dup |= (lst[ax][t] == k);
      if (!dup) lst[ax][cnt[ax]++] = k;
    }
  }
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int n = 0, npn = 0, nh1 = 0;
  for (int zz = 0; zz < cnt[2]; ++zz)
    for (int yy = 0; yy < cnt[1]; ++yy)
      for (int xx = 0; xx < cnt[0]; ++xx) {
        const int cc = (lst[2][zz] * a.g.nc[1] + lst[1][yy]) * a.g.nc[0] + lst[0][xx];
        const int beg = a.cstart[cc], end = a.cstart[cc + 1];
        for (int base = beg; base < end; base += 64) {
          const int t = base + lane;
          bool ok = false;
          int j = -1;
          double d2 = 0.0;
          if (t < end) {
            j = a.order[t];
            double d0 = a.X[4 * j] - qi0, d1 = a.X[4 * j + 1] - qi1, dz = a.X[4 * j + 2] - qi2;
            if (a.g.periodic[0]) d0 = d0 - a.g.box[0] * rint(d0 / a.g.box[0]);
            if (a.g.periodic[1]) d1 = d1 - a.g.box[1] * rint(d1 / a.g.box[1]);
            if (a.g.periodic[2]) dz = dz - a.g.box[2] * rint(dz / a.g.box[2]);
            d2 = d0 * d0 + d1 * d1 + dz * dz;
            ok = (j != s) && (d2 <= a.rc2);
          }
          const unsigned long long m = __ballot(ok);
          if (FILL && ok) {
            const int pos = n + __popcll(m & lt);
            if (pos < NBR_CAP) { sj[pos] = j; sd[pos] = d2; }
          }
          n += __popcll(m);
          npn += __popcll(__ballot(ok && j < a.npd));
          nh1 += __popcll(__ballot(ok && j >= a.npd && d2 < a.sh1));
        }
      }
  const bool pd = s < a.npd;
  const int row = pd ? s : s - a.npd;
  if (n > a.cap) {
    if (lane == 0) atomicOr(a.overflow, 1);
    return;
  }
  const int nhn = n - npn;
  if (!FILL) {
    if (lane == 0) {  // list widths: k_width (one atomic per block, not per site)
      if (pd) {
        a.cnt[0][row] = npn; a.cnt[1][row] = nhn;
      } else {
        a.cnt[2][row] = npn; a.cnt[3][row] = nhn; a.cnt[4][row] = nh1;
      }
    }
    return;
  }
  __syncthreads();  // LDS hits visible to every lane of the (single-wave) block
  const int rowsP = pd ? a.npd : a.N - a.npd;  // rows of this site's lists
  int32_t* outP = pd ? a.idx[0] : a.idx[2];
  int32_t* outH = pd ? a.idx[1] : a.idx[3];
  int sh0 = 0, sh1n = 0, sh2n = 0;
  for (int e0 = 0; e0 < n; e0 += 64) {
    const int e = e0 + lane;
    bool isp = false, s0 = false, s1 = false;
    if (e < n) {
      const int key = sj[e];
      const double d2 = sd[e];
      isp = key < a.npd;
      int rank = 0, rank1 = 0;
      for (int m = 0; m < n; ++m) {
        const int o = sj[m];
        const bool less = o < key;
        rank += less;
        rank1 += less && (o >= a.npd) && (sd[m] < a.sh1);
      }
      if (isp) {
        outP[(size_t)rank * rowsP + row] = key;  // Pd neighbors occupy ranks 0..npn-1
        s0 = d2 < a.sh1;
        s1 = d2 > a.sh1 && d2 < a.sh2;
      } else {
        const int h = key - a.npd;
        outH[(size_t)(rank - npn) * rowsP + row] = h;
        if (!pd && d2 < a.sh1) a.idx[4][(size_t)rank1 * rowsP + row] = h;
      }
    }
    sh0 += __popcll(__ballot(isp && s0));
    sh1n += __popcll(__ballot(isp && s1));
    sh2n += __popcll(__ballot(isp && !s0 && !s1));
  }
  if (pd && lane == 0) {
    a.pdpdnum[3 * row] = sh0;
    a.pdpdnum[3 * row + 1] = sh1n;
    a.pdpdnum[3 * row + 2] = sh2n;
  }
}

// Reverse slots: for entry (row r, slot t) of list L with neighbor j, the slot
// of r in row j of the partner list P (PdPd<->PdPd, PdH<->HPd, HH<->HH).
// Rows are sorted, so a binary search finds it; lists are symmetric.
// Widest row of each list and the largest total count over all sites:
// block-local maxima, then one atomicMax per block and quantity.
__global__ __launch_bounds__(256) void k_width(NbrArgs a) {
  __shared__ int32_t red[6][256];
  int m[6] = {0, 0, 0, 0, 0, 0};
  for (int s = blockIdx.x * 256 + threadIdx.x; s < a.N; s += gridDim.x * 256) {
    if (s < a.npd) {
      const int p = a.cnt[0][s], h = a.cnt[1][s];
      m[0] = max(m[0], p); m[1] = max(m[1], h); m[5] = max(m[5], p + h);
    } else {
      const int r = s - a.npd;
      const int p = a.cnt[2][r], h = a.cnt[3][r], h1 = a.cnt[4][r];
      m[2] = max(m[2], p); m[3] = max(m[3], h); m[4] = max(m[4], h1); m[5] = max(m[5], p + h);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = m[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = max(red[k][threadIdx.x], red[k][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x < 6) atomicMax(&a.maxw[threadIdx.x], red[threadIdx.x][0]);
}

// Split of each pair-list row into owned and received halves (pair-once
// evaluation).  Ownership of {s, j} (global site indices):
// (s < j) XOR ((s ^ j) & 1), which hands each site about half of every list.
// Both halves keep the row's (sorted) order.
__device__ __forceinline__ bool pair_owner(int s, int j) { return (s < j) != (((s ^ j) & 1) != 0); }

__global__ void k_own_count(Ell L, int s_off, int j_off) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= L.rows) return;
  const int n = L.cnt[r];
  const int s = s_off + r;
  int no = 0;
  for (int t = 0; t < n; ++t) no += pair_owner(s, j_off + L.idx[(size_t)t * L.rows + r]);
  L.ocnt[r] = no;
  L.rcnt[r] = n - no;
}

// max over rows of a[r], one atomicMax per block
__global__ __launch_bounds__(256) void k_rowmax(const int32_t* __restrict__ a, int n, int32_t* out) {
  __shared__ int32_t red[256];
  int m = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) m = max(m, a[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, red[0]);
}

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

}  // namespace

// Host driver: grid from the positions, bin, scan, count, allocate, fill.
int nbr_build(a2c_fe_ctx* c) {
  const int N = c->N, npd = c->desc.n_pd, nh = c->desc.n_h;
  const double rc = c->desc.rc;
  Grid g;
  for (int ax = 0; ax < 3; ++ax) {
    g.periodic[ax] = c->desc.periodic[ax];
    g.box[ax] = c->desc.box[ax];
    double lo, hi;
    if (g.periodic[ax]) {
      lo = 0.0; hi = g.box[ax];
    } else {
      lo = 1e300; hi = -1e300;
      for (int s = 0; s < N; ++s) { lo = std::min(lo, c->hX[4 * s + ax]); hi = std::max(hi, c->hX[4 * s + ax]); }
      if (N == 0) { lo = 0.0; hi = 1.0; }
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
  // scratch buffers persist across rebuilds (grow-only)
  if (!c->nb_cid) {
    A2C_HIP(c, dmalloc(&c->nb_cid, sizeof(int) * std::max(N, 1)));
    A2C_HIP(c, dmalloc(&c->nb_order, sizeof(int) * std::max(N, 1)));
    A2C_HIP(c, dmalloc(&c->nb_misc, sizeof(int) * 8));
  }
  if (ncell + 1 > c->nb_cap_cells) {
    (void)dfree(c->nb_cc);
    (void)dfree(c->nb_cursor);
    c->nb_cc = c->nb_cursor = nullptr;
    A2C_HIP(c, dmalloc(&c->nb_cc, sizeof(int) * (ncell + 1)));
    A2C_HIP(c, dmalloc(&c->nb_cursor, sizeof(int) * (ncell + 1)));
    c->nb_cap_cells = ncell + 1;
  }
  int *dcid = c->nb_cid, *dcc = c->nb_cc, *dorder = c->nb_order, *dmisc = c->nb_misc, *dcursor = c->nb_cursor;
  A2C_HIP(c, hipMemsetAsync(dcc, 0, sizeof(int) * (ncell + 1), c->stream));
  A2C_HIP(c, hipMemsetAsync(dmisc, 0, sizeof(int) * 8, c->stream));
  const int TB = 256;
  if (N > 0) {
    hipLaunchKernelGGL(k_bin, dim3((N + TB - 1) / TB), dim3(TB), 0, c->stream, c->dX, N, g, dcid, dcc);
    A2C_HIP(c, hipGetLastError());
  }
  std::vector<int> hc(ncell + 1);
  A2C_HIP(c, hipMemcpyAsync(hc.data(), dcc, sizeof(int) * ncell, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<int> start(ncell + 1, 0);
  for (long long k = 0; k < ncell; ++k) start[k + 1] = start[k] + hc[k];
  A2C_HIP(c, hipMemcpyAsync(dcc, start.data(), sizeof(int) * (ncell + 1), hipMemcpyHostToDevice, c->stream));
  A2C_HIP(c, hipMemcpyAsync(dcursor, start.data(), sizeof(int) * (ncell + 1), hipMemcpyHostToDevice, c->stream));
  if (N > 0) {
    hipLaunchKernelGGL(k_fill, dim3((N + TB - 1) / TB), dim3(TB), 0, c->stream, N, dcid, dcursor, dorder);
    A2C_HIP(c, hipGetLastError());
  }
  // row-length arrays (rows are fixed for a context)
  for (int w = 0; w < 5; ++w) {
    Ell& L = c->ell[w];
    L.rows = (w < 2) ? npd : nh;
    if (!L.cnt) A2C_HIP(c, dmalloc(&L.cnt, sizeof(int32_t) * std::max(L.rows, 1)));
    A2C_HIP(c, hipMemsetAsync(L.cnt, 0, sizeof(int32_t) * std::max(L.rows, 1), c->stream));
  }
  NbrArgs a{};
  a.X = c->dX; a.npd = npd; a.N = N; a.g = g; a.cid = dcid; a.cstart = dcc; a.order = dorder;
  a.rc2 = rc * rc;
  a.sh1 = c->desc.a_pd * c->desc.a_pd * 0.75;
  a.sh2 = c->desc.a_pd * c->desc.a_pd * 1.25;
  a.cap = std::min(c->desc.max_neighbors > 0 ? c->desc.max_neighbors : NBR_CAP, NBR_CAP);
  for (int w = 0; w < 5; ++w) a.cnt[w] = c->ell[w].cnt;
  a.maxw = dmisc;
  a.overflow = dmisc + 6;
  if (N > 0) {
    hipLaunchKernelGGL(k_nbr<false>, dim3(N), dim3(64), 0, c->stream, a);
    A2C_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_width, dim3(std::min((N + 255) / 256, 1024)), dim3(256), 0, c->stream, a);
    A2C_HIP(c, hipGetLastError());
  }
  int hm[8];
  A2C_HIP(c, hipMemcpyAsync(hm, dmisc, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  if (hm[6]) {
    c->have_nbr = false;
    return set_err(c, A2C_ENBR_OVERFLOW, "a site has more neighbors than max_neighbors");
  }
  for (int w = 0; w < 5; ++w) {
    Ell& L = c->ell[w];
    L.width = hm[w];
    const int64_t n = (int64_t)std::max(hm[w], 1) * std::max(L.rows, 1);
    if (n > L.cap) {  // grow with slack so small changes of the widest row reuse the arrays
      (void)dfree(L.idx);
      L.idx = nullptr;
      L.cap = (int64_t)(std::max(hm[w], 1) + 4) * std::max(L.rows, 1);
      A2C_HIP(c, dmalloc(&L.idx, sizeof(int32_t) * L.cap));
    }
    A2C_HIP(c, hipMemsetAsync(L.idx, 0xff, sizeof(int32_t) * n, c->stream));
    a.idx[w] = L.idx;
    a.width[w] = hm[w];
  }
  if (!c->dpdpdnum) A2C_HIP(c, dmalloc(&c->dpdpdnum, sizeof(int32_t) * 3 * std::max(npd, 1)));
  a.pdpdnum = c->dpdpdnum;
  if (N > 0) {
    hipLaunchKernelGGL(k_nbr<true>, dim3(N), dim3(64), 0, c->stream, a);
    A2C_HIP(c, hipGetLastError());
  }
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  c->stats.max_neib = hm[5];
  for (int w = 0; w < 5; ++w) c->stats.width[w] = hm[w];
  c->have_nbr = true;
  return ensure_pair_buffers(c);
}

// Owned/received split and the received-share buffers of the four pair
// lists; recomputed after every build.  Arrays grow only.
int ensure_pair_buffers(a2c_fe_ctx* c) {
  if (!c->have_nbr) return A2C_OK;
  static const int partner[4] = {0, 2, 1, 3};  // PdPd, PdH<->HPd, HH
  static const int s_sp[4] = {0, 0, 1, 1}, j_sp[4] = {0, 1, 0, 1};  // species of rows / of neighbors
  const int npd = c->desc.n_pd;
  const int TB = 256;
  // 1. per-row owned/received counts and their maxima
  for (int w = 0; w < 4; ++w) {
    Ell& L = c->ell[w];
    if (!L.ocnt) {
      A2C_HIP(c, dmalloc(&L.ocnt, sizeof(int32_t) * std::max(L.rows, 1)));
      A2C_HIP(c, dmalloc(&L.rcnt, sizeof(int32_t) * std::max(L.rows, 1)));
    }
  }
  A2C_HIP(c, hipMemsetAsync(c->nb_misc, 0, sizeof(int) * 8, c->stream));
  for (int w = 0; w < 4; ++w) {
    const Ell& L = c->ell[w];
    if (L.rows == 0) continue;
    const int g = (L.rows + TB - 1) / TB;
    hipLaunchKernelGGL(k_own_count, dim3(g), dim3(TB), 0, c->stream, L, s_sp[w] * npd, j_sp[w] * npd);
    A2C_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_rowmax, dim3(std::min(g, 1024)), dim3(TB), 0, c->stream, L.ocnt, L.rows, c->nb_misc + 2 * w);
    A2C_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_rowmax, dim3(std::min(g, 1024)), dim3(TB), 0, c->stream, L.rcnt, L.rows,
                       c->nb_misc + 2 * w + 1);
    A2C_HIP(c, hipGetLastError());
  }
  int wm[8];
  A2C_HIP(c, hipMemcpyAsync(wm, c->nb_misc, sizeof(int) * 8, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  // 2. storage
  for (int w = 0; w < 4; ++w) {
    Ell& L = c->ell[w];
    const int64_t rows = std::max(L.rows, 1);
    L.owidth = wm[2 * w];
    L.rwidth = wm[2 * w + 1];
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
  // 3. split, then the partner slots (needs every list's split)

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
  if (!c->depair) A2C_HIP(c, dmalloc(&c->depair, sizeof(double) * std::max(c->N, 1)));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
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
