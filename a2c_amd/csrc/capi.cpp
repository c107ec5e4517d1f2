// C-ABI entry points of liba2c_fe.so (include/a2c_fe.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "a2c_fe.h"
#include "fe_internal.hpp"

namespace a2c {

int set_err(a2c_fe_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Pinned staging of the host-array entry points (X in, G out), allocated by
// the first call that needs it: a context driven through device arrays, or
// as a batch's lattice, pins no host memory.
int host_staging(a2c_fe_ctx* c) {
  if (c->hX && c->hG) return A2C_OK;
  const size_t N = std::max(c->N, 1);
  if ((!c->hX && hmalloc(&c->hX, 32 * N) != hipSuccess) || (!c->hG && hmalloc(&c->hG, 32 * N) != hipSuccess))
    return set_err(c, A2C_ENOMEM, "pinned host staging");
  return A2C_OK;
}

namespace {

int check_ctx(a2c_fe_ctx* c) {
  if (!c) return A2C_EINVAL;
  c->err.clear();
  return A2C_OK;
}

int upload_state(a2c_fe_ctx* c, const double* X) {
  const size_t n = 4 * (size_t)c->N;
  if (n == 0) return A2C_OK;
  if (int rc = host_staging(c)) return rc;
  std::memcpy(c->hX, X, n * sizeof(double));
  A2C_HIP(c, hipMemcpyAsync(c->dX, c->hX, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return A2C_OK;
}

bool aligned32(const void* p) { return ((uintptr_t)p & 31u) == 0; }

}  // namespace

// evaluation dispatcher shared by eval / eval_device / timing
int eval_dispatch(a2c_fe_ctx* c, const double* dX, double* dG, double* dF, bool timed, Timing* tm) {
  if (!c->have_nbr) return set_err(c, A2C_ESTATE, "build_neighbors must be called before eval");
  if (!c->have_x) return set_err(c, A2C_ESTATE, "set_occupancy must be called before eval");
  if (c->gen_on) return launch_eval_gen(c, dX, dG, dF, timed, tm);
  if (c->desc.precision == A2C_PREC_MIXED) return launch_eval32(c, dX, dG, dF, timed, tm);
  return launch_eval64(c, dX, dG, dF, timed, tm);
}

}  // namespace a2c

using namespace a2c;

extern "C" {

void a2c_fe_desc_default(a2c_fe_desc* d) {
  std::memset(d, 0, sizeof(*d));
  d->rc = 5.35;          // IoData.cpp:518
  d->a_pd = 3.89;
  d->temperature = 300.0;  // IoData.cpp:628
  d->quad_rule = A2C_QUAD_R12;
  d->deriv_mode = A2C_DERIV_ANALYTIC;
  d->fd_eps = 1e-5;      // IoData.cpp:517
  d->precision = A2C_PREC_FP64;
  d->device = 0;
  d->max_neighbors = NBR_CAP;
}

int a2c_fe_create(const a2c_fe_desc* desc, a2c_fe_ctx** out) { return a2c::fe_create(desc, nullptr, out); }

}  // extern "C"

int a2c::fe_create(const a2c_fe_desc* desc, hipStream_t shared, a2c_fe_ctx** out) {
  if (!desc || !out) return A2C_EINVAL;
  *out = nullptr;
  if (desc->n_pd < 0 || desc->n_h < 0 || (int64_t)desc->n_pd + desc->n_h > (1 << 30)) return A2C_EINVAL;
  if (!(desc->rc > 0) || !(desc->temperature > 0)) return A2C_EINVAL;
  if (desc->quad_rule != A2C_QUAD_R12 && desc->quad_rule != A2C_QUAD_GH3) return A2C_EINVAL;
  if (desc->deriv_mode != A2C_DERIV_ANALYTIC && desc->deriv_mode != A2C_DERIV_FD) return A2C_EINVAL;
  if (desc->deriv_mode == A2C_DERIV_FD && !(desc->fd_eps > 0)) return A2C_EINVAL;
  if (desc->precision != A2C_PREC_FP64 && desc->precision != A2C_PREC_MIXED) return A2C_EINVAL;
  if (desc->max_neighbors < 0 || desc->max_neighbors > NBR_CAP) return A2C_EINVAL;
  for (int c = 0; c < 3; ++c)
    if (desc->periodic[c] && !(desc->box[c] >= 2.0 * desc->rc)) return A2C_EINVAL;
  a2c_fe_ctx* c = new (std::nothrow) a2c_fe_ctx;
  if (!c) return A2C_ENOMEM;
  c->desc = *desc;
  if (c->desc.max_neighbors == 0) c->desc.max_neighbors = NBR_CAP;
  c->N = desc->n_pd + desc->n_h;
  auto fail = [&](int code) { a2c_fe_destroy(c); return code; };
  if (hipSetDevice(desc->device) != hipSuccess) return fail(A2C_EHIP);
  if (shared) {
    c->stream = shared;
    c->own_stream = false;
  } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    return fail(A2C_EHIP);
  }
  const size_t N = std::max(c->N, 1), nh = std::max(desc->n_h, 1);
  c->nblk_red = (int)((N + 255) / 256);
  bool ok = dmalloc(&c->dX, 32 * N) == hipSuccess && dmalloc(&c->dG, 32 * N) == hipSuccess &&
            dmalloc(&c->dx, 8 * nh) == hipSuccess && dmalloc(&c->drho, 8 * N) == hipSuccess &&
            dmalloc(&c->dFp, 8 * N) == hipSuccess && dmalloc(&c->demb, 8 * N) == hipSuccess &&
            dmalloc(&c->dpart, 8 * (size_t)c->nblk_red) == hipSuccess && dmalloc(&c->dF, 64) == hipSuccess &&
            dmalloc(&c->dlat, sizeof(LatDev)) == hipSuccess &&
            dmalloc(&c->db2l, 4 * (size_t)c->nblk_red * MAX_LANES_PER_SITE) == hipSuccess;
  if (!ok) return fail(A2C_ENOMEM);
  {
    DeviceTables t;
    if (device_tables(desc->device, &t) != hipSuccess) return fail(A2C_EHIP);
    c->dtab64 = t.t64;  // shared, not freed with the context
  }
  // all-zero block -> lattice map, long enough for the pair-owner launches too
  if (hipMemset(c->db2l, 0, 4 * (size_t)c->nblk_red * MAX_LANES_PER_SITE) != hipSuccess) return fail(A2C_EHIP);
  c->lat_dirty = true;
  if (desc->precision == A2C_PREC_MIXED && build_tables(c) != A2C_OK) return fail(A2C_EHIP);
  *out = c;
  return A2C_OK;
}

extern "C" {

void a2c_fe_destroy(a2c_fe_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  dfree(c->dX); dfree(c->dG); dfree(c->dx); dfree(c->drho); dfree(c->dFp);
  dfree(c->demb); dfree(c->dpart); dfree(c->dF); dfree(c->dpdpdnum); dfree(c->daux);  // dtab, dtab64: shared
  gen_free(c);
  for (int w = 0; w < 5; ++w) {
    dfree(c->ell[w].idx); dfree(c->ell[w].cnt);
    a2c::free_pair_buffers(c->ell[w]);
    dfree(c->nb_sup[w].idx); dfree(c->nb_sup[w].cnt);
    a2c::free_pair_buffers(c->nb_sup[w]);
  }
  dfree(c->nb_sup_q);
  dfree(c->depair);
  dfree(c->occ_g); dfree(c->occ_full); dfree(c->occ_corr); dfree(c->occ_uv);
  dfree(c->occ_xn);
  dfree(c->occ_part); dfree(c->occ_bal);
  dfree(c->nb_cid); dfree(c->nb_order); dfree(c->nb_cc); dfree(c->nb_cursor); dfree(c->nb_misc);
  if (c->hX) hfree(c->hX);
  if (c->hG) hfree(c->hG);
  if (c->hlat) hfree(c->hlat);
  dfree(c->dlat);
  dfree(c->db2l);
  if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* a2c_fe_last_error(const a2c_fe_ctx* c) { return c ? c->err.c_str() : "null context"; }

int a2c_fe_set_precision(a2c_fe_ctx* c, int32_t precision) {
  if (int rc = check_ctx(c)) return rc;
  if (precision != A2C_PREC_FP64 && precision != A2C_PREC_MIXED) return set_err(c, A2C_EINVAL, "bad precision");
  if (precision == A2C_PREC_MIXED && !c->dtab) {
    if (int rc = build_tables(c)) return rc;
    c->lat_dirty = true;
  }
  c->desc.precision = precision;
  return A2C_OK;
}

int a2c_fe_set_occupancy(a2c_fe_ctx* c, const double* x) {
  if (int rc = check_ctx(c)) return rc;
  if (!x && c->desc.n_h > 0) return set_err(c, A2C_EINVAL, "x is null");
  if (c->desc.n_h > 0)
    A2C_HIP(c, hipMemcpy(c->dx, x, 8 * (size_t)c->desc.n_h, hipMemcpyHostToDevice));
  c->have_x = true;
  return A2C_OK;
}

int a2c_fe_build_neighbors(a2c_fe_ctx* c, const double* X, a2c_nbr_stats* stats) {
  if (int rc = check_ctx(c)) return rc;
  if (!X && c->N > 0) return set_err(c, A2C_EINVAL, "X is null");
  // once per build: straight from the caller's array (ordered on the stream;
  // the build synchronises before it returns), no pinned staging
  if (c->N > 0) A2C_HIP(c, hipMemcpyAsync(c->dX, X, 32 * (size_t)c->N, hipMemcpyHostToDevice, c->stream));
  c->lat_dirty = true;
  if (int rc = nbr_build(c)) return rc;
  if (stats) *stats = c->stats;
  return A2C_OK;
}

int a2c_fe_build_neighbors_device(a2c_fe_ctx* c, const double* dX, a2c_nbr_stats* stats) {
  if (int rc = check_ctx(c)) return rc;
  if (!dX && c->N > 0) return set_err(c, A2C_EINVAL, "dX is null");
  if (c->N > 0 && dX != c->dX)
    A2C_HIP(c, hipMemcpyAsync(c->dX, dX, 32 * (size_t)c->N, hipMemcpyDeviceToDevice, c->stream));
  c->lat_dirty = true;
  if (int rc = nbr_build(c)) return rc;
  if (stats) *stats = c->stats;
  return A2C_OK;
}

int a2c_fe_get_neighbors(a2c_fe_ctx* c, int32_t which, int32_t* offsets, int32_t* indices) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->have_nbr) return set_err(c, A2C_ESTATE, "no neighbor lists");
  if (which < 0 || which > 4 || !offsets) return set_err(c, A2C_EINVAL, "bad list id");
  const Ell& L = c->ell[which];
  std::vector<int32_t> cnt(L.rows), idx((size_t)L.width * L.rows);
  if (L.rows) A2C_HIP(c, hipMemcpy(cnt.data(), L.cnt, 4 * cnt.size(), hipMemcpyDeviceToHost));
  if (!idx.empty()) A2C_HIP(c, hipMemcpy(idx.data(), L.idx, 4 * idx.size(), hipMemcpyDeviceToHost));
  offsets[0] = 0;
  for (int r = 0; r < L.rows; ++r) {
    offsets[r + 1] = offsets[r] + cnt[r];
    if (indices)
      for (int t = 0; t < cnt[r]; ++t) indices[offsets[r] + t] = idx[(size_t)t * L.rows + r];
  }
  return A2C_OK;
}

int a2c_fe_get_pdpdnum(a2c_fe_ctx* c, int32_t* pdpdnum) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->have_nbr) return set_err(c, A2C_ESTATE, "no neighbor lists");
  if (c->desc.n_pd)
    A2C_HIP(c, hipMemcpy(pdpdnum, c->dpdpdnum, 12 * (size_t)c->desc.n_pd, hipMemcpyDeviceToHost));
  return A2C_OK;
}

int a2c_fe_density(a2c_fe_ctx* c, const double* X, double* rho) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->have_nbr || !c->have_x) return set_err(c, A2C_ESTATE, "neighbors/occupancy not set");
  if (int rc = upload_state(c, X)) return rc;
  if (int rc = launch_density64(c, c->dX)) return rc;
  if (c->N) A2C_HIP(c, hipMemcpyAsync(rho, c->drho, 8 * (size_t)c->N, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}

int a2c_fe_eval(a2c_fe_ctx* c, const double* X, double* f, double* G) {
  if (int rc = check_ctx(c)) return rc;
  if ((!X || !G) && c->N > 0) return set_err(c, A2C_EINVAL, "null X or G");
  if (!f) return set_err(c, A2C_EINVAL, "null f");
  if (int rc = upload_state(c, X)) return rc;
  if (int rc = eval_dispatch(c, c->dX, c->dG, c->dF, false, nullptr)) return rc;
  const size_t n = 4 * (size_t)c->N;  // upload_state allocated hG with hX
  if (n) A2C_HIP(c, hipMemcpyAsync(c->hG, c->dG, 8 * n, hipMemcpyDeviceToHost, c->stream));
  double fh = 0.0;
  A2C_HIP(c, hipMemcpyAsync(&fh, c->dF, 8, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  if (n) std::memcpy(G, c->hG, 8 * n);
  *f = fh;
  if (!std::isfinite(fh)) return set_err(c, A2C_ENONFINITE, "non-finite objective");
  return A2C_OK;
}

int a2c_fe_eval_device(a2c_fe_ctx* c, const double* dX, double* dG, double* df) {
  if (int rc = check_ctx(c)) return rc;
  if (!dX || !dG || !df || !aligned32(dX)) return set_err(c, A2C_EINVAL, "device pointers must be non-null, X 32-byte aligned");
  return eval_dispatch(c, dX, dG, df, false, nullptr);
}

int a2c_tao_objgrad(void* /*solver*/, const double* X, double* f, double* G, void* ctx) {
  return a2c_fe_eval(static_cast<a2c_fe_ctx*>(ctx), X, f, G);
}

int a2c_dev_alloc(a2c_fe_ctx* c, size_t bytes, void** ptr) {
  if (int rc = check_ctx(c)) return rc;
  A2C_HIP(c, hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return A2C_OK;
}
int a2c_dev_free(a2c_fe_ctx* c, void* ptr) {
  if (int rc = check_ctx(c)) return rc;
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  A2C_HIP(c, hipFree(ptr));
  return A2C_OK;
}
int a2c_dev_upload(a2c_fe_ctx* c, void* dst, const void* src, size_t bytes) {
  if (int rc = check_ctx(c)) return rc;
  A2C_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}
int a2c_dev_download(a2c_fe_ctx* c, void* dst, const void* src, size_t bytes) {
  if (int rc = check_ctx(c)) return rc;
  A2C_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}
int a2c_fe_sync(a2c_fe_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  A2C_HIP(c, hipStreamSynchronize(c->stream));
  return A2C_OK;
}

int a2c_fe_time_eval(a2c_fe_ctx* c, const double* dX, double* dG, double* df, int32_t iters, a2c_timing* out) {
  if (int rc = check_ctx(c)) return rc;
  if (!out || iters <= 0 || iters > 100000 || !dX || !dG || !df || !aligned32(dX))
    return set_err(c, A2C_EINVAL, "bad timing args");
  std::memset(out, 0, sizeof(*out));
  const int per = 9;  // marks per evaluation (<= 8 kernels + end)
  std::vector<hipEvent_t> ev((size_t)per * iters + 2);
  for (auto& e : ev) A2C_HIP(c, hipEventCreate(&e));
  struct Destroy {
    std::vector<hipEvent_t>& v;
    ~Destroy() { for (auto e : v) hipEventDestroy(e); }
  } ds{ev};
  std::vector<int> nmarks(iters, 0);
  Timing tm;
  const char* names[8] = {};
  A2C_HIP(c, hipEventRecord(ev[0], c->stream));
  for (int it = 0; it < iters; ++it) {
    tm.ev = ev.data() + 1 + (size_t)per * it;
    tm.cap = per;
    tm.n = 0;
    tm.on = true;
    if (int rc = eval_dispatch(c, dX, dG, df, true, &tm)) return rc;
    nmarks[it] = tm.n;
    if (it == 0)
      for (int k = 0; k < 8; ++k) names[k] = tm.name[k];
  }
  hipEvent_t tend = ev[1 + (size_t)per * iters];
  A2C_HIP(c, hipEventRecord(tend, c->stream));
  A2C_HIP(c, hipEventSynchronize(tend));
  float tot = 0.f;
  A2C_HIP(c, hipEventElapsedTime(&tot, ev[0], tend));
  const int nk = std::min(nmarks[0] - 1, 8);
  double acc[8] = {0};
  for (int it = 0; it < iters; ++it) {
    hipEvent_t* e = ev.data() + 1 + (size_t)per * it;
    for (int k = 0; k < nk; ++k) {
      float ms = 0.f;
      A2C_HIP(c, hipEventElapsedTime(&ms, e[k], e[k + 1]));
      acc[k] += ms;
    }
  }
  out->n_kernels = nk;
  for (int k = 0; k < nk; ++k) {
    std::strncpy(out->names[k], names[k] ? names[k] : "?", 31);
    out->avg_ms[k] = acc[k] / iters;
  }
  out->total_ms = tot;
  return A2C_OK;
}

}  // extern "C"
