// Handle lifecycle, error reporting, workspace and launch profiling.
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "mh_internal.h"

namespace mh {

int fail(mh_handle h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (h) h->last_error = buf;
  return code;
}

static int grow(mh_handle h, void **p, size_t *have, size_t want) {
  if (*have >= want) return MH_OK;
  // Work already queued may still read the old buffer.
  MH_HIP(h, hipStreamSynchronize(h->stream));
  if (*p) MH_HIP(h, hipFree(*p));
  *p = nullptr;
  *have = 0;
  size_t bytes = (want + 255) & ~size_t(255);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(h, MH_ERR_OOM, "device allocation of %zu bytes failed: %s",
                bytes, hipGetErrorString(e));
  }
  *have = bytes;
  return MH_OK;
}

int ensure_ws(mh_handle h, size_t bytes) {
  return grow(h, &h->ws, &h->ws_bytes, bytes);
}
int ensure_small(mh_handle h, size_t bytes) {
  return grow(h, &h->small, &h->small_bytes, bytes);
}
int ensure_red(mh_handle h, size_t n) {
  if (h->red_len >= n) return MH_OK;
  void *p = h->red;
  size_t have = h->red_len * sizeof(double);
  MH_TRY(grow(h, &p, &have, n * sizeof(double)));
  h->red = (double *)p;
  h->red_len = have / sizeof(double);
  if (h->red_host) MH_HIP(h, hipHostFree(h->red_host));
  h->red_host = nullptr;
  MH_HIP(h, hipHostMalloc((void **)&h->red_host, have));
  return MH_OK;
}

static hipEvent_t take_event(mh_handle h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void prof_begin(mh_handle h, int kclass, double flops, int *slot) {
  *slot = -1;
  if (!h->profiling || h->noalloc) return;
  ProfRec r;
  r.start = take_event(h);
  r.stop = take_event(h);
  if (!r.start || !r.stop) return;
  r.kclass = kclass;
  r.flops = flops;
  (void)hipEventRecord(r.start, h->stream);
  h->prof_pending.push_back(r);
  *slot = (int)h->prof_pending.size() - 1;
}

void prof_end(mh_handle h, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(h->prof_pending[slot].stop, h->stream);
}

int check_launch(mh_handle h, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return fail(h, MH_ERR_HIP, "launch of %s failed: %s", what,
                hipGetErrorString(e));
  return MH_OK;
}

static int prof_collect(mh_handle h) {
  for (auto &r : h->prof_pending) {
    MH_HIP(h, hipEventSynchronize(r.stop));
    float ms = 0.f;
    MH_HIP(h, hipEventElapsedTime(&ms, r.start, r.stop));
    h->prof_ms[r.kclass] += ms;
    h->prof_n[r.kclass] += 1;
    h->prof_flops[r.kclass] += r.flops;
    h->event_pool.push_back(r.start);
    h->event_pool.push_back(r.stop);
  }
  h->prof_pending.clear();
  return MH_OK;
}

}  // namespace mh

using namespace mh;

extern "C" {

int mh_version(void) { return 1; }

int mh_create(int device, mh_handle *out) {
  if (!out) return MH_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return MH_ERR_HIP;
  }
  if (device < 0 || device >= n) return MH_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return MH_ERR_HIP;
  mh_handle h = new mh_context();
  h->device = device;
  h->stream = nullptr;
  // MH_AUTOTUNE=0 disables per-shape GEMM tuning; MH_TUNE_LOG=1 prints it.
  if (const char *s = getenv("MH_AUTOTUNE")) h->autotune = atoi(s) != 0;
  if (const char *s = getenv("MH_TUNE_LOG")) h->tune_log = atoi(s) != 0;
  // MH_GRAPHS=0: run the Krylov recurrences eagerly instead of as a hipGraph;
  // MH_GRAPHS=2: capture the short-vector steps too.
  if (const char *s = getenv("MH_GRAPHS")) h->graphs = std::min(2, std::max(0, atoi(s)));
  // MH_FUSION=0: run the first H_eff step as GEMM + separate mode product.
  if (const char *s = getenv("MH_FUSION")) h->fusion = atoi(s) != 0;
  // MH_KRYLOV_SMALL=0: multi-block vector kernels for every Krylov step.
  if (const char *s = getenv("MH_KRYLOV_SMALL")) h->krylov_small = atoi(s) != 0;
  // MH_BS_PLAIN=1: block plans run one GEMM launch per sector product.
  if (const char *s = getenv("MH_BS_PLAIN")) h->bs_plain = atoi(s) != 0;
  // The allocation-free entry points find their static scratch array's
  // address on this device now, with the device current.
  const int st = mh::noalloc_setup(h);
  if (st != MH_OK) {
    fprintf(stderr, "mh_create: %s\n", h->last_error.c_str());
    delete h;
    return st;
  }
  *out = h;
  return MH_OK;
}

int mh_destroy(mh_handle h) {
  if (!h) return MH_ERR_ARG;
  (void)hipSetDevice(h->device);
  (void)hipStreamSynchronize(h->stream);
  for (auto &r : h->prof_pending) {
    (void)hipEventDestroy(r.start);
    (void)hipEventDestroy(r.stop);
  }
  for (auto e : h->event_pool) (void)hipEventDestroy(e);
  if (h->ws) (void)hipFree(h->ws);
  if (h->small) (void)hipFree(h->small);
  if (h->pack) (void)hipFree(h->pack);
  if (h->red) (void)hipFree(h->red);
  if (h->red_host) (void)hipHostFree(h->red_host);
  if (h->cap_stream) (void)hipStreamDestroy(h->cap_stream);
  mh::eig_release(h);
  mh::dc_release(h);
  mh::bg_release(h);
  delete h;
  return MH_OK;
}

int mh_set_graphs(mh_handle h, int enable) {
  if (!h) return MH_ERR_ARG;
  h->graphs = enable < 0 ? 0 : (enable > 2 ? 2 : enable);
  return MH_OK;
}

int mh_set_fusion(mh_handle h, int enable) {
  if (!h) return MH_ERR_ARG;
  h->fusion = enable ? 1 : 0;
  return MH_OK;
}

int mh_set_krylov_small(mh_handle h, int enable) {
  if (!h) return MH_ERR_ARG;
  h->krylov_small = enable ? 1 : 0;
  return MH_OK;
}

int mh_set_bs_plain(mh_handle h, int enable) {
  if (!h) return MH_ERR_ARG;
  h->bs_plain = enable ? 1 : 0;
  return MH_OK;
}

int mh_set_stream(mh_handle h, void *stream) {
  if (!h) return MH_ERR_ARG;
  h->stream = (hipStream_t)stream;
  return MH_OK;
}

int mh_synchronize(mh_handle h) {
  if (!h) return MH_ERR_ARG;
  MH_HIP(h, hipStreamSynchronize(h->stream));
  return MH_OK;
}

const char *mh_last_error(mh_handle h) {
  if (!h) return "null handle";
  return h->last_error.c_str();
}

int mh_malloc(mh_handle h, size_t bytes, void **ptr) {
  if (!h || !ptr) return MH_ERR_ARG;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, MH_ERR_OOM, "mh_malloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  return MH_OK;
}

int mh_free(mh_handle h, void *ptr) {
  if (!h) return MH_ERR_ARG;
  MH_HIP(h, hipStreamSynchronize(h->stream));
  MH_HIP(h, hipFree(ptr));
  return MH_OK;
}

int mh_upload(mh_handle h, void *dst, const void *src, size_t bytes) {
  if (!h || (!dst && bytes) || (!src && bytes)) return MH_ERR_ARG;
  MH_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
  MH_HIP(h, hipStreamSynchronize(h->stream));
  return MH_OK;
}

int mh_download(mh_handle h, void *dst, const void *src, size_t bytes) {
  if (!h || (!dst && bytes) || (!src && bytes)) return MH_ERR_ARG;
  MH_HIP(h, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
  MH_HIP(h, hipStreamSynchronize(h->stream));
  return MH_OK;
}

int mh_profile_enable(mh_handle h, int enable) {
  if (!h) return MH_ERR_ARG;
  h->profiling = enable ? 1 : 0;
  return MH_OK;
}

int mh_profile_reset(mh_handle h) {
  if (!h) return MH_ERR_ARG;
  MH_TRY(prof_collect(h));
  for (int i = 0; i < MH_K_NCLASS; ++i) {
    h->prof_ms[i] = 0;
    h->prof_n[i] = 0;
    h->prof_flops[i] = 0;
  }
  return MH_OK;
}

int mh_profile_read(mh_handle h, int kclass, double *ms, int64_t *launches,
                    double *flops) {
  if (!h || kclass < 0 || kclass >= MH_K_NCLASS) return MH_ERR_ARG;
  MH_TRY(prof_collect(h));
  if (ms) *ms = h->prof_ms[kclass];
  if (launches) *launches = h->prof_n[kclass];
  if (flops) *flops = h->prof_flops[kclass];
  return MH_OK;
}

}  // extern "C"
