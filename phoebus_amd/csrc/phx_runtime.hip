// phx_runtime.hip -- context lifetime (include/phx.h) and device plumbing
// (include/phx_runtime.h) of libphx.so.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "phx.h"
#include "phx_internal.hpp"
#include "phx_runtime.h"

static inline int rt(hipError_t e) { return e == hipSuccess ? PHX_OK : PHX_ERR_DEVICE; }

extern "C" {

int phx_create(phx_ctx **out, int device) {
  if (!out) return PHX_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return PHX_ERR_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return PHX_ERR_DEVICE;
  phx_ctx *c = new (std::nothrow) phx_ctx();
  if (!c) return PHX_ERR_NOMEM;
  c->device = device;
  // the device's copy of the allocation-free stage's scratch, looked up at set-up
  double *scratch_ptr = nullptr;
  if (phx::static_scratch(&scratch_ptr) != PHX_OK) {
    delete c;
    return PHX_ERR_DEVICE;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess ||
      hipMalloc(&c->dt_partial, sizeof(double) * phx::kDtPartials) != hipSuccess) {
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return PHX_ERR_DEVICE;
  }
  *out = c;
  return PHX_OK;
}

int phx_destroy(phx_ctx *c) {
  if (!c) return PHX_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->profiling) (void)phx_rt_profile_stop(c);
  if (c->dt_partial) (void)hipFree(c->dt_partial);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PHX_OK;
}

const char *phx_last_error(const phx_ctx *c) {
  if (c) return c->last_error.c_str();
  const char *m = phx::noalloc_last_error();  // phx_fluid_stage_noalloc has no context
  return m[0] ? m : "null context";
}

int phx_rt_device_count(int *n) { return rt(hipGetDeviceCount(n)); }
int phx_rt_set_device(int d) { return rt(hipSetDevice(d)); }
int phx_rt_malloc(void **p, size_t bytes) { return rt(hipMalloc(p, bytes)); }
int phx_rt_free(void *p) { return rt(hipFree(p)); }
int phx_rt_memcpy_htod(void *d, const void *s, size_t n) {
  return rt(hipMemcpy(d, s, n, hipMemcpyHostToDevice));
}
int phx_rt_memcpy_dtoh(void *d, const void *s, size_t n) {
  return rt(hipMemcpy(d, s, n, hipMemcpyDeviceToHost));
}
int phx_rt_memcpy_dtod(void *d, const void *s, size_t n) {
  return rt(hipMemcpy(d, s, n, hipMemcpyDeviceToDevice));
}
int phx_rt_memcpy_dtod_async(void *d, const void *s, size_t n, phx_stream st) {
  return rt(hipMemcpyAsync(d, s, n, hipMemcpyDeviceToDevice, (hipStream_t)st));
}
int phx_rt_memset(void *p, int v, size_t n) { return rt(hipMemset(p, v, n)); }
int phx_rt_device_synchronize(void) { return rt(hipDeviceSynchronize()); }
int phx_rt_stream_create(phx_stream *s) {
  hipStream_t h = nullptr;
  const int r = rt(hipStreamCreate(&h));
  *s = h;
  return r;
}
int phx_rt_stream_destroy(phx_stream s) { return rt(hipStreamDestroy((hipStream_t)s)); }
int phx_rt_stream_synchronize(phx_stream s) { return rt(hipStreamSynchronize((hipStream_t)s)); }
int phx_rt_event_create(phx_event *e) {
  hipEvent_t h = nullptr;
  const int r = rt(hipEventCreate(&h));
  *e = h;
  return r;
}
int phx_rt_event_destroy(phx_event e) { return rt(hipEventDestroy((hipEvent_t)e)); }
int phx_rt_event_record(phx_event e, phx_stream s) {
  return rt(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
}
int phx_rt_event_elapsed_ms(phx_event a, phx_event b, float *ms) {
  return rt(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
}
int phx_rt_profile_start(phx_ctx *c) {
  if (!c) return PHX_ERR_INVALID;
  if (!c->profiling) {
    for (int i = 0; i < phx_ctx::kMaxTimed; ++i) {
      if (hipEventCreate(&c->timed[i].start) != hipSuccess ||
          hipEventCreate(&c->timed[i].stop) != hipSuccess)
        return PHX_ERR_DEVICE;
    }
    c->profiling = true;
  }
  c->n_timed = 0;
  return PHX_OK;
}

int phx_rt_profile_stop(phx_ctx *c) {
  if (!c || !c->profiling) return PHX_ERR_INVALID;
  for (int i = 0; i < phx_ctx::kMaxTimed; ++i) {
    (void)hipEventDestroy(c->timed[i].start);
    (void)hipEventDestroy(c->timed[i].stop);
  }
  c->profiling = false;
  c->n_timed = 0;
  return PHX_OK;
}

int phx_rt_profile_count(phx_ctx *c, int *n) {
  if (!c || !n) return PHX_ERR_INVALID;
  *n = c->n_timed;
  return PHX_OK;
}

int phx_rt_profile_get(phx_ctx *c, int idx, char *name, int name_len, float *ms) {
  if (!c || idx < 0 || idx >= c->n_timed) return PHX_ERR_INVALID;
  if (hipEventSynchronize(c->timed[idx].stop) != hipSuccess) return PHX_ERR_DEVICE;
  if (hipEventElapsedTime(ms, c->timed[idx].start, c->timed[idx].stop) != hipSuccess)
    return PHX_ERR_DEVICE;
  if (name && name_len > 0) {
    std::strncpy(name, c->timed[idx].name, name_len - 1);
    name[name_len - 1] = 0;
  }
  return PHX_OK;
}

int phx_rt_profile_reset(phx_ctx *c) {
  if (!c) return PHX_ERR_INVALID;
  c->n_timed = 0;
  return PHX_OK;
}

int phx_rt_device_info(char *name, int len, int *cus, double *mem_gb) {
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return PHX_ERR_DEVICE;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, d) != hipSuccess) return PHX_ERR_DEVICE;
  if (name && len > 0) {
    std::strncpy(name, p.gcnArchName, len - 1);
    name[len - 1] = 0;
  }
  if (cus) *cus = p.multiProcessorCount;
  if (mem_gb) *mem_gb = p.totalGlobalMem / 1073741824.0;
  return PHX_OK;
}

}  // extern "C"
