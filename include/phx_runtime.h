/*
 * phx_runtime.h -- device plumbing exported by libphx.so for hosts that do not bring
 * their own HIP runtime bindings (the Python tests and bench.py use it through ctypes).
 * No counterpart in the reference: parthenon/Kokkos own memory there.  Thin wrappers
 * over the HIP runtime; every call returns PHX_OK or PHX_ERR_DEVICE.
 */
#ifndef PHX_RUNTIME_H_
#define PHX_RUNTIME_H_

#include <stddef.h>
#include <stdint.h>

#include "phx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void *phx_event;

int phx_rt_device_count(int *n);
int phx_rt_set_device(int device);
int phx_rt_malloc(void **ptr, size_t bytes);
int phx_rt_free(void *ptr);
int phx_rt_memcpy_htod(void *dst, const void *src, size_t bytes);
int phx_rt_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int phx_rt_memcpy_dtod(void *dst, const void *src, size_t bytes);
int phx_rt_memcpy_dtod_async(void *dst, const void *src, size_t bytes, phx_stream s);
int phx_rt_memset(void *ptr, int value, size_t bytes);
int phx_rt_device_synchronize(void);
int phx_rt_stream_create(phx_stream *s);
int phx_rt_stream_destroy(phx_stream s);
int phx_rt_stream_synchronize(phx_stream s);
int phx_rt_event_create(phx_event *e);
int phx_rt_event_destroy(phx_event e);
int phx_rt_event_record(phx_event e, phx_stream s);
int phx_rt_event_elapsed_ms(phx_event start, phx_event stop, float *ms);
/* Live per-launch timing of phx_fluid_stage: while started, the context brackets each
 * kernel launch of a stage with HIP events on the stream the launch uses (up to 64 per
 * reset).  Records are read back after the stream has finished. */
int phx_rt_profile_start(phx_ctx *ctx);
int phx_rt_profile_stop(phx_ctx *ctx);
int phx_rt_profile_reset(phx_ctx *ctx);
int phx_rt_profile_count(phx_ctx *ctx, int *n);
int phx_rt_profile_get(phx_ctx *ctx, int idx, char *name, int name_len, float *ms);
/* device name and compute units, for reports */
int phx_rt_device_info(char *name, int name_len, int *compute_units, double *mem_gb);

#ifdef __cplusplus
}
#endif
#endif /* PHX_RUNTIME_H_ */
