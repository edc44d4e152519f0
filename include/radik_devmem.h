/*
 * radik_devmem.h -- device-memory, stream and event helpers exported by
 * libradik_hip.so for hosts that have no GPU array type of their own (the
 * ctypes test harness and bench.py).  No reference counterpart: a Julia
 * caller uses AMDGPU.jl's ROCArray and HIPStream instead.  All return 0 or a
 * RADIK_E* code (radik.h) with radik_last_error() set.
 */
#ifndef RADIK_DEVMEM_H
#define RADIK_DEVMEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int radik_device_count(int *n);
int radik_set_device(int dev);
/* Current device: compute units, LDS bytes per block, HBM bytes, arch name. */
int radik_device_info(int *cus, int64_t *lds_per_block, int64_t *hbm_bytes, char *name, int name_len);
/* LDS capacities the kernels use on the current device (elements). */
int radik_lds_caps(uint32_t *small_cap, uint32_t *refine_cap, uint32_t *sort_cap);

int radik_malloc(void **p, size_t bytes);
int radik_free(void *p);
int radik_memcpy_h2d(void *dst, const void *src, size_t bytes);
int radik_memcpy_d2h(void *dst, const void *src, size_t bytes);
int radik_memcpy_d2d(void *dst, const void *src, size_t bytes);
int radik_memset(void *p, int value, size_t bytes);
int radik_device_sync(void);

int radik_stream_create(void **stream);
int radik_stream_destroy(void *stream);
int radik_stream_sync(void *stream);

int radik_event_create(void **event);
int radik_event_destroy(void *event);
int radik_event_record(void *event, void *stream);
int radik_event_sync(void *event);
int radik_event_elapsed(void *start, void *stop, float *ms);

#ifdef __cplusplus
}
#endif

#endif /* RADIK_DEVMEM_H */
