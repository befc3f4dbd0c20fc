/* hyd_hip_runtime.h - HBM residency for hyd_array_t (libhydhip.so).
 *
 * Every storage block allocated by hyd_array_alloc_bytes carries an HBM
 * mirror next to the host bytes the reference ABI exposes through
 * arr->data (runtime/hydrangea_runtime.h:220-235 keeps its layout).  Kernels
 * read and write the mirror; the host bytes are refreshed on demand:
 *
 *   - an input array that is not device-resident is uploaded on every call
 *     (the host copy is authoritative, so a caller that rewrites arr->data
 *     between calls gets the reference's semantics);
 *   - arrays returned by hyd_export_* and arrays from hyd_array_alloc_device
 *     are device-resident: later kernels read them in place, without PCIe;
 *   - with host sync on (the default), hyd_export_* copy their results back
 *     before returning, so arr->data is valid exactly as with the C backend;
 *     with host sync off the caller calls hyd_array_sync_host before reading
 *     arr->data.
 */
#ifndef HYD_HIP_RUNTIME_H
#define HYD_HIP_RUNTIME_H

#include "hydrangea_runtime.h"
#include "hyd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Device-resident allocation (zero-filled in HBM; host bytes zero too). */
hyd_array_t* hyd_array_alloc_device(hyd_tuple_t shape, size_t elem_size);
/* HBM address of arr->data; uploads the host bytes first when the array is
 * not device-resident.  Exits with a message on failure (drop-in errors). */
void* hyd_array_device_data(hyd_array_t* arr);
/* hyd_array_device_data for code that reports its host writes (the kernels
 * the HIP backend generates, whose printed C marks every host write with
 * hyd_array_mark_host_dirty): after an upload both copies count as current,
 * so later kernels read the HBM copy without uploading again. */
void* hyd_array_device_data_keep(hyd_array_t* arr);
/* The host may have rewritten arr->data since the last kernel (an array
 * handed in by a caller): when the host copy is current, the next kernel
 * uploads it again; an HBM copy that is newer stays authoritative. */
int hyd_array_enter_host(hyd_array_t* arr);
/* Copies the HBM contents back into arr->data if they are newer. */
int hyd_array_sync_host(hyd_array_t* arr);
/* The caller rewrote arr->data of a device-resident array: the next kernel
 * re-uploads it. */
int hyd_array_mark_host_dirty(hyd_array_t* arr);
/* A kernel called through hyd_hip.h wrote the HBM copy (the pointer from
 * hyd_array_device_data): it becomes the authoritative one, and
 * hyd_array_sync_host copies it back. */
int hyd_array_mark_device_written(hyd_array_t* arr);
/* 1 if the HBM copy is the authoritative one. */
int hyd_array_is_device_resident(hyd_array_t* arr);
/* Host sync of hyd_export_* results: 1 (default) or 0 (leave in HBM). */
void hyd_hip_set_host_sync(int on);
int hyd_hip_get_host_sync(void);
/* Stream used by hyd_export_* (NULL = the library's stream). */
void hyd_hip_set_stream(void* stream);
/* The stream hyd_export_*, the uploads of hyd_array_device_data and the
 * generated launchers run on (the library's stream unless set). */
void* hyd_hip_get_stream(void);
/* hyd_array_alloc_bytes(shape, 8) zero in HBM and on the host, both copies
 * valid: the allocation the HIP backend prints for an array of 8-byte
 * elements (int64, fp64, bool), so that a kernel writing it first uploads
 * nothing. */
hyd_array_t* hyd_array_alloc_device8(hyd_tuple_t shape);

#ifdef __cplusplus
}
#endif

#endif /* HYD_HIP_RUNTIME_H */
