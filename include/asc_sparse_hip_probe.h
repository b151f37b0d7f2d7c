/*
 * Measurement probes beside the ASC::sparse_hip library (built as
 * libasc_sparse_probe.so, used by the benchmark).  No reference counterpart:
 * they measure the ceilings the SpMV rooflines are quoted against.  Both
 * take a context of the library (asc_sparse_hip.h) and enqueue on its
 * stream.
 */
#ifndef ASC_SPARSE_HIP_PROBE_H_
#define ASC_SPARSE_HIP_PROBE_H_

#include <stddef.h>
#include <stdint.h>

#include "asc_sparse_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Work-groups of the probe launches: the number of partials they write. */
#define ASC_HIP_GATHER_PROBE_BLOCKS 2048
#define ASC_HIP_ROW_GATHER_PROBE_BLOCKS 8192
#define ASC_HIP_STREAM_PROBE_BLOCKS 16384

/*
 * Gather rate: for p in [0, count) streams indices[p] (int32, clamped on the
 * device to [0, x_extent)) and values[p], gathers x[indices[p]] through the
 * caches as the SpMV kernels do, and sums values[p] * x[indices[p]]; one
 * FP64 partial per work-group goes to `partials` (device, at least
 * ASC_HIP_GATHER_PROBE_BLOCKS values, overlapping no input).  Timing it over
 * random indices with x of a given size gives the gather rate a
 * random-column SpMV can reach when x is resident at that cache level.
 */
asc_hip_status asc_sparse_hip_gather_probe(asc_hip_context* context, const int32_t* indices,
                                           const double* values, int64_t count, const double* x,
                                           int64_t x_extent, double* partials,
                                           size_t partials_count);

/*
 * Row gather rate: for p in [0, count) streams indices[p] (int32, clamped on
 * the device to [0, table_rows)) and gathers row indices[p] of `table`
 * (table_rows rows of 32 FP64, 256 bytes each, 32-byte aligned), an 8-lane
 * group per row as the plain SpMM kernel gathers B; every gathered element
 * is summed, one FP64 partial per work-group to `partials` (device, at least
 * ASC_HIP_ROW_GATHER_PROBE_BLOCKS values, overlapping no input).  Timing it
 * over random indices with tables of a given size gives the B-row gather
 * rate a random-column SpMM with k = 32 can reach at that cache level.
 */
asc_hip_status asc_sparse_hip_row_gather_probe(asc_hip_context* context, const int32_t* indices,
                                               int64_t count, const double* table,
                                               int64_t table_rows, double* partials,
                                               size_t partials_count);

/*
 * HBM stream rate: reads `bytes` (a multiple of 16, 16-byte aligned) of
 * `src` with nontemporal 16-byte loads and, when `dst` is NULL, sums its
 * 32-bit words into one partial per work-group (`partials`, at least
 * ASC_HIP_STREAM_PROBE_BLOCKS values); otherwise copies them to `dst` with
 * nontemporal stores.
 */
asc_hip_status asc_sparse_hip_stream_probe(asc_hip_context* context, const void* src, void* dst,
                                           int64_t bytes, uint32_t* partials,
                                           size_t partials_count);

#ifdef __cplusplus
}
#endif

#endif /* ASC_SPARSE_HIP_PROBE_H_ */
