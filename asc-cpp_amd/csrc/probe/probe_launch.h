// Launch interface of the measurement probes (libasc_sparse_probe.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asc_hip {

// Work-groups of the gather probes (one FP64 partial each) and of the
// stream probes (one 32-bit partial each for the read).
constexpr int64_t kGatherProbeBlocks = 2048;
constexpr int64_t kRowGatherProbeBlocks = 8192;
constexpr int64_t kStreamProbeBlocks = 16384;

hipError_t gather_probe_launch(const int32_t* indices, const double* values, int64_t n,
                               const double* x, int64_t x_extent, double* partials,
                               hipStream_t stream);
// Rows are 256 bytes (32 FP64); indices clamped to table_rows - 1.
hipError_t row_gather_probe_launch(const int32_t* indices, int64_t n, const double* table,
                                   int64_t table_rows, double* partials, hipStream_t stream);
// dst == nullptr: read `words` 16-byte words of src (word sums to partials);
// else copy them to dst.
hipError_t stream_probe_launch(const void* src, void* dst, int64_t words, uint32_t* partials,
                               hipStream_t stream);

}  // namespace asc_hip
