// Device conversions for gfx950: canonical COO -> CSR/CSC and the stable
// CSR <-> CSC transpose.
//
// The reference builds every compressed array by scanning the source once
// per output outer index (compressed.h:1082-1131, 1151-1221); its output is
// the unique canonical order, i.e. a *stable* sort of the source entries by
// the new outer index.  Here that sort is an LSD radix sort:
//   - 8-bit digits, one histogram + scan + scatter round per digit;
//   - tiles of kTile entries; inside a tile each wave64 ranks its entries
//     with a ballot "match" on the digit bits, which keeps equal digits in
//     source order (stability), so the result is bit-exact with the
//     reference on indices and values;
//   - keys travel with their payload (other coordinate + raw value bits);
//   - offsets come from the sorted keys by a boundary pass
//     (offsets[c] = first position whose key >= c), no atomics anywhere.
// Canonical COO -> CSR is already sorted by row: one streaming pass copies
// columns and values and derives the offsets by the same boundary rule.
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kBlock = 256;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kWaves = kBlock / kWave;
// Radix scatter blocks: 16 waves of 8 entries per lane, so a digit's run of
// one tile averages 32 entries (longer contiguous stores) and one block per
// CU holds the LDS reorder buffer.
constexpr int kSortBlock = 1024;
constexpr int kIpt = 16;
constexpr int kTile = kSortBlock * kIpt;  // entries per radix tile
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kWaveItems = kWave * kIpt;

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

unsigned blocks_for(int64_t n, int64_t per) {
  return static_cast<unsigned>((n + per - 1) / per);
}

// ---- exclusive scan of uint32 counts (reduce, scan partials, downsweep) ----
constexpr int kScanIpt = 16;
constexpr int kScanTile = kBlock * kScanIpt;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v,
                                                         uint32_t* s_tmp,
                                                         uint32_t* total) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) s_tmp[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < kWaves; ++w) {
    if (w < wave) base += s_tmp[w];
    all += s_tmp[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

__global__ __launch_bounds__(kBlock) void scan_reduce_kernel(
    const uint32_t* __restrict__ in, int64_t n, uint32_t* __restrict__ partial) {
  __shared__ uint32_t s_tmp[kWaves];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  uint32_t sum = 0;
#pragma unroll
  for (int u = 0; u < kScanIpt; ++u) {
    const int64_t i = base + u * kBlock + threadIdx.x;
    if (i < n) sum += in[i];
  }
  uint32_t total;
  block_exclusive_scan(sum, s_tmp, &total);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Exclusive scan in place of a short array by one block.
__global__ __launch_bounds__(kBlock) void scan_single_kernel(uint32_t* data,
                                                             int64_t n) {
  __shared__ uint32_t s_tmp[kWaves];
  uint32_t carry = 0;
  for (int64_t base = 0; base < n; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const uint32_t v = i < n ? data[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, s_tmp, &total);
    if (i < n) data[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kBlock) void scan_downsweep_kernel(
    const uint32_t* __restrict__ in, int64_t n,
    const uint32_t* __restrict__ partial, uint32_t* __restrict__ out) {
  __shared__ uint32_t s_tmp[kWaves];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  // Each thread owns kScanIpt consecutive entries.
  uint32_t v[kScanIpt];
  uint32_t sum = 0;
#pragma unroll
  for (int u = 0; u < kScanIpt; ++u) {
    const int64_t i = base + threadIdx.x * kScanIpt + u;
    v[u] = i < n ? in[i] : 0u;
    sum += v[u];
  }
  uint32_t total;
  uint32_t run = partial[blockIdx.x] + block_exclusive_scan(sum, s_tmp, &total);
#pragma unroll
  for (int u = 0; u < kScanIpt; ++u) {
    const int64_t i = base + threadIdx.x * kScanIpt + u;
    if (i < n) out[i] = run;
    run += v[u];
  }
}

size_t scan_workspace_bytes(int64_t n) {
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  return align256(sizeof(uint32_t) * (nb > 0 ? nb : 1));
}

// Exclusive scan of n uint32 (n * kScanTile-sized partials must fit one
// block pass; the single-block step loops, so any n works).
hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, int64_t n,
                          uint32_t* partial, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned nb = blocks_for(n, kScanTile);
  scan_reduce_kernel<<<nb, kBlock, 0, s>>>(in, n, partial);
  scan_single_kernel<<<1, kBlock, 0, s>>>(partial, nb);
  scan_downsweep_kernel<<<nb, kBlock, 0, s>>>(in, n, partial, out);
  return hipGetLastError();
}

// ---- row expansion: rows[p] = outer index owning position p --------------
template <int G>
__global__ __launch_bounds__(kBlock) void expand_kernel(
    const int64_t* __restrict__ offsets, int64_t outer,
    uint32_t* __restrict__ rows) {
  const int64_t o = xcd_remap(blockIdx.x, gridDim.x) * (kBlock / G) + threadIdx.x / G;
  if (o >= outer) return;
  const int64_t b = offsets[o], e = offsets[o + 1];
  for (int64_t p = b + threadIdx.x % G; p < e; p += G)
    rows[p] = static_cast<uint32_t>(o);
}

hipError_t expand_launch(const int64_t* offsets, int64_t outer, int64_t nnz,
                         uint32_t* rows, hipStream_t s) {
  if (outer == 0 || nnz == 0) return hipSuccess;
  const int64_t avg = nnz / outer;
  if (avg >= 32) {
    expand_kernel<64><<<blocks_for(outer, kBlock / 64), kBlock, 0, s>>>(offsets, outer, rows);
  } else if (avg >= 4) {
    expand_kernel<8><<<blocks_for(outer, kBlock / 8), kBlock, 0, s>>>(offsets, outer, rows);
  } else {
    expand_kernel<1><<<blocks_for(outer, kBlock), kBlock, 0, s>>>(offsets, outer, rows);
  }
  return hipGetLastError();
}

// ---- radix passes ---------------------------------------------------------
// Key sources: a uint32 array, or an int32/int64 array read with a stride
// (COO coordinates are interleaved).
template <typename K>
__device__ __forceinline__ uint32_t read_key(const K* keys, int64_t stride,
                                             int64_t i) {
  return static_cast<uint32_t>(ld_stream(keys + i * stride));
}

template <typename K>
__global__ __launch_bounds__(kBlock) void radix_hist_kernel(
    const K* __restrict__ keys, int64_t key_stride, int64_t n, int shift,
    uint32_t* __restrict__ counts, int64_t ntiles) {
  __shared__ uint32_t s_hist[kRadix];
  s_hist[threadIdx.x] = 0;  // kBlock == kRadix
  __syncthreads();
  const int64_t tile = blockIdx.x;
  const int64_t base = tile * kTile;
#pragma unroll 8
  for (int u = 0; u < kTile / kBlock; ++u) {
    const int64_t i = base + u * kBlock + threadIdx.x;
    if (i < n) {
      const uint32_t d = (read_key(keys, key_stride, i) >> shift) & (kRadix - 1);
      atomicAdd(&s_hist[d], 1u);
    }
  }
  __syncthreads();
  counts[static_cast<int64_t>(threadIdx.x) * ntiles + tile] = s_hist[threadIdx.x];
}

template <typename T>
__device__ __forceinline__ T ld_plain(const T* p) { return *p; }

// Value bits through one 8-byte LDS slot.
template <typename VB>
__device__ __forceinline__ uint2 pack_bits(VB v) {
  static_assert(sizeof(VB) <= 8, "value payloads are at most 8 bytes");
  uint2 r{0u, 0u};
  __builtin_memcpy(&r, &v, sizeof(VB));
  return r;
}
template <typename VB>
__device__ __forceinline__ VB unpack_bits(uint2 r) {
  VB v;
  __builtin_memcpy(&v, &r, sizeof(VB));
  return v;
}

// One stable scatter round.  Entry i of the tile is handled by wave
// (i / kWaveItems), chunk ((i % kWaveItems) / 64), lane (i % 64), so chunk
// order and lane order are source order.  Keys, rows and values are all
// loaded first, so their latency overlaps the ranking.  The tile is then
// reordered by digit in LDS (stable: digit, then wave, then rank within the
// wave) and written out by consecutive threads, so each digit's run of the
// tile leaves as contiguous stores instead of one scattered store per entry.
// The reorder goes through one 8-byte-per-entry LDS buffer in two rounds
// (key and row, then value), which keeps the block at 42 KiB of LDS
// (three blocks per CU).
template <typename K, typename R, typename VB, typename RO>
__global__ __launch_bounds__(kSortBlock) void radix_scatter_kernel(
    const K* __restrict__ keys, int64_t key_stride, const R* __restrict__ rows,
    int64_t row_stride, const VB* __restrict__ vals, int64_t n, int shift,
    const uint32_t* __restrict__ offsets, int64_t ntiles,
    uint32_t* __restrict__ keys_out, RO* __restrict__ rows_out,
    VB* __restrict__ vals_out) {
  // s_count holds each wave's digit counts, then (in place) the counts of
  // the earlier waves per digit.
  __shared__ uint32_t s_count[kSortWaves][kRadix];
  __shared__ uint32_t s_local[kRadix];   // tile-local start of each digit
  __shared__ uint32_t s_global[kRadix];  // output start of each digit
  __shared__ uint32_t s_tmp[kRadix / kWave];
  __shared__ uint2 s_stage[kTile];
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  for (int d = threadIdx.x; d < kSortWaves * kRadix; d += kSortBlock)
    (&s_count[0][0])[d] = 0;
  const int64_t tile = blockIdx.x;
  const int64_t tile_base = tile * kTile;
  const int64_t base = tile_base + static_cast<int64_t>(wave) * kWaveItems;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  uint32_t key[kIpt];
  uint32_t row[kIpt];
  VB val[kIpt];
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int64_t i = base + u * kWave + lane;
    if (i < n) {
      key[u] = read_key(keys, key_stride, i);
      row[u] = static_cast<uint32_t>(ld_stream(rows + i * row_stride));
      val[u] = ld_stream(vals + i);
    } else {
      key[u] = 0u;
      row[u] = 0u;
      val[u] = VB{};
    }
  }
  __syncthreads();  // s_count cleared

  uint32_t pos[kIpt];
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int64_t i = base + u * kWave + lane;
    const bool valid = i < n;
    const uint32_t d = (key[u] >> shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t lower = __popcll(peers & lt_mask);
    const uint32_t total = __popcll(peers);
    uint32_t r = 0;
    if (valid) r = s_count[wave][d] + lower;
    pos[u] = r;
    // The highest peer publishes the new running count after every peer
    // has read the old one (LDS operations of one wave retire in order).
    if (valid && lower == total - 1) s_count[wave][d] += total;
  }
  __syncthreads();
  // Digit scan by the first kRadix threads: per digit, the earlier waves'
  // counts (in place) and the digit's tile-local start.
  uint32_t run = 0, incl = 0;
  if (threadIdx.x < kRadix) {
    const int d = threadIdx.x;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
      const uint32_t c = s_count[w][d];
      s_count[w][d] = run;
      run += c;
    }
    incl = run;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (lane == kWave - 1) s_tmp[wave] = incl;
    s_global[d] = offsets[static_cast<int64_t>(d) * ntiles + tile];
  }
  __syncthreads();
  if (threadIdx.x < kRadix) {
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += s_tmp[w];
    s_local[threadIdx.x] = before + incl - run;
  }
  __syncthreads();
  // Round 1: keys and rows in digit order.
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int64_t i = base + u * kWave + lane;
    if (i < n) {
      const uint32_t d = (key[u] >> shift) & (kRadix - 1);
      pos[u] += s_local[d] + s_count[wave][d];
      s_stage[pos[u]] = make_uint2(key[u], row[u]);
    }
  }
  __syncthreads();
  // Consecutive threads write consecutive entries of each digit's run; the
  // destinations are kept for the value round (`key` is reused for them).
  const int64_t valid_in_tile = n - tile_base < kTile ? n - tile_base : kTile;
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int t = u * kSortBlock + threadIdx.x;
    if (t < valid_in_tile) {
      const uint2 kr = s_stage[t];
      const uint32_t d = (kr.x >> shift) & (kRadix - 1);
      const uint32_t dst = s_global[d] + (static_cast<uint32_t>(t) - s_local[d]);
      keys_out[dst] = kr.x;
      rows_out[dst] = static_cast<RO>(kr.y);
      key[u] = dst;
    }
  }
  __syncthreads();
  // Round 2: values in digit order.
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int64_t i = base + u * kWave + lane;
    if (i < n) s_stage[pos[u]] = pack_bits(val[u]);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kIpt; ++u) {
    const int t = u * kSortBlock + threadIdx.x;
    if (t < valid_in_tile) vals_out[key[u]] = unpack_bits<VB>(s_stage[t]);
  }
}

// offsets[c] = first sorted position with key >= c, for c in [0, outer].
__global__ __launch_bounds__(kBlock) void boundary_kernel(
    const uint32_t* __restrict__ keys, int64_t n, int64_t outer,
    int64_t* __restrict__ offsets) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
       p <= n; p += stride) {
    const int64_t k = p < n ? static_cast<int64_t>(keys[p]) : outer;
    const int64_t kprev = p > 0 ? static_cast<int64_t>(keys[p - 1]) : -1;
    for (int64_t c = kprev + 1; c <= k; ++c) offsets[c] = p;
  }
}

// Same rule over interleaved COO rows (row of entry p = coords[2p]).
__global__ __launch_bounds__(kBlock) void coo_rows_kernel(
    const int64_t* __restrict__ coords, const uint8_t* __restrict__ vals,
    int value_bytes, int64_t nnz, int64_t outer, int64_t* __restrict__ offsets,
    void* __restrict__ out_indices, int index_bytes,
    uint8_t* __restrict__ out_vals) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
       p <= nnz; p += stride) {
    const int64_t k = p < nnz ? coords[2 * p] : outer;
    const int64_t kprev = p > 0 ? coords[2 * p - 2] : -1;
    for (int64_t c = kprev + 1; c <= k; ++c) offsets[c] = p;
    if (p < nnz) {
      const int64_t col = coords[2 * p + 1];
      if (index_bytes == 4) static_cast<int32_t*>(out_indices)[p] = static_cast<int32_t>(col);
      else static_cast<int64_t*>(out_indices)[p] = col;
      if (value_bytes == 8) reinterpret_cast<uint64_t*>(out_vals)[p] = reinterpret_cast<const uint64_t*>(vals)[p];
      else if (value_bytes == 4) reinterpret_cast<uint32_t*>(out_vals)[p] = reinterpret_cast<const uint32_t*>(vals)[p];
      else reinterpret_cast<uint16_t*>(out_vals)[p] = reinterpret_cast<const uint16_t*>(vals)[p];
    }
  }
}

unsigned grid_stride_blocks(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

int key_bits(int64_t outer) {
  int bits = 0;
  while (bits < 63 && (int64_t(1) << bits) < outer) ++bits;
  return bits;
}

int passes_for(int64_t outer) {
  const int bits = key_bits(outer);
  return bits == 0 ? 1 : (bits + kRadixBits - 1) / kRadixBits;
}

// Workspace layout of a stable sort of n entries keyed below `outer`.
struct SortLayout {
  size_t keys[2], rows[2], vals[2], counts, scanned, partial, total;
};

SortLayout sort_layout(int64_t n, int64_t outer, size_t vbytes) {
  SortLayout L{};
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int64_t ncounts = ntiles * kRadix;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  L.keys[0] = take(sizeof(uint32_t) * n);
  L.keys[1] = take(sizeof(uint32_t) * n);
  L.rows[0] = take(sizeof(uint32_t) * n);
  L.rows[1] = take(sizeof(uint32_t) * n);
  L.vals[0] = take(vbytes * n);
  L.vals[1] = take(vbytes * n);
  L.counts = take(sizeof(uint32_t) * ncounts);
  L.scanned = take(sizeof(uint32_t) * ncounts);
  L.partial = take(scan_workspace_bytes(ncounts));
  L.total = off;
  (void)outer;
  return L;
}

// Stable sort by key (< outer) with payload (row, value).  Round 0 reads the
// caller's arrays (key/row strided); later rounds ping-pong in the
// workspace; the last round writes rows/values to the outputs.
template <typename K0, typename R0, typename VB>
hipError_t stable_sort(const K0* key0, int64_t key_stride, const R0* row0,
                       int64_t row_stride, const VB* val0, int64_t n,
                       int64_t outer, void* out_rows, int out_row_bytes,
                       VB* out_vals, int64_t* out_offsets, char* ws,
                       hipStream_t s) {
  const SortLayout L = sort_layout(n, outer, sizeof(VB));
  const int64_t ntiles = (n + kTile - 1) / kTile;
  const int passes = passes_for(outer);
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + L.counts);
  uint32_t* scanned = reinterpret_cast<uint32_t*>(ws + L.scanned);
  uint32_t* partial = reinterpret_cast<uint32_t*>(ws + L.partial);
  uint32_t* kbuf[2] = {reinterpret_cast<uint32_t*>(ws + L.keys[0]),
                       reinterpret_cast<uint32_t*>(ws + L.keys[1])};
  uint32_t* rbuf[2] = {reinterpret_cast<uint32_t*>(ws + L.rows[0]),
                       reinterpret_cast<uint32_t*>(ws + L.rows[1])};
  VB* vbuf[2] = {reinterpret_cast<VB*>(ws + L.vals[0]),
                 reinterpret_cast<VB*>(ws + L.vals[1])};
  hipError_t err = hipSuccess;
  for (int pass = 0; pass < passes && n > 0; ++pass) {
    const int shift = pass * kRadixBits;
    const int out = pass & 1;
    const bool last = pass == passes - 1;
    if (pass == 0) {
      radix_hist_kernel<K0><<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(
          key0, key_stride, n, shift, counts, ntiles);
    } else {
      radix_hist_kernel<uint32_t><<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(
          kbuf[out ^ 1], 1, n, shift, counts, ntiles);
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if ((err = scan_exclusive(counts, scanned, ntiles * kRadix, partial, s)) != hipSuccess)
      return err;
    const unsigned g = static_cast<unsigned>(ntiles);
#define ASC_SCATTER(KT, KP, KS, RT, RP, RS, VP, ROT, ROP)                        \
  radix_scatter_kernel<KT, RT, VB, ROT><<<g, kSortBlock, 0, s>>>(                \
      KP, KS, RP, RS, VP, n, shift, scanned, ntiles, kbuf[out], ROP,             \
      last ? out_vals : vbuf[out])
    if (pass == 0) {
      if (last) {
        if (out_row_bytes == 4)
          ASC_SCATTER(K0, key0, key_stride, R0, row0, row_stride, val0, int32_t,
                      static_cast<int32_t*>(out_rows));
        else
          ASC_SCATTER(K0, key0, key_stride, R0, row0, row_stride, val0, int64_t,
                      static_cast<int64_t*>(out_rows));
      } else {
        ASC_SCATTER(K0, key0, key_stride, R0, row0, row_stride, val0, uint32_t,
                    rbuf[out]);
      }
    } else {
      if (last) {
        if (out_row_bytes == 4)
          ASC_SCATTER(uint32_t, kbuf[out ^ 1], 1, uint32_t, rbuf[out ^ 1], 1,
                      vbuf[out ^ 1], int32_t, static_cast<int32_t*>(out_rows));
        else
          ASC_SCATTER(uint32_t, kbuf[out ^ 1], 1, uint32_t, rbuf[out ^ 1], 1,
                      vbuf[out ^ 1], int64_t, static_cast<int64_t*>(out_rows));
      } else {
        ASC_SCATTER(uint32_t, kbuf[out ^ 1], 1, uint32_t, rbuf[out ^ 1], 1,
                    vbuf[out ^ 1], uint32_t, rbuf[out]);
      }
    }
#undef ASC_SCATTER
    if ((err = hipGetLastError()) != hipSuccess) return err;
  }
  if (out_offsets == nullptr) return hipSuccess;  // sorted keys stay in the workspace
  const uint32_t* sorted_keys = kbuf[(passes - 1) & 1];
  boundary_kernel<<<grid_stride_blocks(n + 1), kBlock, 0, s>>>(
      sorted_keys, n, outer, out_offsets);
  return hipGetLastError();
}

// Where stable_sort leaves the sorted keys (uint32) inside its workspace.
const uint32_t* sorted_keys_in(const char* ws, int64_t n, int64_t outer, size_t vbytes) {
  const SortLayout L = sort_layout(n, outer, vbytes);
  return reinterpret_cast<const uint32_t*>(ws + L.keys[(passes_for(outer) - 1) & 1]);
}

// ---- COO finalize (CoordinateBuilder::Finalize) ----------------------------
// Raw triplets are checked against the shape, stably sorted by column and
// then by row (two LSD sorts: equal coordinates keep insertion order), and
// compacted: each run of equal coordinates is summed left to right in
// insertion order, as the reference's Compact does (coordinate.h:1462-1497).

__global__ __launch_bounds__(kBlock) void finalize_bounds_kernel(
    const int64_t* __restrict__ coords, int64_t n, int64_t rows, int64_t cols,
    uint32_t* __restrict__ bad) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    const int64_t r = coords[2 * i], c = coords[2 * i + 1];
    if (r < 0 || r >= rows || c < 0 || c >= cols) *bad = 1u;  // any writer: same value
  }
}

// Compaction of the sorted triplets, in two passes over them: a count of
// the surviving runs whose head lies in each tile of kFinTile entries, then
// (after a scan of the tile counts) each tile's runs written at their place.
// Both passes find heads and sum runs the same way, so they agree.  Thread t
// takes entries t, t + kBlock, ... of its tile (kFinPer rounds), so loads
// and stores are coalesced; positions inside a tile follow the rounds' ballots.
constexpr int kFinPer = 8;
constexpr int64_t kFinTile = int64_t(kBlock) * kFinPer;

// The run starting at head i, summed left to right in insertion order (no
// FMA), and whether it survives the zero policy (exact zeros, -0.0
// included, are dropped; NaN is kept).
template <typename T>
__device__ __forceinline__ bool finalize_run(const uint32_t* __restrict__ rows,
                                             const uint32_t* __restrict__ cols,
                                             const T* __restrict__ vals, int64_t n, int64_t i,
                                             uint32_t r, uint32_t c, int drop_zeros, T* sum) {
#pragma clang fp contract(off)
  T v = vals[i];
  for (int64_t j = i + 1; j < n && rows[j] == r && cols[j] == c; ++j) v = v + vals[j];
  *sum = v;
  return !(drop_zeros && v == T(0));
}

// Entry i's coordinates, and whether it heads a run (differs from i - 1).
__device__ __forceinline__ bool finalize_head(const uint32_t* __restrict__ rows,
                                              const uint32_t* __restrict__ cols, int64_t i,
                                              uint32_t* r, uint32_t* c) {
  *r = rows[i];
  *c = cols[i];
  return i == 0 || *r != rows[i - 1] || *c != cols[i - 1];
}

// Pass 1: per tile, the surviving runs that start in it; any entry that is
// not a head marks a duplicate coordinate.
template <typename T>
__global__ __launch_bounds__(kBlock) void finalize_count_kernel(
    const uint32_t* __restrict__ rows, const uint32_t* __restrict__ cols,
    const T* __restrict__ vals, int64_t n, int drop_zeros, uint32_t* __restrict__ counts,
    uint32_t* __restrict__ duplicate) {
  __shared__ uint32_t s_tmp[kWaves];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kFinTile + threadIdx.x;
  uint32_t kept = 0;
  bool dup = false;
#pragma unroll
  for (int u = 0; u < kFinPer; ++u) {
    const int64_t i = base + int64_t(u) * kBlock;
    if (i >= n) break;
    uint32_t r, c;
    if (!finalize_head(rows, cols, i, &r, &c)) {
      dup = true;
      continue;
    }
    T v;
    if (finalize_run(rows, cols, vals, n, i, r, c, drop_zeros, &v)) ++kept;
  }
  if (dup) *duplicate = 1u;  // any writer: same value
  uint32_t total;
  block_exclusive_scan(kept, s_tmp, &total);
  if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Pass 2: each surviving run written at its tile's scanned offset plus the
// runs before it in the tile (round by round, wave by wave, lane by lane:
// entry order); the last tile records the total.
template <typename T>
__global__ __launch_bounds__(kBlock) void finalize_compact_kernel(
    const uint32_t* __restrict__ rows, const uint32_t* __restrict__ cols,
    const T* __restrict__ vals, int64_t n, int drop_zeros, const uint32_t* __restrict__ offsets,
    int64_t* __restrict__ out_coords, T* __restrict__ out_vals, uint32_t* __restrict__ count) {
  __shared__ uint32_t s_cnt[kFinPer][kWaves];
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kFinTile + threadIdx.x;
  T sum[kFinPer];
  uint32_t rr[kFinPer], cc[kFinPer], rank[kFinPer];
  uint32_t keep = 0;
#pragma unroll
  for (int u = 0; u < kFinPer; ++u) {
    const int64_t i = base + int64_t(u) * kBlock;
    sum[u] = T(0);
    rr[u] = cc[u] = 0u;
    bool k = false;
    if (i < n && finalize_head(rows, cols, i, &rr[u], &cc[u]))
      k = finalize_run(rows, cols, vals, n, i, rr[u], cc[u], drop_zeros, &sum[u]);
    const uint64_t m = __ballot(k);
    rank[u] = static_cast<uint32_t>(__popcll(m & below));
    if (lane == 0) s_cnt[u][wave] = static_cast<uint32_t>(__popcll(m));
    if (k) keep |= 1u << u;
  }
  __syncthreads();
  int64_t d = offsets[blockIdx.x];
#pragma unroll
  for (int u = 0; u < kFinPer; ++u) {
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      before += w < wave ? s_cnt[u][w] : 0u;
      all += s_cnt[u][w];
    }
    if (keep & (1u << u)) {
      const int64_t p = d + before + rank[u];
      out_coords[2 * p] = rr[u];
      out_coords[2 * p + 1] = cc[u];
      out_vals[p] = sum[u];
    }
    d += all;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count = static_cast<uint32_t>(d);
}

struct FinalizeLayout {
  size_t sort_a, sort_b, rows_a, vals_a, cols_b, vals_b, tcount, toff, scan, flags, total;
};

FinalizeLayout finalize_layout(int64_t n, int64_t rows, int64_t cols, size_t vbytes) {
  FinalizeLayout L{};
  const size_t un = static_cast<size_t>(n > 0 ? n : 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  L.sort_a = take(sort_layout(n, cols, vbytes).total);
  L.sort_b = take(sort_layout(n, rows, vbytes).total);
  L.rows_a = take(4 * un);
  L.vals_a = take(vbytes * un);
  L.cols_b = take(4 * un);
  L.vals_b = take(vbytes * un);
  const int64_t tiles = (n + kFinTile - 1) / kFinTile;
  const size_t ut = static_cast<size_t>(tiles > 0 ? tiles : 1);
  L.tcount = take(4 * ut);
  L.toff = take(4 * ut);
  L.scan = take(scan_workspace_bytes(tiles));
  L.flags = take(16);
  L.total = off;
  return L;
}

template <typename T>
hipError_t finalize_typed(const int64_t* coords, const T* vals, int64_t rows, int64_t cols,
                          int64_t n, bool sum_duplicates, bool drop_zeros,
                          int64_t* out_coords, T* out_vals, int64_t* out_nnz,
                          int* out_failure, char* ws, hipStream_t s) {
  using VB = T;
  const FinalizeLayout L = finalize_layout(n, rows, cols, sizeof(T));
  uint32_t* flags = reinterpret_cast<uint32_t*>(ws + L.flags);  // [0] bounds, [1] dup, [2] count
  uint32_t host_flags[3] = {0u, 0u, 0u};
  *out_nnz = 0;
  *out_failure = 0;
  hipError_t e = fill_launch(flags, 0, 16, s);
  if (e != hipSuccess || n == 0) return e;
  const unsigned g = grid_stride_blocks(n);
  finalize_bounds_kernel<<<g, kBlock, 0, s>>>(coords, n, rows, cols, flags);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(host_flags, flags, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if (host_flags[0]) {
    *out_failure = 1;  // coordinate out of range
    return hipSuccess;
  }
  // Sort by column, then (stably) by row.  Rows and columns are below 2^32,
  // so the sorts carry them as 32-bit words.
  char* wa = ws + L.sort_a;
  char* wb = ws + L.sort_b;
  uint32_t* rows_a = reinterpret_cast<uint32_t*>(ws + L.rows_a);
  VB* vals_a = reinterpret_cast<VB*>(ws + L.vals_a);
  uint32_t* cols_b = reinterpret_cast<uint32_t*>(ws + L.cols_b);
  VB* vals_b = reinterpret_cast<VB*>(ws + L.vals_b);
  if ((e = stable_sort<int64_t, int64_t, VB>(coords + 1, 2, coords, 2, vals, n, cols, rows_a, 4,
                                             vals_a, nullptr, wa, s)) != hipSuccess)
    return e;
  const uint32_t* cols_a = sorted_keys_in(wa, n, cols, sizeof(VB));
  if ((e = stable_sort<uint32_t, uint32_t, VB>(rows_a, 1, cols_a, 1, vals_a, n, rows, cols_b, 4,
                                               vals_b, nullptr, wb, s)) != hipSuccess)
    return e;
  const uint32_t* rows_b = sorted_keys_in(wb, n, rows, sizeof(VB));
  const unsigned tiles = blocks_for(n, kFinTile);
  uint32_t* tcount = reinterpret_cast<uint32_t*>(ws + L.tcount);
  uint32_t* toff = reinterpret_cast<uint32_t*>(ws + L.toff);
  const int drop = drop_zeros ? 1 : 0;
  finalize_count_kernel<T><<<tiles, kBlock, 0, s>>>(rows_b, cols_b, vals_b, n, drop, tcount,
                                                    flags + 1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (!sum_duplicates) {
    if ((e = hipMemcpyAsync(host_flags + 1, flags + 1, 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    if (host_flags[1]) {
      *out_failure = 2;  // duplicate coordinates under the reject policy
      return hipSuccess;
    }
  }
  if ((e = scan_exclusive(tcount, toff, tiles, reinterpret_cast<uint32_t*>(ws + L.scan), s)) !=
      hipSuccess)
    return e;
  finalize_compact_kernel<T><<<tiles, kBlock, 0, s>>>(rows_b, cols_b, vals_b, n, drop, toff,
                                                      out_coords, out_vals, flags + 2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(host_flags + 2, flags + 2, 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  *out_nnz = host_flags[2];
  return hipSuccess;
}

template <typename VB>
hipError_t transpose_typed(const Compressed& src, int64_t* out_offsets,
                           void* out_indices, int out_index_bytes,
                           void* out_values, char* ws, hipStream_t s) {
  const int64_t n = src.nnz;
  const SortLayout L = sort_layout(n, src.inner, sizeof(VB));
  // Source rows for every entry go in the second row buffer: round 0 reads
  // them, later rounds use both buffers only after round 0 consumed them...
  // so give round 0 its own region at the tail of the workspace.
  uint32_t* rows0 = reinterpret_cast<uint32_t*>(ws + L.total);
  hipError_t err = expand_launch(src.offsets, src.outer, n, rows0, s);
  if (err != hipSuccess) return err;
  const VB* vals = static_cast<const VB*>(src.values);
  if (src.index_bytes == 4) {
    return stable_sort<int32_t, uint32_t, VB>(
        static_cast<const int32_t*>(src.indices), 1, rows0, 1, vals, n,
        src.inner, out_indices, out_index_bytes, static_cast<VB*>(out_values),
        out_offsets, ws, s);
  }
  return stable_sort<int64_t, uint32_t, VB>(
      static_cast<const int64_t*>(src.indices), 1, rows0, 1, vals, n, src.inner,
      out_indices, out_index_bytes, static_cast<VB*>(out_values), out_offsets,
      ws, s);
}

// CSR -> canonical COO: coordinates (row, col) interleaved, values moved as
// raw bits; G lanes walk each row (compressed.h:1781-1837).
template <int G, typename I, typename VB>
__global__ __launch_bounds__(kBlock) void csr_to_coo_kernel(
    const int64_t* __restrict__ offsets, const I* __restrict__ idx,
    const VB* __restrict__ vals, int64_t rows, int64_t* __restrict__ coords,
    VB* __restrict__ out_vals) {
  const int64_t r = xcd_remap(blockIdx.x, gridDim.x) * (kBlock / G) + threadIdx.x / G;
  if (r >= rows) return;
  const int64_t b = offsets[r], e = offsets[r + 1];
  for (int64_t p = b + threadIdx.x % G; p < e; p += G) {
    coords[2 * p] = r;
    coords[2 * p + 1] = static_cast<int64_t>(ld_stream(idx + p));
    out_vals[p] = ld_stream(vals + p);
  }
}

template <typename I, typename VB>
hipError_t csr_to_coo_typed(const int64_t* offsets, const I* idx, const VB* vals,
                            int64_t rows, int64_t nnz, int64_t* coords, VB* out_vals,
                            hipStream_t s) {
  const int64_t avg = rows > 0 ? nnz / rows : 0;
  if (avg >= 32)
    csr_to_coo_kernel<64, I, VB><<<blocks_for(rows, kBlock / 64), kBlock, 0, s>>>(
        offsets, idx, vals, rows, coords, out_vals);
  else if (avg >= 4)
    csr_to_coo_kernel<8, I, VB><<<blocks_for(rows, kBlock / 8), kBlock, 0, s>>>(
        offsets, idx, vals, rows, coords, out_vals);
  else
    csr_to_coo_kernel<1, I, VB><<<blocks_for(rows, kBlock), kBlock, 0, s>>>(
        offsets, idx, vals, rows, coords, out_vals);
  return hipGetLastError();
}

template <typename VB>
hipError_t csr_to_coo_by_index(const Compressed& a, int64_t* coords, void* out_vals,
                               hipStream_t s) {
  if (a.index_bytes == 4)
    return csr_to_coo_typed<int32_t, VB>(a.offsets, static_cast<const int32_t*>(a.indices),
                                         static_cast<const VB*>(a.values), a.outer, a.nnz,
                                         coords, static_cast<VB*>(out_vals), s);
  return csr_to_coo_typed<int64_t, VB>(a.offsets, static_cast<const int64_t*>(a.indices),
                                       static_cast<const VB*>(a.values), a.outer, a.nnz,
                                       coords, static_cast<VB*>(out_vals), s);
}

}  // namespace

size_t scan_u32_scratch_bytes(int64_t n) { return scan_workspace_bytes(n); }

hipError_t scan_u32_exclusive(const uint32_t* in, uint32_t* out, int64_t n, void* scratch,
                              hipStream_t s) {
  return scan_exclusive(in, out, n, static_cast<uint32_t*>(scratch), s);
}

hipError_t csr_to_coo_launch(const Compressed& a, int64_t* coords, void* out_values,
                             hipStream_t s) {
  if (a.outer == 0 || a.nnz == 0) return hipSuccess;
  switch (elem_bytes(a.elem)) {
    case 8: return csr_to_coo_by_index<uint64_t>(a, coords, out_values, s);
    case 4: return csr_to_coo_by_index<uint32_t>(a, coords, out_values, s);
    default: return csr_to_coo_by_index<uint16_t>(a, coords, out_values, s);
  }
}

size_t transpose_workspace_bytes(const Compressed& src) {
  if (transpose2_applies(src)) return transpose2_workspace_bytes(src);
  const SortLayout L = sort_layout(src.nnz, src.inner, elem_bytes(src.elem));
  return L.total + align256(sizeof(uint32_t) * src.nnz);
}

hipError_t transpose_launch(const Compressed& src, int64_t* out_offsets,
                            void* out_indices, int out_index_bytes,
                            void* out_values, void* workspace, hipStream_t s) {
  char* ws = static_cast<char*>(workspace);
  if (src.nnz == 0) {
    // No entries: every output offset is zero.
    return fill_launch(out_offsets, 0, sizeof(int64_t) * (src.inner + 1), s);
  }
  if (transpose2_applies(src))
    return transpose2_launch(src, out_offsets, out_indices, out_index_bytes, out_values, workspace, s);
  switch (elem_bytes(src.elem)) {
    case 8: return transpose_typed<uint64_t>(src, out_offsets, out_indices, out_index_bytes, out_values, ws, s);
    case 4: return transpose_typed<uint32_t>(src, out_offsets, out_indices, out_index_bytes, out_values, ws, s);
    default: return transpose_typed<uint16_t>(src, out_offsets, out_indices, out_index_bytes, out_values, ws, s);
  }
}

size_t coo_convert_workspace_bytes(int64_t rows, int64_t cols, int64_t nnz,
                                   bool to_csc, Elem elem) {
  if (!to_csc) return 0;
  (void)rows;
  return sort_layout(nnz, cols, elem_bytes(elem)).total;
}

hipError_t coo_convert_launch(const int64_t* coords, const void* values,
                              Elem elem, int64_t rows, int64_t cols,
                              int64_t nnz, bool to_csc, int64_t* out_offsets,
                              void* out_indices, int out_index_bytes,
                              void* out_values, void* workspace,
                              hipStream_t s) {
  const int64_t outer = to_csc ? cols : rows;
  if (nnz == 0) return fill_launch(out_offsets, 0, sizeof(int64_t) * (outer + 1), s);
  if (!to_csc) {
    coo_rows_kernel<<<grid_stride_blocks(nnz + 1), kBlock, 0, s>>>(
        coords, static_cast<const uint8_t*>(values),
        static_cast<int>(elem_bytes(elem)), nnz, rows, out_offsets, out_indices,
        out_index_bytes, static_cast<uint8_t*>(out_values));
    return hipGetLastError();
  }
  // CSC: stable sort by column; the row travels as payload.  Canonical COO
  // is row-major, so equal columns keep increasing rows.
  char* ws = static_cast<char*>(workspace);
  switch (elem_bytes(elem)) {
    case 8:
      return stable_sort<int64_t, int64_t, uint64_t>(
          coords + 1, 2, coords, 2, static_cast<const uint64_t*>(values), nnz,
          cols, out_indices, out_index_bytes, static_cast<uint64_t*>(out_values),
          out_offsets, ws, s);
    case 4:
      return stable_sort<int64_t, int64_t, uint32_t>(
          coords + 1, 2, coords, 2, static_cast<const uint32_t*>(values), nnz,
          cols, out_indices, out_index_bytes, static_cast<uint32_t*>(out_values),
          out_offsets, ws, s);
    default:
      return stable_sort<int64_t, int64_t, uint16_t>(
          coords + 1, 2, coords, 2, static_cast<const uint16_t*>(values), nnz,
          cols, out_indices, out_index_bytes, static_cast<uint16_t*>(out_values),
          out_offsets, ws, s);
  }
}

namespace {

__global__ __launch_bounds__(kBlock) void iota_kernel(uint32_t* __restrict__ out, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
    out[i] = static_cast<uint32_t>(i);
}

}  // namespace

size_t sort_by_key_workspace_bytes(int64_t n, int64_t key_bound) {
  return align256(sizeof(uint32_t) * static_cast<size_t>(n > 0 ? n : 1)) * 2 +
         sort_layout(n, key_bound, sizeof(uint32_t)).total;
}

hipError_t sort_by_key_launch(const uint32_t* keys, int64_t n, int64_t key_bound,
                              uint32_t* out_index, int64_t* out_key_offsets, void* workspace,
                              hipStream_t s) {
  if (n == 0) return fill_launch(out_key_offsets, 0, sizeof(int64_t) * (key_bound + 1), s);
  char* ws = static_cast<char*>(workspace);
  uint32_t* iota = reinterpret_cast<uint32_t*>(ws);
  uint32_t* dummy = reinterpret_cast<uint32_t*>(ws + align256(sizeof(uint32_t) * n));
  char* sws = ws + 2 * align256(sizeof(uint32_t) * n);
  iota_kernel<<<grid_stride_blocks(n), kBlock, 0, s>>>(iota, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return stable_sort<uint32_t, uint32_t, uint32_t>(keys, 1, iota, 1, iota, n, key_bound, out_index,
                                                   4, dummy, out_key_offsets, sws, s);
}

size_t coo_finalize_workspace_bytes(int64_t rows, int64_t cols, int64_t nnz, Elem elem) {
  return finalize_layout(nnz, rows, cols, elem_bytes(elem)).total;
}

hipError_t coo_finalize_launch(const int64_t* coords, const void* values, Elem elem,
                               int64_t rows, int64_t cols, int64_t nnz, bool sum_duplicates,
                               bool drop_zeros, int64_t* out_coords, void* out_values,
                               int64_t* out_nnz, int* out_failure, void* workspace,
                               hipStream_t s) {
  char* ws = static_cast<char*>(workspace);
  if (elem == Elem::F64)
    return finalize_typed<double>(coords, static_cast<const double*>(values), rows, cols, nnz,
                                  sum_duplicates, drop_zeros, out_coords,
                                  static_cast<double*>(out_values), out_nnz, out_failure, ws, s);
  return finalize_typed<float>(coords, static_cast<const float*>(values), rows, cols, nnz,
                               sum_duplicates, drop_zeros, out_coords,
                               static_cast<float*>(out_values), out_nnz, out_failure, ws, s);
}

}  // namespace asc_hip
