// Two-pass stable transpose of a compressed operand (CSR <-> CSC) for gfx950.
//
// The output of the reference's compressed conversion (compressed.h:1151-1221)
// is the unique canonical order: entries sorted by the new outer index (the
// old inner index, "column" below), ties kept in source order, i.e. by the old
// outer index ("row").  That is a stable sort by column.  Here it is an LSD
// radix sort of the column split into two digits, lo and hi bits (at most
// 11 each), instead of three 8-bit rounds:
//
//   pass 1  the CSR arrays -> an intermediate sorted by the low digit.  The
//           row of each entry is derived in the tile from the offsets (no
//           expanded row array), and the high digit and the row are packed
//           into one 32-bit word.  When hi + row bits exceed 32 by s bits,
//           pass-1 tiles are cut at the row-segment boundaries
//           offsets[k << (row bits - s)], so inside each low-digit run the
//           segments follow one another and their split points recover the
//           dropped row bits in pass 2.
//   pass 2  intermediate -> output sorted by the high digit.  Its tiles are
//           cut at low-digit run boundaries, so every tile holds one low
//           digit; the scanned (digit, tile) counts then give each column's
//           first output position directly (no boundary pass over the keys).
//
// A tile is 32768 entries, one 1024-thread work-group.  Inside a tile each
// wave ranks its 2048 entries by digit with ballots (stable: lane, then
// chunk, then wave order is source order); the tile's entries are then
// written in digit order by consecutive threads through LDS, so each digit's
// run of the tile (16 entries on average at 2048 digits) leaves as
// contiguous stores.  Counts, positions and offsets are uint32 (nnz < 2^32).
#include "device_common.h"
#include "launch.h"

namespace asc_hip {
namespace {

constexpr int kTrBlock = 1024;
constexpr int kTrIpt = 32;
constexpr int kTrTile = kTrBlock * kTrIpt;    // entries per tile
constexpr int kTrWaves = kTrBlock / kWave;     // 8
constexpr int kTrWaveItems = kWave * kTrIpt;   // 2048 entries per wave
constexpr int kTrTableBlock = 1024;            // second-pass tile table: two digits a thread
constexpr int kTrDpt = (1 << 11) / kTrBlock;   // digits a thread owns in the digit scan
static_assert(kTrDpt >= 2 && kTrDpt % 2 == 0, "whole 32-bit words of 16-bit counts");
constexpr int kTrMaxDigitBits = 11;
constexpr int kTrMaxDigits = 1 << kTrMaxDigitBits;
constexpr int kTrMinKeyBits = 9;
constexpr int kTrMaxKeyBits = 2 * kTrMaxDigitBits;
constexpr int kTrMaxSegBits = 4;
constexpr int kTrMaxSegs = 1 << kTrMaxSegBits;
constexpr int kTrSmall = 256;
constexpr int kTrEntryBits = __builtin_ctz(kTrTile);  // entry index inside a tile
static_assert(kTrTile == 1 << kTrEntryBits, "entry index field");
static_assert(kTrEntryBits + kTrMaxDigitBits <= 32, "slot word holds digit and entry");

inline size_t tr_align(size_t b) { return (b + 255) & ~size_t(255); }

int tr_key_bits(int64_t bound) {
  int bits = 0;
  while (bits < 63 && (int64_t(1) << bits) < bound) ++bits;
  return bits;
}

struct TrShape {
  int lo = 0, hi = 0;  // digit bits (pass 1, pass 2)
  int rlow = 0;        // row bits kept in the packed word
  int segs = 1;        // row segments (2^s)
  int64_t s1 = 0, s2 = 0;  // tile slots of each pass (upper bounds)
};

bool tr_shape(const Compressed& src, TrShape* sh) {
  const int64_t n = src.nnz;
  if (n <= 0 || n > int64_t(UINT32_MAX) || src.outer > int64_t(UINT32_MAX)) return false;
  const int kb = tr_key_bits(src.inner);
  if (kb < kTrMinKeyBits || kb > kTrMaxKeyBits) return false;
  sh->lo = (kb + 1) / 2;
  sh->hi = kb - sh->lo;
  const int rb = tr_key_bits(src.outer);
  const int s = rb + sh->hi > 32 ? rb + sh->hi - 32 : 0;
  if (s > kTrMaxSegBits) return false;
  sh->rlow = rb - s;
  sh->segs = 1 << s;
  const int64_t full = (n + kTrTile - 1) / kTrTile;
  sh->s1 = full + sh->segs;
  sh->s2 = full + (int64_t(1) << sh->lo);
  return true;
}

struct TrLayout {
  size_t packed, vals, cnt1, cnt2, tiles1, tiles2, seg, tstart, split, scan, total;
};

TrLayout tr_layout(const Compressed& src, const TrShape& sh) {
  TrLayout L{};
  const size_t n = static_cast<size_t>(src.nnz);
  const size_t d1 = size_t(1) << sh.lo, d2 = size_t(1) << sh.hi;
  const size_t c1 = d1 * static_cast<size_t>(sh.s1) + 1, c2 = d2 * static_cast<size_t>(sh.s2) + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += tr_align(bytes);
    return o;
  };
  L.packed = take(4 * n);
  L.vals = take(elem_bytes(src.elem) * n);
  L.cnt1 = take(4 * c1);
  L.cnt2 = take(4 * c2);
  L.tiles1 = take(16 * static_cast<size_t>(sh.s1));
  L.tiles2 = take(16 * static_cast<size_t>(sh.s2));
  L.seg = take(4 * (2 * (kTrMaxSegs + 1)));
  L.tstart = take(4 * (d1 + 1));
  L.split = take(4 * d1 * kTrMaxSegs);
  L.scan = take(scan_u32_scratch_bytes(static_cast<int64_t>(c1 > c2 ? c1 : c2)));
  L.total = off;
  return L;
}

// ---- tile tables ------------------------------------------------------------
// seg[k] (k = 0..segs): first CSR position of row segment k (seg[segs] = n);
// seg[kTrMaxSegs + 1 + k]: first pass-1 tile of segment k (the last: tile count).
__global__ __launch_bounds__(64) void tr_seg_kernel(const int64_t* __restrict__ offsets,
                                                    int64_t outer, int segs, int rlow,
                                                    uint32_t* __restrict__ seg) {
  __shared__ uint32_t s_pos[kTrMaxSegs + 1];
  const int k = threadIdx.x;
  if (k <= segs) {
    const int64_t r = (int64_t(k) << rlow) < outer ? (int64_t(k) << rlow) : outer;
    s_pos[k] = static_cast<uint32_t>(offsets[r]);
    seg[k] = s_pos[k];
  }
  __syncthreads();
  if (k == 0) {
    uint32_t t = 0;
    for (int j = 0; j < segs; ++j) {
      seg[kTrMaxSegs + 1 + j] = t;
      t += static_cast<uint32_t>((uint64_t(s_pos[j + 1]) - s_pos[j] + kTrTile - 1) / kTrTile);
    }
    seg[kTrMaxSegs + 1 + segs] = t;
  }
}

// tiles1[t] = {start, end, r0 (row holding start), segment}; unused slots empty.
__global__ __launch_bounds__(kTrSmall) void tr_tiles1_kernel(
    const int64_t* __restrict__ offsets, int64_t outer, int64_t n,
    const uint32_t* __restrict__ seg, int segs, uint32_t* __restrict__ tiles, int64_t slots) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * kTrSmall + threadIdx.x;
  if (t >= slots) return;
  const uint32_t* first = seg + kTrMaxSegs + 1;
  uint32_t* out = tiles + 4 * t;
  if (t >= first[segs]) {
    out[0] = static_cast<uint32_t>(n);
    out[1] = static_cast<uint32_t>(n);
    out[2] = static_cast<uint32_t>(outer - 1);
    out[3] = 0;
    return;
  }
  int k = 0;
  while (k + 1 < segs && first[k + 1] <= t) ++k;
  const int64_t start = static_cast<int64_t>(seg[k]) + (t - static_cast<int64_t>(first[k])) * kTrTile;
  const int64_t end = start + kTrTile < static_cast<int64_t>(seg[k + 1]) ? start + kTrTile
                                                                          : static_cast<int64_t>(seg[k + 1]);
  // r0: the last row whose first entry is at or before start.
  int64_t lo = 0, hi = outer - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) / 2;
    if (offsets[mid] <= start) lo = mid;
    else hi = mid - 1;
  }
  out[0] = static_cast<uint32_t>(start);
  out[1] = static_cast<uint32_t>(end);
  out[2] = static_cast<uint32_t>(lo);
  out[3] = static_cast<uint32_t>(k);
}

// Block-wide exclusive scan of one uint32 per thread (1024 threads).
__device__ __forceinline__ uint32_t tr_block_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == kWave - 1) s_tmp[wave] = x;
  __syncthreads();
  uint32_t base = 0, all = 0;
  for (int w = 0; w < static_cast<int>(blockDim.x / kWave); ++w) {
    if (w < wave) base += s_tmp[w];
    all += s_tmp[w];
  }
  __syncthreads();
  *total = all;
  return base + x - v;
}

// Pass-2 tiles per low digit l: its run [L_l, L_{l+1}) of the intermediate
// (L_l = scanned pass-1 count of (l, tile 0)) in ceil(len / tile) tiles;
// tstart[l] = first tile of l (tstart[d1] = tile count).  split[l][k] = first
// intermediate position of row segment k inside l's run.
__global__ __launch_bounds__(kTrTableBlock) void tr_tiles2a_kernel(
    const uint32_t* __restrict__ cnt1, int64_t s1, int d1, int64_t n,
    const uint32_t* __restrict__ seg, int segs, uint32_t* __restrict__ tstart,
    uint32_t* __restrict__ split) {
  __shared__ uint32_t s_tmp[kTrTableBlock / kWave];
  const uint32_t* first = seg + kTrMaxSegs + 1;
  uint32_t cnt[2] = {0u, 0u};
  for (int j = 0; j < 2; ++j) {
    const int l = 2 * threadIdx.x + j;
    if (l < d1) {
      const uint32_t a = cnt1[static_cast<int64_t>(l) * s1];
      const uint32_t b = l + 1 < d1 ? cnt1[static_cast<int64_t>(l + 1) * s1] : static_cast<uint32_t>(n);
      cnt[j] = static_cast<uint32_t>((uint64_t(b) - a + kTrTile - 1) / kTrTile);
      for (int k = 0; k < segs; ++k)
        split[static_cast<int64_t>(l) * kTrMaxSegs + k] = cnt1[static_cast<int64_t>(l) * s1 + first[k]];
    }
  }
  uint32_t total;
  const uint32_t ex = tr_block_scan(cnt[0] + cnt[1], s_tmp, &total);
  if (2 * static_cast<int>(threadIdx.x) < d1) {
    tstart[2 * threadIdx.x] = ex;
    tstart[2 * threadIdx.x + 1] = ex + cnt[0];
  }
  if (threadIdx.x == 0) tstart[d1] = total;
}

// tiles2[t] = {start, end, l}; slots past the tile count are empty.
__global__ __launch_bounds__(kTrSmall) void tr_tiles2b_kernel(
    const uint32_t* __restrict__ cnt1, int64_t s1, int d1, int64_t n,
    const uint32_t* __restrict__ tstart, uint32_t* __restrict__ tiles, int64_t slots) {
  const int64_t id = static_cast<int64_t>(blockIdx.x) * kTrSmall + threadIdx.x;
  const int64_t nthreads = static_cast<int64_t>(gridDim.x) * kTrSmall;
  if (id < d1) {
    const int l = static_cast<int>(id);
    const uint32_t a = cnt1[static_cast<int64_t>(l) * s1];
    const uint32_t b = l + 1 < d1 ? cnt1[static_cast<int64_t>(l + 1) * s1] : static_cast<uint32_t>(n);
    int64_t t = tstart[l];
    for (int64_t p = a; p < static_cast<int64_t>(b); p += kTrTile, ++t) {
      uint32_t* out = tiles + 4 * t;
      out[0] = static_cast<uint32_t>(p);
      out[1] = static_cast<uint32_t>(p + kTrTile < static_cast<int64_t>(b) ? p + kTrTile : b);
      out[2] = static_cast<uint32_t>(l);
      out[3] = 0;
    }
  }
  for (int64_t t = tstart[d1] + id; t < slots; t += nthreads) {
    uint32_t* out = tiles + 4 * t;
    out[0] = static_cast<uint32_t>(n);
    out[1] = static_cast<uint32_t>(n);
    out[2] = 0;
    out[3] = 0;
  }
}

// ---- histograms ---------------------------------------------------------------
// cnt[d * slots + t] = entries of tile t whose digit is d, digit = key bits
// [shift, shift + bits).  Block 0 also clears the trailing element the scan
// turns into the total.
constexpr int kTrHistBlock = 1024;
template <typename K>
__global__ __launch_bounds__(kTrHistBlock) void tr_hist_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ tiles, int shift, int bits,
    uint32_t* __restrict__ cnt, int64_t slots) {
  __shared__ uint32_t s_hist[kTrMaxDigits];
  const int digits = 1 << bits;
  for (int d = threadIdx.x; d < digits; d += kTrHistBlock) s_hist[d] = 0u;
  __syncthreads();
  // Consecutive tiles on one XCD: the counts of 32 adjacent tiles share a
  // line of each digit's row, which then fills in that XCD's L2.
  const int64_t t = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t start = tiles[4 * t], end = tiles[4 * t + 1];
  const uint32_t mask = static_cast<uint32_t>(digits - 1);
  constexpr int kU = 16;
  uint32_t i = start + threadIdx.x;
  for (; i + (kU - 1) * kTrHistBlock < end; i += kU * kTrHistBlock) {
    uint32_t k[kU];
#pragma unroll
    for (int j = 0; j < kU; ++j) k[j] = static_cast<uint32_t>(ld_stream(keys + i + j * kTrHistBlock));
#pragma unroll
    for (int j = 0; j < kU; ++j) atomicAdd(&s_hist[(k[j] >> shift) & mask], 1u);
  }
  for (; i < end; i += kTrHistBlock) atomicAdd(&s_hist[(static_cast<uint32_t>(ld_stream(keys + i)) >> shift) & mask], 1u);
  __syncthreads();
  for (int d = threadIdx.x; d < digits; d += kTrHistBlock) cnt[static_cast<int64_t>(d) * slots + t] = s_hist[d];
  if (t == 0 && threadIdx.x == 0) cnt[static_cast<int64_t>(digits) * slots] = 0u;
}

// ---- the scatter passes -------------------------------------------------------
struct TrShared {
  union {
    uint32_t rowmap[kTrTile];                // pass 1: row of each entry, relative to r0
    uint16_t cnt[kTrWaves][kTrMaxDigits];    // per-wave digit counts, then earlier waves' counts
    uint32_t inv[kTrTile];                   // sorted slot -> digit << 15 | entry
    uint32_t stage32[kTrTile];               // one 32-bit word per entry, entry order
    uint64_t stage64[kTrTile / 2];           // 8-byte values of half a tile, entry order
  } u;
  uint16_t local[kTrMaxDigits];   // tile-local first slot of each digit
  uint32_t global[kTrMaxDigits];  // output position of that slot
  uint32_t tmp[kTrWaves];
  uint32_t split[kTrMaxSegs];
};

// Entry e of a tile is held by wave e / 2048, chunk (e % 2048) / 64, lane e % 64.
__device__ __forceinline__ uint32_t tr_entry(int wave, int u, int lane) {
  return static_cast<uint32_t>(wave * kTrWaveItems + u * kWave + lane);
}

// The scanned counts of this tile's digits owned by thread t, loaded
// early so their latency overlaps the tile's first phases.
__device__ __forceinline__ void tr_load_counts(uint32_t (&gcnt)[kTrDpt], const uint32_t* __restrict__ scanned,
                                               int64_t slots, int64_t tile, int bits) {
#pragma unroll
  for (int j = 0; j < kTrDpt; ++j) {
    const int d = kTrDpt * static_cast<int>(threadIdx.x) + j;
    gcnt[j] = d < (1 << bits) ? scanned[static_cast<int64_t>(d) * slots + tile] : 0u;
  }
}

// Stable ranking of the tile's entries by digit.  dig[u] holds the digit of
// entry (wave, u, lane) on entry and rank << kTrMaxDigitBits | digit inside
// the tr_rank call.  On return sh.local / sh.global hold each digit's first
// slot and that slot's output position (gcnt: the scanned counts from
// tr_load_counts), and sh.u.inv maps every sorted slot to its digit << 14 |
// entry.
__device__ __forceinline__ void tr_rank(TrShared& sh, uint32_t (&dig)[kTrIpt], uint32_t len,
                                        int bits, const uint32_t (&gcnt)[kTrDpt]) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int digits = 1 << bits;
  constexpr uint32_t kDigitMask = kTrMaxDigits - 1u;
  for (int i = threadIdx.x; i < kTrWaves * kTrMaxDigits / 2; i += kTrBlock)
    reinterpret_cast<uint32_t*>(&sh.u.cnt[0][0])[i] = 0u;
  __syncthreads();
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const bool valid = tr_entry(wave, u, lane) < len;
    uint32_t d = dig[u];
    // Keeps the chunks' rankings in program order: hoisting the later
    // chunks' ballots above the earlier chunks' count updates would hold
    // every chunk's bit masks live at once.
    asm volatile("" : "+v"(d));
    // The lanes holding the same digit: for each digit bit, the ballot of
    // that bit XOR the lane's own bit (spread over the word) is zero exactly
    // at the lanes that agree.  Always kTrMaxDigitBits rounds: a digit's bits
    // at and above `bits` are zero in every lane and clear nothing.
    const uint64_t vmask = __ballot(valid);
    uint32_t p_lo = static_cast<uint32_t>(vmask), p_hi = static_cast<uint32_t>(vmask >> 32);
#pragma unroll
    for (int b = 0; b < kTrMaxDigitBits; ++b) {
      const uint32_t own = static_cast<uint32_t>(static_cast<int32_t>(d << (31 - b)) >> 31);
      const uint64_t bal = __ballot(own != 0u);
      p_lo &= ~(own ^ static_cast<uint32_t>(bal));
      p_hi &= ~(own ^ static_cast<uint32_t>(bal >> 32));
    }
    const uint64_t peers = (static_cast<uint64_t>(p_hi) << 32) | p_lo;
    const uint32_t lower = __popcll(peers & lt);
    const uint32_t total = __popcll(peers);
    uint32_t r = 0;
    if (valid) r = sh.u.cnt[wave][d] + lower;
    dig[u] = (r << kTrMaxDigitBits) | d;
    // The last peer records the new count after every peer read the old one
    // (one wave's LDS operations take effect in issue order).
    if (valid && lower == total - 1) sh.u.cnt[wave][d] = static_cast<uint16_t>(r + 1);
  }
  __syncthreads();
  // Thread t owns digits kTrDpt*t .. kTrDpt*t + kTrDpt-1 (kTrDpt/2 32-bit
  // words of each wave's row): per digit the counts of the earlier waves (in
  // place), then the digit's first slot by a block scan of the digit totals.
  const int t = threadIdx.x;
  uint32_t tot[kTrDpt];
#pragma unroll
  for (int j = 0; j < kTrDpt; ++j) tot[j] = 0u;
  if (kTrDpt * t < digits) {
#pragma unroll
    for (int half = 0; half < kTrDpt / 2; ++half) {
      uint32_t run0 = 0, run1 = 0;
#pragma unroll
      for (int w = 0; w < kTrWaves; ++w) {
        uint32_t* word = reinterpret_cast<uint32_t*>(&sh.u.cnt[w][kTrDpt * t + 2 * half]);
        const uint32_t c = *word;
        *word = run0 | (run1 << 16);
        run0 += c & 0xFFFFu;
        run1 += c >> 16;
      }
      tot[2 * half] = run0;
      tot[2 * half + 1] = run1;
    }
  }
  uint32_t sum = 0;
#pragma unroll
  for (int j = 0; j < kTrDpt; ++j) sum += tot[j];
  uint32_t all;
  const uint32_t ex = tr_block_scan(sum, sh.tmp, &all);
  if (kTrDpt * t < digits) {
    uint32_t run = ex;
#pragma unroll
    for (int j = 0; j < kTrDpt; ++j) {
      sh.local[kTrDpt * t + j] = static_cast<uint16_t>(run);
      run += tot[j];
      sh.global[kTrDpt * t + j] = gcnt[j];
    }
  }
  __syncthreads();
  // Sorted slot = digit's first slot + earlier waves' count + rank in wave.
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t d = dig[u] & kDigitMask;
    if (tr_entry(wave, u, lane) < len)
      dig[u] = (((dig[u] >> kTrMaxDigitBits) + sh.local[d] + sh.u.cnt[wave][d]) << kTrMaxDigitBits) | d;
  }
  __syncthreads();  // every count read before the slot map overwrites them
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    if (e < len)
      sh.u.inv[dig[u] >> kTrMaxDigitBits] = ((dig[u] & kDigitMask) << kTrEntryBits) | e;
  }
  __syncthreads();
}

// The slot side: thread j writes sorted slots j + 1024*v; sw[v] is the
// slot's digit << 15 | entry (0 past the tile's end).
__device__ __forceinline__ void tr_slots(TrShared& sh, uint32_t len, uint32_t (&sw)[kTrIpt]) {
#pragma unroll
  for (int v = 0; v < kTrIpt; ++v) {
    const uint32_t s = static_cast<uint32_t>(v * kTrBlock + threadIdx.x);
    sw[v] = s < len ? sh.u.inv[s] : 0u;
  }
  __syncthreads();  // the slot map is read before staging reuses its space
}

// Output position of sorted slot s holding slot word w.
__device__ __forceinline__ uint32_t tr_dst(const TrShared& sh, uint32_t s, uint32_t w) {
  const uint32_t d = w >> kTrEntryBits;
  return sh.global[d] + (s - sh.local[d]);
}

// One 32-bit word per entry (held by the entry side) to out[dst] in slot order.
template <typename O>
__device__ __forceinline__ void tr_write_words(TrShared& sh, uint32_t len,
                                               const uint32_t (&word)[kTrIpt],
                                               const uint32_t (&sw)[kTrIpt], O* __restrict__ out) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    if (e < len) sh.u.stage32[e] = word[u];
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < kTrIpt; ++v) {
    const uint32_t s = static_cast<uint32_t>(v * kTrBlock + threadIdx.x);
    if (s < len)
      out[tr_dst(sh, s, sw[v])] = static_cast<O>(sh.u.stage32[sw[v] & (kTrTile - 1u)]);
  }
  __syncthreads();
}

// The tile's values go through LDS in entry order, in kTrValRounds rounds of
// kTrValPer values a thread (8-byte values: two halves of the tile).
template <typename VB> struct TrVals {
  static constexpr int kRounds = sizeof(VB) == 8 ? 2 : 1;
  static constexpr int kPer = kTrTile / kRounds / kTrBlock;
  VB v[kPer];
};

template <typename VB>
__device__ __forceinline__ void tr_load_values(TrVals<VB>& r, const VB* __restrict__ in, uint32_t len,
                                               int round) {
  const uint32_t base = static_cast<uint32_t>(round * (kTrTile / TrVals<VB>::kRounds));
#pragma unroll
  for (int k = 0; k < TrVals<VB>::kPer; ++k) {
    const uint32_t e = base + static_cast<uint32_t>(k * kTrBlock + threadIdx.x);
    r.v[k] = e < len ? ld_stream(in + e) : VB{};
  }
}

template <typename VB>
__device__ __forceinline__ void tr_store_values(TrShared& sh, const TrVals<VB>& r, uint32_t len, int round,
                                                const uint32_t (&sw)[kTrIpt], VB* __restrict__ out) {
  constexpr uint32_t kSpan = kTrTile / TrVals<VB>::kRounds;
  const uint32_t base = static_cast<uint32_t>(round) * kSpan;
#pragma unroll
  for (int k = 0; k < TrVals<VB>::kPer; ++k) {
    const uint32_t e = base + static_cast<uint32_t>(k * kTrBlock + threadIdx.x);
    if (e < len) {
      if constexpr (sizeof(VB) == 8) sh.u.stage64[e - base] = static_cast<uint64_t>(r.v[k]);
      else sh.u.stage32[e - base] = static_cast<uint32_t>(r.v[k]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kTrIpt; ++q) {
    const uint32_t s = static_cast<uint32_t>(q * kTrBlock + threadIdx.x);
    const uint32_t e = sw[q] & (kTrTile - 1u);
    if (s < len && e - base < kSpan) {
      if constexpr (sizeof(VB) == 8) out[tr_dst(sh, s, sw[q])] = static_cast<VB>(sh.u.stage64[e - base]);
      else out[tr_dst(sh, s, sw[q])] = static_cast<VB>(sh.u.stage32[e - base]);
    }
  }
  __syncthreads();
}

// The rest of a tile once ranked: the slot map, the 32-bit words, then the
// values.  Each value round's loads go out one step ahead (round 0 during
// the words, round 1 during round 0), so their latency overlaps LDS work.
template <typename VB, typename O>
__device__ __forceinline__ void tr_finish(TrShared& sh, uint32_t len, const uint32_t (&word)[kTrIpt],
                                          const VB* __restrict__ in, O* __restrict__ wout,
                                          VB* __restrict__ vout) {
  uint32_t sw[kTrIpt];
  tr_slots(sh, len, sw);
  TrVals<VB> r0;
  tr_load_values<VB>(r0, in, len, 0);
  tr_write_words<O>(sh, len, word, sw, wout);
  if constexpr (TrVals<VB>::kRounds == 2) {
    TrVals<VB> r1;
    tr_load_values<VB>(r1, in, len, 1);
    tr_store_values<VB>(sh, r0, len, 0, sw, vout);
    tr_store_values<VB>(sh, r1, len, 1, sw, vout);
  } else {
    tr_store_values<VB>(sh, r0, len, 0, sw, vout);
  }
}

// Inclusive max-scan over the wave (DPP row shifts and broadcasts; VALU only).
__device__ __forceinline__ uint32_t tr_wave_max_scan(uint32_t v) {
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xF, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xF, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xF, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xF, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xA, 0xF, false)));
  v = max(v, static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xC, 0xF, false)));
  return v;
}

// Pass 1: CSR tile -> intermediate (packed high digit | row low bits, value),
// positions by the low digit.
template <typename I, typename VB>
__global__ __launch_bounds__(kTrBlock, 4) void tr_pass1_kernel(
    const int64_t* __restrict__ offsets, int64_t outer, const I* __restrict__ cols,
    const VB* __restrict__ vals, const uint32_t* __restrict__ tiles, int64_t slots,
    const uint32_t* __restrict__ scanned, int lo, int rlow, uint32_t* __restrict__ packed_out,
    VB* __restrict__ vals_out) {
  __shared__ TrShared sh;
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t start = tiles[4 * tile], end = tiles[4 * tile + 1];
  const uint32_t len = end - start;
  if (len == 0) return;
  const uint32_t r0 = tiles[4 * tile + 2];
  const int64_t r_hi = tile + 1 < slots ? static_cast<int64_t>(tiles[4 * tile + 6]) : outer - 1;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  // Every load the tile's first phases need goes out at once: its columns,
  // the first row starts it marks, and its scanned digit counts.
  uint32_t col[kTrIpt];
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    col[u] = e < len ? static_cast<uint32_t>(ld_stream(cols + start + e)) : 0u;
  }
  const int64_t rfirst = static_cast<int64_t>(r0) + 1 + threadIdx.x;
  const int64_t o_first = rfirst <= r_hi ? offsets[rfirst] : 0;
  const int64_t o1_first = rfirst <= r_hi ? offsets[rfirst + 1] : 0;
  uint32_t gcnt[kTrDpt];
  tr_load_counts(gcnt, scanned, slots, tile, lo);
  // Rows: mark each row starting inside the tile at its first entry (relative
  // to r0, whose start is at or before the tile's), then a running max over
  // the entries.
  for (uint32_t i = threadIdx.x; i < len; i += kTrBlock) sh.u.rowmap[i] = 0u;
  __syncthreads();
  for (int64_t r = rfirst; r <= r_hi; r += kTrBlock) {
    const int64_t o = r == rfirst ? o_first : offsets[r];
    const int64_t o1 = r == rfirst ? o1_first : offsets[r + 1];
    if (o > static_cast<int64_t>(start) && o < static_cast<int64_t>(end) && o1 > o)
      sh.u.rowmap[o - start] = static_cast<uint32_t>(r - r0);
  }
  __syncthreads();
  uint32_t m = 0;
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    if (e < len) m = max(m, sh.u.rowmap[e]);
  }
  m = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(tr_wave_max_scan(m)), kWave - 1));
  if (lane == 0) sh.tmp[wave] = m;
  __syncthreads();
  uint32_t carry = 0;
  for (int w = 0; w < wave; ++w) carry = max(carry, sh.tmp[w]);
  const uint32_t rmask = rlow >= 32 ? ~0u : (1u << rlow) - 1u;
  uint32_t word[kTrIpt];
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    uint32_t v = e < len ? sh.u.rowmap[e] : 0u;
    v = tr_wave_max_scan(max(v, carry));
    carry = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), kWave - 1));
    const uint32_t row = r0 + v;
    word[u] = ((col[u] >> lo) << rlow) | (row & rmask);
    col[u] &= (1u << lo) - 1u;  // the pass-1 digit
  }
  __syncthreads();  // row map read before the counts reuse its space
  tr_rank(sh, col, len, lo, gcnt);
  tr_finish<VB, uint32_t>(sh, len, word, vals + start, packed_out, vals_out);
}

// Pass 2: intermediate tile (one low digit) -> output rows and values,
// positions by the high digit.  The row's dropped top bits are its segment:
// the number of segment starts of this low digit's run at or before the
// entry.
template <typename VB, typename RO>
__global__ __launch_bounds__(kTrBlock, 4) void tr_pass2_kernel(
    const uint32_t* __restrict__ packed, const VB* __restrict__ ivals,
    const uint32_t* __restrict__ tiles, int64_t slots, const uint32_t* __restrict__ scanned,
    const uint32_t* __restrict__ split, int segs, int hi, int rlow, RO* __restrict__ rows_out,
    VB* __restrict__ vals_out) {
  __shared__ TrShared sh;
  const int64_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t start = tiles[4 * tile], end = tiles[4 * tile + 1];
  const uint32_t len = end - start;
  if (len == 0) return;
  const uint32_t l = tiles[4 * tile + 2];
  // Segment starts past the last segment never compare below a position.
  if (threadIdx.x < kTrMaxSegs)
    sh.split[threadIdx.x] = static_cast<int>(threadIdx.x) < segs
                                ? split[static_cast<int64_t>(l) * kTrMaxSegs + threadIdx.x]
                                : UINT32_MAX;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  uint32_t word[kTrIpt], dig[kTrIpt];
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t e = tr_entry(wave, u, lane);
    word[u] = e < len ? ld_stream(packed + start + e) : 0u;
  }
  uint32_t gcnt[kTrDpt];
  tr_load_counts(gcnt, scanned, slots, tile, hi);
  __syncthreads();  // split points visible
  const uint32_t rmask = rlow >= 32 ? ~0u : (1u << rlow) - 1u;
#pragma unroll
  for (int u = 0; u < kTrIpt; ++u) {
    const uint32_t q = start + tr_entry(wave, u, lane);
    uint32_t sgm = 0;
#pragma unroll
    for (int k = 1; k < kTrMaxSegs; ++k) sgm += q >= sh.split[k] ? 1u : 0u;
    dig[u] = word[u] >> rlow;
    word[u] = (sgm << rlow) | (word[u] & rmask);
  }
  tr_rank(sh, dig, len, hi, gcnt);
  tr_finish<VB, RO>(sh, len, word, ivals + start, rows_out, vals_out);
}

// offsets[c] = first output position of column c: with h, l its digits, the
// scanned pass-2 count of (h, first tile of l); offsets[inner] = nnz.
__global__ __launch_bounds__(kTrSmall) void tr_offsets_kernel(
    const uint32_t* __restrict__ scanned, int64_t slots, const uint32_t* __restrict__ tstart,
    int lo, int64_t inner, int64_t n, int64_t* __restrict__ offsets) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kTrSmall;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * kTrSmall + threadIdx.x; c <= inner; c += stride) {
    if (c == inner) {
      offsets[c] = n;
    } else {
      const int64_t h = c >> lo, l = c & ((int64_t(1) << lo) - 1);
      offsets[c] = scanned[h * slots + tstart[l]];
    }
  }
}

template <typename VB>
hipError_t tr_run(const Compressed& src, const TrShape& sh, int64_t* out_offsets,
                  void* out_rows, int out_row_bytes, VB* out_vals, char* ws, hipStream_t s) {
  const TrLayout L = tr_layout(src, sh);
  const int64_t n = src.nnz;
  const int d1 = 1 << sh.lo, d2 = 1 << sh.hi;
  uint32_t* packed = reinterpret_cast<uint32_t*>(ws + L.packed);
  VB* ivals = reinterpret_cast<VB*>(ws + L.vals);
  uint32_t* cnt1 = reinterpret_cast<uint32_t*>(ws + L.cnt1);
  uint32_t* cnt2 = reinterpret_cast<uint32_t*>(ws + L.cnt2);
  uint32_t* tiles1 = reinterpret_cast<uint32_t*>(ws + L.tiles1);
  uint32_t* tiles2 = reinterpret_cast<uint32_t*>(ws + L.tiles2);
  uint32_t* seg = reinterpret_cast<uint32_t*>(ws + L.seg);
  uint32_t* tstart = reinterpret_cast<uint32_t*>(ws + L.tstart);
  uint32_t* split = reinterpret_cast<uint32_t*>(ws + L.split);
  void* scan = ws + L.scan;
  const unsigned g1 = static_cast<unsigned>(sh.s1), g2 = static_cast<unsigned>(sh.s2);
  const VB* vals = static_cast<const VB*>(src.values);
  hipError_t e;

  tr_seg_kernel<<<1, 64, 0, s>>>(src.offsets, src.outer, sh.segs, sh.rlow, seg);
  tr_tiles1_kernel<<<static_cast<unsigned>((sh.s1 + kTrSmall - 1) / kTrSmall), kTrSmall, 0, s>>>(
      src.offsets, src.outer, n, seg, sh.segs, tiles1, sh.s1);
  if (src.index_bytes == 4)
    tr_hist_kernel<int32_t><<<g1, kTrHistBlock, 0, s>>>(static_cast<const int32_t*>(src.indices), tiles1,
                                                    0, sh.lo, cnt1, sh.s1);
  else
    tr_hist_kernel<int64_t><<<g1, kTrHistBlock, 0, s>>>(static_cast<const int64_t*>(src.indices), tiles1,
                                                    0, sh.lo, cnt1, sh.s1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t c1 = int64_t(d1) * sh.s1 + 1, c2 = int64_t(d2) * sh.s2 + 1;
  if ((e = scan_u32_exclusive(cnt1, cnt1, c1, scan, s)) != hipSuccess) return e;
  if (src.index_bytes == 4)
    tr_pass1_kernel<int32_t, VB><<<g1, kTrBlock, 0, s>>>(
        src.offsets, src.outer, static_cast<const int32_t*>(src.indices), vals, tiles1, sh.s1, cnt1,
        sh.lo, sh.rlow, packed, ivals);
  else
    tr_pass1_kernel<int64_t, VB><<<g1, kTrBlock, 0, s>>>(
        src.offsets, src.outer, static_cast<const int64_t*>(src.indices), vals, tiles1, sh.s1, cnt1,
        sh.lo, sh.rlow, packed, ivals);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  tr_tiles2a_kernel<<<1, kTrTableBlock, 0, s>>>(cnt1, sh.s1, d1, n, seg, sh.segs, tstart, split);
  const int64_t fill = sh.s2 > d1 ? sh.s2 : d1;
  tr_tiles2b_kernel<<<static_cast<unsigned>((fill + kTrSmall - 1) / kTrSmall), kTrSmall, 0, s>>>(
      cnt1, sh.s1, d1, n, tstart, tiles2, sh.s2);
  tr_hist_kernel<uint32_t><<<g2, kTrHistBlock, 0, s>>>(packed, tiles2, sh.rlow, sh.hi, cnt2, sh.s2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_u32_exclusive(cnt2, cnt2, c2, scan, s)) != hipSuccess) return e;
  if (out_row_bytes == 4)
    tr_pass2_kernel<VB, int32_t><<<g2, kTrBlock, 0, s>>>(packed, ivals, tiles2, sh.s2, cnt2, split,
                                                         sh.segs, sh.hi, sh.rlow,
                                                         static_cast<int32_t*>(out_rows), out_vals);
  else
    tr_pass2_kernel<VB, int64_t><<<g2, kTrBlock, 0, s>>>(packed, ivals, tiles2, sh.s2, cnt2, split,
                                                         sh.segs, sh.hi, sh.rlow,
                                                         static_cast<int64_t*>(out_rows), out_vals);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  int64_t og = (src.inner + 1 + kTrSmall - 1) / kTrSmall;
  if (og > 65536) og = 65536;
  tr_offsets_kernel<<<static_cast<unsigned>(og), kTrSmall, 0, s>>>(cnt2, sh.s2, tstart, sh.lo,
                                                                   src.inner, n, out_offsets);
  return hipGetLastError();
}

}  // namespace

bool transpose2_applies(const Compressed& src) {
  TrShape sh;
  return tr_shape(src, &sh);
}

size_t transpose2_workspace_bytes(const Compressed& src) {
  TrShape sh;
  if (!tr_shape(src, &sh)) return 0;
  return tr_layout(src, sh).total;
}

hipError_t transpose2_launch(const Compressed& src, int64_t* out_offsets, void* out_indices,
                             int out_index_bytes, void* out_values, void* workspace,
                             hipStream_t s) {
  TrShape sh;
  if (!tr_shape(src, &sh)) return hipErrorInvalidValue;
  char* ws = static_cast<char*>(workspace);
  switch (elem_bytes(src.elem)) {
    case 8:
      return tr_run<uint64_t>(src, sh, out_offsets, out_indices, out_index_bytes,
                              static_cast<uint64_t*>(out_values), ws, s);
    case 4:
      return tr_run<uint32_t>(src, sh, out_offsets, out_indices, out_index_bytes,
                              static_cast<uint32_t*>(out_values), ws, s);
    default:
      return tr_run<uint16_t>(src, sh, out_offsets, out_indices, out_index_bytes,
                              static_cast<uint16_t*>(out_values), ws, s);
  }
}

}  // namespace asc_hip
