// Seeded synthetic CSR generators (host, multi-threaded, deterministic).
#include "generators.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

inline uint64_t mix64(uint64_t z) {
  // SplitMix64 finalizer (Steele, Lea, Flood 2014).
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Independent stream per (seed, row, tag); counter indexes draws in it.
struct Stream {
  uint64_t key;
  Stream(uint64_t seed, uint64_t row, uint64_t tag)
      : key(mix64(mix64(seed ^ 0xA5C05EEDull) + row * 0xD1B54A32D192ED03ull + tag)) {}
  uint64_t u64(uint64_t counter) const { return mix64(key ^ (counter * 0x9E3779B97F4A7C15ull)); }
  double unit(uint64_t counter) const {  // [0, 1)
    return static_cast<double>(u64(counter) >> 11) * 0x1.0p-53;
  }
  double open_unit(uint64_t counter) const {  // (0, 1]
    return static_cast<double>((u64(counter) >> 11) + 1) * 0x1.0p-53;
  }
  int64_t below(uint64_t counter, int64_t n) const {  // [0, n)
    // 128-bit multiply-shift; bias is below 2^-64 * n.
    const unsigned __int128 m = static_cast<unsigned __int128>(u64(counter)) * static_cast<uint64_t>(n);
    return static_cast<int64_t>(m >> 64);
  }
  double normal(uint64_t counter) const {  // Box-Muller, one of the pair
    const double u1 = open_unit(2 * counter), u2 = unit(2 * counter + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
  }
};

enum : uint64_t { kTagLength = 1, kTagColumns = 2, kTagValues = 3 };

double zeta_approx(double s) {
  // Euler-Maclaurin: sum_{k<M} k^-s + M^(1-s)/(s-1) + M^-s/2 + s M^(-s-1)/12.
  const int64_t M = 4096;
  double sum = 0.0;
  for (int64_t k = M - 1; k >= 1; --k) sum += std::pow(static_cast<double>(k), -s);
  const double m = static_cast<double>(M);
  return sum + std::pow(m, 1.0 - s) / (s - 1.0) + 0.5 * std::pow(m, -s) +
         s * std::pow(m, -s - 1.0) / 12.0;
}

struct ZipfTable {
  std::vector<double> cdf;  // cdf[k-1] = P(K <= k), k = 1..max-1
  int64_t max = 1;
  void build(double s, int64_t lmax) {
    max = lmax;
    cdf.resize(lmax > 1 ? lmax - 1 : 0);
    const double z = zeta_approx(s);
    double acc = 0.0;
    for (int64_t k = 1; k < lmax; ++k) {
      acc += std::pow(static_cast<double>(k), -s) / z;
      cdf[k - 1] = acc;
    }
  }
  int64_t sample(double u) const {
    const auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
    return it == cdf.end() ? max : static_cast<int64_t>(it - cdf.begin()) + 1;
  }
};

template <typename F>
void parallel_rows(int64_t rows, int threads, F&& f) {
  if (threads < 1) threads = 1;
  if (rows < 4096 || threads == 1) {
    f(0, rows);
    return;
  }
  std::vector<std::thread> pool;
  const int64_t chunk = (rows + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const int64_t b = t * chunk, e = std::min(rows, b + chunk);
    if (b >= e) break;
    pool.emplace_back([&, b, e] { f(b, e); });
  }
  for (auto& th : pool) th.join();
}

struct Gen {
  asc_gen_params p;
  ZipfTable zipf;

  explicit Gen(const asc_gen_params& params) : p(params) {
    if (p.family == ASC_GEN_ZIPF) zipf.build(p.zipf_s, p.zipf_max);
  }

  int64_t var_length(const Stream& s) const {
    const double z = s.normal(0);
    int64_t l = static_cast<int64_t>(std::llround(static_cast<double>(p.per_row) * std::exp(0.5 * z)));
    return std::max<int64_t>(1, l);
  }

  int64_t length(int64_t row) const {
    const Stream s(p.seed, row, kTagLength);
    switch (p.family) {
      case ASC_GEN_UNIFORM_FIXED: return std::min(p.per_row, p.cols);
      case ASC_GEN_ZIPF: return std::min(zipf.sample(s.unit(0)), p.cols);
      case ASC_GEN_PRECISION: {
        const int64_t lo = std::max<int64_t>(0, row - 64), hi = std::min(p.cols - 1, row + 64);
        const int64_t w = hi >= lo ? hi - lo + 1 : 0;
        const int64_t near = std::min<int64_t>(16, w);
        return near + std::min<int64_t>(16, p.cols - near);
      }
      case ASC_GEN_STENCIL2D: {
        const int64_t k = p.grid, a = row / k, b = row % k;
        return 1 + (a > 0) + (a < k - 1) + (b > 0) + (b < k - 1);
      }
      case ASC_GEN_STENCIL3D: {
        const int64_t k = p.grid, a = row / (k * k), b = (row / k) % k, c = row % k;
        return 1 + (a > 0) + (a < k - 1) + (b > 0) + (b < k - 1) + (c > 0) + (c < k - 1);
      }
      case ASC_GEN_BANDED: {
        int64_t l = 0;
        for (int64_t j = row - p.band; j <= row + p.band; ++j) {
          if (j < 0 || j >= p.cols) continue;
          if (j == row || s.unit(static_cast<uint64_t>(j - row + p.band)) < p.fill) ++l;
        }
        return l;
      }
      case ASC_GEN_BLOCKDIAG8: {
        const int64_t b0 = (row / 8) * 8;
        return std::max<int64_t>(0, std::min(p.cols, b0 + 8) - b0);
      }
      case ASC_GEN_BIPARTITE: {
        const int64_t half = p.cols / 2;
        const int64_t width = row < p.rows / 2 ? p.cols - half : half;
        return std::min(var_length(s), width);
      }
      case ASC_GEN_UNIFORM_VAR: return std::min(var_length(s), p.cols);
    }
    return -1;
  }

  double value(const Stream& s, uint64_t k) const {
    switch (p.value_dist) {
      case ASC_GEN_VALUES_UNIFORM: return 2.0 * s.unit(k) - 1.0;
      case ASC_GEN_VALUES_NORMAL: return s.normal(k);
      case ASC_GEN_VALUES_LOGNORMAL: {
        const double mag = std::exp(2.0 * s.normal(k));
        return (s.u64(0x8000000000000000ull + k) & 1) ? -mag : mag;
      }
    }
    return 0.0;
  }

  // Sorted distinct columns: L sorted iid draws in [0, n-L] plus 0..L-1.
  static void sorted_distinct(const Stream& s, uint64_t base, int64_t L, int64_t lo,
                              int64_t n, std::vector<int64_t>& out) {
    out.resize(L);
    for (int64_t k = 0; k < L; ++k) out[k] = s.below(base + k, n - L + 1);
    std::sort(out.begin(), out.end());
    for (int64_t k = 0; k < L; ++k) out[k] += k + lo;
  }

  void columns(int64_t row, int64_t L, std::vector<int64_t>& out) const {
    const Stream s(p.seed, row, kTagColumns);
    out.clear();
    switch (p.family) {
      case ASC_GEN_UNIFORM_FIXED:
      case ASC_GEN_ZIPF:
      case ASC_GEN_UNIFORM_VAR:
        if (p.window > 0 && p.window < p.cols) {
          const int64_t w = std::max(p.window, L);
          const int64_t centre = p.rows > 0 ? static_cast<int64_t>(
              (static_cast<double>(row) + 0.5) * static_cast<double>(p.cols) / static_cast<double>(p.rows)) : 0;
          const int64_t lo = std::min(std::max<int64_t>(0, centre - w / 2), p.cols - w);
          sorted_distinct(s, 0, L, lo, w, out);
        } else {
          sorted_distinct(s, 0, L, 0, p.cols, out);
        }
        return;
      case ASC_GEN_BIPARTITE: {
        const int64_t half = p.cols / 2;
        if (row < p.rows / 2) sorted_distinct(s, 0, L, half, p.cols - half, out);
        else sorted_distinct(s, 0, L, 0, half, out);
        return;
      }
      case ASC_GEN_PRECISION: {
        const int64_t lo = std::max<int64_t>(0, row - 64), hi = std::min(p.cols - 1, row + 64);
        const int64_t w = hi >= lo ? hi - lo + 1 : 0;
        const int64_t near = std::min<int64_t>(16, w);
        std::vector<int64_t> window(w);
        for (int64_t j = 0; j < w; ++j) window[j] = lo + j;
        for (int64_t j = 0; j < near; ++j) {  // partial Fisher-Yates
          const int64_t r = j + s.below(static_cast<uint64_t>(j), w - j);
          std::swap(window[j], window[r]);
          out.push_back(window[j]);
        }
        uint64_t c = 1000;
        while (static_cast<int64_t>(out.size()) < L) {
          const int64_t cand = s.below(c++, p.cols);
          if (std::find(out.begin(), out.end(), cand) == out.end()) out.push_back(cand);
        }
        std::sort(out.begin(), out.end());
        return;
      }
      case ASC_GEN_STENCIL2D: {
        const int64_t k = p.grid, a = row / k, b = row % k;
        if (a > 0) out.push_back(row - k);
        if (b > 0) out.push_back(row - 1);
        out.push_back(row);
        if (b < k - 1) out.push_back(row + 1);
        if (a < k - 1) out.push_back(row + k);
        return;
      }
      case ASC_GEN_STENCIL3D: {
        const int64_t k = p.grid, kk = k * k;
        const int64_t a = row / kk, b = (row / k) % k, c = row % k;
        if (a > 0) out.push_back(row - kk);
        if (b > 0) out.push_back(row - k);
        if (c > 0) out.push_back(row - 1);
        out.push_back(row);
        if (c < k - 1) out.push_back(row + 1);
        if (b < k - 1) out.push_back(row + k);
        if (a < k - 1) out.push_back(row + kk);
        return;
      }
      case ASC_GEN_BANDED: {
        const Stream ls(p.seed, row, kTagLength);
        for (int64_t j = row - p.band; j <= row + p.band; ++j) {
          if (j < 0 || j >= p.cols) continue;
          if (j == row || ls.unit(static_cast<uint64_t>(j - row + p.band)) < p.fill) out.push_back(j);
        }
        return;
      }
      case ASC_GEN_BLOCKDIAG8: {
        const int64_t b0 = (row / 8) * 8;
        for (int64_t j = b0; j < std::min(p.cols, b0 + 8); ++j) out.push_back(j);
        return;
      }
    }
  }
};

bool valid(const asc_gen_params* p) {
  if (p == nullptr || p->rows < 0 || p->cols < 0) return false;
  switch (p->family) {
    case ASC_GEN_STENCIL2D: return p->grid >= 1 && p->rows == p->grid * p->grid && p->cols == p->rows;
    case ASC_GEN_STENCIL3D: return p->grid >= 1 && p->rows == p->grid * p->grid * p->grid && p->cols == p->rows;
    case ASC_GEN_ZIPF: return p->zipf_s > 1.0 && p->zipf_max >= 1;
    case ASC_GEN_BANDED: return p->band >= 0;
    case ASC_GEN_UNIFORM_FIXED:
    case ASC_GEN_UNIFORM_VAR:
    case ASC_GEN_BIPARTITE: return p->per_row >= 0;
    case ASC_GEN_PRECISION:
    case ASC_GEN_BLOCKDIAG8: return true;
  }
  return false;
}

}  // namespace

extern "C" {

int64_t asc_gen_lengths(const asc_gen_params* p, int64_t* offsets, int threads) {
  if (!valid(p)) return -1;
  const Gen g(*p);
  offsets[0] = 0;
  parallel_rows(p->rows, threads, [&](int64_t b, int64_t e) {
    for (int64_t r = b; r < e; ++r) offsets[r + 1] = g.length(r);
  });
  for (int64_t r = 0; r < p->rows; ++r) offsets[r + 1] += offsets[r];
  return offsets[p->rows];
}

int asc_gen_fill(const asc_gen_params* p, const int64_t* offsets, void* indices,
                 int index_bytes, double* values, int threads) {
  if (!valid(p) || (index_bytes != 4 && index_bytes != 8)) return -1;
  const Gen g(*p);
  parallel_rows(p->rows, threads, [&](int64_t b, int64_t e) {
    std::vector<int64_t> cols;
    for (int64_t r = b; r < e; ++r) {
      const int64_t L = offsets[r + 1] - offsets[r];
      g.columns(r, L, cols);
      const Stream vs(p->seed, r, kTagValues);
      const int64_t o = offsets[r];
      for (int64_t k = 0; k < L; ++k) {
        if (index_bytes == 4) static_cast<int32_t*>(indices)[o + k] = static_cast<int32_t>(cols[k]);
        else static_cast<int64_t*>(indices)[o + k] = cols[k];
        double v;
        if (p->value_dist == ASC_GEN_VALUES_LAPLACIAN) {
          v = cols[k] == r ? (p->family == ASC_GEN_STENCIL3D ? 6.0 : 4.0) : -1.0;
        } else {
          v = g.value(vs, static_cast<uint64_t>(k));
        }
        values[o + k] = v;
      }
    }
  });
  return 0;
}

void asc_gen_vector(int64_t n, uint64_t seed, int32_t value_dist, double* out,
                    int threads) {
  asc_gen_params p{};
  p.value_dist = value_dist;
  p.seed = seed;
  const Gen g(p);
  parallel_rows(n, threads, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const Stream s(seed, static_cast<uint64_t>(i), 7);
      out[i] = g.value(s, 0);
    }
  });
}

double asc_gen_zipf_mean(double s, int64_t zipf_max) {
  ZipfTable t;
  t.build(s, zipf_max);
  double mean = 0.0, prev = 0.0;
  for (int64_t k = 1; k < zipf_max; ++k) {
    mean += static_cast<double>(k) * (t.cdf[k - 1] - prev);
    prev = t.cdf[k - 1];
  }
  return mean + static_cast<double>(zipf_max) * (1.0 - prev);
}

}  // extern "C"
