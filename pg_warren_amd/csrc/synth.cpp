// Synthetic R-MAT graphs and edge weights (host, multi-threaded).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/warren_gpu.h"
#include "../../include/warren_gpu_synth.h"

namespace {

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// A uniform double in [0, 1) from the top 53 bits.
double unit(uint64_t bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }

int nthreads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)std::min(h, 64u) : 1;
}

template <class F>
void parallel_for(uint64_t n, int threads, F f) {
  int nt = (int)std::min<uint64_t>(nthreads(threads), std::max<uint64_t>(n / 4096, 1));
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (int i = 0; i < nt; ++i) {
    uint64_t a = n * i / nt, b = n * (i + 1) / nt;
    pool.emplace_back([=, &f] { f(a, b); });
  }
  for (auto& t : pool) t.join();
}

}  // namespace

extern "C" {

// The raw R-MAT edge list: edge e's endpoints from its own SplitMix64 stream.
static void rmat_edges(uint32_t scale, uint64_t m, double a, double b, double c, uint64_t seed,
                       uint32_t* src, uint32_t* dst, int threads) {
  const double ab = a + b, abc = a + b + c;
  const uint64_t key = mix64(seed ^ 0x5eed5eed5eed5eedull);
  parallel_for(m, threads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t e = lo; e < hi; ++e) {
      uint32_t s = 0, t = 0;
      const uint64_t base = mix64(key ^ mix64(e * 0x9e3779b97f4a7c15ull));
      for (uint32_t level = 0; level < scale; ++level) {
        double u = unit(mix64(base ^ ((uint64_t)(level + 1) * 0xd1b54a32d192ed03ull)));
        uint32_t sb = 0, tb = 0;
        if (u < a) {
        } else if (u < ab) {
          tb = 1;
        } else if (u < abc) {
          sb = 1;
        } else {
          sb = 1;
          tb = 1;
        }
        s = (s << 1) | sb;
        t = (t << 1) | tb;
      }
      src[e] = s;
      dst[e] = t;
    }
  });
}

int wg_rmat_edges(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
                  uint32_t* sources, uint32_t* targets, int threads) {
  if (scale == 0 || scale > 31 || !sources || !targets) return WG_INVALID_ARGUMENT;
  rmat_edges(scale, (uint64_t)ef << scale, a, b, c, seed, sources, targets, threads);
  return WG_OK;
}

int wg_rmat_csr(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed,
                int symmetrize, uint64_t* offsets, uint32_t* targets, uint64_t capacity,
                uint64_t* nnz_out, int threads) {
  if (scale == 0 || scale > 31) return WG_INVALID_ARGUMENT;
  const uint64_t n = 1ull << scale;
  const uint64_t m = (uint64_t)ef << scale;
  const uint64_t total = symmetrize ? 2 * m : m;
  if (capacity < total) return WG_INVALID_ARGUMENT;
  std::vector<uint32_t> src(m), dst(m);
  rmat_edges(scale, m, a, b, c, seed, src.data(), dst.data(), threads);
  // Counting sort by source (optionally with reversed copies), then sort
  // and deduplicate each row, dropping self loops.
  std::vector<uint64_t> cnt(n + 1, 0);
  for (uint64_t e = 0; e < m; ++e) {
    cnt[src[e] + 1]++;
    if (symmetrize) cnt[dst[e] + 1]++;
  }
  for (uint64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
  std::vector<uint32_t> adj(total);
  {
    std::vector<uint64_t> cur(cnt.begin(), cnt.end() - 1);
    for (uint64_t e = 0; e < m; ++e) {
      adj[cur[src[e]]++] = dst[e];
      if (symmetrize) adj[cur[dst[e]]++] = src[e];
    }
  }
  src.clear();
  src.shrink_to_fit();
  dst.clear();
  dst.shrink_to_fit();
  std::vector<uint64_t> kept(n, 0);
  parallel_for(n, threads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t r = lo; r < hi; ++r) {
      uint32_t* row = adj.data() + cnt[r];
      uint64_t len = cnt[r + 1] - cnt[r];
      std::sort(row, row + len);
      uint64_t k = 0;
      for (uint64_t i = 0; i < len; ++i) {
        if (row[i] == r) continue;
        if (k > 0 && row[k - 1] == row[i]) continue;
        row[k++] = row[i];
      }
      kept[r] = k;
    }
  });
  offsets[0] = 0;
  for (uint64_t r = 0; r < n; ++r) offsets[r + 1] = offsets[r] + kept[r];
  parallel_for(n, threads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t r = lo; r < hi; ++r)
      memcpy(targets + offsets[r], adj.data() + cnt[r], kept[r] * sizeof(uint32_t));
  });
  *nnz_out = offsets[n];
  return WG_OK;
}

int wg_edge_weights(uint64_t n, const uint64_t* offsets, const uint32_t* targets, uint64_t seed,
                    int kind, double p0, double p1, double* weights, int threads) {
  if (kind < 0 || kind > 2) return WG_INVALID_ARGUMENT;
  const uint64_t key = mix64(seed ^ 0x77e1647e1647e16ull);
  parallel_for(n, threads, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t r = lo; r < hi; ++r)
      for (uint64_t p = offsets[r]; p < offsets[r + 1]; ++p) {
        uint64_t h = mix64(key ^ mix64((r << 32) ^ targets[p]));
        double w;
        if (kind == 0) {
          w = (double)((h >> 40) + 1) * (1.0 / 16777216.0);
        } else if (kind == 1) {
          uint64_t span = (uint64_t)(p1 - p0) + 1;
          w = p0 + (double)(h % span);
        } else {
          double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);
          double u2 = unit(mix64(h));
          double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
          w = std::exp(p0 + p1 * z);
        }
        weights[p] = w;
      }
  });
  return WG_OK;
}

}  // extern "C"
