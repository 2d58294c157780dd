// Device construction of the CSR pair on the context stream: the forward
// CSR of an edge list, and the reverse CSR by transposition.
//
// Forward: GraphCsr::build_as (graph/csr/graph_csr.rs:210-221) for a graph
// of one relation, or its merged all-relations CSR for several
// (graph/csr/build/mod.rs:248-288). The edges are ordered by
// (source, target, relation) with a stable LSD radix sort, so edges with
// equal keys keep their input order. That is the order the reference's
// stable counting sort by source (build/columns.rs:263-303, 322-343) and
// stable per-row sort by target (build/deduplication.rs:83-117) leave them
// in. Each run of equal (source, target) becomes one entry. Its weight is
// the sum, in input order, of each relation's weights; those per-relation
// sums are then added in relation order (deduplication.rs:122-152,
// relation_sum at :196-206). With one relation that is the plain
// left-to-right sum of fold_weighted (:83-117). Self loops stay.
//
// Reverse: graph/csr/transpose.rs:24-44. A stable radix sort of the
// forward entries by target keeps each reverse row's sources ascending,
// as the reference's scan of the sources in order does.
//
// Row offsets come from one binary search per row over the sorted row
// keys (no atomics, no scan). The sorts are rocPRIM's radix sort.

#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <string>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr unsigned kThreads = 256;

unsigned grid_of(uint64_t items) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kThreads - 1) / kThreads, 8192));
}

// The bits that hold every id below n.
unsigned id_bits(uint64_t n) { return n <= 1 ? 1 : 64 - __builtin_clzll(n - 1); }

__device__ __forceinline__ uint64_t first_index() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t stride() { return (uint64_t)gridDim.x * blockDim.x; }

// The lowest index of an edge outside the id space or naming a relation
// beyond the dictionary (the edge check of BuildPlan::new,
// build/mod.rs:329-352). A thread's first hit is its lowest.
__global__ __launch_bounds__(kThreads) void first_bad_edge(const uint32_t* __restrict__ src,
                                                           const uint32_t* __restrict__ dst,
                                                           const uint32_t* __restrict__ rel, uint64_t m,
                                                           uint64_t n, uint32_t nrel,
                                                           unsigned long long* __restrict__ first) {
  for (uint64_t e = first_index(); e < m; e += stride())
    if (src[e] >= n || dst[e] >= n || (rel && rel[e] >= nrel)) {
      atomicMin(first, (unsigned long long)e);
      return;
    }
}

// keys[e] = (source, target) packed with the source in the high bits;
// idx[e] = e.
template <class V>
__global__ __launch_bounds__(kThreads) void pack_keys(const uint32_t* __restrict__ src,
                                                      const uint32_t* __restrict__ dst, uint64_t m,
                                                      unsigned b, uint64_t* __restrict__ keys,
                                                      V* __restrict__ idx) {
  for (uint64_t e = first_index(); e < m; e += stride()) {
    keys[e] = ((uint64_t)src[e] << b) | dst[e];
    idx[e] = (V)e;
  }
}

template <class V>
__global__ __launch_bounds__(kThreads) void iota(V* __restrict__ idx, uint64_t m) {
  for (uint64_t e = first_index(); e < m; e += stride()) idx[e] = (V)e;
}

template <class V>
__global__ __launch_bounds__(kThreads) void gather_keys(const uint64_t* __restrict__ keys,
                                                        const V* __restrict__ perm, uint64_t m,
                                                        uint64_t* __restrict__ out) {
  for (uint64_t k = first_index(); k < m; k += stride()) out[k] = keys[perm[k]];
}

// heads[k] = 1 where a run of equal (source, target) starts; *runs += the
// runs of equal (source, target, relation), the edges that stay after a
// relation folds its parallel edges.
template <class V>
__global__ __launch_bounds__(kThreads) void mark_runs(const uint64_t* __restrict__ skeys,
                                                      const V* __restrict__ perm,
                                                      const uint32_t* __restrict__ rel, uint64_t m,
                                                      bool keep, V* __restrict__ heads,
                                                      unsigned long long* __restrict__ runs) {
  __shared__ unsigned long long part[kThreads / 64];
  unsigned long long cnt = 0;
  for (uint64_t k = first_index(); k < m; k += stride()) {
    const bool h = keep || k == 0 || skeys[k] != skeys[k - 1];
    const bool hr = h || (rel && rel[perm[k]] != rel[perm[k - 1]]);
    heads[k] = h ? 1 : 0;
    cnt += hr ? 1 : 0;
  }
  for (int d = 32; d > 0; d >>= 1) cnt += __shfl_down(cnt, d, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (unsigned w = 0; w < kThreads / 64; ++w) t += part[w];
    if (t) atomicAdd(runs, t);
  }
}

// One forward entry per run of equal (source, target), at position
// pos[k] - 1 (pos = inclusive count of run heads): its target, its row, and
// its weight, each relation's weights summed in input order and those
// sums added in relation order (relation_sum and merge_weighted,
// build/deduplication.rs:122-152, 196-206). With keep, every edge is an
// entry of its own (a multigraph: Product::per_relation).
template <class V>
__global__ __launch_bounds__(kThreads) void fold_runs(const uint64_t* __restrict__ skeys,
                                                      const V* __restrict__ perm,
                                                      const V* __restrict__ pos,
                                                      const uint32_t* __restrict__ rel,
                                                      const double* __restrict__ w, uint64_t m,
                                                      unsigned b, bool keep, uint32_t* __restrict__ tgt,
                                                      uint32_t* __restrict__ rowkey,
                                                      double* __restrict__ wout) {
  const uint64_t mask = (1ull << b) - 1;
  for (uint64_t k = first_index(); k < m; k += stride()) {
    const uint64_t key = skeys[k];
    if (!keep && k > 0 && skeys[k - 1] == key) continue;
    const uint64_t u = (uint64_t)pos[k] - 1;
    tgt[u] = (uint32_t)(key & mask);
    rowkey[u] = (uint32_t)(key >> b);
    if (wout && keep) {
      wout[u] = w[perm[k]];
    } else if (wout) {
      double total = 0.0;
      bool first = true;
      uint64_t j = k;
      while (j < m && skeys[j] == key) {
        const uint32_t r0 = rel ? rel[perm[j]] : 0u;
        double sum = w[perm[j]];
        ++j;
        while (j < m && skeys[j] == key && (!rel || rel[perm[j]] == r0)) {
          sum += w[perm[j]];
          ++j;
        }
        total = first ? sum : total + sum;
        first = false;
      }
      wout[u] = total;
    }
  }
}

// off[r] = the first position whose row key is at least r, r in [0, n]:
// the row starts of entries sorted by row.
__global__ __launch_bounds__(kThreads) void row_starts(const uint32_t* __restrict__ rowkey, uint64_t cnt,
                                                       uint64_t n, uint64_t* __restrict__ off) {
  for (uint64_t r = first_index(); r <= n; r += stride()) {
    uint64_t lo = 0, hi = cnt;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)rowkey[mid] < r)
        lo = mid + 1;
      else
        hi = mid;
    }
    off[r] = lo;
  }
}

// The row holding entry p: the last r in [lo, n) with off[r] <= p (so
// off[r] <= p < off[r + 1]), given off[lo] <= p < off[n].
__device__ __forceinline__ uint64_t row_holding(const uint64_t* __restrict__ off, uint64_t n, uint64_t p,
                                                uint64_t lo) {
  uint64_t hi = n;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (off[mid] <= p) lo = mid;
    else hi = mid;
  }
  return lo;
}

// row_of[p] = the row of entry p, a tile of kExpandTile entries at a time.
// anchor[k] = the row of entry k * kExpandTile (a thread each, a search over
// all rows); a tile's rows then lie in [anchor[k], anchor[k + 1]] (the
// last tile's up to n - 1), so a thread's run of kExpandRun consecutive
// entries searches only those: the first entry by binary search, a row
// change inside the run again from there (empty rows are skipped however
// many). Runs are stored as 16-byte writes when row_of is so aligned.
constexpr uint32_t kExpandRun = 16;
constexpr uint64_t kExpandTile = (uint64_t)kThreads * kExpandRun;

__global__ __launch_bounds__(kThreads) void expand_anchors(const uint64_t* __restrict__ off, uint64_t n,
                                                           uint64_t m, uint32_t* __restrict__ anchor) {
  const uint64_t tiles = (m + kExpandTile - 1) / kExpandTile;
  for (uint64_t k = first_index(); k < tiles; k += stride())
    anchor[k] = (uint32_t)row_holding(off, n, k * kExpandTile, 0);
}

__global__ __launch_bounds__(kThreads) void expand_entries(const uint64_t* __restrict__ off, uint64_t n,
                                                           uint64_t m, const uint32_t* __restrict__ anchor,
                                                           uint32_t* __restrict__ row_of, bool vec) {
  const uint64_t tiles = (m + kExpandTile - 1) / kExpandTile;
  for (uint64_t k = blockIdx.x; k < tiles; k += gridDim.x) {
    const uint64_t p0 = k * kExpandTile + (uint64_t)threadIdx.x * kExpandRun;
    if (p0 >= m) continue;
    const uint32_t cnt = (uint32_t)min<uint64_t>(kExpandRun, m - p0);
    // Rows [lo, hi) hold the tile: off[lo] <= p0 < off[hi].
    const uint64_t lo = anchor[k];
    const uint64_t hi = k + 1 < tiles ? (uint64_t)anchor[k + 1] + 1 : n;
    uint64_t r = row_holding(off, hi, p0, lo);
    uint64_t next = off[r + 1];
    uint32_t v[kExpandRun];
#pragma unroll
    for (uint32_t j = 0; j < kExpandRun; ++j) {
      if (j < cnt && p0 + j >= next && r + 1 < hi) {
        r = row_holding(off, hi, p0 + j, r + 1);
        next = off[r + 1];
      }
      v[j] = (uint32_t)r;
    }
    if (vec && cnt == kExpandRun) {
      uint4* dst = reinterpret_cast<uint4*>(row_of + p0);
#pragma unroll
      for (uint32_t q = 0; q < kExpandRun / 4; ++q)
        dst[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
      for (uint32_t j = 0; j < kExpandRun; ++j)
        if (j < cnt) row_of[p0 + j] = v[j];
    }
  }
}

// Reverse position k holds forward entry perm[k]: its source, its weight.
template <class V>
__global__ __launch_bounds__(kThreads) void place_reverse(const V* __restrict__ perm,
                                                          const uint32_t* __restrict__ row_of,
                                                          const double* __restrict__ w, uint64_t m,
                                                          uint32_t* __restrict__ rtgt,
                                                          double* __restrict__ rw) {
  for (uint64_t k = first_index(); k < m; k += stride()) {
    const uint64_t p = perm[k];
    rtgt[k] = row_of[p];
    if (rw) rw[k] = w[p];
  }
}

// pos[p] = the reverse position of forward entry p (source s, target t):
// where s sits in the ascending reverse row of t.
__global__ __launch_bounds__(kThreads) void reverse_positions(const uint32_t* __restrict__ row_of,
                                                              const uint32_t* __restrict__ ftgt,
                                                              const uint64_t* __restrict__ roff,
                                                              const uint32_t* __restrict__ rtgt, uint64_t m,
                                                              uint64_t* __restrict__ pos) {
  for (uint64_t p = first_index(); p < m; p += stride()) {
    const uint32_t s = row_of[p], t = ftgt[p];
    uint64_t lo = roff[t], hi = roff[t + 1];
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (rtgt[mid] < s)
        lo = mid + 1;
      else
        hi = mid;
    }
    pos[p] = lo;
  }
}

// The edges regrouped by relation: out[k] = in[perm[k]].
template <class V>
__global__ __launch_bounds__(kThreads) void gather_edges(const V* __restrict__ perm,
                                                         const uint32_t* __restrict__ src,
                                                         const uint32_t* __restrict__ dst,
                                                         const double* __restrict__ w, uint64_t m,
                                                         uint32_t* __restrict__ osrc,
                                                         uint32_t* __restrict__ odst,
                                                         double* __restrict__ ow) {
  for (uint64_t k = first_index(); k < m; k += stride()) {
    const uint64_t e = perm[k];
    osrc[k] = src[e];
    odst[k] = dst[e];
    if (ow) ow[k] = w[e];
  }
}

// The weight of an entry of an unweighted relation read with weights
// (compute/product/matrix.rs:270-272).
__global__ __launch_bounds__(kThreads) void fill_ones(double* __restrict__ w, uint64_t m) {
  for (uint64_t k = first_index(); k < m; k += stride()) w[k] = 1.0;
}

// *flag |= 1 if a[i] != b[i] for some i (compared with ==).
template <class T>
__global__ __launch_bounds__(kThreads) void differs(const T* __restrict__ a, const T* __restrict__ b,
                                                    uint64_t count, unsigned* __restrict__ flag) {
  int d = 0;
  for (uint64_t i = first_index(); i < count; i += stride()) d |= !(a[i] == b[i]);
  if (__syncthreads_or(d) && threadIdx.x == 0) atomicOr(flag, 1u);
}

// Library temporary storage rounded up to a power of two (at least 64 KB),
// so that released blocks are reused across calls of varying size.
inline size_t reusable_size(size_t b) {
  size_t r = size_t(64) << 10;
  while (r < b) r <<= 1;
  return r;
}

// A device buffer charged to the context, released on scope exit.
template <class T>
struct Scratch {
  wg_context* ctx = nullptr;
  T* p = nullptr;
  size_t bytes = 0;
  int alloc(wg_context* c, uint64_t count) {
    ctx = c;
    bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    return dalloc(ctx, (void**)&p, bytes);
  }
  void reset() {
    if (p) dfree(ctx, p, bytes);
    p = nullptr;
  }
  ~Scratch() { reset(); }
};

// Stable ascending radix sort of (kin, vin) on bits [0, bits) of the keys.
template <class K, class V>
int sort_pairs(wg_context* ctx, const K* kin, K* kout, const V* vin, V* vout, uint64_t m, unsigned bits) {
  size_t tb = 0;
  WG_HIP(ctx, rocprim::radix_sort_pairs(nullptr, tb, kin, kout, vin, vout, (size_t)m, 0u, bits, ctx->stream));
  Scratch<char> tmp;
  if (int s = tmp.alloc(ctx, reusable_size(tb))) return s;
  WG_HIP(ctx, rocprim::radix_sort_pairs((void*)tmp.p, tb, kin, kout, vin, vout, (size_t)m, 0u, bits,
                                        ctx->stream));
  // The scratch is released on return, after the sort (dfree waits for it).
  return WG_OK;
}

template <class V>
int inclusive_count(wg_context* ctx, const V* in, V* out, uint64_t m) {
  size_t tb = 0;
  WG_HIP(ctx, rocprim::inclusive_scan(nullptr, tb, in, out, (size_t)m, rocprim::plus<V>(), ctx->stream));
  Scratch<char> tmp;
  if (int s = tmp.alloc(ctx, reusable_size(tb))) return s;
  WG_HIP(ctx, rocprim::inclusive_scan((void*)tmp.p, tb, in, out, (size_t)m, rocprim::plus<V>(), ctx->stream));
  return WG_OK;
}

// The forward CSR of a device edge list into f (offsets, targets, weights;
// not yet partitioned); *folded = edges folded into an edge of the same
// relation. keep: no folding, parallel edges stay entries of their own.
template <class V>
int build_forward(wg_context* ctx, uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                  const double* w, const uint32_t* rel, uint32_t nrel, DevCsr& f, uint64_t* folded,
                  bool keep = false) {
  hipStream_t st = ctx->stream;
  const unsigned b = id_bits(n);
  const unsigned grid = grid_of(m);
  int s;
  f.n = n;
  if (m == 0) {
    *folded = 0;
    f.nnz = 0;
    if ((s = dalloc(ctx, (void**)&f.targets, 0))) return s;
    if (w && (s = dalloc(ctx, (void**)&f.weights, 0))) return s;
    if ((s = dalloc(ctx, (void**)&f.offsets, (n + 1) * 8))) return s;
    row_starts<<<grid_of(n + 1), kThreads, 0, st>>>(nullptr, 0, n, f.offsets);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
    return WG_OK;
  }
  Scratch<uint64_t> skeys;
  Scratch<V> perm;
  if ((s = skeys.alloc(ctx, m)) || (s = perm.alloc(ctx, m))) return s;
  {
    Scratch<uint64_t> keys;
    Scratch<V> idx;
    if ((s = keys.alloc(ctx, m)) || (s = idx.alloc(ctx, m))) return s;
    pack_keys<V><<<grid, kThreads, 0, st>>>(src, dst, m, b, keys.p, idx.p);
    WG_HIP(ctx, hipGetLastError());
    if (rel && nrel > 1) {
      // Relation first, then (source, target): the LSD composition orders
      // by (source, target, relation), equal keys in input order.
      Scratch<uint32_t> rsorted;
      Scratch<V> perm1;
      Scratch<uint64_t> keys2;
      if ((s = rsorted.alloc(ctx, m)) || (s = perm1.alloc(ctx, m))) return s;
      if ((s = sort_pairs(ctx, rel, rsorted.p, (const V*)idx.p, perm1.p, m, id_bits(nrel)))) return s;
      rsorted.reset();
      idx.reset();
      if ((s = keys2.alloc(ctx, m))) return s;
      gather_keys<V><<<grid, kThreads, 0, st>>>(keys.p, perm1.p, m, keys2.p);
      WG_HIP(ctx, hipGetLastError());
      keys.reset();
      if ((s = sort_pairs(ctx, (const uint64_t*)keys2.p, skeys.p, (const V*)perm1.p, perm.p, m, 2 * b)))
        return s;
    } else {
      if ((s = sort_pairs(ctx, (const uint64_t*)keys.p, skeys.p, (const V*)idx.p, perm.p, m, 2 * b)))
        return s;
    }
  }
  const uint32_t* frel = (rel && nrel > 1) ? rel : nullptr;
  Scratch<V> pos;
  Scratch<unsigned long long> runs;
  if ((s = pos.alloc(ctx, m)) || (s = runs.alloc(ctx, 1))) return s;
  WG_HIP(ctx, hipMemsetAsync(runs.p, 0, 8, st));
  {
    Scratch<V> heads;
    if ((s = heads.alloc(ctx, m))) return s;
    mark_runs<V><<<grid, kThreads, 0, st>>>(skeys.p, perm.p, frel, m, keep, heads.p, runs.p);
    WG_HIP(ctx, hipGetLastError());
    if ((s = inclusive_count(ctx, (const V*)heads.p, pos.p, m))) return s;
  }
  V last = 0;
  unsigned long long nruns = 0;
  WG_HIP(ctx, hipMemcpyAsync(&last, pos.p + (m - 1), sizeof(V), hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipMemcpyAsync(&nruns, runs.p, 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  const uint64_t nnz = (uint64_t)last;
  *folded = m - nruns;
  f.nnz = nnz;
  if ((s = dalloc(ctx, (void**)&f.targets, nnz * 4))) return s;
  if (w && (s = dalloc(ctx, (void**)&f.weights, nnz * 8))) return s;
  if ((s = dalloc(ctx, (void**)&f.offsets, (n + 1) * 8))) return s;
  Scratch<uint32_t> rowkey;
  if ((s = rowkey.alloc(ctx, nnz))) return s;
  fold_runs<V><<<grid, kThreads, 0, st>>>(skeys.p, perm.p, pos.p, frel, w, m, b, keep, f.targets,
                                          rowkey.p, w ? f.weights : nullptr);
  WG_HIP(ctx, hipGetLastError());
  row_starts<<<grid_of(n + 1), kThreads, 0, st>>>(rowkey.p, nnz, n, f.offsets);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(st));
  return WG_OK;
}

}  // namespace

// row_of[p] for every entry of a CSR.
int expand_rows(wg_context* ctx, const DevCsr& c, uint32_t* row_of) {
  if (c.nnz == 0 || c.n == 0) return WG_OK;
  const uint64_t tiles = (c.nnz + kExpandTile - 1) / kExpandTile;
  Scratch<uint32_t> anchor;
  int s;
  if ((s = anchor.alloc(ctx, tiles))) return s;
  expand_anchors<<<grid_of(tiles), kThreads, 0, ctx->stream>>>(c.offsets, c.n, c.nnz, anchor.p);
  expand_entries<<<(unsigned)std::min<uint64_t>(tiles, 65536), kThreads, 0, ctx->stream>>>(
      c.offsets, c.n, c.nnz, anchor.p, row_of, (uintptr_t)row_of % 16 == 0);
  WG_HIP(ctx, hipGetLastError());
  return WG_OK;
}

namespace {

// The reverse CSR of the partitioned forward CSR f into r (offsets,
// targets, weights; not yet partitioned).
template <class V>
int transpose_into(wg_context* ctx, const DevCsr& f, DevCsr& r) {
  hipStream_t st = ctx->stream;
  const uint64_t n = f.n, m = f.nnz;
  int s;
  r.n = n;
  r.nnz = m;
  if ((s = dalloc(ctx, (void**)&r.offsets, (n + 1) * 8))) return s;
  if ((s = dalloc(ctx, (void**)&r.targets, m * 4))) return s;
  if (f.weights && (s = dalloc(ctx, (void**)&r.weights, m * 8))) return s;
  Scratch<uint32_t> skeys;
  if ((s = skeys.alloc(ctx, m))) return s;
  if (m) {
    Scratch<uint32_t> row_of;
    Scratch<V> idx, perm;
    if ((s = row_of.alloc(ctx, m)) || (s = idx.alloc(ctx, m)) || (s = perm.alloc(ctx, m))) return s;
    if ((s = expand_rows(ctx, f, row_of.p))) return s;
    iota<V><<<grid_of(m), kThreads, 0, st>>>(idx.p, m);
    WG_HIP(ctx, hipGetLastError());
    if ((s = sort_pairs(ctx, (const uint32_t*)f.targets, skeys.p, (const V*)idx.p, perm.p, m, id_bits(n))))
      return s;
    place_reverse<V><<<grid_of(m), kThreads, 0, st>>>(perm.p, row_of.p, f.weights, m, r.targets,
                                                      f.weights ? r.weights : nullptr);
    WG_HIP(ctx, hipGetLastError());
  }
  row_starts<<<grid_of(n + 1), kThreads, 0, st>>>(skeys.p, m, n, r.offsets);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(st));
  return WG_OK;
}

// Each relation's CSR pair (GraphCsr's relation CSRs, graph/csr/build/
// mod.rs:252-279) as a graph owned by g: the edges grouped by relation with
// a stable sort (input order kept within a relation), then one build per
// relation.
template <class V>
int build_relations(wg_context* ctx, wg_graph* g, uint64_t n, uint64_t m, const uint32_t* src,
                    const uint32_t* dst, const double* w, const uint32_t* rel, uint32_t R) {
  hipStream_t st = ctx->stream;
  int s;
  Scratch<uint32_t> bsrc, bdst;
  Scratch<double> bw;
  std::vector<uint64_t> starts(R + 1, 0);
  if ((s = bsrc.alloc(ctx, m)) || (s = bdst.alloc(ctx, m))) return s;
  if (w && (s = bw.alloc(ctx, m))) return s;
  if (m) {
    Scratch<uint32_t> rsorted;
    Scratch<V> idx, perm;
    Scratch<uint64_t> dstarts;
    if ((s = rsorted.alloc(ctx, m)) || (s = idx.alloc(ctx, m)) || (s = perm.alloc(ctx, m)) ||
        (s = dstarts.alloc(ctx, R + 1)))
      return s;
    iota<V><<<grid_of(m), kThreads, 0, st>>>(idx.p, m);
    WG_HIP(ctx, hipGetLastError());
    if ((s = sort_pairs(ctx, rel, rsorted.p, (const V*)idx.p, perm.p, m, id_bits(R)))) return s;
    row_starts<<<grid_of(R + 1), kThreads, 0, st>>>(rsorted.p, m, R, dstarts.p);
    gather_edges<V><<<grid_of(m), kThreads, 0, st>>>(perm.p, src, dst, w, m, bsrc.p, bdst.p,
                                                     w ? bw.p : nullptr);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(starts.data(), dstarts.p, (R + 1) * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
  }
  for (uint32_t r = 0; r < R; ++r) {
    wg_graph* c = new (std::nothrow) wg_graph();
    if (!c) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
    c->ctx = ctx;
    c->n = n;
    c->weighted = w != nullptr;
    c->owner = g;
    g->relations.push_back(c);
    const uint64_t b0 = starts[r], cnt = starts[r + 1] - starts[r];
    uint64_t folded = 0;
    s = cnt < (1ull << 32)
            ? build_forward<uint32_t>(ctx, n, cnt, bsrc.p + b0, bdst.p + b0, w ? bw.p + b0 : nullptr,
                                      nullptr, 0, c->fwd, &folded)
            : build_forward<uint64_t>(ctx, n, cnt, bsrc.p + b0, bdst.p + b0, w ? bw.p + b0 : nullptr,
                                      nullptr, 0, c->fwd, &folded);
    if (!s) {
      c->nnz = c->fwd.nnz;
      s = adopt_device_csr(ctx, c->fwd);
    }
    if (!s) s = build_reverse(ctx, c);
    if (s) return s;
  }
  return WG_OK;
}

}  // namespace

// Transposes g->fwd (partitioned) into g->rev on the device, partitions it
// and records whether the two are equal (A = A^T).
int build_reverse(wg_context* ctx, wg_graph* g) {
  int s = g->fwd.nnz < (1ull << 32) ? transpose_into<uint32_t>(ctx, g->fwd, g->rev)
                                    : transpose_into<uint64_t>(ctx, g->fwd, g->rev);
  if (!s) s = adopt_device_csr(ctx, g->rev);
  if (s) return s;
  // A = A^T exactly when the reverse arrays equal the forward ones (an
  // undirected graph's do; spectral/operator.rs:420-436).
  bool same = false;
  if ((s = same_csr(ctx, g->fwd, g->rev, &same))) return s;
  g->rev_is_fwd = same;
  return WG_OK;
}

// Whether two CSRs over one id space hold equal arrays (weights compared
// with ==; both weighted or neither).
int same_csr(wg_context* ctx, const DevCsr& a, const DevCsr& b, bool* same) {
  *same = false;
  if (a.n != b.n || a.nnz != b.nnz || (a.weights == nullptr) != (b.weights == nullptr)) return WG_OK;
  Scratch<unsigned> flag;
  int s;
  if ((s = flag.alloc(ctx, 1))) return s;
  hipStream_t st = ctx->stream;
  WG_HIP(ctx, hipMemsetAsync(flag.p, 0, 4, st));
  const uint64_t n = a.n, m = a.nnz;
  differs<uint64_t><<<grid_of(n + 1), kThreads, 0, st>>>(a.offsets, b.offsets, n + 1, flag.p);
  if (m) {
    differs<uint32_t><<<grid_of(m), kThreads, 0, st>>>(a.targets, b.targets, m, flag.p);
    if (a.weights) differs<double><<<grid_of(m), kThreads, 0, st>>>(a.weights, b.weights, m, flag.p);
  }
  WG_HIP(ctx, hipGetLastError());
  unsigned h = 0;
  WG_HIP(ctx, hipMemcpyAsync(&h, flag.p, 4, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  *same = h == 0;
  return WG_OK;
}

int sort_u32_pairs(wg_context* ctx, const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                   uint64_t m, unsigned bits) {
  return m ? sort_pairs(ctx, kin, kout, vin, vout, m, bits) : WG_OK;
}

int sort_u64_keyed(wg_context* ctx, const uint64_t* kin, uint64_t* kout, const uint64_t* vin, uint64_t* vout,
                   uint64_t m, unsigned bits) {
  return m ? sort_pairs(ctx, kin, kout, vin, vout, m, bits) : WG_OK;
}

int sort_u32_keyed(wg_context* ctx, const uint32_t* kin, uint32_t* kout, const uint64_t* vin, uint64_t* vout,
                   uint64_t m, unsigned bits) {
  return m ? sort_pairs(ctx, kin, kout, vin, vout, m, bits) : WG_OK;
}

size_t exclusive_scan_u64_bytes(uint64_t m) {
  size_t tb = 0;
  if (m && rocprim::exclusive_scan(nullptr, tb, (const uint64_t*)nullptr, (uint64_t*)nullptr, (uint64_t)0,
                                   (size_t)m, rocprim::plus<uint64_t>(), (hipStream_t)0) != hipSuccess)
    return 0;
  return tb;
}

int exclusive_scan_u64(wg_context* ctx, const uint64_t* in, uint64_t* out, uint64_t m, void* temp,
                       size_t temp_bytes) {
  if (!m) return WG_OK;
  size_t tb = 0;
  WG_HIP(ctx, rocprim::exclusive_scan(nullptr, tb, in, out, (uint64_t)0, (size_t)m, rocprim::plus<uint64_t>(),
                                      ctx->stream));
  Scratch<char> tmp;
  if (!temp || temp_bytes < tb) {
    if (int s = tmp.alloc(ctx, reusable_size(tb))) return s;
    temp = (void*)tmp.p;
  }
  WG_HIP(ctx, rocprim::exclusive_scan(temp, tb, in, out, (uint64_t)0, (size_t)m, rocprim::plus<uint64_t>(),
                                      ctx->stream));
  return WG_OK;
}

// The edges of a CSR as (row, column) pairs, into src/dst at `base`.
int csr_edges(wg_context* ctx, const DevCsr& c, uint32_t* src, uint32_t* dst) {
  if (!c.nnz) return WG_OK;
  int s = expand_rows(ctx, c, src);
  if (s) return s;
  WG_HIP(ctx, hipMemcpyAsync(dst, c.targets, c.nnz * 4, hipMemcpyDeviceToDevice, ctx->stream));
  return WG_OK;
}

// An unweighted CSR (forward only, partitioned) of an edge list on the
// device: parallel edges folded (keep = false) or kept, rows sorted by
// column, equal columns in input order.
int csr_of_edges(wg_context* ctx, uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, bool keep,
                 DevCsr& out) {
  uint64_t folded = 0;
  int s = m < (1ull << 32) ? build_forward<uint32_t>(ctx, n, m, src, dst, nullptr, nullptr, 0, out, &folded, keep)
                           : build_forward<uint64_t>(ctx, n, m, src, dst, nullptr, nullptr, 0, out, &folded, keep);
  if (!s) s = adopt_device_csr(ctx, out);
  return s;
}

// pos[p] = the reverse position of forward entry p, for every entry.
int reverse_position_map(wg_context* ctx, const wg_graph* g, uint64_t* pos) {
  const uint64_t m = g->nnz;
  if (m == 0) return WG_OK;
  Scratch<uint32_t> row_of;
  int s;
  if ((s = row_of.alloc(ctx, m))) return s;
  if ((s = expand_rows(ctx, g->fwd, row_of.p))) return s;
  reverse_positions<<<grid_of(m), kThreads, 0, ctx->stream>>>(row_of.p, g->fwd.targets, g->rev.offsets,
                                                              g->rev.targets, m, pos);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

// The matrix a product over relations reads (compute/product/matrix.rs:
// 126-157, 247-294): one selected relation's CSR pair as stored; several
// relations' rows merged by column, an edge in several of them one entry
// whose weight is their weights' sum in relation order (f64 sums start
// from the first term: Rust's f64 Sum folds from -0.0), or with
// per_relation an entry in each, in relation order. An unweighted relation
// read with weights contributes 1 per entry. The selection is built once
// and kept with the graph.
int select_graph(wg_context* ctx, wg_graph* g, const uint32_t* rels, uint32_t k, int per_relation,
                 wg_graph** out) {
  *out = g;
  std::vector<uint32_t> sel;
  if (rels) {
    sel.assign(rels, rels + k);
    std::sort(sel.begin(), sel.end());
    sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
    // GraphCsr::selected (graph_csr.rs:320-340): the smallest unknown one.
    for (uint32_t r : sel)
      if (r >= g->relation_count)
        return fail(ctx, WG_UNKNOWN_RELATION, "relation " + std::to_string(r) + " is not in the graph");
  } else {
    // Every relation: the merged CSR, unless read per relation.
    if (!per_relation || g->relations.empty()) return WG_OK;
    for (uint32_t r = 0; r < g->relation_count; ++r) sel.push_back(r);
  }
  if (sel.size() == 1) {
    if (!g->relations.empty()) *out = g->relations[sel[0]];
    return WG_OK;
  }
  std::string key = per_relation ? "p" : "m";
  for (uint32_t r : sel) key += ":" + std::to_string(r);
  auto it = g->selections.find(key);
  if (it != g->selections.end()) {
    *out = it->second;
    return WG_OK;
  }
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  uint64_t total = 0;
  bool weighted = false;
  for (uint32_t r : sel) {
    total += g->relations[r]->nnz;
    weighted |= g->relations[r]->weighted;
  }
  int s;
  Scratch<uint32_t> src, dst;
  Scratch<double> w;
  if ((s = src.alloc(ctx, total)) || (s = dst.alloc(ctx, total))) return s;
  if (weighted && (s = w.alloc(ctx, total))) return s;
  uint64_t base = 0;
  for (uint32_t r : sel) {
    const DevCsr& c = g->relations[r]->fwd;
    if (c.nnz) {
      if ((s = expand_rows(ctx, c, src.p + base))) return s;
      WG_HIP(ctx, hipMemcpyAsync(dst.p + base, c.targets, c.nnz * 4, hipMemcpyDeviceToDevice, st));
      if (weighted && c.weights)
        WG_HIP(ctx, hipMemcpyAsync(w.p + base, c.weights, c.nnz * 8, hipMemcpyDeviceToDevice, st));
      else if (weighted)
        fill_ones<<<grid_of(c.nnz), kThreads, 0, st>>>(w.p + base, c.nnz);
      WG_HIP(ctx, hipGetLastError());
    }
    base += c.nnz;
  }
  wg_graph* v = new (std::nothrow) wg_graph();
  if (!v) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  v->ctx = ctx;
  v->n = g->n;
  v->weighted = weighted;
  v->owner = g;
  uint64_t folded = 0;
  // Relation order is input order here, and each relation holds a pair at
  // most once: the build's runs are the selection's merged entries.
  s = total < (1ull << 32)
          ? build_forward<uint32_t>(ctx, g->n, total, src.p, dst.p, weighted ? w.p : nullptr, nullptr, 0,
                                    v->fwd, &folded, per_relation != 0)
          : build_forward<uint64_t>(ctx, g->n, total, src.p, dst.p, weighted ? w.p : nullptr, nullptr, 0,
                                    v->fwd, &folded, per_relation != 0);
  src.reset();
  dst.reset();
  w.reset();
  if (!s) {
    v->nnz = v->fwd.nnz;
    s = adopt_device_csr(ctx, v->fwd);
  }
  if (!s) s = build_reverse(ctx, v);
  if (s) {
    destroy_graph(v);
    return s;
  }
  g->selections[key] = v;
  *out = v;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_graph_select(wg_context* ctx, wg_graph* g, const uint32_t* relations,
                               uint32_t nrelations, int per_relation, wg_graph** out) {
  if (!ctx || !g || !out) return WG_INVALID_ARGUMENT;
  if (int cs_ = check_graph(ctx, g)) return cs_;
  *out = nullptr;
  if (nrelations && !relations) return fail(ctx, WG_INVALID_ARGUMENT, "null relation list");
  int s = select_graph(ctx, g, relations, nrelations, per_relation, out);
  return s ? s : ok(ctx);
}

extern "C" int wg_graph_relation_count(const wg_graph* g, uint32_t* count) {
  if (!g || !count) return WG_INVALID_ARGUMENT;
  *count = g->relation_count;
  return WG_OK;
}

extern "C" int wg_graph_build(wg_context* ctx, uint64_t n, uint64_t m, const uint32_t* sources,
                              const uint32_t* targets, const double* weights, const uint32_t* relations,
                              uint32_t relation_count, int device_pointers, wg_graph** out,
                              uint64_t* deduplicated) {
  if (!ctx || !out) return WG_INVALID_ARGUMENT;
  *out = nullptr;
  if (m && (!sources || !targets)) return fail(ctx, WG_INVALID_ARGUMENT, "null edge list");
  if (n >= (1ull << 32)) return fail(ctx, WG_UNSUPPORTED, "id space beyond u32 targets");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int s;
  // The edge list in device memory: the caller's, or a copy of it.
  Scratch<uint32_t> dsrc, ddst, drel;
  Scratch<double> dw;
  const uint32_t *src = sources, *dst = targets, *rel = relations;
  const double* w = weights;
  if (!device_pointers && m) {
    if ((s = dsrc.alloc(ctx, m)) || (s = ddst.alloc(ctx, m))) return s;
    WG_HIP(ctx, hipMemcpyAsync(dsrc.p, sources, m * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(ddst.p, targets, m * 4, hipMemcpyHostToDevice, st));
    if (relations) {
      if ((s = drel.alloc(ctx, m))) return s;
      WG_HIP(ctx, hipMemcpyAsync(drel.p, relations, m * 4, hipMemcpyHostToDevice, st));
    }
    if (weights) {
      if ((s = dw.alloc(ctx, m))) return s;
      WG_HIP(ctx, hipMemcpyAsync(dw.p, weights, m * 8, hipMemcpyHostToDevice, st));
    }
    src = dsrc.p;
    dst = ddst.p;
    rel = relations ? drel.p : nullptr;
    w = weights ? dw.p : nullptr;
  }
  if (m) {
    Scratch<unsigned long long> first;
    if ((s = first.alloc(ctx, 1))) return s;
    WG_HIP(ctx, hipMemsetAsync(first.p, 0xff, 8, st));
    first_bad_edge<<<grid_of(m), kThreads, 0, st>>>(src, dst, rel, m, n, relation_count, first.p);
    WG_HIP(ctx, hipGetLastError());
    unsigned long long bad = 0;
    WG_HIP(ctx, hipMemcpyAsync(&bad, first.p, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    if (bad != ~0ull) {
      uint32_t e3[3] = {0, 0, 0};
      WG_HIP(ctx, hipMemcpy(&e3[0], src + bad, 4, hipMemcpyDeviceToHost));
      WG_HIP(ctx, hipMemcpy(&e3[1], dst + bad, 4, hipMemcpyDeviceToHost));
      if (rel) WG_HIP(ctx, hipMemcpy(&e3[2], rel + bad, 4, hipMemcpyDeviceToHost));
      if (e3[0] >= n || e3[1] >= n)
        return fail(ctx, WG_EDGE_OUTSIDE_ID_SPACE,
                    "edge " + std::to_string(bad) + " from " + std::to_string(e3[0]) + " to " +
                        std::to_string(e3[1]) + " lies outside the id space " + std::to_string(n));
      return fail(ctx, WG_RELATION_OUTSIDE_DICTIONARY,
                  "edge " + std::to_string(bad) + " names relation " + std::to_string(e3[2]) +
                      " outside a dictionary of " + std::to_string(relation_count));
    }
  }
  wg_graph* g = new (std::nothrow) wg_graph();
  if (!g) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  g->ctx = ctx;
  g->n = n;
  g->weighted = weights != nullptr;
  uint64_t folded = 0;
  // Two relations or more: each relation's CSR pair, then the merged one
  // (GraphCsr::build stores both, graph_csr.rs:194-202).
  if (rel && relation_count >= 2) {
    g->relation_count = relation_count;
    s = m < (1ull << 32) ? build_relations<uint32_t>(ctx, g, n, m, src, dst, w, rel, relation_count)
                         : build_relations<uint64_t>(ctx, g, n, m, src, dst, w, rel, relation_count);
    if (s) {
      destroy_graph(g);
      return s;
    }
  }
  s = m < (1ull << 32) ? build_forward<uint32_t>(ctx, n, m, src, dst, w, rel, relation_count, g->fwd, &folded)
                       : build_forward<uint64_t>(ctx, n, m, src, dst, w, rel, relation_count, g->fwd, &folded);
  // The edge list copies are no longer needed.
  dsrc.reset();
  ddst.reset();
  drel.reset();
  dw.reset();
  if (!s) {
    g->nnz = g->fwd.nnz;
    s = adopt_device_csr(ctx, g->fwd);
  }
  if (!s) s = build_reverse(ctx, g);
  if (s) {
    destroy_graph(g);
    return s;
  }
  if (deduplicated) *deduplicated = folded;
  ctx->graphs.insert(g);
  *out = g;
  return ok(ctx);
}
