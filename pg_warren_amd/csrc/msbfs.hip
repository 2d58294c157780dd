// Multi-source BFS with one bit per source (compute/centrality/
// multi_source.rs:34-193, which packs 64 sources in a word; here up to 512
// in eight u64 words per node).
//
// Every level is a pull over the scanned (reverse) CSR: a node not yet
// reached by every source ORs the frontier words of its in-neighbours, and
// the bits it did not have are the sources that reach it at this distance:
//   next[v] = (OR_{u -> v} frontier[u]) & ~seen[v];  seen[v] |= next[v].
// Eight lanes hold a node's eight words. The lanes first read eight
// consecutive entries of the row (one each, coalesced) and test each
// entry's bit in the one-bit "frontier is non-empty" map; then, entry by
// entry, each lane ORs its word of that neighbour's frontier (a coalesced
// 64-byte read), skipped for inactive neighbours -- eight independent loads
// in flight per lane. A 256-thread workgroup holds 32 nodes, exactly one word
// of the next non-empty map, which it stores whole. Rows longer than
// kMsShort entries are cut into pieces of kMsPiece entries, a workgroup each,
// whose ORs meet in a per-row accumulator (an atomic OR on distinct words)
// that msbfs_settle_long then settles. The result of a level is a pure
// function of the previous one, so runs are deterministic.
#include <algorithm>
#include <vector>

#include "wg_internal.hpp"

namespace wg {

constexpr uint32_t kMsWords = 8;     // 512 sources
constexpr uint32_t kMsShort = 128;   // rows up to this many entries: 8 lanes each
constexpr uint32_t kMsPiece = 1024;  // entries per workgroup for longer rows

struct MsArgs {
  const uint64_t* roff;
  const uint32_t* rtgt;
  uint64_t n;
  uint32_t nwords;          // words per node in use (ceil(ns / 64))
  const uint64_t* front;    // n x 8
  uint64_t* next;           // n x 8
  uint64_t* seen;           // n x 8
  const uint32_t* act;      // frontier non-empty, one bit per node
  uint32_t* act_next;
  const uint64_t* full;     // the words of a node every source has reached
  uint32_t* depths;         // n x ns or null
  uint32_t ns;
  unsigned long long* farness;  // n or null: sum of distances
  uint32_t* reach;              // n or null: sources reaching the node
  uint32_t* changed;            // set to 1 by any node that gained a source
  unsigned long long* long_acc; // split rows: 8 words each, the OR of their pieces
  uint32_t d1;
};

__device__ __forceinline__ bool ms_all_done(const MsArgs& a, uint64_t v) {
  bool done = true;
  for (uint32_t j = 0; j < a.nwords; ++j) done = done && a.seen[v * kMsWords + j] == a.full[j];
  return done;
}

// ORs into acc (lane k's word) the frontier words of the entries
// [i0, i1) of the scanned row, eight entries at a time per 8-lane group.
__device__ __forceinline__ uint64_t ms_gather(const MsArgs& a, uint64_t i0, uint64_t i1,
                                              uint64_t stride, uint32_t k, bool lane_on) {
  uint64_t acc = 0;
  for (uint64_t base = i0; base < i1; base += stride) {
    const uint64_t i = base + k;
    uint32_t u = 0, on = 0;
    if (i < i1) {
      u = a.rtgt[i];
      on = (a.act[u >> 5] >> (u & 31)) & 1u;
    }
    uint64_t w[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t uj = __shfl(u, j, 8);
      const uint32_t oj = __shfl(on, j, 8);
      w[j] = (oj && lane_on) ? a.front[(uint64_t)uj * kMsWords + k] : 0;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) acc |= w[j];
  }
  return acc;
}

// Takes a node's gain at this level (lane k's word of acc) and reports
// whether the lane gained a source; returns the lane's popcount in c.
__device__ __forceinline__ bool ms_settle(const MsArgs& a, uint64_t v, uint32_t k, uint64_t acc,
                                          uint32_t& c) {
  const uint64_t seen = a.seen[v * kMsWords + k];
  const uint64_t nw = acc & ~seen;
  a.next[v * kMsWords + k] = nw;
  c = __popcll(nw);
  if (!nw) return false;
  a.seen[v * kMsWords + k] = seen | nw;
  if (a.depths) {
    uint64_t b = nw;
    while (b) {
      const uint32_t s = k * 64 + (__ffsll((long long)b) - 1);
      b &= b - 1;
      a.depths[v * a.ns + s] = a.d1;
    }
  }
  return true;
}

__device__ __forceinline__ void ms_account(const MsArgs& a, uint64_t v, uint32_t k, uint32_t c) {
  c += __shfl_xor(c, 1, 8);
  c += __shfl_xor(c, 2, 8);
  c += __shfl_xor(c, 4, 8);
  if (k == 0 && c) {
    if (a.reach) a.reach[v] += c;
    if (a.farness) a.farness[v] += (unsigned long long)c * a.d1;
  }
}

// Rows up to kMsShort entries: 8 lanes per node, 32 nodes per workgroup.
// The eight lanes of a node share its control flow (lanes past the sources
// in use only skip their memory accesses), so the shuffles see the group.
__global__ __launch_bounds__(256) void msbfs_short(MsArgs a) {
  __shared__ uint32_t word_bits;
  const uint32_t tid = threadIdx.x, k = tid & 7;
  const uint64_t v = (uint64_t)blockIdx.x * 32 + (tid >> 3);
  if (tid == 0) word_bits = 0;
  __syncthreads();
  bool any = false;
  if (v < a.n) {
    const uint64_t b = a.roff[v], e = a.roff[v + 1];
    if (e - b <= kMsShort) {
      const bool lane_on = k < a.nwords;
      const uint64_t acc = ms_all_done(a, v) ? 0 : ms_gather(a, b, e, 8, k, lane_on);
      uint32_t c = 0;
      if (lane_on) any = ms_settle(a, v, k, acc, c);
      ms_account(a, v, k, c);
    }
  }
  const unsigned long long m = __ballot(any);
  if ((tid & 63) == 0) {
    uint32_t bits = 0;
    for (int j = 0; j < 8; ++j)
      if ((m >> (8 * j)) & 0xffull) bits |= 1u << j;
    atomicOr(&word_bits, bits << (((tid >> 6) & 3) * 8));
  }
  __syncthreads();
  if (tid == 0) {
    a.act_next[blockIdx.x] = word_bits;
    if (word_bits) *a.changed = 1;
  }
}

// One piece of a split row: 32 groups of 8 lanes take the piece's entries
// 256 at a time; the groups' ORs meet in LDS, then in the row's accumulator.
__global__ __launch_bounds__(256) void msbfs_piece(MsArgs a, const uint32_t* __restrict__ pieces) {
  __shared__ uint64_t part[32][kMsWords];
  const uint32_t tid = threadIdx.x, k = tid & 7, g = tid >> 3;
  const uint32_t row_idx = pieces[2 * blockIdx.x], piece = pieces[2 * blockIdx.x + 1];
  const uint32_t v = pieces[2 * gridDim.x + row_idx];  // the rows follow the pairs
  const uint64_t b = a.roff[v], e = a.roff[v + 1];
  const uint64_t p0 = b + (uint64_t)piece * kMsPiece;
  const uint64_t p1 = p0 + kMsPiece < e ? p0 + kMsPiece : e;
  uint64_t acc = 0;
  if (!ms_all_done(a, v)) acc = ms_gather(a, p0 + g * 8, p1, 256, k, k < a.nwords);
  part[g][k] = acc;
  __syncthreads();
  if (tid < kMsWords && k < a.nwords) {
    uint64_t all = 0;
    for (int j = 0; j < 32; ++j) all |= part[j][k];
    if (all) atomicOr(&a.long_acc[(uint64_t)row_idx * kMsWords + k], (unsigned long long)all);
  }
}

// Settles the split rows from their accumulators (8 lanes a row).
__global__ __launch_bounds__(256) void msbfs_settle_long(MsArgs a, const uint32_t* __restrict__ rows,
                                                         uint32_t nrows) {
  const uint32_t tid = blockIdx.x * 256 + threadIdx.x, k = threadIdx.x & 7;
  const uint32_t r = tid >> 3;
  bool any = false;
  uint64_t v = 0;
  if (r < nrows) {
    v = rows[r];
    const bool lane_on = k < a.nwords;
    uint32_t c = 0;
    if (lane_on) {
      const uint64_t acc = a.long_acc[(uint64_t)r * kMsWords + k];
      any = ms_settle(a, v, k, acc, c);
    }
    ms_account(a, v, k, c);
  }
  // One lane per row publishes the row's bit.
  const unsigned long long m = __ballot(any);
  const uint32_t lane = threadIdx.x & 63;
  if (r < nrows && k == 0 && ((m >> (lane & ~7u)) & 0xffull)) {
    atomicOr(&a.act_next[v >> 5], 1u << (v & 31));
    *a.changed = 1;
  }
}

__global__ __launch_bounds__(256) void msbfs_init(const uint32_t* __restrict__ sources, uint32_t ns,
                                                  uint64_t* __restrict__ front, uint64_t* __restrict__ seen,
                                                  uint32_t* __restrict__ act, uint32_t* __restrict__ depths) {
  for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < ns; s += gridDim.x * 256) {
    const uint64_t v = sources[s];
    // Sources may repeat: each bit is its own source.
    atomicOr((unsigned long long*)&front[v * kMsWords + s / 64], 1ull << (s % 64));
    atomicOr((unsigned long long*)&seen[v * kMsWords + s / 64], 1ull << (s % 64));
    atomicOr(&act[v >> 5], 1u << (v & 31));
    if (depths) depths[v * ns + s] = 0;
  }
}

}  // namespace wg

using namespace wg;

extern "C" int wg_msbfs(wg_context* ctx, const wg_graph* gc, int direction, const uint64_t* sources,
                        uint32_t nsources, int64_t max_depth, wg_vector* depths, wg_vector* reach,
                        wg_vector* farness, uint32_t* levels) {
  if (!ctx || !gc || !sources) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n;
  if (direction != WG_OUT && direction != WG_IN) return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  if (nsources == 0 || nsources > 64 * kMsWords)
    return fail(ctx, WG_INVALID_ARGUMENT, "between 1 and 512 sources");
  for (uint32_t s = 0; s < nsources; ++s)
    if (sources[s] >= n)
      return fail(ctx, WG_NODE_OUTSIDE_ID_SPACE, "node " + std::to_string(sources[s]) + " outside the id space");
  if (depths && (depths->dtype != WG_U32 || depths->id_space != n || depths->width != nsources))
    return fail(ctx, WG_INVALID_ARGUMENT, "depths must be a u32 vector over the graph, one column per source");
  if (reach && (reach->dtype != WG_U32 || reach->id_space != n || reach->width != 1))
    return fail(ctx, WG_INVALID_ARGUMENT, "reach must be a u32 vector over the graph");
  if (farness && (farness->dtype != WG_I64 || farness->id_space != n || farness->width != 1))
    return fail(ctx, WG_INVALID_ARGUMENT, "farness must be an i64 vector over the graph");
  if (levels) *levels = 0;
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // Pull over the reverse of the adjacency followed.
  DevCsr& scan = direction == WG_IN ? g->fwd : g->rev;
  const uint32_t nwords = (nsources + 63) / 64;
  const uint64_t nblk = std::max<uint64_t>(1, (n + 31) / 32);
  // Split rows and their pieces: (row index, piece) pairs, then the rows.
  std::vector<uint32_t> lrows, pieces;
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t deg = scan.h_offsets[v + 1] - scan.h_offsets[v];
    if (deg <= kMsShort) continue;
    const uint32_t np = (uint32_t)((deg + kMsPiece - 1) / kMsPiece);
    for (uint32_t p = 0; p < np; ++p) {
      pieces.push_back((uint32_t)lrows.size());
      pieces.push_back(p);
    }
    lrows.push_back((uint32_t)v);
  }
  const uint32_t npieces = (uint32_t)(pieces.size() / 2);
  pieces.insert(pieces.end(), lrows.begin(), lrows.end());
  std::vector<std::pair<void*, size_t>> held;
  struct Free {
    wg_context* ctx;
    std::vector<std::pair<void*, size_t>>& h;
    ~Free() {
      (void)hipStreamSynchronize(ctx->stream);
      for (auto& x : h) dfree(ctx, x.first, x.second);
    }
  } release{ctx, held};
  auto get = [&](void** p, size_t bytes) {
    int s = dalloc(ctx, p, bytes);
    if (!s) held.push_back({*p, bytes});
    return s;
  };
  uint64_t *front, *next, *seen, *full;
  unsigned long long* long_acc;
  uint32_t *act, *act_next, *src_d, *changed, *pieces_d, *rows_d;
  const size_t words = std::max<uint64_t>(n, 1) * kMsWords * 8;
  int s;
  if ((s = get((void**)&front, words)) || (s = get((void**)&next, words)) || (s = get((void**)&seen, words)) ||
      (s = get((void**)&full, kMsWords * 8)) || (s = get((void**)&act, nblk * 4)) ||
      (s = get((void**)&act_next, nblk * 4)) || (s = get((void**)&src_d, nsources * 4)) ||
      (s = get((void**)&changed, 4)) ||
      (s = get((void**)&long_acc, std::max<size_t>(lrows.size(), 1) * kMsWords * 8)) ||
      (s = get((void**)&pieces_d, std::max<size_t>(pieces.size(), 1) * 4)))
    return s;
  rows_d = pieces_d + 2 * npieces;
  std::vector<uint32_t> src(nsources);
  for (uint32_t i = 0; i < nsources; ++i) src[i] = (uint32_t)sources[i];
  std::vector<uint64_t> hfull(kMsWords, 0);
  for (uint32_t i = 0; i < nsources; ++i) hfull[i / 64] |= 1ull << (i % 64);
  WG_HIP(ctx, hipMemsetAsync(front, 0, words, st));
  WG_HIP(ctx, hipMemsetAsync(next, 0, words, st));
  WG_HIP(ctx, hipMemsetAsync(seen, 0, words, st));
  WG_HIP(ctx, hipMemsetAsync(act, 0, nblk * 4, st));
  WG_HIP(ctx, hipMemcpyAsync(full, hfull.data(), kMsWords * 8, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemcpyAsync(src_d, src.data(), nsources * 4, hipMemcpyHostToDevice, st));
  if (!pieces.empty())
    WG_HIP(ctx, hipMemcpyAsync(pieces_d, pieces.data(), pieces.size() * 4, hipMemcpyHostToDevice, st));
  if (depths) WG_HIP(ctx, hipMemsetAsync(depths->data, 0xff, depths->bytes, st));
  if (farness) WG_HIP(ctx, hipMemsetAsync(farness->data, 0, farness->bytes, st));
  if (reach) {
    // Distance 0: each source reaches its own node (multi_source.rs:92-116).
    std::vector<uint32_t> r0(std::max<uint64_t>(n, 1), 0);
    for (uint32_t v : src) r0[v]++;
    WG_HIP(ctx, hipMemcpyAsync(reach->data, r0.data(), n * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
  }
  msbfs_init<<<(nsources + 255) / 256, 256, 0, st>>>(src_d, nsources, front, seen, act,
                                                     depths ? (uint32_t*)depths->data : nullptr);
  MsArgs a{scan.offsets, scan.targets, n, nwords, front, next, seen, act, act_next, full,
           depths ? (uint32_t*)depths->data : nullptr, nsources,
           farness ? (unsigned long long*)farness->data : nullptr,
           reach ? (uint32_t*)reach->data : nullptr, changed, long_acc, 0};
  uint32_t d = 0;
  while (!(max_depth >= 0 && (int64_t)d >= max_depth)) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    a.d1 = d + 1;
    WG_HIP(ctx, hipMemsetAsync(changed, 0, 4, st));
    msbfs_short<<<(unsigned)nblk, 256, 0, st>>>(a);
    if (npieces) {
      WG_HIP(ctx, hipMemsetAsync(long_acc, 0, lrows.size() * kMsWords * 8, st));
      msbfs_piece<<<npieces, 256, 0, st>>>(a, pieces_d);
      msbfs_settle_long<<<(unsigned)((lrows.size() * 8 + 255) / 256), 256, 0, st>>>(a, rows_d,
                                                                                    (uint32_t)lrows.size());
    }
    WG_HIP(ctx, hipGetLastError());
    uint32_t hc = 0;
    WG_HIP(ctx, hipMemcpyAsync(&hc, changed, 4, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    uint64_t* f = const_cast<uint64_t*>(a.front);
    a.front = a.next;
    a.next = f;
    uint32_t* t = const_cast<uint32_t*>(a.act);
    a.act = a.act_next;
    a.act_next = t;
    ++d;
    if (!hc) break;
  }
  if (levels) *levels = d;
  return ok(ctx);
}
