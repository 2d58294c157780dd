// Level-synchronous, direction-optimising BFS on gfx950
// (traversal/level.rs:88-337, bfs.rs:34-111, reachable.rs:79-90).
//
// Each level is pushed or pulled by Beamer's switch with the reference's
// constants (alpha = 1, beta = 24), decided on the host from three counters
// the kernels keep on the device: the next frontier's size, its row entries
// (m_f) and the reverse-row entries of every visited node (so the unvisited
// nodes' entries m_u follow by subtraction).
//   push  a wavefront per frontier node, lanes striding its row; an unvisited
//         target takes atomicMin(via, frontier node) and is claimed once by
//         an atomicOr on the next-frontier bitmap, then appended to the next
//         queue with one wave-aggregated atomic.
//   pull  over the scanned CSR's row blocks (the SpMV partition): a group of
//         T lanes per unvisited row reads T consecutive entries at a time
//         (coalesced), and the lowest lane whose entry is in the frontier
//         bitmap (a group ballot) gives the first frontier hit in ascending
//         order; rows longer than a block search their pieces in parallel
//         and bfs_pull_long keeps the first piece that hit. The block's
//         next-frontier bits are gathered in LDS and written word by word.
// Either way a node's via is its smallest frontier predecessor, which is
// what the reference reports in both directions, its depth is exact, and
// the row entries scanned are counted as the reference counts them.
#include <algorithm>
#include <numeric>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

constexpr uint32_t kUnvisited = 0xffffffffu;

struct BfsCounters {
  unsigned long long next_size;    // nodes claimed this level
  unsigned long long next_entries; // their rows' entries (m_f of the next level)
  unsigned long long visited_rev;  // reverse-row entries of the nodes claimed
  unsigned long long edges;        // row entries scanned
};

__device__ __forceinline__ bool bit_of(const uint32_t* bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Appends the lanes with `take` set to the queue with one atomic per wave.
__device__ __forceinline__ void wave_append(bool take, uint32_t v, uint32_t* queue,
                                            unsigned long long* size) {
  const unsigned long long mask = __ballot(take);
  if (!mask) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t leader = __ffsll((long long)mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(size, (unsigned long long)__popcll(mask));
  base = __shfl(base, leader, 64);
  if (take) {
    const uint32_t rank = __popcll(mask & ((1ull << lane) - 1));
    queue[base + rank] = v;
  }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  return v;
}

__device__ __forceinline__ void add_counters(BfsCounters* c, unsigned long long edges,
                                             unsigned long long entries, unsigned long long vrev) {
  edges = wave_sum(edges);
  entries = wave_sum(entries);
  vrev = wave_sum(vrev);
  if ((threadIdx.x & 63) == 0) {
    if (edges) atomicAdd(&c->edges, edges);
    if (entries) atomicAdd(&c->next_entries, entries);
    if (vrev) atomicAdd(&c->visited_rev, vrev);
  }
}

__global__ __launch_bounds__(256) void bfs_seed(const uint32_t* __restrict__ starts, uint32_t ns,
                                                uint32_t* __restrict__ depth,
                                                uint32_t* __restrict__ visited,
                                                uint32_t* __restrict__ frontier,
                                                const uint64_t* __restrict__ adj_off,
                                                const uint64_t* __restrict__ rev_off,
                                                BfsCounters* c) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ns; i += gridDim.x * 256) {
    const uint32_t s = starts[i];
    depth[s] = 0;
    atomicOr(&visited[s >> 5], 1u << (s & 31));
    atomicOr(&frontier[s >> 5], 1u << (s & 31));
    atomicAdd(&c->next_entries, (unsigned long long)(adj_off[s + 1] - adj_off[s]));
    atomicAdd(&c->visited_rev, (unsigned long long)(rev_off[s + 1] - rev_off[s]));
  }
}

__global__ __launch_bounds__(256) void bfs_push(
    const uint32_t* __restrict__ queue, uint64_t qn, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tgt, const uint64_t* __restrict__ rev_off,
    const uint32_t* __restrict__ visited, const uint8_t* __restrict__ filter,
    uint32_t* __restrict__ depth, uint32_t* __restrict__ via, uint32_t* __restrict__ next_bm,
    uint32_t* __restrict__ next_queue, uint32_t d1, BfsCounters* c) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint64_t nwaves = (uint64_t)gridDim.x * 4;
  unsigned long long edges = 0, entries = 0, vrev = 0;
  for (uint64_t q = wave; q < qn; q += nwaves) {
    const uint32_t u = queue[q];
    const uint64_t a = off[u], b = off[u + 1];
    if (lane == 0) edges += b - a;
    for (uint64_t base = a; base < b; base += 64) {
      const uint64_t i = base + lane;
      bool take = false;
      uint32_t v = 0;
      if (i < b) {
        v = tgt[i];
        if (!bit_of(visited, v) && !(filter && !filter[v])) {
          atomicMin(&via[v], u);
          const uint32_t bit = 1u << (v & 31);
          const uint32_t old = atomicOr(&next_bm[v >> 5], bit);
          if (!(old & bit)) {
            take = true;
            depth[v] = d1;
            entries += off[v + 1] - off[v];
            vrev += rev_off[v + 1] - rev_off[v];
          }
        }
      }
      wave_append(take, v, next_queue, &c->next_size);
    }
  }
  add_counters(c, edges, entries, vrev);
}

// Pull over the scanned CSR's row blocks. `off` are the adjacency's offsets
// (the found node's next-level entries), roff/rtgt the scanned (reverse)
// CSR. Pieces of split rows record their first hit (offset within the row)
// or ~0 in piece_hit for bfs_pull_long.
__global__ __launch_bounds__(256) void bfs_pull_blocks(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ off,
    const uint64_t* __restrict__ roff, const uint32_t* __restrict__ rtgt,
    const uint32_t* __restrict__ visited, const uint32_t* __restrict__ frontier,
    const uint8_t* __restrict__ filter, uint32_t* __restrict__ depth, uint32_t* __restrict__ via,
    uint32_t* __restrict__ next_bm, uint32_t* __restrict__ next_queue, uint32_t* __restrict__ piece_hit,
    uint32_t d1, BfsCounters* c) {
  __shared__ uint32_t lbm[kBlockRows / 32 + 2];
  __shared__ uint32_t whit[kBlockThreads / 64];
  const Block b = blocks[blockIdx.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  unsigned long long edges = 0, entries = 0, vrev = 0;
  if (b.nrows) {
    const uint32_t w0 = b.row0 >> 5, nw = ((b.row0 + b.nrows - 1) >> 5) - w0 + 1;
    for (uint32_t k = tid; k < nw; k += kBlockThreads) lbm[k] = 0;
    __syncthreads();
    const uint32_t nr = b.nrows;
    uint32_t T = 1;
    if (nr < kBlockThreads) {
      uint32_t q = kBlockThreads / nr;
      T = 1u << (31 - __builtin_clz(q));
      if (T > 64) T = 64;
    }
    const uint32_t groups = kBlockThreads / T, g = tid / T, l = tid % T;
    const uint32_t gbase = lane & ~(T - 1);
    const unsigned long long gmask = T == 64 ? ~0ull : ((1ull << T) - 1);
    for (uint32_t r = g; r < nr; r += groups) {
      const uint32_t v = b.row0 + r;
      bool take = false;
      if (!bit_of(visited, v) && !(filter && !filter[v])) {
        const uint64_t a = roff[v], e = roff[v + 1];
        bool found = false;
        for (uint64_t base = a; base < e; base += T) {
          const uint64_t i = base + l;
          bool hit = false;
          uint32_t u = 0;
          if (i < e) {
            u = rtgt[i];
            hit = bit_of(frontier, u);
          }
          const unsigned long long gm = (__ballot(hit) >> gbase) & gmask;
          if (gm) {
            const uint32_t first = __ffsll((long long)gm) - 1;
            if (l == first) {
              via[v] = u;
              depth[v] = d1;
              take = true;
              atomicOr(&lbm[(v >> 5) - w0], 1u << (v & 31));
              entries += off[v + 1] - off[v];
              vrev += e - a;
            }
            if (l == 0) edges += (base - a) + first + 1;
            found = true;
            break;
          }
        }
        if (!found && l == 0) edges += e - a;
      }
      wave_append(take, v, next_queue, &c->next_size);
    }
    __syncthreads();
    // Words wholly inside the block are its own; the two edge words may be
    // shared with the neighbouring blocks.
    for (uint32_t k = tid; k < nw; k += kBlockThreads) {
      const uint32_t word = w0 + k, bits = lbm[k];
      const bool whole = word * 32 >= b.row0 && word * 32 + 31 < b.row0 + nr;
      if (whole)
        next_bm[word] = bits;
      else if (bits)
        atomicOr(&next_bm[word], bits);
    }
  } else {
    const uint32_t v = b.row0;
    uint32_t hit_at = ~0u;
    if (!bit_of(visited, v) && !(filter && !filter[v])) {
      const uint64_t a = roff[v], re = roff[v + 1];
      const uint64_t e0 = a + (uint64_t)b.piece * kBlockNnz;
      const uint64_t e1 = e0 + kBlockNnz < re ? e0 + kBlockNnz : re;
      for (uint64_t base = e0; base < e1; base += kBlockThreads) {
        const uint64_t i = base + tid;
        const bool hit = i < e1 && bit_of(frontier, rtgt[i]);
        const unsigned long long m = __ballot(hit);
        if (lane == 0) whit[tid >> 6] = m ? (uint32_t)(tid + __ffsll((long long)m) - 1) : ~0u;
        __syncthreads();
        uint32_t best = ~0u;
        for (int k = 0; k < (int)(kBlockThreads / 64); ++k) best = whit[k] < best ? whit[k] : best;
        __syncthreads();
        if (best != ~0u) {
          hit_at = (uint32_t)(base - a) + best;
          break;
        }
      }
    }
    if (tid == 0) piece_hit[b.pbase + b.piece] = hit_at;
  }
  add_counters(c, edges, entries, vrev);
}

// A split row found in the pull: the first piece that met the frontier.
__global__ __launch_bounds__(256) void bfs_pull_long(
    const uint32_t* __restrict__ long_rows, const uint32_t* __restrict__ long_pbase, uint32_t nlong,
    const uint64_t* __restrict__ off, const uint64_t* __restrict__ roff,
    const uint32_t* __restrict__ rtgt, const uint32_t* __restrict__ visited,
    const uint8_t* __restrict__ filter, const uint32_t* __restrict__ piece_hit,
    uint32_t* __restrict__ depth, uint32_t* __restrict__ via, uint32_t* __restrict__ next_bm,
    uint32_t* __restrict__ next_queue, uint32_t d1, BfsCounters* c) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long edges = 0, entries = 0, vrev = 0;
  bool take = false;
  uint32_t v = 0;
  if (i < nlong) {
    v = long_rows[i];
    if (!bit_of(visited, v) && !(filter && !filter[v])) {
      const uint64_t a = roff[v], e = roff[v + 1];
      uint32_t at = ~0u;
      for (uint32_t p = long_pbase[i]; p < long_pbase[i + 1] && at == ~0u; ++p) at = piece_hit[p];
      if (at != ~0u) {
        via[v] = rtgt[a + at];
        depth[v] = d1;
        atomicOr(&next_bm[v >> 5], 1u << (v & 31));
        take = true;
        edges = (unsigned long long)at + 1;
        entries = off[v + 1] - off[v];
        vrev = e - a;
      } else {
        edges = e - a;
      }
    }
  }
  wave_append(take, v, next_queue, &c->next_size);
  add_counters(c, edges, entries, vrev);
}

// visited |= next; the old frontier bitmap is cleared to become the next
// level's claim bitmap.
__global__ __launch_bounds__(256) void bfs_merge(uint64_t words, uint32_t* __restrict__ visited,
                                                 const uint32_t* __restrict__ next_bm,
                                                 uint32_t* __restrict__ old_frontier) {
  for (uint64_t w = (uint64_t)blockIdx.x * 256 + threadIdx.x; w < words; w += (uint64_t)gridDim.x * 256) {
    visited[w] |= next_bm[w];
    old_frontier[w] = 0;
  }
}

__global__ void bfs_clear_counters(BfsCounters* c) {
  c->next_size = 0;
  c->next_entries = 0;
  c->visited_rev = 0;
  c->edges = 0;
}

struct BfsRun {
  std::vector<uint32_t> level_pulled;  // per expanded level: 1 if pulled
  std::vector<uint64_t> level_size;    // nodes admitted at each depth (0 = starts)
  uint64_t visited = 0, edges = 0;
  uint32_t levels = 0;
};

// The search's workspace, kept with the graph: three bitmaps, two queues,
// the counters, the split rows' piece hits and a starts buffer.
struct BfsWorkspace {
  uint32_t *visited, *bm_a, *bm_b, *q_a, *q_b, *piece_hit, *starts;
  BfsCounters* cnt;
};

static int bfs_workspace(wg_context* ctx, wg_graph* g, uint64_t words, BfsWorkspace& w) {
  if (!g->bfs_ws[0]) {
    const uint64_t n = std::max<uint64_t>(g->n, 1);
    const size_t sizes[8] = {words * 4, words * 4, words * 4, n * 4, n * 4,
                             std::max<uint32_t>(std::max(g->fwd.npartials, g->rev.npartials), 1) * 4,
                             n * 4, sizeof(BfsCounters)};
    for (int k = 0; k < 8; ++k) {
      void* p = nullptr;
      int s = dalloc(ctx, &p, sizes[k]);
      if (s) return s;
      g->workspace.push_back({p, sizes[k]});
      g->bfs_ws[k] = p;
    }
  }
  w = BfsWorkspace{(uint32_t*)g->bfs_ws[0], (uint32_t*)g->bfs_ws[1], (uint32_t*)g->bfs_ws[2],
                   (uint32_t*)g->bfs_ws[3], (uint32_t*)g->bfs_ws[4], (uint32_t*)g->bfs_ws[5],
                   (uint32_t*)g->bfs_ws[6], (BfsCounters*)g->bfs_ws[7]};
  return WG_OK;
}

// Runs the search, leaving depth and via (kUnvisited where unreached) in
// device buffers. starts must be in range and deduplicated, in order.
// Stops once more than `stop_after` nodes are visited (to cut a search
// whose output is truncated anyway) unless stop_after < 0.
int bfs_run(wg_context* ctx, wg_graph* g, int direction, const std::vector<uint32_t>& starts,
            const uint8_t* filter_d, int64_t max_depth, int64_t stop_after, bool pulling,
            uint32_t* depth, uint32_t* via, BfsRun& run) {
  const uint64_t n = g->n;
  DevCsr& adj = direction == WG_IN ? g->rev : g->fwd;
  DevCsr& rev = direction == WG_IN ? g->fwd : g->rev;
  hipStream_t st = ctx->stream;
  const uint64_t words = (n + 63) / 64 * 2;  // 32-bit words, whole 64-node waves
  BfsWorkspace w;
  int s = bfs_workspace(ctx, g, words, w);
  if (s) return s;
  WG_HIP(ctx, hipMemsetAsync(depth, 0xff, std::max<uint64_t>(n, 1) * 4, st));
  WG_HIP(ctx, hipMemsetAsync(via, 0xff, std::max<uint64_t>(n, 1) * 4, st));
  WG_HIP(ctx, hipMemsetAsync(w.visited, 0, words * 4, st));
  WG_HIP(ctx, hipMemsetAsync(w.bm_a, 0, words * 4, st));
  WG_HIP(ctx, hipMemsetAsync(w.bm_b, 0, words * 4, st));
  if (!starts.empty()) {
    WG_HIP(ctx, hipMemcpyAsync(w.starts, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(w.q_a, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, st));
  }
  bfs_clear_counters<<<1, 1, 0, st>>>(w.cnt);
  uint32_t* frontier = w.bm_a;  // bitmap of the current frontier
  uint32_t* claim = w.bm_b;     // bitmap of the nodes claimed this level
  if (!starts.empty())
    bfs_seed<<<(unsigned)std::min<size_t>((starts.size() + 255) / 256, 1024), 256, 0, st>>>(
        w.starts, (uint32_t)starts.size(), depth, w.visited, frontier, adj.offsets, rev.offsets, w.cnt);
  BfsCounters hc{};
  WG_HIP(ctx, hipMemcpyAsync(&hc, w.cnt, sizeof hc, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  uint64_t qn = starts.size(), mf = hc.next_entries, visited_rev = hc.visited_rev;
  run.visited = qn;
  run.level_size.assign(1, qn);
  uint32_t d = 0;
  uint32_t *q_cur = w.q_a, *q_next = w.q_b;
  while (qn > 0 && !(max_depth >= 0 && (int64_t)d >= max_depth) &&
         !(stop_after >= 0 && run.visited > (uint64_t)stop_after)) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    // Beamer's switch (level.rs:184-211): pull iff |F| * 24 > n and the
    // unvisited nodes' reverse entries are fewer than alpha * m_f.
    const uint64_t m_u = rev.nnz - visited_rev;
    const bool pull = pulling && qn * 24 > n && m_u < mf;
    bfs_clear_counters<<<1, 1, 0, st>>>(w.cnt);
    if (pull) {
      if (rev.nblocks)
        bfs_pull_blocks<<<rev.nblocks, kBlockThreads, 0, st>>>(
            rev.blocks, adj.offsets, rev.offsets, rev.targets, w.visited, frontier, filter_d, depth, via,
            claim, q_next, w.piece_hit, d + 1, w.cnt);
      if (rev.nlong)
        bfs_pull_long<<<(rev.nlong + 255) / 256, 256, 0, st>>>(
            rev.long_rows, rev.long_pbase, rev.nlong, adj.offsets, rev.offsets, rev.targets, w.visited,
            filter_d, w.piece_hit, depth, via, claim, q_next, d + 1, w.cnt);
    } else {
      const unsigned grid = (unsigned)std::min<uint64_t>((qn + 3) / 4, 256 * 64);
      bfs_push<<<grid, 256, 0, st>>>(q_cur, qn, adj.offsets, adj.targets, rev.offsets, w.visited,
                                     filter_d, depth, via, claim, q_next, d + 1, w.cnt);
    }
    bfs_merge<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((words + 255) / 256, 4096)), 256, 0, st>>>(
        words, w.visited, claim, frontier);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(&hc, w.cnt, sizeof hc, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    std::swap(frontier, claim);
    std::swap(q_cur, q_next);
    qn = hc.next_size;
    mf = hc.next_entries;
    visited_rev += hc.visited_rev;
    run.edges += hc.edges;
    run.visited += qn;
    run.level_pulled.push_back(pull ? 1 : 0);
    run.level_size.push_back(qn);
    ++d;
  }
  run.levels = d;
  return WG_OK;
}

// Starts checked and deduplicated in order (LevelSearch::seed, level.rs:128-147).
static int prepare_starts(wg_context* ctx, uint64_t n, const uint64_t* starts, uint64_t ns,
                          std::vector<uint32_t>& out) {
  for (uint64_t i = 0; i < ns; ++i)
    if (starts[i] >= n)
      return fail(ctx, WG_NODE_OUTSIDE_ID_SPACE, "node " + std::to_string(starts[i]) + " outside the id space");
  std::vector<uint64_t> seen((n + 63) / 64, 0);
  for (uint64_t i = 0; i < ns; ++i) {
    const uint64_t s = starts[i];
    if (seen[s >> 6] >> (s & 63) & 1) continue;
    seen[s >> 6] |= 1ull << (s & 63);
    out.push_back((uint32_t)s);
  }
  return WG_OK;
}

struct DevBytes {
  wg_context* ctx;
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBytes() { dfree(ctx, p, bytes); }
};

}  // namespace wg

using namespace wg;

extern "C" {

int wg_bfs_levels(wg_context* ctx, const wg_graph* gc, int direction, const uint64_t* starts,
                  uint64_t nstarts, const uint8_t* filter, int64_t max_depth, int pulling,
                  wg_vector* depth, wg_vector* via, wg_traversal_stats* stats) {
  if (!ctx || !gc || !depth || !via) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  wg_graph* g = const_cast<wg_graph*>(gc);
  if (direction != WG_OUT && direction != WG_IN) return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  for (wg_vector* v : {depth, via})
    if (v->dtype != WG_U32 || v->width != 1 || v->id_space != g->n)
      return fail(ctx, WG_INVALID_ARGUMENT, "depth and via must be u32 vectors over the graph");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<uint32_t> st;
  int s = prepare_starts(ctx, g->n, starts, nstarts, st);
  if (s) return s;
  DevBytes fb{ctx};
  if (filter) {
    fb.bytes = std::max<uint64_t>(g->n, 1);
    if ((s = dalloc(ctx, &fb.p, fb.bytes))) return s;
    WG_HIP(ctx, hipMemcpy(fb.p, filter, g->n, hipMemcpyHostToDevice));
  }
  BfsRun run;
  s = bfs_run(ctx, g, direction, st, (const uint8_t*)fb.p, max_depth, -1, pulling != 0,
              (uint32_t*)depth->data, (uint32_t*)via->data, run);
  if (s) return s;
  if (stats) {
    uint32_t pulled = 0;
    for (uint32_t p : run.level_pulled) pulled += p;
    *stats = wg_traversal_stats{run.visited, run.edges, 0, pulled};
  }
  return ok(ctx);
}

// The visits in the reference's order: by depth; the starts in the order
// given; a pushed level by (via, node), the order reached when frontier
// nodes expand in ascending order; a pulled level ascending by node.
static int emit(wg_context* ctx, wg_graph* g, int direction, const uint64_t* starts, uint64_t ns,
                const uint8_t* filter, int64_t max_depth, int64_t max_results, bool pulling,
                bool pulled_order, std::vector<wg_visit>& visits, wg_traversal_stats* stats) {
  const uint64_t n = g->n;
  std::vector<uint32_t> st;
  int s = prepare_starts(ctx, n, starts, ns, st);
  if (s) return s;
  DevBytes fb{ctx}, dd{ctx}, vd{ctx};
  if (filter) {
    fb.bytes = std::max<uint64_t>(n, 1);
    if ((s = dalloc(ctx, &fb.p, fb.bytes))) return s;
    WG_HIP(ctx, hipMemcpy(fb.p, filter, n, hipMemcpyHostToDevice));
  }
  dd.bytes = vd.bytes = std::max<uint64_t>(n, 1) * 4;
  if ((s = dalloc(ctx, &dd.p, dd.bytes)) || (s = dalloc(ctx, &vd.p, vd.bytes))) return s;
  BfsRun run;
  s = bfs_run(ctx, g, direction, st, (const uint8_t*)fb.p, max_depth, max_results, pulling,
              (uint32_t*)dd.p, (uint32_t*)vd.p, run);
  if (s) return s;
  std::vector<uint32_t> depth(n), via(n);
  if (n) {
    WG_HIP(ctx, hipMemcpy(depth.data(), dd.p, n * 4, hipMemcpyDeviceToHost));
    WG_HIP(ctx, hipMemcpy(via.data(), vd.p, n * 4, hipMemcpyDeviceToHost));
  }
  // Counting sort by depth (nodes ascending within), then per level the
  // pushed order (via, node) by a stable counting pass over via.
  const uint32_t levels = run.levels;
  std::vector<uint64_t> start_of(levels + 2, 0);
  for (uint64_t v = 0; v < n; ++v)
    if (depth[v] != kUnvisited && depth[v] >= 1 && depth[v] <= levels) start_of[depth[v] + 1]++;
  for (uint32_t l = 1; l < levels + 2; ++l) start_of[l] += start_of[l - 1];
  std::vector<uint32_t> by_depth(start_of[levels + 1]);
  {
    std::vector<uint64_t> cur(start_of);
    for (uint64_t v = 0; v < n; ++v)
      if (depth[v] != kUnvisited && depth[v] >= 1 && depth[v] <= levels) by_depth[cur[depth[v]]++] = (uint32_t)v;
  }
  visits.clear();
  for (uint32_t sv : st) visits.push_back(wg_visit{sv, UINT64_MAX, 0, 0});
  for (uint32_t l = 1; l <= levels; ++l) {
    uint32_t* a = by_depth.data() + start_of[l];
    uint32_t* b = by_depth.data() + start_of[l + 1];
    const bool keep_ascending = pulled_order && run.level_pulled[l - 1];
    if (!keep_ascending)
      std::stable_sort(a, b, [&](uint32_t x, uint32_t y) { return via[x] < via[y]; });
    for (uint32_t* p = a; p < b; ++p) visits.push_back(wg_visit{*p, via[*p], l, 0});
  }
  bool truncated = false;
  if (max_results >= 0 && visits.size() > (uint64_t)max_results) {
    visits.resize(max_results);
    truncated = true;
  }
  if (stats) {
    uint32_t pulled = 0;
    for (uint32_t p : run.level_pulled) pulled += p;
    *stats = wg_traversal_stats{visits.size(), run.edges, truncated ? 1u : 0u, pulled};
  }
  return WG_OK;
}

int wg_bfs(wg_context* ctx, const wg_graph* gc, int direction, const uint64_t* starts,
           uint64_t nstarts, const uint8_t* filter, int64_t max_depth, int64_t max_results,
           wg_visit* out, uint64_t* count, wg_traversal_stats* stats) {
  if (!ctx || !gc || !out || !count) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (direction != WG_OUT && direction != WG_IN) return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<wg_visit> v;
  int s = emit(ctx, const_cast<wg_graph*>(gc), direction, starts, nstarts, filter, max_depth,
               max_results, true, false, v, stats);
  if (s) return s;
  std::copy(v.begin(), v.end(), out);
  *count = v.size();
  return ok(ctx);
}

int wg_reachable_set(wg_context* ctx, const wg_graph* gc, int direction, const uint64_t* starts,
                     uint64_t nstarts, int64_t max_depth, uint64_t* out, uint64_t* count,
                     wg_traversal_stats* stats) {
  if (!ctx || !gc || !out || !count) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (direction != WG_OUT && direction != WG_IN) return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<wg_visit> v;
  int s = emit(ctx, const_cast<wg_graph*>(gc), direction, starts, nstarts, nullptr, max_depth, -1,
               true, true, v, stats);
  if (s) return s;
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].node;
  *count = v.size();
  return ok(ctx);
}

}  // extern "C"
