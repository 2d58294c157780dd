// Level-synchronous, direction-optimising BFS on gfx950
// (traversal/level.rs:88-337, bfs.rs:34-111, reachable.rs:79-90).
//
// Each level is pushed or pulled by Beamer's switch with the reference's
// constants (alpha = 1, beta = 24), decided on the host from counters of
// the frontier: its size, its row entries (m_f) and the reverse-row entries
// of every visited node (so the unvisited nodes' entries m_u follow by
// subtraction).
//   push  bfs_mark: a wavefront per frontier node strides its row and marks
//         each unvisited target a candidate (a plain byte store: every
//         writer writes 1, so no atomic is needed); the pull kernel then
//         resolves only the candidates.
//   pull  bfs_pull_blocks over the scanned (reverse) CSR's row blocks, the
//         SpMV partition: a group of T lanes per row reads T consecutive
//         entries at a time (coalesced), and the lowest lane whose entry is
//         in the frontier bitmap (a group ballot) is the first frontier hit
//         in ascending order -- the smallest frontier predecessor, which is
//         the via the reference reports both ways. Rows longer than a block
//         search their pieces in parallel and bfs_pull_long keeps the first
//         piece that hit. A block's next-frontier bits are gathered in LDS
//         and stored a word at a time.
// The next frontier's queue and counters come from the claim bitmap by a
// count / scan / write compaction with per-block partials. The hot path has
// no device-scope atomics (on gfx950 they resolve past the per-XCD L2 and
// serialise on a shared address).
#include <algorithm>
#include <numeric>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

constexpr uint32_t kUnvisited = 0xffffffffu;
constexpr uint32_t kWordsPerCountBlock = 512;  // bfs_count / bfs_write: 2 words a thread

struct BfsCounters {
  unsigned long long next_size;    // nodes claimed this level
  unsigned long long next_entries; // their rows' entries (m_f of the next level)
  unsigned long long visited_rev;  // reverse-row entries of the nodes claimed
  unsigned long long edges;        // row entries scanned by a pulled level
};

__device__ __forceinline__ bool bit_of(const uint32_t* bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

template <class T>
__device__ __forceinline__ T block_reduce_sum(T v, T* sh) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  const uint32_t tid = threadIdx.x;
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  T t = 0;
  if (tid == 0)
    for (uint32_t k = 0; k < blockDim.x / 64; ++k) t += sh[k];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void bfs_seed(const uint32_t* __restrict__ starts, uint32_t ns,
                                                uint32_t* __restrict__ depth,
                                                uint32_t* __restrict__ visited,
                                                uint32_t* __restrict__ frontier) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < ns; i += gridDim.x * 256) {
    const uint32_t s = starts[i];
    depth[s] = 0;
    atomicOr(&visited[s >> 5], 1u << (s & 31));
    atomicOr(&frontier[s >> 5], 1u << (s & 31));
  }
}

// Push, first half: candidates of the next level.
__global__ __launch_bounds__(256) void bfs_mark(const uint32_t* __restrict__ queue, uint64_t qn,
                                                const uint64_t* __restrict__ off,
                                                const uint32_t* __restrict__ tgt,
                                                const uint32_t* __restrict__ visited,
                                                const uint8_t* __restrict__ filter,
                                                uint8_t* __restrict__ cand) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint64_t nwaves = (uint64_t)gridDim.x * 4;
  for (uint64_t q = wave; q < qn; q += nwaves) {
    const uint32_t u = queue[q];
    const uint64_t a = off[u], b = off[u + 1];
    for (uint64_t i = a + lane; i < b; i += 64) {
      const uint32_t v = tgt[i];
      if (!bit_of(visited, v) && !(filter && !filter[v])) cand[v] = 1;
    }
  }
}

// Pull over the scanned CSR's row blocks (roff/rtgt). candidates == 0: every
// unvisited row admitted by the filter searches; otherwise only rows marked
// in cand (the push's second half), which are cleared. Pieces of split rows
// record their first hit (offset within the row) or ~0 in piece_hit for
// bfs_pull_long. A pulled level's scanned entries go to part_edges.
__global__ __launch_bounds__(256) void bfs_pull_blocks(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ roff,
    const uint32_t* __restrict__ rtgt, const uint32_t* __restrict__ visited,
    const uint32_t* __restrict__ frontier, const uint8_t* __restrict__ filter,
    uint8_t* __restrict__ cand, int candidates, uint32_t* __restrict__ depth,
    uint32_t* __restrict__ via, uint32_t* __restrict__ next_bm, uint32_t* __restrict__ piece_hit,
    unsigned long long* __restrict__ part_edges, uint32_t d1) {
  __shared__ uint32_t lbm[kBlockRows / 32 + 2];
  __shared__ uint32_t whit[kBlockThreads / 64];
  __shared__ unsigned long long red[kBlockThreads / 64];
  const Block b = blocks[blockIdx.x];
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  unsigned long long edges = 0;
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
      bool search;
      if (candidates) {
        search = cand[v] != 0;
      } else {
        search = !bit_of(visited, v) && !(filter && !filter[v]);
      }
      if (search) {
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
              atomicOr(&lbm[(v >> 5) - w0], 1u << (v & 31));
            }
            if (l == 0) edges += (base - a) + first + 1;
            found = true;
            break;
          }
        }
        if (!found && l == 0) edges += e - a;
        if (candidates && l == 0) cand[v] = 0;
      }
    }
    __syncthreads();
    // Words wholly inside the block are its own; the edge words may be
    // shared with neighbouring blocks (or hold a split row's bit).
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
    const bool search = candidates ? cand[v] != 0 : (!bit_of(visited, v) && !(filter && !filter[v]));
    if (search) {
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
  if (!candidates) {
    const unsigned long long t = block_reduce_sum(edges, red);
    if (tid == 0) part_edges[blockIdx.x] = t;
  }
}

// A split row in the pull: the first piece that met the frontier.
__global__ __launch_bounds__(256) void bfs_pull_long(
    const uint32_t* __restrict__ long_rows, const uint32_t* __restrict__ long_pbase, uint32_t nlong,
    const uint64_t* __restrict__ roff, const uint32_t* __restrict__ rtgt,
    const uint32_t* __restrict__ visited, const uint8_t* __restrict__ filter, uint8_t* __restrict__ cand,
    int candidates, const uint32_t* __restrict__ piece_hit, uint32_t* __restrict__ depth,
    uint32_t* __restrict__ via, uint32_t* __restrict__ next_bm,
    unsigned long long* __restrict__ part_edges, uint32_t d1) {
  __shared__ unsigned long long red[4];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long edges = 0;
  if (i < nlong) {
    const uint32_t v = long_rows[i];
    const bool search = candidates ? cand[v] != 0 : (!bit_of(visited, v) && !(filter && !filter[v]));
    if (search) {
      const uint64_t a = roff[v], e = roff[v + 1];
      uint32_t at = ~0u;
      for (uint32_t p = long_pbase[i]; p < long_pbase[i + 1] && at == ~0u; ++p) at = piece_hit[p];
      if (at != ~0u) {
        via[v] = rtgt[a + at];
        depth[v] = d1;
        atomicOr(&next_bm[v >> 5], 1u << (v & 31));
        edges = (unsigned long long)at + 1;
      } else {
        edges = e - a;
      }
      if (candidates) cand[v] = 0;
    }
  }
  if (!candidates) {
    const unsigned long long t = block_reduce_sum(edges, red);
    if (threadIdx.x == 0) part_edges[blockIdx.x] = t;
  }
}

// Compaction, first pass: per block of words, the claimed nodes and their
// forward / reverse row entries; also folds the claims into visited and
// clears the old frontier bitmap (it becomes the next level's claim map).
__global__ __launch_bounds__(256) void bfs_count(uint64_t words, const uint32_t* __restrict__ claim,
                                                 uint32_t* __restrict__ visited,
                                                 uint32_t* __restrict__ old_frontier,
                                                 const uint64_t* __restrict__ off,
                                                 const uint64_t* __restrict__ roff,
                                                 unsigned long long* __restrict__ part) {
  __shared__ unsigned long long red[4];
  unsigned long long cnt = 0, ent = 0, vr = 0;
  const uint64_t w0 = (uint64_t)blockIdx.x * kWordsPerCountBlock + threadIdx.x * 2;
  for (uint64_t w = w0; w < w0 + 2 && w < words; ++w) {
    uint32_t bits = claim[w];
    visited[w] |= bits;
    old_frontier[w] = 0;
    cnt += __popc(bits);
    while (bits) {
      const uint32_t v = (uint32_t)(w * 32) + (__ffs(bits) - 1);
      bits &= bits - 1;
      ent += off[v + 1] - off[v];
      vr += roff[v + 1] - roff[v];
    }
  }
  cnt = block_reduce_sum(cnt, red);
  ent = block_reduce_sum(ent, red);
  vr = block_reduce_sum(vr, red);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = cnt;
    part[3 * blockIdx.x + 1] = ent;
    part[3 * blockIdx.x + 2] = vr;
  }
}

// Compaction, second pass (one workgroup): the blocks' queue offsets by an
// exclusive scan of their counts, and the level's totals.
__global__ __launch_bounds__(1024) void bfs_scan(uint32_t nb, unsigned long long* __restrict__ part,
                                                 unsigned long long* __restrict__ offsets,
                                                 const unsigned long long* __restrict__ part_edges,
                                                 uint32_t nedges, BfsCounters* c) {
  __shared__ unsigned long long sh[1024];
  __shared__ unsigned long long red[16];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (nb + 1023) / 1024, b0 = tid * per;
  unsigned long long local = 0, ent = 0, vr = 0, ed = 0;
  for (uint32_t b = b0; b < b0 + per && b < nb; ++b) {
    local += part[3 * b];
    ent += part[3 * b + 1];
    vr += part[3 * b + 2];
  }
  for (uint32_t k = tid; k < nedges; k += 1024) ed += part_edges[k];
  sh[tid] = local;
  __syncthreads();
  for (uint32_t d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan
    unsigned long long t = tid >= d ? sh[tid - d] : 0;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  unsigned long long run = sh[tid] - local;
  for (uint32_t b = b0; b < b0 + per && b < nb; ++b) {
    offsets[b] = run;
    run += part[3 * b];
  }
  ent = block_reduce_sum(ent, red);
  vr = block_reduce_sum(vr, red);
  ed = block_reduce_sum(ed, red);
  if (tid == 0) {
    c->next_size = sh[1023];
    c->next_entries = ent;
    c->visited_rev = vr;
    c->edges = ed;
  }
}

// Compaction, third pass: the claimed nodes in ascending order into the
// queue, each block at its scanned offset.
__global__ __launch_bounds__(256) void bfs_write(uint64_t words, const uint32_t* __restrict__ claim,
                                                 const unsigned long long* __restrict__ offsets,
                                                 uint32_t* __restrict__ queue) {
  __shared__ uint32_t sh[256];
  const uint64_t w0 = (uint64_t)blockIdx.x * kWordsPerCountBlock + threadIdx.x * 2;
  uint32_t bits[2] = {0, 0};
  for (int k = 0; k < 2; ++k)
    if (w0 + k < words) bits[k] = claim[w0 + k];
  const uint32_t mine = __popc(bits[0]) + __popc(bits[1]);
  sh[threadIdx.x] = mine;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    uint32_t t = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  unsigned long long pos = offsets[blockIdx.x] + sh[threadIdx.x] - mine;
  for (int k = 0; k < 2; ++k) {
    uint32_t b = bits[k];
    while (b) {
      queue[pos++] = (uint32_t)((w0 + k) * 32) + (__ffs(b) - 1);
      b &= b - 1;
    }
  }
}

struct BfsRun {
  std::vector<uint32_t> level_pulled;  // per expanded level: 1 if pulled
  std::vector<uint64_t> level_size;    // nodes admitted at each depth (0 = starts)
  uint64_t visited = 0, edges = 0;
  uint32_t levels = 0;
};

// The search's workspace, kept with the graph: bitmaps, queues, candidate
// bytes, split-row piece hits, compaction partials and counters.
struct BfsWorkspace {
  uint32_t *visited, *bm_a, *bm_b, *q_a, *q_b, *piece_hit;
  uint8_t* cand;
  unsigned long long *part, *offsets, *part_edges;
  BfsCounters* cnt;
};

static uint32_t edge_parts(const DevCsr& c) { return c.nblocks + (c.nlong + 255) / 256; }

static int bfs_workspace(wg_context* ctx, wg_graph* g, uint64_t words, BfsWorkspace& w) {
  const uint64_t nbc = (words + kWordsPerCountBlock - 1) / kWordsPerCountBlock;
  if (!g->bfs_ws[0]) {
    const uint64_t n = std::max<uint64_t>(g->n, 1);
    const size_t sizes[8] = {3 * words * 4,
                             2 * n * 4,
                             std::max<uint32_t>(std::max(g->fwd.npartials, g->rev.npartials), 1) * 4,
                             n,
                             std::max<uint64_t>(nbc, 1) * 3 * 8,
                             std::max<uint64_t>(nbc, 1) * 8,
                             std::max<uint32_t>(std::max(edge_parts(g->fwd), edge_parts(g->rev)), 1) * 8,
                             sizeof(BfsCounters)};
    for (int k = 0; k < 8; ++k) {
      void* p = nullptr;
      int s = dalloc(ctx, &p, sizes[k]);
      if (s) return s;
      g->workspace.push_back({p, sizes[k]});
      g->bfs_ws[k] = p;
    }
    // Candidate bytes start clear and every resolve clears what it reads.
    WG_HIP(ctx, hipMemsetAsync(g->bfs_ws[3], 0, n, ctx->stream));
  }
  uint32_t* bms = (uint32_t*)g->bfs_ws[0];
  uint32_t* qs = (uint32_t*)g->bfs_ws[1];
  w = BfsWorkspace{bms, bms + words, bms + 2 * words, qs, qs + std::max<uint64_t>(g->n, 1),
                   (uint32_t*)g->bfs_ws[2], (uint8_t*)g->bfs_ws[3],
                   (unsigned long long*)g->bfs_ws[4], (unsigned long long*)g->bfs_ws[5],
                   (unsigned long long*)g->bfs_ws[6], (BfsCounters*)g->bfs_ws[7]};
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
  const uint32_t nbc = (uint32_t)((words + kWordsPerCountBlock - 1) / kWordsPerCountBlock);
  BfsWorkspace w;
  int s = bfs_workspace(ctx, g, words, w);
  if (s) return s;
  WG_HIP(ctx, hipMemsetAsync(depth, 0xff, std::max<uint64_t>(n, 1) * 4, st));
  WG_HIP(ctx, hipMemsetAsync(via, 0xff, std::max<uint64_t>(n, 1) * 4, st));
  WG_HIP(ctx, hipMemsetAsync(w.visited, 0, 3 * words * 4, st));  // visited, bm_a, bm_b
  uint64_t mf = 0, visited_rev = 0;
  if (!starts.empty()) {
    WG_HIP(ctx, hipMemcpyAsync(w.q_a, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, st));
    bfs_seed<<<(unsigned)std::min<size_t>((starts.size() + 255) / 256, 1024), 256, 0, st>>>(
        w.q_a, (uint32_t)starts.size(), depth, w.visited, w.bm_a);
    for (uint32_t v : starts) {
      mf += adj.h_offsets[v + 1] - adj.h_offsets[v];
      visited_rev += rev.h_offsets[v + 1] - rev.h_offsets[v];
    }
  }
  uint64_t qn = starts.size();
  run.visited = qn;
  run.level_size.assign(1, qn);
  uint32_t d = 0;
  uint32_t* frontier = w.bm_a;  // bitmap of the current frontier
  uint32_t* claim = w.bm_b;     // bitmap of the nodes claimed this level
  bool queue_ready = true;      // q_a holds the frontier as a list
  BfsCounters hc{};
  while (qn > 0 && !(max_depth >= 0 && (int64_t)d >= max_depth) &&
         !(stop_after >= 0 && run.visited > (uint64_t)stop_after)) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    // Beamer's switch (level.rs:184-211): pull iff |F| * 24 > n and the
    // unvisited nodes' reverse entries are fewer than alpha * m_f.
    const uint64_t m_u = rev.nnz - visited_rev;
    const bool pull = pulling && qn * 24 > n && m_u < mf;
    if (!pull) {
      if (!queue_ready)
        bfs_write<<<nbc, 256, 0, st>>>(words, frontier, w.offsets, w.q_a);
      bfs_mark<<<(unsigned)std::min<uint64_t>((qn + 3) / 4, 256 * 64), 256, 0, st>>>(
          w.q_a, qn, adj.offsets, adj.targets, w.visited, filter_d, w.cand);
    }
    const int cands = pull ? 0 : 1;
    if (rev.nblocks)
      bfs_pull_blocks<<<rev.nblocks, kBlockThreads, 0, st>>>(
          rev.blocks, rev.offsets, rev.targets, w.visited, frontier, filter_d, w.cand, cands, depth,
          via, claim, w.piece_hit, w.part_edges, d + 1);
    if (rev.nlong)
      bfs_pull_long<<<(rev.nlong + 255) / 256, 256, 0, st>>>(
          rev.long_rows, rev.long_pbase, rev.nlong, rev.offsets, rev.targets, w.visited, filter_d,
          w.cand, cands, w.piece_hit, depth, via, claim, w.part_edges + rev.nblocks, d + 1);
    bfs_count<<<nbc, 256, 0, st>>>(words, claim, w.visited, frontier, adj.offsets, rev.offsets, w.part);
    bfs_scan<<<1, 1024, 0, st>>>(nbc, w.part, w.offsets, w.part_edges, pull ? edge_parts(rev) : 0, w.cnt);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(&hc, w.cnt, sizeof hc, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    std::swap(frontier, claim);
    queue_ready = false;
    // A pushed level scans every entry of every frontier row.
    run.edges += pull ? hc.edges : mf;
    qn = hc.next_size;
    mf = hc.next_entries;
    visited_rev += hc.visited_rev;
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
  // Counting sort by depth (nodes ascending within), then per pushed level
  // a stable sort by via, which gives (via, node).
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
