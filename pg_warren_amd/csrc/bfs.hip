// Level-synchronous, direction-optimising BFS on gfx950
// (traversal/level.rs:88-337, bfs.rs:34-111, reachable.rs:79-90).
//
// Each level is pushed or pulled by Beamer's switch with the reference's
// constants (alpha = 1, beta = 24), decided on the host from counters of
// the frontier: its size, its row entries (m_f) and the reverse-row entries
// of every visited node (so the unvisited nodes' entries m_u follow by
// subtraction).
//   push  bfs_push: a wavefront per frontier node strides its row and marks
//         each unvisited target a candidate (a plain byte store: every
//         writer writes 1, so no atomic is needed); the pull kernel then
//         resolves only the candidates. Rows longer than a block are pushed
//         a piece per workgroup. A small push sets the vias as it marks
//         (atomicMin) and the compaction's first pass claims the candidates.
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
// count / scan / write compaction with per-block partials; the scan posts
// the counters to the context's mailbox (coherent host memory), which the
// host polls, so a level costs no copy or stream synchronisation.
// Small push levels of a search that lists no levels run in one workgroup,
// several to a launch (bfs_tiny_levels), without the compaction's passes
// over the whole bitmap or a return to the host between them. The hot
// path has no device-scope atomics on a shared address (on gfx950 they
// resolve past the per-XCD L2 and serialise: an arrival counter bumped by
// each of the count pass's 2048 blocks cost more than the scan's launch).
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

constexpr uint32_t kUnvisited = 0xffffffffu;
constexpr uint32_t kWordsPerCountBlock = 64;  // bfs_count / bfs_write: 8 nodes a thread
// A push level whose frontier rows hold fewer than n / kSmallPushRatio
// entries sets vias by atomicMin while marking (bfs_push<true>).
constexpr uint64_t kSmallPushRatio = 4;

// The CSRs a traversal reads, in order (a node's row entries are its rows
// concatenated, Adjacency::row_entries): passed by value to the kernels that
// sum a node's degrees or find where a push meets a node.
constexpr uint32_t kMaxRowSets = 32;
struct RowSets {
  const uint64_t* off[kMaxRowSets];
  const uint32_t* tgt[kMaxRowSets];
  uint32_t k;
};

__device__ __forceinline__ uint64_t degree_in(const RowSets& r, uint32_t v) {
  uint64_t d = 0;
  for (uint32_t j = 0; j < r.k; ++j) d += r.off[j][v + 1] - r.off[j][v];
  return d;
}

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

// A search's set-up in one launch for at most kInitStarts starts: depth and
// via unvisited, the three bitmaps (visited, frontier, claims) clear, the
// starts at depth 0 in the visited and frontier bitmaps, and the starts, in
// order, as the first queue. Each element is written by one thread, its
// start bits included.
constexpr uint32_t kInitStarts = 8;
struct InitStarts {
  uint32_t v[kInitStarts];
  uint32_t count;
};

__global__ __launch_bounds__(256) void bfs_init(uint64_t n, uint64_t words, uint32_t* __restrict__ depth,
                                                uint32_t* __restrict__ via, uint32_t* __restrict__ bitmaps,
                                                uint32_t* __restrict__ queue, InitStarts st) {
  const uint64_t total = std::max<uint64_t>(n, 3 * words);
  if (blockIdx.x == 0 && threadIdx.x < st.count) queue[threadIdx.x] = st.v[threadIdx.x];
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (uint64_t)gridDim.x * 256) {
    if (i < n) {
      uint32_t dv = kUnvisited;
      for (uint32_t k = 0; k < st.count; ++k)
        if (st.v[k] == i) dv = 0;
      depth[i] = dv;
      via[i] = kUnvisited;
    }
    if (i < 3 * words) {
      // Words [0, words) visited, [words, 2 words) the frontier: a start's
      // bit in both; [2 words, 3 words) the claims, clear.
      uint32_t bits = 0;
      if (i < 2 * words) {
        const uint64_t wi = i < words ? i : i - words;
        for (uint32_t k = 0; k < st.count; ++k)
          if ((st.v[k] >> 5) == wi) bits |= 1u << (st.v[k] & 31);
      }
      bitmaps[i] = bits;
    }
  }
}

// Entries i, i + step, ... below b of frontier node u's row, pushed four
// at a time: the targets' loads, then their visited words' loads, go out
// together (a lane past the end reads node 0's word and drops it). An
// unvisited target admitted by the filter becomes a candidate (a plain
// byte store: every writer writes 1); kMin: it also takes the smallest
// frontier node with an edge to it as its via (atomicMin: the first in
// ascending order, as the reference's push reaches it).
template <bool kMin>
__device__ __forceinline__ void push_run(uint64_t i, uint64_t b, uint32_t step, uint32_t u,
                                         const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ visited,
                                         const uint8_t* __restrict__ filter, uint8_t* __restrict__ cand,
                                         uint32_t* __restrict__ via) {
  for (; i < b; i += 4 * (uint64_t)step) {
    uint32_t v[4], word[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i + k * step < b ? tgt[i + k * step] : 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) word[k] = visited[v[k] >> 5];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (i + k * step < b && !((word[k] >> (v[k] & 31)) & 1u) && !(filter && !filter[v[k]])) {
        if (kMin) atomicMin(&via[v[k]], u);
        cand[v[k]] = 1;
      }
  }
}

// Push: blocks [0, qblocks) give a wavefront to each queued frontier node
// in turn, which strides its row; rows of more than kBlockNnz entries are
// left to the blocks after them, one per split-row piece of the CSR, which
// return at once unless the piece's row is in the frontier bitmap. Splitting
// keeps a hub's row (R-MAT scale 22 has one of 161K entries) from
// serialising a level on one wavefront. kMin (a small frontier, whose
// targets have few predecessors, so the atomics seldom meet): the vias are
// set here and bfs_claim_count settles the candidates; otherwise the pull
// kernel resolves them.
template <bool kMin>
__global__ __launch_bounds__(256) void bfs_push(const uint32_t* __restrict__ queue, uint64_t qn, uint32_t qblocks,
                                                const uint2* __restrict__ pieces,
                                                const uint64_t* __restrict__ off,
                                                const uint32_t* __restrict__ tgt,
                                                const uint32_t* __restrict__ frontier,
                                                const uint32_t* __restrict__ visited,
                                                const uint8_t* __restrict__ filter,
                                                uint8_t* __restrict__ cand, uint32_t* __restrict__ via) {
  if (blockIdx.x >= qblocks) {
    const uint2 pc = pieces[blockIdx.x - qblocks];
    const uint32_t u = pc.x;
    if (!bit_of(frontier, u)) return;
    const uint64_t a = off[u] + (uint64_t)pc.y * kBlockNnz;
    const uint64_t b = min(a + kBlockNnz, off[u + 1]);
    push_run<kMin>(a + threadIdx.x, b, 256, u, tgt, visited, filter, cand, via);
    return;
  }
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint64_t nwaves = (uint64_t)qblocks * 4;
  for (uint64_t q = wave; q < qn; q += nwaves) {
    const uint32_t u = queue[q];
    const uint64_t a = off[u], b = off[u + 1];
    if (b - a > kBlockNnz) continue;  // a split row: its pieces' blocks
    push_run<kMin>(a + lane, b, 64, u, tgt, visited, filter, cand, via);
  }
}

// The claim bits of the eight nodes from v0 (a multiple of 8) from the
// candidate bytes, which are cleared.
__device__ __forceinline__ uint32_t claim_byte(uint64_t n, uint64_t v0, uint8_t* __restrict__ cand) {
  uint32_t bits = 0;
  if (v0 + 8 <= n) {
    const uint2 c = *(const uint2*)(cand + v0);
    for (int j = 0; j < 4; ++j) {
      if ((c.x >> (8 * j)) & 0xffu) bits |= 1u << j;
      if ((c.y >> (8 * j)) & 0xffu) bits |= 1u << (4 + j);
    }
    if (bits) *(uint2*)(cand + v0) = make_uint2(0, 0);
  } else {
    for (uint64_t v = v0; v < n && v < v0 + 8; ++v)
      if (cand[v]) {
        bits |= 1u << (v - v0);
        cand[v] = 0;
      }
  }
  return bits;
}

// A pull over several CSRs (a traversal reading several row sets) runs one
// pass per CSR in the reversed adjacency's order. Pass kind of a later pass:
constexpr int kPassFirst = 1;  // rows claimed by an earlier pass are done:
                               // the via is the first frontier node in the
                               // rows' concatenated order (a pulled level)
constexpr int kPassMin = 2;    // every pass searches every unvisited row and
                               // the via is the smallest frontier node met
                               // (the via a push gives)

// Whether row v searches in a pull pass.
__device__ __forceinline__ bool pull_searches(uint32_t v, int candidates, int pass, const uint8_t* cand,
                                              const uint32_t* visited, const uint8_t* filter,
                                              const uint32_t* depth, uint32_t d1) {
  if (candidates) return cand[v] != 0;
  if (bit_of(visited, v) || (filter && !filter[v])) return false;
  return pass != kPassFirst || depth[v] != d1;
}

// Row v met frontier node u: claimed at depth d1 with via u, or, in a
// kPassMin pass when an earlier pass claimed it, its via lowered to u.
__device__ __forceinline__ void pull_claims(uint32_t v, uint32_t u, int pass, uint32_t* depth, uint32_t* via,
                                            uint32_t d1) {
  if (pass == kPassMin && depth[v] == d1) {
    if (u < via[v]) via[v] = u;
  } else {
    via[v] = u;
    depth[v] = d1;
  }
}

// Pull over the scanned CSR's row blocks (roff/rtgt). candidates == 0: every
// unvisited row admitted by the filter searches (as `pass` allows);
// otherwise only rows marked in cand (the push's second half), which are
// cleared. Pieces of split rows record their first hit (offset within the
// row) or ~0 in piece_hit for bfs_pull_long. A pulled level's scanned
// entries go to part_edges. Claims are ORed into next_bm (zero at the
// level's start).
__global__ __launch_bounds__(256) void bfs_pull_blocks(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ roff,
    const uint32_t* __restrict__ rtgt, const uint32_t* __restrict__ visited,
    const uint32_t* __restrict__ frontier, const uint8_t* __restrict__ filter,
    uint8_t* __restrict__ cand, int candidates, int pass, uint32_t* __restrict__ depth,
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
    if (T == 1) {
      // A thread per row (rows tid and tid + 256): both rows' offsets, the
      // search test and then their first entries are read together, so a
      // row whose first entry meets the frontier costs three dependent
      // reads; a row that goes on reads four entries at a time.
      constexpr int R = kBlockRows / kBlockThreads;
      uint32_t v[R], u[R];
      uint64_t a[R], e[R], at[R];
      bool act[R], hit[R];
#pragma unroll
      for (int k = 0; k < R; ++k) {
        const uint32_t r = tid + kBlockThreads * k;
        v[k] = b.row0 + (r < nr ? r : 0);
        a[k] = roff[v[k]];
        e[k] = roff[v[k] + 1];
      }
#pragma unroll
      for (int k = 0; k < R; ++k)
        act[k] = tid + kBlockThreads * k < nr && pull_searches(v[k], candidates, pass, cand, visited, filter,
                                                                depth, d1);
#pragma unroll
      for (int k = 0; k < R; ++k) u[k] = act[k] && a[k] < e[k] ? rtgt[a[k]] : 0u;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        hit[k] = act[k] && a[k] < e[k] && bit_of(frontier, u[k]);
        at[k] = a[k];
      }
#pragma unroll
      for (int k = 0; k < R; ++k) {
        if (!act[k]) continue;
        for (uint64_t i = a[k] + 1; !hit[k] && i < e[k]; i += 4) {
          uint32_t q[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) q[j] = i + j < e[k] ? rtgt[i + j] : 0u;
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (!hit[k] && i + j < e[k] && bit_of(frontier, q[j])) {
              hit[k] = true;
              u[k] = q[j];
              at[k] = i + j;
            }
        }
        if (hit[k]) {
          pull_claims(v[k], u[k], pass, depth, via, d1);
          atomicOr(&lbm[(v[k] >> 5) - w0], 1u << (v[k] & 31));
          edges += at[k] - a[k] + 1;
        } else {
          edges += e[k] - a[k];
        }
        if (candidates) cand[v[k]] = 0;
      }
    }
    for (uint32_t r = g; T > 1 && r < nr; r += groups) {
      const uint32_t v = b.row0 + r;
      if (pull_searches(v, candidates, pass, cand, visited, filter, depth, d1)) {
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
              pull_claims(v, u, pass, depth, via, d1);
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
      if (!bits) continue;
      if (whole)
        next_bm[word] |= bits;
      else
        atomicOr(&next_bm[word], bits);
    }
  } else {
    const uint32_t v = b.row0;
    uint32_t hit_at = ~0u;
    if (pull_searches(v, candidates, pass, cand, visited, filter, depth, d1)) {
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
    int candidates, int pass, const uint32_t* __restrict__ piece_hit, uint32_t* __restrict__ depth,
    uint32_t* __restrict__ via, uint32_t* __restrict__ next_bm,
    unsigned long long* __restrict__ part_edges, uint32_t d1) {
  __shared__ unsigned long long red[4];
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  unsigned long long edges = 0;
  if (i < nlong) {
    const uint32_t v = long_rows[i];
    if (pull_searches(v, candidates, pass, cand, visited, filter, depth, d1)) {
      const uint64_t a = roff[v], e = roff[v + 1];
      uint32_t at = ~0u;
      for (uint32_t p = long_pbase[i]; p < long_pbase[i + 1] && at == ~0u; ++p) at = piece_hit[p];
      if (at != ~0u) {
        pull_claims(v, rtgt[a + at], pass, depth, via, d1);
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

// Compaction, second pass (one workgroup of 256): an exclusive scan of the
// blocks' counts into offsets, and the sums of their entries, reverse
// entries and (a pulled level's) scanned entries, posted to the context's
// mailbox (values first, then the sequence number with release).
__global__ __launch_bounds__(256) void bfs_scan(uint32_t nb, const unsigned long long* __restrict__ part,
                                                unsigned long long* __restrict__ offsets,
                                                unsigned long long* __restrict__ mail, unsigned long long seq) {
  __shared__ unsigned long long sh[256];
  __shared__ unsigned long long red[4];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (nb + 255) / 256, b0 = tid * per, b1 = min(b0 + per, nb);
  unsigned long long local = 0, ent = 0, vr = 0, ed = 0;
  // A thread's blocks eight at a time, their loads in flight together.
  constexpr uint32_t U = 8;
  for (uint32_t b = b0; b < b1; b += U) {
    unsigned long long p[U][4];
#pragma unroll
    for (uint32_t j = 0; j < U; ++j)
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) p[j][i] = b + j < b1 ? part[4 * (b + j) + i] : 0ull;
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) {
      local += p[j][0];
      ent += p[j][1];
      vr += p[j][2];
      ed += p[j][3];
    }
  }
  sh[tid] = local;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {  // inclusive Hillis-Steele scan
    unsigned long long t = tid >= d ? sh[tid - d] : 0;
    __syncthreads();
    sh[tid] += t;
    __syncthreads();
  }
  unsigned long long run = sh[tid] - local;
  for (uint32_t b = b0; b < b1; b += U) {
    unsigned long long c[U];
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) c[j] = b + j < b1 ? part[4 * (b + j)] : 0ull;
#pragma unroll
    for (uint32_t j = 0; j < U; ++j) {
      if (b + j < b1) offsets[b + j] = run;
      run += c[j];
    }
  }
  ent = block_reduce_sum(ent, red);
  vr = block_reduce_sum(vr, red);
  ed = block_reduce_sum(ed, red);
  if (tid == 0) {
    // BfsCounters' order.
    const unsigned long long v[4] = {sh[255], ent, vr, ed};
    for (int i = 0; i < 4; ++i) __hip_atomic_store(&mail[1 + i], v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&mail[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Compaction, first pass: per block of words, the claimed nodes and their
// row entries in the rows read and in the reversed rows (summed over the
// CSRs of each); also folds the claims into visited and clears the old
// frontier bitmap (it becomes the next level's claim map). A thread takes
// eight nodes, their offsets read together (2048 blocks at 2^22 nodes, so
// that enough loads are in flight).
// kClaim (after a small push, whose vias are set): the claims come from the
// candidate bytes here, and the claimed nodes take depth d1. A pulled
// level's per-block scanned entries (nedges of them) are summed too, an
// equal slice a block, into each block's fourth partial.
template <bool kClaim>
__global__ __launch_bounds__(256) void bfs_count(uint64_t n, uint64_t words, uint32_t* __restrict__ claim,
                                                 uint32_t* __restrict__ visited,
                                                 uint32_t* __restrict__ old_frontier, const RowSets adj,
                                                 const RowSets rev, unsigned long long* __restrict__ part,
                                                 uint8_t* __restrict__ cand, uint32_t* __restrict__ depth,
                                                 uint32_t d1, const unsigned long long* __restrict__ part_edges,
                                                 uint32_t nedges) {
  __shared__ unsigned long long red[4];
  unsigned long long cnt = 0, ent = 0, vr = 0, ed = 0;
  {
    const uint32_t per = (nedges + gridDim.x - 1) / gridDim.x, e0 = blockIdx.x * per;
    for (uint32_t e = e0 + threadIdx.x; e < e0 + per && e < nedges; e += 256) ed += part_edges[e];
  }
  // Four threads a word, eight nodes each (the lanes of a word are adjacent
  // in the wave).
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x, w = t >> 2;
  const uint32_t sub = (uint32_t)(t & 3);
  uint32_t word = 0;
  if (kClaim) {
    if (w < words) word = claim_byte(n, w * 32 + 8 * sub, cand) << (8 * sub);
    word |= __shfl_xor(word, 1, 64);
    word |= __shfl_xor(word, 2, 64);
  } else if (w < words) {
    word = claim[w];
  }
  if (w < words && sub == 0) {
    if (kClaim) claim[w] = word;
    visited[w] |= word;
    old_frontier[w] = 0;
  }
  uint32_t bits = (word >> (8 * sub)) & 0xffu;
  cnt = __popc(bits);
  if (bits) {
    constexpr int B = 8;
    uint32_t v[B];
    uint32_t live = 0;
#pragma unroll
    for (int k = 0; k < B; ++k) {
      v[k] = (uint32_t)(w * 32 + 8 * sub) + (bits ? __ffs(bits) - 1 : 0);
      live |= (bits ? 1u : 0u) << k;
      bits &= bits - 1;
    }
#pragma unroll
    for (int k = 0; k < B; ++k)
      if (kClaim && ((live >> k) & 1)) depth[v[k]] = d1;
    for (uint32_t j = 0; j < adj.k; ++j) {
      uint64_t lo[B], hi[B];
#pragma unroll
      for (int k = 0; k < B; ++k) {
        lo[k] = adj.off[j][v[k]];
        hi[k] = adj.off[j][v[k] + 1];
      }
#pragma unroll
      for (int k = 0; k < B; ++k)
        if ((live >> k) & 1) ent += hi[k] - lo[k];
    }
    for (uint32_t j = 0; j < rev.k; ++j) {
      uint64_t lo[B], hi[B];
#pragma unroll
      for (int k = 0; k < B; ++k) {
        lo[k] = rev.off[j][v[k]];
        hi[k] = rev.off[j][v[k] + 1];
      }
#pragma unroll
      for (int k = 0; k < B; ++k)
        if ((live >> k) & 1) vr += hi[k] - lo[k];
    }
  }
  cnt = block_reduce_sum(cnt, red);
  ent = block_reduce_sum(ent, red);
  vr = block_reduce_sum(vr, red);
  ed = block_reduce_sum(ed, red);
  if (threadIdx.x == 0) {
    part[4 * blockIdx.x] = cnt;
    part[4 * blockIdx.x + 1] = ent;
    part[4 * blockIdx.x + 2] = vr;
    part[4 * blockIdx.x + 3] = ed;
  }
}

// Compaction, third pass: the claimed nodes in ascending order into the
// queue, each block at its scanned offset.
__global__ __launch_bounds__(256) void bfs_write(uint64_t words, const uint32_t* __restrict__ claim,
                                                 const unsigned long long* __restrict__ offsets,
                                                 uint32_t* __restrict__ queue) {
  __shared__ uint32_t sh[256];
  // Four threads a word, eight nodes each, in node order.
  const uint64_t t = (uint64_t)blockIdx.x * 256 + threadIdx.x, w = t >> 2;
  const uint32_t sub = (uint32_t)(t & 3);
  uint32_t b = w < words ? (claim[w] >> (8 * sub)) & 0xffu : 0u;
  const uint32_t mine = __popc(b);
  sh[threadIdx.x] = mine;
  __syncthreads();
  for (uint32_t d = 1; d < 256; d <<= 1) {
    uint32_t x = threadIdx.x >= d ? sh[threadIdx.x - d] : 0;
    __syncthreads();
    sh[threadIdx.x] += x;
    __syncthreads();
  }
  unsigned long long pos = offsets[blockIdx.x] + sh[threadIdx.x] - mine;
  while (b) {
    queue[pos++] = (uint32_t)(w * 32 + 8 * sub) + (__ffs(b) - 1);
    b &= b - 1;
  }
}

// Small push levels in one workgroup, level after level without a return
// to the host: while a level's frontier rows hold at most kTinyEntries
// entries and the direction switch keeps pushing (the same test as
// BfsSearch::step), the 1024 threads push it as bfs_push<true> and
// bfs_count<true> do. An unvisited target admitted by the filter takes the
// smallest frontier node with an edge to it as its via (atomicMin) and is
// claimed once (the claim bitmap's atomicOr returns its old bit), into the
// next queue; the claimed nodes take depth d + 1 and join the visited
// bitmap, their row and reverse-row entries are summed, and the old
// frontier's bits are cleared (the bitmap becomes the next claim map). The
// queue a level leaves is in claim order: a push reads its frontier in any
// order. Each level's counters are posted to the mailbox (levels first).
// visited is read with device-scope loads: the claims' atomics update it in
// L2, past this CU's L1.
constexpr uint64_t kTinyEntries = 16384, kTinyNodes = 4096;
constexpr uint32_t kTinyLevels = 9;  // 1 + 3 per level in the mailbox's 30 values

// Exclusive scan of x over a 1024-thread workgroup; *total the sum.
__device__ __forceinline__ uint32_t scan_1024(uint32_t x, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = x;
#pragma unroll
  for (uint32_t k = 1; k < 64; k <<= 1) {
    const uint32_t y = __shfl_up(inc, k, 64);
    if (lane >= k) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
  for (uint32_t w = 0; w < 16; ++w) {
    const uint32_t t = wsum[w];
    if (w < wave) base += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return base + inc - x;
}
struct TinyState {
  unsigned long long frontier_size, mf, visited_rev, visited;
  uint32_t d;
};

__global__ __launch_bounds__(1024) void bfs_tiny_levels(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ tgt, const uint64_t* __restrict__ roff,
    uint64_t n, uint64_t rr_nnz, int pulling, int64_t max_depth, uint32_t* visited, uint32_t* fa, uint32_t* fb,
    uint32_t* qa, uint32_t* qb, const uint8_t* __restrict__ filter, uint32_t* __restrict__ depth,
    uint32_t* __restrict__ via, TinyState st, unsigned long long* __restrict__ mail, unsigned long long seq) {
  __shared__ uint32_t qcount;
  __shared__ unsigned long long red[16], tot[2];
  __shared__ unsigned long long recs[3 * kTinyLevels];
  // The frontier's rows: node, first entry, and entry offset (exclusive
  // prefix of the row lengths, pre[F] the level's entries).
  __shared__ uint32_t nod[kTinyNodes], pre[kTinyNodes + 1], wsum[16];
  __shared__ uint64_t rowa[kTinyNodes];
  const uint32_t tid = threadIdx.x;
  unsigned long long F = st.frontier_size, mf = st.mf, vrev = st.visited_rev, vis = st.visited;
  uint32_t d = st.d, L = 0;
  while (L < kTinyLevels) {
    if (F == 0 || (max_depth >= 0 && (int64_t)d >= max_depth)) break;
    const unsigned long long m_u = rr_nnz - vrev;
    const bool pull = pulling && F * 24 > n && (m_u < mf || vis == n);
    if (pull || mf > kTinyEntries || F > kTinyNodes || mf * kSmallPushRatio >= n) break;
    if (tid == 0) qcount = 0;
    {
      // Up to four frontier nodes a thread, their rows' lengths scanned.
      const uint32_t per = (uint32_t)((F + 1023) / 1024), i0 = tid * per;
      uint32_t len[4], un[4], my = 0;
      uint64_t ra[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        len[k] = 0;
        un[k] = 0;
        ra[k] = 0;
        if (k < per && i0 + k < F) {
          un[k] = qa[i0 + k];
          ra[k] = off[un[k]];
          len[k] = (uint32_t)(off[un[k] + 1] - ra[k]);
          my += len[k];
        }
      }
      uint32_t total;
      uint32_t base = scan_1024(my, wsum, &total);
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (k < per && i0 + k < F) {
          nod[i0 + k] = un[k];
          rowa[i0 + k] = ra[k];
          pre[i0 + k] = base;
          base += len[k];
        }
      if (tid == 0) pre[F] = total;
    }
    __syncthreads();
    const uint32_t ne = pre[F];
    // The level's entries, four a thread in flight: entry e lies in the row
    // of the last frontier node whose offset is at most e.
    for (uint32_t e0 = tid; e0 < ne; e0 += 4 * 1024) {
      uint32_t v[4], u[4], word[4];
      bool ok[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t e = e0 + k * 1024;
        ok[k] = e < ne;
        v[k] = u[k] = 0;
        if (ok[k]) {
          uint32_t lo = 0, hi = (uint32_t)F;  // pre[lo] <= e < pre[hi]
          while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (pre[mid] <= e) lo = mid;
            else hi = mid;
          }
          u[k] = nod[lo];
          v[k] = tgt[rowa[lo] + (e - pre[lo])];
        }
      }
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        word[k] = ok[k] ? __hip_atomic_load(&visited[v[k] >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0u;
      bool go[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) go[k] = ok[k] && !((word[k] >> (v[k] & 31)) & 1u);
      if (filter) {
        uint8_t f[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) f[k] = go[k] ? filter[v[k]] : 0;
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) go[k] = go[k] && f[k];
      }
      // The four entries' atomics go out together; the claim's old bit
      // decides which thread queues the node.
      uint32_t old[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (go[k]) atomicMin(&via[v[k]], u[k]);
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) old[k] = go[k] ? atomicOr(&fb[v[k] >> 5], 1u << (v[k] & 31)) : ~0u;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (!((old[k] >> (v[k] & 31)) & 1u)) qb[atomicAdd(&qcount, 1u)] = v[k];
    }
    __syncthreads();
    const uint32_t Fn = qcount;
    unsigned long long ent = 0, vr = 0;
    // The claimed nodes, four a thread in flight.
    for (uint32_t i0 = tid; i0 < Fn; i0 += 4 * 1024) {
      uint32_t v[4];
      uint64_t a0[4], a1[4], r0[4], r1[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) v[k] = i0 + k * 1024 < Fn ? qb[i0 + k * 1024] : ~0u;
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (v[k] != ~0u) {
          depth[v[k]] = d + 1;
          atomicOr(&visited[v[k] >> 5], 1u << (v[k] & 31));
          a0[k] = off[v[k]];
          a1[k] = off[v[k] + 1];
          r0[k] = roff[v[k]];
          r1[k] = roff[v[k] + 1];
        }
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k)
        if (v[k] != ~0u) {
          ent += a1[k] - a0[k];
          vr += r1[k] - r0[k];
        }
    }
    for (uint64_t i = tid; i < F; i += 1024) {
      const uint32_t u = qa[i];
      atomicAnd(&fa[u >> 5], ~(1u << (u & 31)));
    }
    ent = block_reduce_sum(ent, red);
    vr = block_reduce_sum(vr, red);
    if (tid == 0) {
      recs[3 * L] = Fn;
      recs[3 * L + 1] = ent;
      recs[3 * L + 2] = vr;
      tot[0] = ent;
      tot[1] = vr;
    }
    __syncthreads();
    F = Fn;
    mf = tot[0];
    vrev += tot[1];
    vis += Fn;
    ++d;
    ++L;
    uint32_t* t = qa;
    qa = qb;
    qb = t;
    t = fa;
    fa = fb;
    fb = t;
  }
  if (tid == 0) {
    __hip_atomic_store(&mail[1], (unsigned long long)L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    for (uint32_t i = 0; i < 3 * L; ++i)
      __hip_atomic_store(&mail[2 + i], recs[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&mail[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

struct BfsRun {
  std::vector<uint32_t> level_pulled;  // per expanded level: 1 if pulled
  std::vector<uint64_t> level_size;    // nodes admitted at each depth (0 = starts)
  uint64_t visited = 0, edges = 0;
  uint64_t push_edges = 0;  // the entries a push of every level would scan
  uint32_t levels = 0;
};

// The CSRs a traversal reads (Adjacency::new and Adjacency::reversed,
// traversal/adjacency.rs:68-152, 288-320): with every relation, the graph's
// own (all-relations) pair; with a subset (sorted, repeats dropped), each
// relation's pair in index order. Out reads the forward rows, In the
// reverse ones, Both the forward then the reverse; the reversed adjacency
// (a pulled level's) reads the opposite direction, Both for Both.
// first_via: a pulled level's via is the first frontier node in the
// reversed rows' concatenated order (reachable sets); otherwise the one a
// push gives, the smallest frontier node with an edge to the node.
struct Traversal {
  std::vector<DevCsr*> rows, rrows;
  bool first_via = false;
};

static int traversal_of(wg_context* ctx, wg_graph* g, int direction, const uint32_t* rels, uint32_t nrel,
                        bool first_via, Traversal& t) {
  if (direction != WG_OUT && direction != WG_IN && direction != WG_BOTH)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  std::vector<wg_graph*> pairs;
  if (!rels) {
    pairs.push_back(g);
  } else {
    std::vector<uint32_t> sel(rels, rels + nrel);
    std::sort(sel.begin(), sel.end());
    sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
    for (uint32_t r : sel)
      if (r >= g->relation_count)
        return fail(ctx, WG_UNKNOWN_RELATION, "relation " + std::to_string(r) + " is not in the graph");
    for (uint32_t r : sel) pairs.push_back(g->relations.empty() ? g : g->relations[r]);
  }
  if (pairs.size() * (direction == WG_BOTH ? 2 : 1) > kMaxRowSets)
    return fail(ctx, WG_INVALID_ARGUMENT, "a traversal reads at most 32 CSRs");
  auto read = [&](int d, std::vector<DevCsr*>& out) {
    for (wg_graph* pr : pairs) {
      if (d != WG_IN) out.push_back(&pr->fwd);
      if (d != WG_OUT) out.push_back(&pr->rev);
    }
  };
  read(direction, t.rows);
  read(direction == WG_OUT ? WG_IN : direction == WG_IN ? WG_OUT : WG_BOTH, t.rrows);
  t.first_via = first_via;
  return WG_OK;
}

static RowSets row_sets(const std::vector<DevCsr*>& rows) {
  RowSets r{};
  r.k = (uint32_t)rows.size();
  for (uint32_t j = 0; j < r.k; ++j) {
    r.off[j] = rows[j]->offsets;
    r.tgt[j] = rows[j]->targets;
  }
  return r;
}

// The search's workspace: bitmaps, queues, candidate bytes, split-row piece
// hits and compaction partials, in kWsSlots buffers.
constexpr int kWsSlots = 7;
struct BfsWorkspace {
  uint32_t *visited, *bm_a, *bm_b, *q_a, *q_b, *piece_hit;
  uint8_t* cand;
  unsigned long long *part, *offsets, *part_edges;
};

static uint32_t edge_parts(const DevCsr& c) { return c.nblocks + (c.nlong + 255) / 256; }

// The workspace's buffer sizes for a search of g reading t (the pieces of
// split rows and the blocks of every reversed CSR a pulled level scans).
static void workspace_sizes(const wg_graph* g, const Traversal& t, size_t sizes[kWsSlots]) {
  const uint64_t words = (g->n + 63) / 64 * 2;
  const uint64_t nbc = (words + kWordsPerCountBlock - 1) / kWordsPerCountBlock;
  const uint64_t n = std::max<uint64_t>(g->n, 1);
  uint64_t pieces = 1, parts = 0;
  for (const DevCsr* c : t.rrows) {
    pieces = std::max<uint64_t>(pieces, c->npartials);
    parts += edge_parts(*c);
  }
  const size_t s[kWsSlots] = {3 * words * 4, 2 * n * 4, pieces * 4, n, std::max<uint64_t>(nbc, 1) * 4 * 8,
                              std::max<uint64_t>(nbc, 1) * 8, std::max<uint64_t>(parts, 1) * 8};
  for (int k = 0; k < kWsSlots; ++k) sizes[k] = s[k];
}

// Allocates a workspace's buffers of the given sizes into slot[], recording
// them in `owned` (charged to the context).
static int alloc_workspace(wg_context* ctx, const size_t sizes[kWsSlots], void* slot[kWsSlots],
                           std::vector<std::pair<void*, size_t>>& owned) {
  for (int k = 0; k < kWsSlots; ++k) {
    void* p = nullptr;
    int s = dalloc(ctx, &p, sizes[k]);
    if (s) return s;
    owned.push_back({p, sizes[k]});
    slot[k] = p;
  }
  // Candidate bytes start clear and every resolve clears what it reads.
  WG_HIP(ctx, hipMemsetAsync(slot[3], 0, sizes[3], ctx->stream));
  return WG_OK;
}

static void view_workspace(const wg_graph* g, void* const slot[kWsSlots], BfsWorkspace& w) {
  const uint64_t words = (g->n + 63) / 64 * 2;
  uint32_t* bms = (uint32_t*)slot[0];
  uint32_t* qs = (uint32_t*)slot[1];
  w = BfsWorkspace{bms, bms + words, bms + 2 * words, qs, qs + std::max<uint64_t>(g->n, 1),
                   (uint32_t*)slot[2], (uint8_t*)slot[3], (unsigned long long*)slot[4],
                   (unsigned long long*)slot[5], (unsigned long long*)slot[6]};
}

// The graph's own workspace, allocated on first use (sized for its own
// pair read both ways) and kept with it; a search that needs more (several
// relations' CSRs) gets one of its own in `own`, freed with it.
struct WorkspaceLease {
  wg_context* ctx = nullptr;
  void* slot[kWsSlots] = {};
  std::vector<std::pair<void*, size_t>> own;
  ~WorkspaceLease() {
    if (own.empty()) return;
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& b : own) dfree(ctx, b.first, b.second);
  }
};

static int bfs_workspace(wg_context* ctx, wg_graph* g, const Traversal& t, BfsWorkspace& w, WorkspaceLease& lease) {
  size_t need[kWsSlots];
  workspace_sizes(g, t, need);
  if (!g->bfs_ws[0]) {
    Traversal both;
    both.rows = both.rrows = {&g->fwd, &g->rev};
    size_t mine[kWsSlots];
    workspace_sizes(g, both, mine);
    for (int k = 0; k < kWsSlots; ++k) g->bfs_ws_bytes[k] = std::max(mine[k], need[k]);
    int s = alloc_workspace(ctx, g->bfs_ws_bytes, g->bfs_ws, g->workspace);
    if (s) return s;
  }
  bool fits = true;
  for (int k = 0; k < kWsSlots; ++k) fits &= need[k] <= g->bfs_ws_bytes[k];
  if (fits) {
    view_workspace(g, g->bfs_ws, w);
    return WG_OK;
  }
  lease.ctx = ctx;
  int s = alloc_workspace(ctx, need, lease.slot, lease.own);
  if (s) return s;
  view_workspace(g, lease.slot, w);
  return WG_OK;
}

// The level loop of the search, one level per step, its state kept between
// steps (LevelSearch, traversal/level.rs:88-337). After init the frontier
// is the starts; each step expands it into the next level, whose nodes are
// then the frontier bitmap (frontier_size of them) and whose depth is d.
// Levels are pushed or pulled by Beamer's switch (level.rs:184-211); a push
// over several CSRs sets vias by atomicMin (the smallest frontier node that
// reaches a node through any of them, as the reference's push); a pull over
// several scans one CSR per pass in the reversed adjacency's order.
struct BfsSearch {
  wg_context* ctx = nullptr;
  wg_graph* g = nullptr;
  Traversal t;
  RowSets ra{}, rr{};
  uint64_t rr_nnz = 0;
  uint64_t n = 0, words = 0;
  uint32_t nbc = 0;
  BfsWorkspace w{};
  const uint8_t* filter = nullptr;
  bool pulling = true;
  int64_t max_depth = -1;
  uint32_t *depth = nullptr, *via = nullptr;
  uint64_t frontier_size = 0, mf = 0, visited_rev = 0;
  uint32_t d = 0;
  uint32_t *frontier = nullptr, *claim = nullptr;
  bool queue_ready = true;  // q_a holds the frontier as an ascending list
  // Small push levels may run several to a launch (bfs_tiny_levels), which
  // leaves the queue in claim order: for searches whose levels are not
  // listed (bfs_run), over one CSR each way.
  bool tiny_levels = false;
  BfsRun run;

  int init(wg_context* c, wg_graph* gr, const Traversal& tr, const std::vector<uint32_t>& starts,
           const uint8_t* filter_d, int64_t maxd, bool pull_ok, uint32_t* depth_d, uint32_t* via_d,
           const BfsWorkspace& ws) {
    ctx = c;
    g = gr;
    n = g->n;
    t = tr;
    ra = row_sets(t.rows);
    rr = row_sets(t.rrows);
    rr_nnz = 0;
    for (const DevCsr* cs : t.rrows) rr_nnz += cs->nnz;
    words = (n + 63) / 64 * 2;  // 32-bit words, whole 64-node waves
    nbc = (uint32_t)((words + kWordsPerCountBlock - 1) / kWordsPerCountBlock);
    w = ws;
    filter = filter_d;
    max_depth = maxd;
    pulling = pull_ok;
    depth = depth_d;
    via = via_d;
    hipStream_t st = ctx->stream;
    mf = visited_rev = 0;
    if (starts.size() <= kInitStarts) {
      // One launch: fills and seeds (w.visited, bm_a, bm_b are contiguous).
      InitStarts is{};
      is.count = (uint32_t)starts.size();
      for (size_t k = 0; k < starts.size(); ++k) is.v[k] = starts[k];
      const uint64_t total = std::max<uint64_t>(n, 3 * words);
      bfs_init<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((total + 255) / 256, 8192)), 256, 0, st>>>(
          n, words, depth, via, w.visited, w.q_a, is);
    } else {
      WG_HIP(ctx, hipMemsetAsync(depth, 0xff, std::max<uint64_t>(n, 1) * 4, st));
      WG_HIP(ctx, hipMemsetAsync(via, 0xff, std::max<uint64_t>(n, 1) * 4, st));
      WG_HIP(ctx, hipMemsetAsync(w.visited, 0, 3 * words * 4, st));  // visited, bm_a, bm_b
      WG_HIP(ctx, hipMemcpyAsync(w.q_a, starts.data(), starts.size() * 4, hipMemcpyHostToDevice, st));
      bfs_seed<<<(unsigned)std::min<size_t>((starts.size() + 255) / 256, 1024), 256, 0, st>>>(
          w.q_a, (uint32_t)starts.size(), depth, w.visited, w.bm_a);
    }
    WG_HIP(ctx, hipGetLastError());
    if (!starts.empty()) {
      for (uint32_t v : starts) {
        for (const DevCsr* cs : t.rows) mf += cs->h_offsets[v + 1] - cs->h_offsets[v];
        for (const DevCsr* cs : t.rrows) visited_rev += cs->h_offsets[v + 1] - cs->h_offsets[v];
      }
    }
    frontier_size = starts.size();
    run = BfsRun{};
    run.visited = frontier_size;
    run.level_size.assign(1, frontier_size);
    d = 0;
    frontier = w.bm_a;
    claim = w.bm_b;
    queue_ready = true;
    return WG_OK;
  }

  // No further level: the frontier is empty or max_depth levels are done.
  bool exhausted() const { return frontier_size == 0 || (max_depth >= 0 && (int64_t)d >= max_depth); }

  int step() {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    hipStream_t st = ctx->stream;
    // Beamer's switch (level.rs:184-211): pull iff |F| * 24 > n and the
    // unvisited nodes' reverse entries are fewer than alpha * m_f (or no
    // node is unvisited).
    const uint64_t m_u = rr_nnz - visited_rev;
    const bool pull = pulling && frontier_size * 24 > n && (m_u < mf || run.visited == n);
    const bool multi = t.rows.size() > 1;
    // A push whose frontier rows hold few entries, or that reads several
    // CSRs, sets vias as it marks; a larger one over one CSR marks, then
    // resolves the candidates by the pull kernel (each scanning its reverse
    // row up to the first frontier node).
    const bool small_push = !pull && (multi || mf * kSmallPushRatio < n);
    if (small_push && tiny_levels && !multi && t.rrows.size() == 1 && mf <= kTinyEntries &&
        frontier_size <= kTinyNodes)
      return tiny_step();
    if (!pull) {
      if (!queue_ready) bfs_write<<<nbc, 256, 0, st>>>(words, frontier, w.offsets, w.q_a);
      const unsigned qblocks = (unsigned)std::min<uint64_t>((frontier_size + 3) / 4, 256 * 64);
      for (DevCsr* a : t.rows) {
        const unsigned grid = qblocks + a->npartials;
        if (small_push)
          bfs_push<true><<<grid, 256, 0, st>>>(w.q_a, frontier_size, qblocks, a->pieces, a->offsets, a->targets,
                                               frontier, w.visited, filter, w.cand, via);
        else
          bfs_push<false><<<grid, 256, 0, st>>>(w.q_a, frontier_size, qblocks, a->pieces, a->offsets,
                                                a->targets, frontier, w.visited, filter, w.cand, via);
      }
    }
    uint32_t nedges = 0;
    const unsigned long long seq = ++ctx->mail_seq;
    if (small_push) {
      bfs_count<true><<<nbc, 256, 0, st>>>(n, words, claim, w.visited, frontier, ra, rr, w.part, w.cand, depth,
                                           d + 1, nullptr, 0);
    } else {
      // One pass per reversed CSR (one after a push: its candidates).
      for (size_t j = 0; j < t.rrows.size(); ++j) {
        const DevCsr* r = t.rrows[j];
        const int cands = pull ? 0 : 1;
        const int pass = j == 0 ? 0 : t.first_via ? kPassFirst : kPassMin;
        if (r->nblocks)
          bfs_pull_blocks<<<r->nblocks, kBlockThreads, 0, st>>>(
              r->blocks, r->offsets, r->targets, w.visited, frontier, filter, w.cand, cands, pass, depth, via,
              claim, w.piece_hit, w.part_edges + nedges, d + 1);
        if (r->nlong)
          bfs_pull_long<<<(r->nlong + 255) / 256, 256, 0, st>>>(
              r->long_rows, r->long_pbase, r->nlong, r->offsets, r->targets, w.visited, filter, w.cand, cands,
              pass, w.piece_hit, depth, via, claim, w.part_edges + nedges + r->nblocks, d + 1);
        nedges += edge_parts(*r);
        if (!pull) break;
      }
      bfs_count<false><<<nbc, 256, 0, st>>>(n, words, claim, w.visited, frontier, ra, rr, w.part, nullptr,
                                            nullptr, 0, w.part_edges, pull ? nedges : 0);
    }
    bfs_scan<<<1, 256, 0, st>>>(nbc, w.part, w.offsets, ctx->d_mail, seq);
    // After a push the next level is most often pushed too: its queue is
    // listed now, while the host waits for the counters.
    if (!pull) bfs_write<<<nbc, 256, 0, st>>>(words, claim, w.offsets, w.q_a);
    WG_HIP(ctx, hipGetLastError());
    BfsCounters hc;
    static_assert(sizeof(BfsCounters) == 4 * sizeof(unsigned long long), "four counters");
    int s = wait_mail(ctx, seq, (unsigned long long*)&hc, 4);
    if (s) return s;
    std::swap(frontier, claim);
    queue_ready = !pull;
    // A pushed level scans every entry of every frontier row.
    run.edges += pull ? hc.edges : mf;
    run.push_edges += mf;
    frontier_size = hc.next_size;
    mf = hc.next_entries;
    visited_rev += hc.visited_rev;
    run.visited += frontier_size;
    run.level_pulled.push_back(pull ? 1 : 0);
    run.level_size.push_back(frontier_size);
    ++d;
    run.levels = d;
    return WG_OK;
  }

  // Small push levels, as many as bfs_tiny_levels takes in one launch
  // (at least this one, whose test step() has made).
  int tiny_step() {
    hipStream_t st = ctx->stream;
    if (!queue_ready) bfs_write<<<nbc, 256, 0, st>>>(words, frontier, w.offsets, w.q_a);
    const unsigned long long seq = ++ctx->mail_seq;
    const TinyState ts{frontier_size, mf, visited_rev, run.visited, d};
    bfs_tiny_levels<<<1, 1024, 0, st>>>(t.rows[0]->offsets, t.rows[0]->targets, t.rrows[0]->offsets, n,
                                        rr_nnz, pulling ? 1 : 0, max_depth, w.visited, frontier, claim, w.q_a,
                                        w.q_b, filter, depth, via, ts, ctx->d_mail, seq);
    WG_HIP(ctx, hipGetLastError());
    unsigned long long rec[1 + 3 * kTinyLevels];
    int s = wait_mail(ctx, seq, rec, 1 + 3 * kTinyLevels);
    if (s) return s;
    const uint32_t L = (uint32_t)rec[0];
    if (L == 0 || L > kTinyLevels) return fail(ctx, WG_DEVICE_ERROR, "small BFS levels: no level ran");
    for (uint32_t l = 0; l < L; ++l) {
      // A pushed level scans every entry of every frontier row.
      run.edges += mf;
      run.push_edges += mf;
      frontier_size = rec[1 + 3 * l];
      mf = rec[2 + 3 * l];
      visited_rev += rec[3 + 3 * l];
      run.visited += frontier_size;
      run.level_pulled.push_back(0);
      run.level_size.push_back(frontier_size);
      ++d;
      // Each level swaps the queues and the frontier / claim bitmaps.
      std::swap(frontier, claim);
      std::swap(w.q_a, w.q_b);
    }
    run.levels = d;
    queue_ready = true;  // in claim order, which only a push reads
    return WG_OK;
  }

  // The current level's nodes, ascending, into q_a (the next push's queue).
  int list_frontier() {
    if (!queue_ready && frontier_size)
      bfs_write<<<nbc, 256, 0, ctx->stream>>>(words, frontier, w.offsets, w.q_a);
    WG_HIP(ctx, hipGetLastError());
    queue_ready = true;
    return WG_OK;
  }
};

// Runs the search, leaving depth and via (kUnvisited where unreached) in
// device buffers. starts must be in range and deduplicated, in order.
int bfs_run(wg_context* ctx, wg_graph* g, const Traversal& t, const std::vector<uint32_t>& starts,
            const uint8_t* filter_d, int64_t max_depth, bool pulling, uint32_t* depth, uint32_t* via,
            BfsRun& run) {
  BfsWorkspace w;
  WorkspaceLease lease;
  int s = bfs_workspace(ctx, g, t, w, lease);
  if (s) return s;
  BfsSearch S;
  if ((s = S.init(ctx, g, t, starts, filter_d, max_depth, pulling, depth, via, w))) return s;
  S.tiny_levels = true;
  while (!S.exhausted())
    if ((s = S.step())) return s;
  run = S.run;
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
  int alloc(size_t b) {
    bytes = std::max<size_t>(b, 1);
    return dalloc(ctx, &p, bytes);
  }
};

// keys[i] = via[nodes[i]]: a level's nodes keyed by their via.
__global__ __launch_bounds__(256) void via_keys(const uint32_t* __restrict__ nodes,
                                                const uint32_t* __restrict__ via, uint64_t count,
                                                uint32_t* __restrict__ keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    keys[i] = via[nodes[i]];
}

// keys[i]: the first of the CSRs read whose row of via[v] holds v =
// nodes[i] (binary search: rows ascend), so a pushed level over several
// CSRs lists each via's nodes in the order its concatenated rows reach them.
__global__ __launch_bounds__(256) void row_set_keys(const uint32_t* __restrict__ nodes,
                                                    const uint32_t* __restrict__ via, uint64_t count,
                                                    const RowSets ra, uint32_t* __restrict__ keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = nodes[i], u = via[v];
    uint32_t key = ra.k;
    for (uint32_t j = 0; j < ra.k && key == ra.k; ++j) {
      uint64_t lo = ra.off[j][u], hi = ra.off[j][u + 1];
      while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ra.tgt[j][mid] < v) lo = mid + 1; else hi = mid;
      }
      if (lo < ra.off[j][u + 1] && ra.tgt[j][lo] == v) key = j;
    }
    keys[i] = key;
  }
}

// The row entries a push scans in a level up to and including the one that
// reaches *refused (the visit max_results turned away), Progress's count
// (level.rs:213-236): every row entry of the frontier nodes (depth dprev)
// below its via, then the via's rows up to the entry. Blocks add their sums
// to *out, which starts at 0.
__global__ __launch_bounds__(256) void push_scan_edges(const uint32_t* __restrict__ refused,
                                                       const uint32_t* __restrict__ via,
                                                       const uint32_t* __restrict__ depth, uint32_t dprev,
                                                       const RowSets ra, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[4];
  const uint32_t v = *refused, u = via[v];
  unsigned long long sum = 0;
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < u; x += (uint64_t)gridDim.x * 256)
    if (depth[x] == dprev) sum += degree_in(ra, (uint32_t)x);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (uint32_t j = 0; j < ra.k; ++j) {
      const uint64_t a = ra.off[j][u], e = ra.off[j][u + 1];
      uint64_t lo = a, hi = e;
      while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (ra.tgt[j][mid] < v) lo = mid + 1; else hi = mid;
      }
      if (lo < e && ra.tgt[j][lo] == v) {
        sum += lo - a + 1;
        break;
      }
      sum += e - a;
    }
  }
  sum = block_reduce_sum(sum, red);
  if (threadIdx.x == 0 && sum) atomicAdd(out, sum);
}

// The level's node ids in emission order (a reachable set's rows).
__global__ __launch_bounds__(256) void write_ids(const uint32_t* __restrict__ nodes, uint64_t count,
                                                 uint64_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = nodes[i];
}

// The level's visits in emission order.
__global__ __launch_bounds__(256) void write_visits(const uint32_t* __restrict__ nodes,
                                                    const uint32_t* __restrict__ via, uint64_t count,
                                                    uint32_t depth, wg_visit* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = nodes[i];
    out[i] = wg_visit{v, via[v], depth, 0};
  }
}

// A search that hands out its visits level by level, in the reference's
// order: the starts in the order given; a pushed level by (via, first CSR
// whose row of the via holds the node, node), the order reached when
// frontier nodes expand in ascending order; a pulled level by node when
// `reachable` (a reachable set), else as if pushed. Visits are sorted and
// written on the device; a level's rows reach the host in one copy. At
// most max_results visits (< 0: no limit); a visit beyond them sets
// truncated. Outside a reachable set (whose levels the reference may pull)
// the stats are a push's: every level's frontier row entries, up to the
// refused visit's entry in a truncated level, and no level pulled.
struct LevelEmitter {
  wg_context* ctx = nullptr;
  wg_graph* g = nullptr;
  std::vector<std::pair<void*, size_t>> owned;  // workspace and buffers
  void* slot[kWsSlots] = {};
  BfsSearch S;
  std::vector<uint32_t> starts;
  uint32_t *depth = nullptr, *via = nullptr, *keys_in = nullptr, *keys_out = nullptr, *nodes_tmp = nullptr;
  unsigned long long* scan_count = nullptr;
  uint8_t* filter = nullptr;
  wg_visit* visits = nullptr;
  int64_t max_results = -1;
  bool reachable = false, seeded = false, finished = false, truncated = false;
  bool ids_only = false;  // rows are node ids (u64), not wg_visit
  uint64_t emitted = 0;
  uint64_t push_edges = 0;  // stats of a push, outside a reachable set

  ~LevelEmitter() {
    if (ctx) {
      (void)hipStreamSynchronize(ctx->stream);
      for (auto& b : owned) dfree(ctx, b.first, b.second);
    }
  }

  int buffer(size_t bytes, void** p) {
    int s = dalloc(ctx, p, std::max<size_t>(bytes, 1));
    if (!s) owned.push_back({*p, std::max<size_t>(bytes, 1)});
    return s;
  }

  int init(wg_context* c, wg_graph* gr, int direction, const uint32_t* rels, uint32_t nrel, const uint64_t* st,
           uint64_t ns, const uint8_t* host_filter, int64_t max_depth, int64_t maxr, bool pulling, bool reach) {
    ctx = c;
    g = gr;
    max_results = maxr;
    reachable = reach;
    Traversal t;
    int s = traversal_of(ctx, g, direction, rels, nrel, reachable, t);
    if (s) return s;
    if ((s = prepare_starts(ctx, g->n, st, ns, starts))) return s;
    const uint64_t n = std::max<uint64_t>(g->n, 1);
    size_t sizes[kWsSlots];
    workspace_sizes(g, t, sizes);
    if ((s = alloc_workspace(ctx, sizes, slot, owned))) return s;
    if ((s = buffer(n * 4, (void**)&depth)) || (s = buffer(n * 4, (void**)&via)) ||
        (s = buffer(n * 4, (void**)&keys_in)) || (s = buffer(n * 4, (void**)&keys_out)) ||
        (s = buffer(n * sizeof(wg_visit), (void**)&visits)) || (s = buffer(8, (void**)&scan_count)))
      return s;
    if (t.rows.size() > 1 && (s = buffer(n * 4, (void**)&nodes_tmp))) return s;
    if (host_filter) {
      if ((s = buffer(n, (void**)&filter))) return s;
      WG_HIP(ctx, hipMemcpy(filter, host_filter, g->n, hipMemcpyHostToDevice));
    }
    BfsWorkspace w;
    view_workspace(g, slot, w);
    return S.init(ctx, g, t, starts, filter, max_depth, pulling, depth, via, w);
  }

  // The room left under max_results.
  uint64_t room() const {
    return max_results < 0 ? UINT64_MAX : (uint64_t)max_results - std::min<uint64_t>(emitted, max_results);
  }

  // The level's nodes in emission order (device; size of them).
  int order_level(uint64_t size, const uint32_t** nodes) {
    int s;
    *nodes = S.w.q_a;
    if (reachable && S.run.level_pulled.back()) return WG_OK;  // by node
    const unsigned grid = (unsigned)std::min<uint64_t>((size + 255) / 256, 4096);
    unsigned bits = 1;
    while (bits < 32 && (1ull << bits) < g->n) ++bits;
    const uint32_t* by_node = S.w.q_a;
    if (S.t.rows.size() > 1) {
      // Stable sorts, least significant key first: by CSR, then by via.
      row_set_keys<<<grid, 256, 0, ctx->stream>>>(S.w.q_a, via, size, S.ra, keys_in);
      WG_HIP(ctx, hipGetLastError());
      unsigned kbits = 1;
      while ((1u << kbits) <= S.ra.k) ++kbits;
      if ((s = sort_u32_pairs(ctx, keys_in, keys_out, S.w.q_a, nodes_tmp, size, kbits))) return s;
      by_node = nodes_tmp;
    }
    via_keys<<<grid, 256, 0, ctx->stream>>>(by_node, via, size, keys_in);
    WG_HIP(ctx, hipGetLastError());
    // Stable: equal vias keep the order above.
    if ((s = sort_u32_pairs(ctx, keys_in, keys_out, by_node, S.w.q_b, size, bits))) return s;
    *nodes = S.w.q_b;
    return WG_OK;
  }

  // The next level's visits into out (host, room for the graph's n: visits,
  // or node ids when ids_only); *count of them. Returns with *more = 0 once
  // nothing follows.
  int next(void* out, uint64_t* count, int* more) {
    *count = 0;
    if (finished) {
      *more = 0;
      return WG_OK;
    }
    int s;
    uint64_t size = 0;
    if (!seeded) {
      seeded = true;
      size = starts.size();
      const uint64_t k = std::min<uint64_t>(size, room());
      for (uint64_t i = 0; i < k; ++i)
        if (ids_only)
          ((uint64_t*)out)[i] = starts[i];
        else
          ((wg_visit*)out)[i] = wg_visit{starts[i], UINT64_MAX, 0, 0};
      *count = k;
    } else {
      const uint64_t level_entries = S.mf;
      if ((s = S.step())) return s;
      size = S.frontier_size;
      const uint64_t k = std::min<uint64_t>(size, room());
      const bool cut = k < size;
      push_edges += level_entries;
      if (k || (cut && !reachable)) {
        if ((s = S.list_frontier())) return s;
        const uint32_t* nodes;
        if ((s = order_level(size, &nodes))) return s;
        if (k) {
          const unsigned grid = (unsigned)std::min<uint64_t>((k + 255) / 256, 4096);
          const size_t row = ids_only ? sizeof(uint64_t) : sizeof(wg_visit);
          if (ids_only)
            write_ids<<<grid, 256, 0, ctx->stream>>>(nodes, k, (uint64_t*)visits);
          else
            write_visits<<<grid, 256, 0, ctx->stream>>>(nodes, via, k, S.d, visits);
          WG_HIP(ctx, hipGetLastError());
          WG_HIP(ctx, hipMemcpyAsync(out, visits, k * row, hipMemcpyDeviceToHost, ctx->stream));
        }
        if (cut && !reachable) {
          // The push stops at the refused visit: its level counts the
          // entries scanned up to it.
          unsigned long long scanned = 0;
          WG_HIP(ctx, hipMemsetAsync(scan_count, 0, 8, ctx->stream));
          push_scan_edges<<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((g->n + 255) / 256, 2048)), 256, 0,
                            ctx->stream>>>(nodes + k, via, depth, S.d - 1, S.ra, scan_count);
          WG_HIP(ctx, hipGetLastError());
          WG_HIP(ctx, hipMemcpyAsync(&scanned, scan_count, 8, hipMemcpyDeviceToHost, ctx->stream));
          WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
          push_edges += scanned - level_entries;
        }
        WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
      }
      *count = k;
    }
    emitted += *count;
    if (*count < size) truncated = true;
    // A level that fills max_results exactly is not yet a truncation: the
    // next call finds out whether a visit follows (Results::admit).
    finished = truncated || S.exhausted();
    *more = finished ? 0 : 1;
    return WG_OK;
  }

  wg_traversal_stats stats() const {
    if (!reachable) return wg_traversal_stats{emitted, push_edges, truncated ? 1u : 0u, 0};
    uint32_t pulled = 0;
    for (uint32_t p : S.run.level_pulled) pulled += p;
    return wg_traversal_stats{emitted, S.run.edges, truncated ? 1u : 0u, pulled};
  }
};

}  // namespace wg

using namespace wg;

struct wg_bfs_cursor {
  wg::LevelEmitter e;
};

namespace wg {
void release_cursor(wg_bfs_cursor* c) {
  LevelEmitter& e = c->e;
  if (!e.ctx) return;
  (void)hipStreamSynchronize(e.ctx->stream);
  for (auto& b : e.owned) dfree(e.ctx, b.first, b.second);
  e.owned.clear();
  e.ctx = nullptr;
  e.g = nullptr;
  e.finished = true;
}

bool cursor_reads(const wg_bfs_cursor* c, const wg_graph* g) {
  for (const wg_graph* r = c->e.g; r; r = r->owner)
    if (r == g) return true;
  return false;
}
}  // namespace wg

extern "C" {

int wg_bfs_levels(wg_context* ctx, const wg_graph* gc, int direction, const uint32_t* relations,
                  uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, const uint8_t* filter,
                  int64_t max_depth, int pulling, wg_vector* depth, wg_vector* via, wg_traversal_stats* stats) {
  if (!ctx || !gc || !depth || !via) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  wg_graph* g = const_cast<wg_graph*>(gc);
  Traversal t;
  int s = traversal_of(ctx, g, direction, relations, nrelations, false, t);
  if (s) return s;
  for (wg_vector* v : {depth, via})
    if (v->dtype != WG_U32 || v->width != 1 || v->id_space != g->n)
      return fail(ctx, WG_INVALID_ARGUMENT, "depth and via must be u32 vectors over the graph");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  std::vector<uint32_t> st;
  if ((s = prepare_starts(ctx, g->n, starts, nstarts, st))) return s;
  DevBytes fb{ctx};
  if (filter) {
    if ((s = fb.alloc(g->n))) return s;
    WG_HIP(ctx, hipMemcpy(fb.p, filter, g->n, hipMemcpyHostToDevice));
  }
  BfsRun run;
  s = bfs_run(ctx, g, t, st, (const uint8_t*)fb.p, max_depth, pulling != 0,
              (uint32_t*)depth->data, (uint32_t*)via->data, run);
  if (s) return s;
  if (stats) {
    uint32_t pulled = 0;
    for (uint32_t p : run.level_pulled) pulled += p;
    *stats = wg_traversal_stats{run.visited, run.edges, 0, pulled};
  }
  return ok(ctx);
}

// Every level of an emitter into out.
static int emit_all(wg_context* ctx, LevelEmitter& e, void* out, uint64_t* count) {
  uint64_t total = 0;
  int more = 1;
  while (more) {
    uint64_t k = 0;
    const size_t row = e.ids_only ? sizeof(uint64_t) : sizeof(wg_visit);
    int s = e.next((char*)out + total * row, &k, &more);
    if (s) return s;
    total += k;
  }
  *count = total;
  return WG_OK;
}

int wg_bfs(wg_context* ctx, const wg_graph* gc, int direction, const uint32_t* relations, uint32_t nrelations,
           const uint64_t* starts, uint64_t nstarts, const uint8_t* filter, int64_t max_depth, int64_t max_results,
           wg_visit* out, uint64_t* count, wg_traversal_stats* stats) {
  if (!ctx || !gc || !out || !count) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  LevelEmitter e;
  int s = e.init(ctx, const_cast<wg_graph*>(gc), direction, relations, nrelations, starts, nstarts, filter,
                 max_depth, max_results, true, false);
  if (!s) s = emit_all(ctx, e, out, count);
  if (s) return s;
  if (stats) *stats = e.stats();
  return ok(ctx);
}

int wg_reachable_set(wg_context* ctx, const wg_graph* gc, int direction, const uint32_t* relations,
                     uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, int64_t max_depth, uint64_t* out,
                     uint64_t* count, wg_traversal_stats* stats) {
  if (!ctx || !gc || !out || !count) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  LevelEmitter e;
  int s = e.init(ctx, const_cast<wg_graph*>(gc), direction, relations, nrelations, starts, nstarts, nullptr,
                 max_depth, -1, true, true);
  e.ids_only = true;
  if (!s) s = emit_all(ctx, e, out, count);
  if (s) return s;
  if (stats) *stats = e.stats();
  return ok(ctx);
}

int wg_bfs_cursor_create(wg_context* ctx, const wg_graph* g, int direction, const uint32_t* relations,
                         uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, const uint8_t* filter,
                         int64_t max_depth, int64_t max_results, int reachable, wg_bfs_cursor** out) {
  if (!ctx || !g || !out) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, g)) return cs_;
  *out = nullptr;
  if (reachable && (filter || max_results >= 0))
    return fail(ctx, WG_INVALID_ARGUMENT, "a reachable-set cursor takes no filter or result limit");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_bfs_cursor* c = new (std::nothrow) wg_bfs_cursor();
  if (!c) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  // Levels are pushed or pulled as wg_bfs does (same visits and stats); a
  // reachable set lists a pulled level by node (bfs_cursor.rs:74-87).
  int s = c->e.init(ctx, const_cast<wg_graph*>(g), direction, relations, nrelations, starts, nstarts, filter,
                    max_depth, max_results, true, reachable != 0);
  if (s) {
    delete c;
    return s;
  }
  ctx->cursors.insert(c);
  *out = c;
  return ok(ctx);
}

int wg_bfs_cursor_next(wg_context* ctx, wg_bfs_cursor* c, wg_visit* out, uint64_t* count, int* more) {
  if (!ctx || !c || !out || !count || !more) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (c->e.ctx != ctx)
    return fail(ctx, WG_INVALID_ARGUMENT,
                c->e.ctx ? "the cursor belongs to another context" : "the cursor's graph or context was destroyed");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  int s = c->e.next(out, count, more);
  return s ? s : ok(ctx);
}

int wg_bfs_cursor_stats(const wg_bfs_cursor* c, wg_traversal_stats* stats) {
  if (!c || !stats) return WG_INVALID_ARGUMENT;
  *stats = c->e.stats();
  return WG_OK;
}

void wg_bfs_cursor_destroy(wg_bfs_cursor* c) {
  if (!c) return;
  if (c->e.ctx) c->e.ctx->cursors.erase(c);
  release_cursor(c);
  delete c;
}

}  // extern "C"
