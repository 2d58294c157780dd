// gather (compute/gather/mod.rs): per-node messages fanned out over the
// graph's edges and combined at each receiver, on the device.
//
// A receiver's senders are its distinct neighbours in the gather's
// direction over the selected relations (in: the reverse CSR's row, out:
// the forward row, both: their union), once each, or with per_relation once
// per relation linking them; each sender's messages arrive in push order,
// the senders in ascending dense id, the receiver's own messages first with
// include_self (mod.rs:1-23). The pieces:
//   messages   a stable sort by sender (push order kept) gives each sender's
//              messages as a contiguous run (the outbox, outbox.rs:41-75);
//   senders    a CSR per receiver: a relation selection's forward or
//              reverse CSR, or for both directions the union of the two
//              built like a selection; kept with the graph;
//   collect    each sender entry's message count, prefix sums over the
//              entries and over the receivers give every (receiver, sender)
//              run its place, so a thread per entry copies its sender's
//              messages there: each receiver's messages contiguous, in
//              collect order, written in parallel;
//   combine    a thread per receiver (a 64-lane block for a segment of 512
//              or more, staging 256 at a time in LDS) folds its contiguous
//              messages left to right with the
//              reference's operator rules (SemiringElement::apply: i64 sums
//              checked, min and max in SQL order, the first operand kept on
//              ties), so float sums are the reference's bit for bit; count
//              needs only the counts;
//   mode       the collected lists sorted by (receiver, value in SQL order)
//              stably, so a run of equal values keeps collect order; a
//              receiver's longest run wins, the smallest value on a tie,
//              and its first message is the answer.
// Receivers that got no message are left out; the rest come in ascending
// dense-id order (gathered.rs:10-16).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr unsigned kThreads = 256;
// Combiner::ALL order (compute/gather/combiner.rs:46-57).
enum { C_COLLECT = 0, C_COUNT, C_SUM, C_MIN, C_MAX, C_FIRST, C_CONCAT, C_BOOL_OR, C_BOOL_AND, C_MODE };

unsigned grid_of(uint64_t items) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kThreads - 1) / kThreads, 8192));
}

__device__ __forceinline__ uint64_t first_index() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t stride() { return (uint64_t)gridDim.x * blockDim.x; }

__global__ __launch_bounds__(kThreads) void first_sender_outside(const uint32_t* __restrict__ senders, uint64_t k,
                                                                 uint64_t n, unsigned long long* __restrict__ first) {
  for (uint64_t i = first_index(); i < k; i += stride())
    if (senders[i] >= n) {
      atomicMin(first, (unsigned long long)i);
      return;
    }
}

template <class E>
__global__ __launch_bounds__(kThreads) void permute(const uint32_t* __restrict__ perm, const E* __restrict__ in,
                                                    uint64_t k, E* __restrict__ out) {
  for (uint64_t i = first_index(); i < k; i += stride()) out[i] = in[perm[i]];
}

__global__ __launch_bounds__(kThreads) void iota32(uint32_t* __restrict__ a, uint64_t k) {
  for (uint64_t i = first_index(); i < k; i += stride()) a[i] = (uint32_t)i;
}

// moff[s] = the first sorted message of sender s, s in [0, n].
__global__ __launch_bounds__(kThreads) void sender_starts(const uint32_t* __restrict__ sorted, uint64_t k,
                                                          uint64_t n, uint64_t* __restrict__ moff) {
  for (uint64_t s = first_index(); s <= n; s += stride()) {
    uint64_t lo = 0, hi = k;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)sorted[mid] < s)
        lo = mid + 1;
      else
        hi = mid;
    }
    moff[s] = lo;
  }
}

struct Receiving {
  const uint64_t* soff;  // senders CSR (null: no senders)
  const uint32_t* stgt;
  const uint64_t* moff;  // each sender's run of messages
  const void* mask;
  int mask_dtype, mask_complement, self;
};

__device__ __forceinline__ bool selected(const Receiving& R, uint64_t r) {
  return !R.mask || mask_selects(R.mask, R.mask_dtype, R.mask_complement, r);
}

// ecnt[p] = the messages of entry p's sender (ecnt[m] = 0 closes the scan).
__global__ __launch_bounds__(kThreads) void entry_counts(const uint32_t* __restrict__ stgt, uint64_t m,
                                                         const uint64_t* __restrict__ moff,
                                                         uint64_t* __restrict__ ecnt) {
  if (m == 0) {  // no entry: only the scan's closing zero (moff may hold one offset)
    if (first_index() == 0) ecnt[0] = 0;
    return;
  }
  // Four entries a step (p, p + stride, ...), their loads in flight together
  // (a slot past the end reads sender 0's offsets and drops them).
  const uint64_t st = stride();
  for (uint64_t p = first_index(); p <= m; p += 4 * st) {
    uint32_t s4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) s4[k] = p + k * st < m ? stgt[p + k * st] : 0;
    uint64_t a4[4], b4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a4[k] = moff[s4[k]];
      b4[k] = moff[s4[k] + 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t q = p + k * st;
      if (q < m) ecnt[q] = b4[k] - a4[k];
      else if (q == m) ecnt[q] = 0;
    }
  }
}

// cnt[r] = the messages receiver r gets: its own (include_self) and its
// senders' (a difference of the entries' prefix sums); 0 outside the mask.
// cnt[n] = 0 closes the scan.
__global__ __launch_bounds__(kThreads) void receiver_counts(Receiving R, uint64_t n,
                                                            const uint64_t* __restrict__ eprefix,
                                                            uint64_t* __restrict__ cnt) {
  for (uint64_t r = first_index(); r <= n; r += stride()) {
    if (r == n || !selected(R, r)) {
      cnt[r] = 0;
      continue;
    }
    uint64_t c = R.self ? R.moff[r + 1] - R.moff[r] : 0;
    if (R.soff) c += eprefix[R.soff[r + 1]] - eprefix[R.soff[r]];
    cnt[r] = c;
  }
}

// The collected messages: a receiver's own first (thread per receiver),
// then each sender entry's run at its prefix position (thread per entry),
// so every receiver's messages lie in collect order at [start[r],
// start[r + 1]). ord, if given, gets the receiver's rank at each message.
// fill_self also leaves base[r], where entry p of receiver r puts its run:
// base[r] + eprefix[p] (its own messages first, then its entries' in order).
template <class E>
__global__ __launch_bounds__(kThreads) void fill_self(Receiving R, uint64_t n, const E* __restrict__ mval,
                                                      const uint64_t* __restrict__ start,
                                                      const uint64_t* __restrict__ eprefix,
                                                      const uint64_t* __restrict__ rank, E* __restrict__ out,
                                                      uint32_t* __restrict__ ord, uint64_t* __restrict__ base) {
  for (uint64_t r = first_index(); r < n; r += stride()) {
    if (R.soff)
      base[r] = start[r] + (R.self ? R.moff[r + 1] - R.moff[r] : 0) - eprefix[R.soff[r]];
    if (!R.self || !selected(R, r)) continue;
    uint64_t at = start[r];
    for (uint64_t k = R.moff[r]; k < R.moff[r + 1]; ++k, ++at) {
      out[at] = mval[k];
      if (ord) ord[at] = (uint32_t)rank[r];
    }
  }
}

template <class E>
__global__ __launch_bounds__(kThreads) void fill_entries(Receiving R, uint64_t m, const uint32_t* __restrict__ row_of,
                                                         const uint64_t* __restrict__ eprefix,
                                                         const E* __restrict__ mval,
                                                         const uint64_t* __restrict__ base,
                                                         const uint64_t* __restrict__ rank, E* __restrict__ out,
                                                         uint32_t* __restrict__ ord) {
  // Four entries a step (p, p + stride, ...): each stage's loads go out
  // together (a slot past the end reads entry 0's and drops it).
  const uint64_t st = stride();
  for (uint64_t p = first_index(); p < m; p += 4 * st) {
    uint32_t r4[4], s4[4];
    uint64_t e4[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint64_t q = p + k * st < m ? p + k * st : 0;
      r4[k] = row_of[q];
      s4[k] = R.stgt[q];
      e4[k] = eprefix[q];
    }
    uint64_t at4[4], m0[4], m1[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t r = r4[k];
      at4[k] = base[r] + e4[k];
      m0[k] = R.moff[s4[k]];
      m1[k] = R.moff[s4[k] + 1];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (p + k * st >= m || !selected(R, r4[k])) continue;
      uint64_t at = at4[k];
      for (uint64_t j = m0[k]; j < m1[k]; ++j, ++at) {
        out[at] = mval[j];
        if (ord) ord[at] = (uint32_t)rank[r4[k]];
      }
    }
  }
}

// Left fold of each receiver's collected messages with OP (SemiringElement::
// apply, in collect order); the lowest failing receiver raises. A thread
// per receiver folds a segment shorter than kLong with sixteen loads ahead
// of the fold; longer segments are listed (in any order: each is folded on
// its own) for fold_long.
constexpr uint64_t kLong = 512;

template <class E, int OP>
__device__ __forceinline__ void fold_into(E& acc, bool& first, uint32_t& e_err, E x) {
  if (first) {
    acc = x;
    first = false;
  } else if (!e_err) {
    acc = apply<OP>(acc, x, e_err);
  }
}

template <class E, int OP>
__global__ __launch_bounds__(kThreads) void fold_segments(uint64_t n, const uint64_t* __restrict__ start,
                                                          const E* __restrict__ coll, E* __restrict__ val,
                                                          uint32_t* __restrict__ long_list,
                                                          unsigned* __restrict__ nlong,
                                                          unsigned long long* __restrict__ err) {
  for (uint64_t r = first_index(); r < n; r += stride()) {
    const uint64_t a = start[r], b = start[r + 1];
    if (a == b) continue;
    if (b - a >= kLong) {
      long_list[atomicAdd(nlong, 1u)] = (uint32_t)r;
      continue;
    }
    E acc{};
    bool first = true;
    uint32_t e_err = 0;
    constexpr int kAhead = 16;
    uint64_t k = a;
    for (; k + kAhead <= b && !e_err; k += kAhead) {
      E v[kAhead];
#pragma unroll
      for (int t = 0; t < kAhead; ++t) v[t] = coll[k + t];
#pragma unroll
      for (int t = 0; t < kAhead; ++t) fold_into<E, OP>(acc, first, e_err, v[t]);
    }
    for (; k < b && !e_err; ++k) fold_into<E, OP>(acc, first, e_err, coll[k]);
    val[r] = acc;
    if (e_err) raise_error(err, r, e_err);
  }
}

// A long segment, one 64-lane block each: the block stages 256 messages at
// a time in LDS, the next 256 loading into registers while lane 0 folds
// the staged ones in order.
constexpr unsigned kLongThreads = 64, kStage = 256;

template <class E, int OP>
__global__ __launch_bounds__(kLongThreads) void fold_long(const uint32_t* __restrict__ long_list,
                                                          const uint64_t* __restrict__ start,
                                                          const E* __restrict__ coll, E* __restrict__ val,
                                                          unsigned long long* __restrict__ err) {
  __shared__ E stage[kStage];
  const uint32_t r = long_list[blockIdx.x];
  const uint64_t a = start[r], b = start[r + 1];
  const unsigned lane = threadIdx.x;
  E next[kStage / kLongThreads];
  auto load = [&](uint64_t base) {
#pragma unroll
    for (unsigned c = 0; c < kStage / kLongThreads; ++c) {
      const uint64_t k = base + c * kLongThreads + lane;
      next[c] = k < b ? coll[k] : E{};
    }
  };
  E acc{};
  bool first = true;
  uint32_t e_err = 0;
  load(a);
  for (uint64_t base = a; base < b; base += kStage) {
#pragma unroll
    for (unsigned c = 0; c < kStage / kLongThreads; ++c) stage[c * kLongThreads + lane] = next[c];
    __syncthreads();
    if (base + kStage < b) load(base + kStage);
    if (lane == 0) {
      const unsigned cnt = (unsigned)(b - base < kStage ? b - base : kStage);
      for (unsigned t = 0; t < cnt; ++t) fold_into<E, OP>(acc, first, e_err, stage[t]);
    }
    __syncthreads();
  }
  if (lane == 0) {
    val[r] = acc;
    if (e_err) raise_error(err, r, e_err);
  }
}

// received[r] = 1 if r got a message.
__global__ __launch_bounds__(kThreads) void received_flags(const uint64_t* __restrict__ cnt, uint64_t n,
                                                           uint64_t* __restrict__ flags) {
  for (uint64_t r = first_index(); r < n; r += stride()) flags[r] = cnt[r] ? 1 : 0;
}

// The receivers that got a message, ascending, their values and (collect)
// their offsets into the collected messages.
template <class E>
__global__ __launch_bounds__(kThreads) void compact_receivers(const uint64_t* __restrict__ cnt,
                                                              const uint64_t* __restrict__ rank, uint64_t n,
                                                              const E* __restrict__ val,
                                                              const uint64_t* __restrict__ start,
                                                              uint32_t* __restrict__ recv, E* __restrict__ oval,
                                                              int64_t* __restrict__ ocnt,
                                                              uint64_t* __restrict__ ooff) {
  for (uint64_t r = first_index(); r < n; r += stride()) {
    if (!cnt[r]) continue;
    const uint64_t i = rank[r];
    recv[i] = (uint32_t)r;
    if (oval) oval[i] = val[r];
    if (ocnt) ocnt[i] = (int64_t)cnt[r];
    if (ooff) ooff[i] = start[r];
  }
}

// A key whose unsigned order is SQL order: NaN above every number and one
// value, -0 equal to 0 (SemiringElement::sql_cmp).
template <class E>
__device__ __forceinline__ uint64_t sql_key(E v) {
  if constexpr (std::is_same<E, double>::value || std::is_same<E, float>::value) {
    const double d = (double)v;
    if (d != d) return ~0ull;
    if (d == 0.0) return 0x8000000000000000ull;
    const uint64_t b = __double_as_longlong(d);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
  } else if constexpr (std::is_same<E, int64_t>::value) {
    return (uint64_t)v ^ 0x8000000000000000ull;
  } else {
    return (uint64_t)(v != 0);
  }
}

template <class E>
__global__ __launch_bounds__(kThreads) void sql_keys(const E* __restrict__ vals, uint64_t k,
                                                     uint64_t* __restrict__ keys, uint64_t* __restrict__ pos) {
  for (uint64_t i = first_index(); i < k; i += stride()) {
    keys[i] = sql_key(vals[i]);
    pos[i] = i;
  }
}

__global__ __launch_bounds__(kThreads) void ord_of(const uint32_t* __restrict__ ord, const uint64_t* __restrict__ pos,
                                                   uint64_t k, uint32_t* __restrict__ out) {
  for (uint64_t i = first_index(); i < k; i += stride()) out[i] = ord[pos[i]];
}

// The mode of receiver i's collected messages, now in (value, collect)
// order at [off[i], off[i+1]): the longest run of one SQL value, the
// first such (the smallest value) on a tie, answered by its first message.
template <class E>
__global__ __launch_bounds__(kThreads) void mode_pick(const uint64_t* __restrict__ off, uint64_t nrecv,
                                                      const uint64_t* __restrict__ pos, const E* __restrict__ vals,
                                                      E* __restrict__ out) {
  for (uint64_t i = first_index(); i < nrecv; i += stride()) {
    uint64_t best_at = off[i], best_len = 0;
    uint64_t k = off[i];
    while (k < off[i + 1]) {
      const uint64_t key = sql_key(vals[pos[k]]);
      uint64_t j = k + 1;
      while (j < off[i + 1] && sql_key(vals[pos[j]]) == key) ++j;
      if (j - k > best_len) {
        best_len = j - k;
        best_at = k;
      }
      k = j;
    }
    out[i] = vals[pos[best_at]];
  }
}

// ---- byte messages ------------------------------------------------------
// Byte-string messages ride through the gather as their indexes (i64
// messages i = 0..count-1); the strings the result holds are then laid out
// by these kernels: each slot's length, its place (an exclusive scan), and
// a thread per output byte that finds its slot by binary search.
__global__ __launch_bounds__(kThreads) void string_lengths(const int64_t* __restrict__ idx, uint64_t k,
                                                           const uint64_t* __restrict__ boff,
                                                           uint64_t* __restrict__ len) {
  for (uint64_t j = first_index(); j < k; j += stride()) {
    const uint64_t m = (uint64_t)idx[j];
    len[j] = boff[m + 1] - boff[m];
  }
  if (first_index() == 0) len[k] = 0;
}

__global__ __launch_bounds__(kThreads) void string_bytes(const int64_t* __restrict__ idx, uint64_t k,
                                                         const uint64_t* __restrict__ pos,
                                                         const uint64_t* __restrict__ boff,
                                                         const uint8_t* __restrict__ in, uint64_t total,
                                                         uint8_t* __restrict__ out) {
  for (uint64_t b = first_index(); b < total; b += stride()) {
    uint64_t lo = 0, hi = k;  // the last slot j with pos[j] <= b
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (pos[mid] <= b) lo = mid;
      else hi = mid;
    }
    out[b] = in[boff[(uint64_t)idx[lo]] + (b - pos[lo])];
  }
}

// concat: each receiver's string spans its messages' strings.
__global__ __launch_bounds__(kThreads) void receiver_spans(const uint64_t* __restrict__ offs, uint64_t nrecv,
                                                           const uint64_t* __restrict__ pos,
                                                           uint64_t* __restrict__ out) {
  for (uint64_t i = first_index(); i <= nrecv; i += stride()) out[i] = pos[offs[i]];
}

// A device buffer charged to the context, freed on scope exit.
template <class T>
struct Buf {
  wg_context* ctx = nullptr;
  T* p = nullptr;
  size_t bytes = 0;
  int alloc(wg_context* c, uint64_t count) {
    ctx = c;
    bytes = std::max<uint64_t>(count, 1) * sizeof(T);
    return dalloc(ctx, (void**)&p, bytes);
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
  void reset() {
    if (p) dfree(ctx, p, bytes);
    p = nullptr;
  }
  ~Buf() { reset(); }
};

// Combiner::check for byte messages (combiner.rs:82-94): collect, count,
// first and concat.
int combiner_accepts_bytes(int combiner) {
  if (combiner < C_COLLECT || combiner > C_MODE) return -1;
  return combiner == C_COLLECT || combiner == C_COUNT || combiner == C_FIRST || combiner == C_CONCAT;
}

int combiner_accepts(int combiner, int dtype) {
  const bool numeric = dtype == WG_F32 || dtype == WG_F64 || dtype == WG_I64;
  switch (combiner) {
    case C_COLLECT: case C_COUNT: case C_FIRST: return 1;
    case C_SUM: case C_MIN: case C_MAX: return numeric;
    case C_BOOL_OR: case C_BOOL_AND: return dtype == WG_BOOL;
    case C_CONCAT: return 0;  // bytes messages only
    case C_MODE: return 1;
  }
  return -1;
}

}  // namespace

// The senders CSR of a gather over g (Sources, neighbourhood.rs): rows are
// receivers. In and out read a relation selection's reverse or forward
// CSR; both builds the union, kept with g.
static int senders_csr(wg_context* ctx, wg_graph* g, int direction, const uint32_t* rels, uint32_t k,
                       int per_relation, const DevCsr** out) {
  wg_graph* v;
  int s;
  if (direction != WG_BOTH) {
    if ((s = select_graph(ctx, g, rels, k, per_relation, &v))) return s;
    *out = direction == WG_IN ? &v->rev : &v->fwd;
    return WG_OK;
  }
  // The relations linking: each once (merged), or each selected relation's
  // own union, those kept side by side in relation order (per relation).
  std::vector<uint32_t> sel;
  if (rels) {
    sel.assign(rels, rels + k);
    std::sort(sel.begin(), sel.end());
    sel.erase(std::unique(sel.begin(), sel.end()), sel.end());
    for (uint32_t r : sel)
      if (r >= g->relation_count)
        return fail(ctx, WG_UNKNOWN_RELATION, "relation " + std::to_string(r) + " is not in the graph");
  } else {
    for (uint32_t r = 0; r < g->relation_count; ++r) sel.push_back(r);
  }
  const bool per = per_relation && g->relations.size() > 1 && sel.size() > 1;
  std::string key = std::string("b") + (per ? "p" : "m");
  for (uint32_t r : sel) key += ":" + std::to_string(r);
  auto it = g->selections.find(key);
  if (it != g->selections.end()) {
    *out = &it->second->fwd;
    return WG_OK;
  }
  auto relation = [&](uint32_t r) -> wg_graph* { return g->relations.empty() ? g : g->relations[r]; };
  // The merged view of all selected relations serves the merged union.
  std::vector<wg_graph*> parts;
  if (per) {
    for (uint32_t r : sel) parts.push_back(relation(r));
  } else {
    wg_graph* merged;
    if ((s = select_graph(ctx, g, rels, k, 0, &merged))) return s;
    parts.push_back(merged);
  }
  // Each part's union of forward and reverse rows, folded.
  std::vector<DevCsr> unions(parts.size());
  uint64_t total = 0;
  auto free_unions = [&]() {
    for (DevCsr& c : unions) free_csr(ctx, c);
  };
  for (size_t i = 0; i < parts.size(); ++i) {
    const wg_graph* p = parts[i];
    const uint64_t m2 = p->fwd.nnz + p->rev.nnz;
    Buf<uint32_t> src, dst;
    if ((s = src.alloc(ctx, m2)) || (s = dst.alloc(ctx, m2)) || (s = csr_edges(ctx, p->fwd, src.p, dst.p)) ||
        (s = csr_edges(ctx, p->rev, src.p + p->fwd.nnz, dst.p + p->fwd.nnz)) ||
        (s = csr_of_edges(ctx, g->n, m2, src.p, dst.p, false, unions[i]))) {
      free_unions();
      return s;
    }
    total += unions[i].nnz;
  }
  wg_graph* view = new (std::nothrow) wg_graph();
  if (!view) {
    free_unions();
    return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  }
  view->ctx = ctx;
  view->n = g->n;
  view->owner = g;
  if (parts.size() == 1) {
    view->fwd = std::move(unions[0]);
    unions.clear();
  } else {
    // Side by side in relation order: the union of each, kept as it is.
    Buf<uint32_t> src, dst;
    if ((s = src.alloc(ctx, total)) || (s = dst.alloc(ctx, total))) {
      free_unions();
      destroy_graph(view);
      return s;
    }
    uint64_t base = 0;
    for (DevCsr& c : unions) {
      if ((s = csr_edges(ctx, c, src.p + base, dst.p + base))) break;
      base += c.nnz;
    }
    if (!s) s = csr_of_edges(ctx, g->n, total, src.p, dst.p, true, view->fwd);
    free_unions();
    if (s) {
      destroy_graph(view);
      return s;
    }
  }
  view->nnz = view->fwd.nnz;
  g->selections[key] = view;
  *out = &view->fwd;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

struct wg_gathered {
  wg_context* ctx = nullptr;
  int combiner = 0, dtype = 0;
  uint64_t nrecv = 0, nvals = 0;
  uint32_t* recv = nullptr;   // nrecv receivers, ascending
  void* vals = nullptr;       // nvals values (count: i64)
  uint64_t* offs = nullptr;   // collect: nrecv + 1 offsets into vals
  size_t recv_bytes = 0, vals_bytes = 0, offs_bytes = 0;
  // Byte messages: the result's strings (collect: one a collected message;
  // first and concat: one a receiver), their bytes and nvals + 1 offsets.
  bool is_bytes = false;
  uint8_t* str = nullptr;
  uint64_t* soffs = nullptr;
  uint64_t nstr_bytes = 0;
  size_t str_bytes = 0, soffs_bytes = 0;
  ~wg_gathered() {
    if (ctx) {
      dfree(ctx, recv, recv_bytes);
      dfree(ctx, vals, vals_bytes);
      dfree(ctx, offs, offs_bytes);
      dfree(ctx, str, str_bytes);
      dfree(ctx, soffs, soffs_bytes);
    }
  }
};

namespace wg {

template <class E>
static int run_gather(wg_context* ctx, const Receiving& R, const DevCsr& S, uint64_t n, int combiner, const E* mval,
                      wg_gathered* out) {
  hipStream_t st = ctx->stream;
  const unsigned grid = grid_of(n);
  const uint64_t m = S.nnz;
  int s;
  // Message counts: per sender entry, prefix sums over the entries, per
  // receiver, then each receiver's start in the collected order.
  Buf<uint64_t> ecnt, eprefix, cnt, start, flags, rank;
  if ((s = ecnt.alloc(ctx, m + 1)) || (s = eprefix.alloc(ctx, m + 1)) || (s = cnt.alloc(ctx, n + 1)) ||
      (s = start.alloc(ctx, n + 1)) || (s = flags.alloc(ctx, n)) || (s = rank.alloc(ctx, n)))
    return s;
  entry_counts<<<grid_of(m + 1), kThreads, 0, st>>>(S.targets, m, R.moff, ecnt.p);
  WG_HIP(ctx, hipGetLastError());
  if ((s = exclusive_scan_u64(ctx, ecnt.p, eprefix.p, m + 1))) return s;
  receiver_counts<<<grid_of(n + 1), kThreads, 0, st>>>(R, n, eprefix.p, cnt.p);
  WG_HIP(ctx, hipGetLastError());
  if ((s = exclusive_scan_u64(ctx, cnt.p, start.p, n + 1))) return s;
  // Ranks of the receivers that got a message.
  received_flags<<<grid, kThreads, 0, st>>>(cnt.p, n, flags.p);
  WG_HIP(ctx, hipGetLastError());
  if ((s = exclusive_scan_u64(ctx, flags.p, rank.p, n))) return s;
  uint64_t last_rank = 0, last_flag = 0, total = 0;
  if (n) {
    WG_HIP(ctx, hipMemcpyAsync(&last_rank, rank.p + n - 1, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(&last_flag, flags.p + n - 1, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(&total, start.p + n, 8, hipMemcpyDeviceToHost, st));
  }
  WG_HIP(ctx, hipStreamSynchronize(st));
  const uint64_t nrecv = last_rank + last_flag;
  out->nrecv = nrecv;
  out->recv_bytes = std::max<uint64_t>(nrecv, 1) * 4;
  if ((s = dalloc(ctx, (void**)&out->recv, out->recv_bytes))) return s;
  if (combiner == C_COUNT) {
    out->nvals = nrecv;
    out->vals_bytes = std::max<uint64_t>(nrecv, 1) * 8;
    if ((s = dalloc(ctx, &out->vals, out->vals_bytes))) return s;
    compact_receivers<E><<<grid, kThreads, 0, st>>>(cnt.p, rank.p, n, nullptr, nullptr, out->recv, nullptr,
                                                    (int64_t*)out->vals, nullptr);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
    return WG_OK;
  }
  // Every receiver's messages in collect order.
  Buf<E> coll;
  Buf<uint32_t> ord, row_of;
  if ((s = coll.alloc(ctx, total)) || (s = row_of.alloc(ctx, m))) return s;
  if (combiner == C_MODE && (s = ord.alloc(ctx, total))) return s;
  if ((s = expand_rows(ctx, S, row_of.p))) return s;
  uint32_t* ordp = combiner == C_MODE ? ord.p : nullptr;
  // flags (n entries) is free once the ranks are scanned: it holds base.
  fill_self<E><<<grid, kThreads, 0, st>>>(R, n, mval, start.p, eprefix.p, rank.p, coll.p, ordp, flags.p);
  fill_entries<E><<<grid_of(m), kThreads, 0, st>>>(R, m, row_of.p, eprefix.p, mval, flags.p, rank.p, coll.p, ordp);
  WG_HIP(ctx, hipGetLastError());
  row_of.reset();
  if (combiner != C_COLLECT && combiner != C_MODE) {
    Buf<E> val;
    if ((s = val.alloc(ctx, n))) return s;
    if ((s = clear_error(ctx))) return s;
    Buf<uint32_t> long_list;
    Buf<unsigned> nlong;
    if ((s = long_list.alloc(ctx, n)) || (s = nlong.alloc(ctx, 1))) return s;
    WG_HIP(ctx, hipMemsetAsync(nlong.p, 0, 4, st));
    unsigned hlong = 0;
    auto fold = [&](auto op_tag) -> int {
      constexpr int OP = decltype(op_tag)::value;
      fold_segments<E, OP><<<grid, kThreads, 0, st>>>(n, start.p, coll.p, val.p, long_list.p, nlong.p, ctx->d_err);
      WG_HIP(ctx, hipGetLastError());
      WG_HIP(ctx, hipMemcpyAsync(&hlong, nlong.p, 4, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      if (hlong)
        fold_long<E, OP><<<hlong, kLongThreads, 0, st>>>(long_list.p, start.p, coll.p, val.p, ctx->d_err);
      return WG_OK;
    };
    switch (combiner) {
      case C_SUM: s = fold(std::integral_constant<int, PLUS>()); break;
      case C_MIN: s = fold(std::integral_constant<int, MIN>()); break;
      case C_MAX: s = fold(std::integral_constant<int, MAX>()); break;
      case C_BOOL_OR: s = fold(std::integral_constant<int, LOR>()); break;
      case C_BOOL_AND: s = fold(std::integral_constant<int, LAND>()); break;
      default: s = fold(std::integral_constant<int, FIRST>()); break;
    }
    if (s) return s;
    WG_HIP(ctx, hipGetLastError());
    if ((s = collect_error(ctx, "gather"))) return s;
    out->nvals = nrecv;
    out->vals_bytes = std::max<uint64_t>(nrecv, 1) * sizeof(E);
    if ((s = dalloc(ctx, &out->vals, out->vals_bytes))) return s;
    compact_receivers<E><<<grid, kThreads, 0, st>>>(cnt.p, rank.p, n, val.p, nullptr, out->recv, (E*)out->vals,
                                                    nullptr, nullptr);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
    return WG_OK;
  }
  out->offs_bytes = (nrecv + 1) * 8;
  if ((s = dalloc(ctx, (void**)&out->offs, out->offs_bytes))) return s;
  compact_receivers<E><<<grid, kThreads, 0, st>>>(cnt.p, rank.p, n, nullptr, start.p, out->recv, nullptr, nullptr,
                                                  out->offs);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipMemcpyAsync(out->offs + nrecv, &total, 8, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  if (combiner == C_COLLECT) {
    out->nvals = total;
    out->vals_bytes = coll.bytes;
    out->vals = coll.release();
    return WG_OK;
  }
  // Mode: (receiver, SQL value) order, collect order kept within.
  Buf<uint64_t> keys, keys2, pos, pos2;
  Buf<uint32_t> ord2, ord3;
  if ((s = keys.alloc(ctx, total)) || (s = keys2.alloc(ctx, total)) || (s = pos.alloc(ctx, total)) ||
      (s = pos2.alloc(ctx, total)) || (s = ord2.alloc(ctx, total)) || (s = ord3.alloc(ctx, total)))
    return s;
  sql_keys<E><<<grid_of(total), kThreads, 0, st>>>(coll.p, total, keys.p, pos.p);
  WG_HIP(ctx, hipGetLastError());
  if ((s = sort_u64_keyed(ctx, keys.p, keys2.p, pos.p, pos2.p, total, 64))) return s;
  ord_of<<<grid_of(total), kThreads, 0, st>>>(ord.p, pos2.p, total, ord2.p);
  WG_HIP(ctx, hipGetLastError());
  unsigned bits = 1;
  while (bits < 32 && (1ull << bits) < nrecv) ++bits;
  if ((s = sort_u32_keyed(ctx, ord2.p, ord3.p, pos2.p, pos.p, total, bits))) return s;
  out->nvals = nrecv;
  out->vals_bytes = std::max<uint64_t>(nrecv, 1) * sizeof(E);
  if ((s = dalloc(ctx, &out->vals, out->vals_bytes))) return s;
  mode_pick<E><<<grid_of(nrecv), kThreads, 0, st>>>(out->offs, nrecv, pos.p, coll.p, (E*)out->vals);
  WG_HIP(ctx, hipGetLastError());
  // The lists were only the means: a mode result is one value a receiver.
  dfree(ctx, out->offs, out->offs_bytes);
  out->offs = nullptr;
  out->offs_bytes = 0;
  WG_HIP(ctx, hipStreamSynchronize(st));
  return WG_OK;
}

}  // namespace wg

// The gather of `count` element messages of dtype under `combiner` (checked
// by the caller) into a new result.
static int gather_impl(wg_context* ctx, const wg_graph* gc, const wg_gather_opts* o, int combiner, int dtype,
                       uint64_t count, const uint32_t* senders, const void* messages, wg_gathered** out);

extern "C" int wg_gather(wg_context* ctx, const wg_graph* gc, const wg_gather_opts* o, int dtype,
                         uint64_t count, const uint32_t* senders, const void* messages, wg_gathered** out) {
  if (!ctx || !gc || !o || !out) return WG_INVALID_ARGUMENT;
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  *out = nullptr;
  if (count && (!senders || !messages)) return fail(ctx, WG_INVALID_ARGUMENT, "null messages");
  if (o->direction != WG_OUT && o->direction != WG_IN && o->direction != WG_BOTH)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  if (o->nrelations && !o->relations) return fail(ctx, WG_INVALID_ARGUMENT, "null relation list");
  const int acc = combiner_accepts(o->combiner, dtype);
  if (acc < 0) return fail(ctx, WG_INVALID_ARGUMENT, "unknown combiner");
  if (dtype != WG_F32 && dtype != WG_F64 && dtype != WG_I64 && dtype != WG_BOOL)
    return fail(ctx, WG_UNSUPPORTED, "messages of this element type are not gathered on the device");
  // Combiner::check (combiner.rs:82-94).
  if (!acc) return fail(ctx, WG_COMBINER_TYPE, "the combiner does not accept messages of this type");
  return gather_impl(ctx, gc, o, o->combiner, dtype, count, senders, messages, out);
}

extern "C" int wg_gather_bytes(wg_context* ctx, const wg_graph* gc, const wg_gather_opts* o, uint64_t count,
                               const uint32_t* senders, const uint8_t* bytes, const uint64_t* offsets,
                               wg_gathered** out) {
  if (!ctx || !gc || !o || !out) return WG_INVALID_ARGUMENT;
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  *out = nullptr;
  if (!offsets || (count && !senders)) return fail(ctx, WG_INVALID_ARGUMENT, "null messages");
  if (o->direction != WG_OUT && o->direction != WG_IN && o->direction != WG_BOTH)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown direction");
  if (o->nrelations && !o->relations) return fail(ctx, WG_INVALID_ARGUMENT, "null relation list");
  const int acc = combiner_accepts_bytes(o->combiner);
  if (acc < 0) return fail(ctx, WG_INVALID_ARGUMENT, "unknown combiner");
  if (!acc) return fail(ctx, WG_COMBINER_TYPE, "the combiner does not accept byte messages");
  if (offsets[0] != 0) return fail(ctx, WG_INVALID_ARGUMENT, "byte offsets must start at 0");
  for (uint64_t i = 0; i < count; ++i)
    if (offsets[i + 1] < offsets[i]) return fail(ctx, WG_INVALID_ARGUMENT, "byte offsets must not decrease");
  const uint64_t total_in = offsets[count];
  if (total_in && !bytes) return fail(ctx, WG_INVALID_ARGUMENT, "null bytes");
  // The messages travel as their indexes; concat gathers them as collect.
  std::vector<int64_t> ids(count);
  for (uint64_t i = 0; i < count; ++i) ids[i] = (int64_t)i;
  const int inner = o->combiner == C_CONCAT ? C_COLLECT : o->combiner;
  wg_gathered* r = nullptr;
  int s;
  if ((s = gather_impl(ctx, gc, o, inner, WG_I64, count, senders, ids.data(), &r))) return s;
  r->combiner = o->combiner;
  r->is_bytes = true;
  if (o->combiner == C_COUNT) {
    *out = r;
    return ok(ctx);
  }
  hipStream_t st = ctx->stream;
  auto done = [&](int status) {
    if (status) wg_gathered_destroy(r);
    return status;
  };
  const uint64_t k = r->nvals;  // collected (or first) message indexes
  Buf<uint8_t> din;
  Buf<uint64_t> dboff, len, pos;
  if ((s = din.alloc(ctx, total_in)) || (s = dboff.alloc(ctx, count + 1)) || (s = len.alloc(ctx, k + 1)) ||
      (s = pos.alloc(ctx, k + 1)))
    return done(s);
  if (total_in) WG_HIP(ctx, hipMemcpyAsync(din.p, bytes, total_in, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemcpyAsync(dboff.p, offsets, (count + 1) * 8, hipMemcpyHostToDevice, st));
  string_lengths<<<grid_of(k + 1), kThreads, 0, st>>>((const int64_t*)r->vals, k, dboff.p, len.p);
  WG_HIP(ctx, hipGetLastError());
  if ((s = exclusive_scan_u64(ctx, len.p, pos.p, k + 1))) return done(s);
  uint64_t total = 0;
  WG_HIP(ctx, hipMemcpyAsync(&total, pos.p + k, 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  r->nstr_bytes = total;
  r->str_bytes = std::max<uint64_t>(total, 1);
  if ((s = dalloc(ctx, (void**)&r->str, r->str_bytes))) return done(s);
  if (total)
    string_bytes<<<grid_of(total), kThreads, 0, st>>>((const int64_t*)r->vals, k, pos.p, dboff.p, din.p, total,
                                                       r->str);
  WG_HIP(ctx, hipGetLastError());
  if (o->combiner == C_CONCAT) {
    // One string a receiver: its messages' strings back to back.
    r->soffs_bytes = (r->nrecv + 1) * 8;
    if ((s = dalloc(ctx, (void**)&r->soffs, r->soffs_bytes))) return done(s);
    receiver_spans<<<grid_of(r->nrecv + 1), kThreads, 0, st>>>(r->offs, r->nrecv, pos.p, r->soffs);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
    dfree(ctx, r->offs, r->offs_bytes);
    r->offs = nullptr;
    r->offs_bytes = 0;
    r->nvals = r->nrecv;
  } else {
    r->soffs_bytes = pos.bytes;
    r->soffs = pos.release();
  }
  WG_HIP(ctx, hipStreamSynchronize(st));
  dfree(ctx, r->vals, r->vals_bytes);
  r->vals = nullptr;
  r->vals_bytes = 0;
  *out = r;
  return ok(ctx);
}

static int gather_impl(wg_context* ctx, const wg_graph* gc, const wg_gather_opts* o, int combiner, int dtype,
                       uint64_t count, const uint32_t* senders, const void* messages, wg_gathered** out) {
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n;
  if (o->mask) {
    if (o->mask->id_space != n) return fail(ctx, WG_ID_SPACE_MISMATCH, "mask id space mismatch");
    if (o->mask->width != 1) return fail(ctx, WG_WIDTH_MISMATCH, "mask width must be 1");
  }
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  int s;
  const size_t es = dtype_size(dtype);
  // The outbox: messages sorted by sender, push order kept (outbox.rs:41-75).
  Buf<uint32_t> dsend, sorted, idx, perm;
  Buf<uint8_t> dval, mval;
  Buf<uint64_t> moff;
  if ((s = dsend.alloc(ctx, count)) || (s = dval.alloc(ctx, count * es)) || (s = moff.alloc(ctx, n + 1)))
    return s;
  if (count) {
    WG_HIP(ctx, hipMemcpyAsync(dsend.p, senders, count * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(dval.p, messages, count * es, hipMemcpyHostToDevice, st));
    Buf<unsigned long long> first;
    if ((s = first.alloc(ctx, 1))) return s;
    WG_HIP(ctx, hipMemsetAsync(first.p, 0xff, 8, st));
    first_sender_outside<<<grid_of(count), kThreads, 0, st>>>(dsend.p, count, n, first.p);
    WG_HIP(ctx, hipGetLastError());
    unsigned long long bad = 0;
    WG_HIP(ctx, hipMemcpyAsync(&bad, first.p, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    if (bad != ~0ull)
      return fail(ctx, WG_SENDER_OUTSIDE_ID_SPACE,
                  "sender " + std::to_string(senders[bad]) + " outside the id space " + std::to_string(n));
  }
  if ((s = sorted.alloc(ctx, count)) || (s = idx.alloc(ctx, count)) || (s = perm.alloc(ctx, count)) ||
      (s = mval.alloc(ctx, count * es)))
    return s;
  if (count) {
    iota32<<<grid_of(count), kThreads, 0, st>>>(idx.p, count);
    unsigned bits = 1;
    while (bits < 32 && (1ull << bits) < n) ++bits;
    if ((s = sort_u32_pairs(ctx, dsend.p, sorted.p, idx.p, perm.p, count, bits))) return s;
    switch (dtype) {
      case WG_F32: permute<float><<<grid_of(count), kThreads, 0, st>>>(perm.p, (const float*)dval.p, count, (float*)mval.p); break;
      case WG_F64: permute<double><<<grid_of(count), kThreads, 0, st>>>(perm.p, (const double*)dval.p, count, (double*)mval.p); break;
      case WG_I64: permute<int64_t><<<grid_of(count), kThreads, 0, st>>>(perm.p, (const int64_t*)dval.p, count, (int64_t*)mval.p); break;
      default: permute<uint8_t><<<grid_of(count), kThreads, 0, st>>>(perm.p, (const uint8_t*)dval.p, count, (uint8_t*)mval.p); break;
    }
    WG_HIP(ctx, hipGetLastError());
  }
  sender_starts<<<grid_of(n + 1), kThreads, 0, st>>>(sorted.p, count, n, moff.p);
  WG_HIP(ctx, hipGetLastError());
  // The senders of each receiver.
  const DevCsr* S = nullptr;
  if ((s = senders_csr(ctx, g, o->direction, o->relations, o->nrelations, o->per_relation, &S))) return s;
  Receiving R{S->offsets, S->targets, moff.p, o->mask ? o->mask->data : nullptr, o->mask ? o->mask->dtype : 0,
              o->mask_complement, o->include_self};
  wg_gathered* res = new (std::nothrow) wg_gathered();
  if (!res) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  res->ctx = ctx;
  res->combiner = combiner;
  res->dtype = dtype;
  switch (dtype) {
    case WG_F32: s = run_gather<float>(ctx, R, *S, n, combiner, (const float*)mval.p, res); break;
    case WG_F64: s = run_gather<double>(ctx, R, *S, n, combiner, (const double*)mval.p, res); break;
    case WG_I64: s = run_gather<int64_t>(ctx, R, *S, n, combiner, (const int64_t*)mval.p, res); break;
    default: s = run_gather<uint8_t>(ctx, R, *S, n, combiner, (const uint8_t*)mval.p, res); break;
  }
  if (s) {
    delete res;
    return s;
  }
  ctx->gathereds.insert(res);
  *out = res;
  return ok(ctx);
}

extern "C" int wg_gathered_info(const wg_gathered* r, uint64_t* receivers, uint64_t* values) {
  if (!r) return WG_INVALID_ARGUMENT;
  if (receivers) *receivers = r->nrecv;
  if (values) *values = r->nvals;
  return WG_OK;
}

extern "C" int wg_gathered_read(wg_context* ctx, const wg_gathered* r, uint64_t* receivers, void* values,
                                uint64_t* offsets) {
  if (!ctx || !r) return WG_INVALID_ARGUMENT;
  if (r->ctx != ctx) return fail(ctx, WG_INVALID_ARGUMENT, "the result belongs to another or a destroyed context");
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (receivers && r->nrecv) {
    std::vector<uint32_t> tmp(r->nrecv);
    WG_HIP(ctx, hipMemcpy(tmp.data(), r->recv, r->nrecv * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < r->nrecv; ++i) receivers[i] = tmp[i];
  }
  const size_t es = r->combiner == 1 ? 8 : dtype_size(r->dtype);
  // A byte result's strings are read with wg_gathered_read_bytes.
  if (values && r->nvals && r->vals) WG_HIP(ctx, hipMemcpy(values, r->vals, r->nvals * es, hipMemcpyDeviceToHost));
  if (offsets && r->offs) WG_HIP(ctx, hipMemcpy(offsets, r->offs, (r->nrecv + 1) * 8, hipMemcpyDeviceToHost));
  return ok(ctx);
}

extern "C" int wg_gathered_bytes_info(const wg_gathered* r, uint64_t* bytes) {
  if (!r) return WG_INVALID_ARGUMENT;
  if (bytes) *bytes = r->nstr_bytes;
  return WG_OK;
}

extern "C" int wg_gathered_read_bytes(wg_context* ctx, const wg_gathered* r, uint8_t* bytes,
                                      uint64_t* string_offsets) {
  if (!ctx || !r) return WG_INVALID_ARGUMENT;
  if (r->ctx != ctx) return fail(ctx, WG_INVALID_ARGUMENT, "the result belongs to another or a destroyed context");
  if (!r->is_bytes || !r->soffs) return fail(ctx, WG_INVALID_ARGUMENT, "the result holds no strings");
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (bytes && r->nstr_bytes) WG_HIP(ctx, hipMemcpy(bytes, r->str, r->nstr_bytes, hipMemcpyDeviceToHost));
  if (string_offsets) WG_HIP(ctx, hipMemcpy(string_offsets, r->soffs, (r->nvals + 1) * 8, hipMemcpyDeviceToHost));
  return ok(ctx);
}

namespace wg {
void release_gathered(wg_gathered* r) {
  if (!r->ctx) return;
  (void)hipStreamSynchronize(r->ctx->stream);
  dfree(r->ctx, r->recv, r->recv_bytes);
  dfree(r->ctx, r->vals, r->vals_bytes);
  dfree(r->ctx, r->offs, r->offs_bytes);
  dfree(r->ctx, r->str, r->str_bytes);
  dfree(r->ctx, r->soffs, r->soffs_bytes);
  r->recv = nullptr;
  r->vals = nullptr;
  r->offs = nullptr;
  r->str = nullptr;
  r->soffs = nullptr;
  r->ctx = nullptr;
}
}  // namespace wg

extern "C" void wg_gathered_destroy(wg_gathered* r) {
  if (!r) return;
  if (r->ctx) r->ctx->gathereds.erase(r);
  release_gathered(r);
  delete r;
}
