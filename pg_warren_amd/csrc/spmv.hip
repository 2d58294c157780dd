// Semiring SpMV and SpMM on gfx950 (compute/product/{mod,kernel}.rs).
//
// SpMV: one 256-thread workgroup per row block. A block stages the products
// of its (at most 2048) entries in LDS after coalesced loads of the target
// and weight streams and a gather of x, then folds each row: short blocks
// give each row a power-of-two group of lanes, each lane folding a
// contiguous run of the row and the group combining the runs in order with
// DPP/bpermute shuffles; rows longer than a block are split into pieces that
// each leave a partial, and a second kernel (spmv_long) folds every split
// row's partials in piece order.
// Folding contiguous runs in order keeps the reference's left-to-right
// result exactly for every monoid but float plus.
//
// SpMM: a group of lanes per row, lanes over the columns of the row-major X,
// each lane folding its columns over the row's entries in ascending order
// (the reference's fold order); long rows split as above.
#include <algorithm>
#include <mutex>

#include "semiring.hpp"
#include "spmv_tiles.hpp"
#include "wide.hpp"
#include "wg_internal.hpp"

namespace wg {

template <class E>
struct DtypeOf;
template <> struct DtypeOf<float> { static constexpr int v = WG_F32; };
template <> struct DtypeOf<double> { static constexpr int v = WG_F64; };
template <> struct DtypeOf<int64_t> { static constexpr int v = WG_I64; };
template <> struct DtypeOf<uint8_t> { static constexpr int v = WG_BOOL; };
template <> struct DtypeOf<uint32_t> { static constexpr int v = WG_U32; };

template <class T>
__device__ __forceinline__ T load_coherent(const T* p) {
  constexpr int W = (sizeof(T) + 3) / 4;
  union U {
    T t;
    int w[W];
    __device__ U() {}
  } u;
  const int* q = reinterpret_cast<const int*>(p);
#pragma unroll
  for (int i = 0; i < W; ++i) u.w[i] = __hip_atomic_load(const_cast<int*>(q + i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return u.t;
}

// x is read unless (x) ignores the value.
template <int MUL>
constexpr bool reads_x() { return MUL != FIRST && MUL != PAIR; }

// Graphs smaller than this use the default launch without tuning.
constexpr uint64_t kTuneMinNnz = 1ull << 22;

// Launch-shape choices the autotuner makes per graph orientation and
// semiring (bit flags).
enum : int {
  kVarDirect = 1,  // spmv_direct instead of the LDS-staged spmv_blocks
  kVarNT = 2,      // non-temporal loads of the target / weight streams
  kVarXcd = 4,     // XCD-contiguous block order
  kVarHot = 8,     // gather from x permuted into the hot-first column order
  kVarLds = 16,    // persistent workgroups with the hottest columns' x in LDS
  kVarTile = 32,   // the tiled kernel (spmv_tiles.hpp), alone
};

// Targets tagged for the LDS cache: rank | bit 31 for the first nhot ranks.
__global__ __launch_bounds__(256) void tag_targets(const uint32_t* __restrict__ tgt,
                                                   const uint32_t* __restrict__ rank, uint32_t nhot,
                                                   uint64_t nnz, uint32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * 256) {
    const uint32_t t = tgt[i], r = rank[t];
    out[i] = r < nhot ? (0x80000000u | r) : t;
  }
}

// x in the hot-first column order: xh[r] = x[inv[r]].
template <class E>
__global__ __launch_bounds__(256) void gather_hot(const E* __restrict__ x, const uint32_t* __restrict__ inv,
                                                  uint64_t nh, E* __restrict__ xh) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nh; i += (uint64_t)gridDim.x * 256)
    xh[i] = x[inv[i]];
}

// The tiled kernel's two x copies in one launch: the hot run in rank order
// (xh[r] = x[inv[r]]) and the warm run in ascending column order
// (xw[j] = x[warm[j]]). A thread copies four elements a grid apart, their
// gathers in flight together. Thread 0 also clears the error word jclear
// when one is given.
template <class E>
__global__ __launch_bounds__(256) void gather_tiers(const E* __restrict__ x, const uint32_t* __restrict__ inv,
                                                    uint32_t nh, E* __restrict__ xh,
                                                    const uint32_t* __restrict__ warm, uint32_t nw,
                                                    E* __restrict__ xw, unsigned long long* __restrict__ jclear) {
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x, T = (uint64_t)gridDim.x * 256;
  const uint64_t total = (uint64_t)nh + nw;
  if (jclear && q == 0) *jclear = ~0ull;
  uint32_t id[4];
  E v[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint64_t i = q + k * T;
    id[k] = i < nh ? inv[i] : i < total ? warm[i - nh] : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) v[k] = x[id[k]];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint64_t i = q + k * T;
    if (i < nh)
      xh[i] = v[k];
    else if (i < total)
      xw[i - nh] = v[k];
  }
}

inline unsigned gather_tiers_grid(uint32_t nh, uint32_t nw) {
  return (unsigned)(((uint64_t)nh + nw + 4 * 256 - 1) / (4 * 256));
}

// A streamed load, non-temporal when nt (the row streams are read once and
// should not evict the gathered x from L2 and the infinity cache).
template <class T>
__device__ __forceinline__ T ld_stream(const T* p, bool nt) {
  return nt ? __builtin_nontemporal_load(p) : *p;
}

// Workgroups are dispatched to the eight XCDs in turn; with xcd set, XCD k
// takes the k-th contiguous run of blocks (a bijection for any count), so
// neighbouring rows share one L2.
__device__ __forceinline__ uint32_t block_index(bool xcd) {
  if (!xcd) return blockIdx.x;
  const uint32_t nb = gridDim.x, q = nb / 8, r = nb % 8;
  const uint32_t k = blockIdx.x % 8, j = blockIdx.x / 8;
  return k * q + (k < r ? k : r) + j;
}

// The LDS a row block stages through.
template <class E, int ADD>
struct BlockLds {
  E prod[kBlockNnz];
  uint32_t rs[kBlockRows + 1];
  typename Fold<E, ADD>::A wpart[kBlockThreads / 64];
  int wpres[kBlockThreads / 64];
};

// One row block: its products staged in LDS (x read through `gather`),
// each row folded in order, split rows' pieces leaving partials.
// A stored weight as the element: the cached cast (WT = E), or the f64
// weight cast per entry (compute/product/kernel.rs:40-47).
template <class E, class WT>
__device__ __forceinline__ E weight_as(WT v, uint32_t& err) {
  if constexpr (std::is_same<WT, E>::value) return v;
  else return weight_cast<E>(v, err);
}

template <class E, int ADD, int MUL, bool HW, class Gather, class WT = E>
__device__ __forceinline__ void block_body(const Block& b, const uint64_t* __restrict__ off,
                                           const uint32_t* __restrict__ tgt, const WT* __restrict__ w,
                                           const Gather& gather, E* __restrict__ y, const WriteArgs& wa,
                                           typename Fold<E, ADD>::A* __restrict__ partials,
                                           unsigned long long* __restrict__ err, bool nt,
                                           BlockLds<E, ADD>& L) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  E* prod = L.prod;
  uint32_t* rs = L.rs;
  A* wpart = L.wpart;
  int* wpres = L.wpres;
  const uint32_t tid = threadIdx.x;
  uint32_t e_err = 0;
  uint64_t e0, e1;
  if (b.nrows) {
    e0 = off[b.row0];
    e1 = off[b.row0 + b.nrows];
  } else {
    const uint64_t rb = off[b.row0], re = off[b.row0 + 1];
    e0 = rb + (uint64_t)b.piece * kBlockNnz;
    e1 = e0 + kBlockNnz < re ? e0 + kBlockNnz : re;
  }
  const uint32_t cnt = (uint32_t)(e1 - e0);

  // Stage the products: every load of the block issued before any use. The
  // gathers are unconditional (a lane past the block's entries reads column
  // 0, which exists, and drops it), so the compiler issues all of them
  // before the first wait instead of one per guarded branch.
  {
    constexpr int kPer = (int)(kBlockNnz / kBlockThreads);
    uint32_t col[kPer];
    E wv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      col[k] = 0;
      wv[k] = E(1);
      if (i < cnt) {
        if constexpr (reads_x<MUL>()) col[k] = ld_stream(tgt + e0 + i, nt);
        if constexpr (HW) wv[k] = weight_as<E, WT>(ld_stream(w + e0 + i, nt), e_err);
      }
    }
    E xv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      xv[k] = E();
      if constexpr (reads_x<MUL>()) xv[k] = gather(col[k]);
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      if (i < cnt) prod[i] = multiply<E, ADD, MUL>(wv[k], xv[k], e_err);
    }
  }
  if (b.nrows)
    for (uint32_t k = tid; k <= b.nrows; k += kBlockThreads) rs[k] = (uint32_t)(off[b.row0 + k] - e0);
  __syncthreads();

  if (b.nrows) {
    const uint32_t nr = b.nrows;
    uint32_t T = 1;
    if (nr < kBlockThreads) {
      uint32_t q = kBlockThreads / nr;
      T = 1u << (31 - __builtin_clz(q));
      if (T > 64) T = 64;
    }
    const uint32_t groups = kBlockThreads / T, g = tid / T, l = tid % T;
    for (uint32_t r = g; r < nr; r += groups) {
      const uint32_t a = rs[r], bnd = rs[r + 1];
      const uint32_t len = bnd - a, chunk = (len + T - 1) / T;
      uint32_t s = a + l * chunk, t = s + chunk < bnd ? s + chunk : bnd;
      A acc{};
      int pres = 0;
      for (uint32_t i = s; i < t; ++i) {
        A v = F::lift(prod[i]);
        acc = pres ? F::combine(acc, v, e_err) : v;
        pres = 1;
      }
      for (uint32_t d = 1; d < T; d <<= 1) {
        A o = shfl_down(acc, d, T);
        int op = __shfl_down(pres, d, T);
        if ((l & (2 * d - 1)) == 0 && op) {
          acc = pres ? F::combine(acc, o, e_err) : o;
          pres = 1;
        }
      }
      if (l == 0) {
        const uint64_t row = (uint64_t)b.row0 + r;
        E res = pres ? F::finish(acc, e_err) : E();
        write_elem<E, ADD>(wa, row, &y[row], pres != 0, res, e_err);
      }
    }
  } else {
    // One piece of a long row: each thread folds a contiguous run, the waves
    // and then the pieces combine in order.
    const uint32_t chunk = (cnt + kBlockThreads - 1) / kBlockThreads;
    uint32_t s = tid * chunk, t = s + chunk < cnt ? s + chunk : cnt;
    A acc{};
    int pres = 0;
    for (uint32_t i = s; i < t; ++i) {
      A v = F::lift(prod[i]);
      acc = pres ? F::combine(acc, v, e_err) : v;
      pres = 1;
    }
    const uint32_t l = tid & 63;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      A o = shfl_down(acc, d, 64);
      int op = __shfl_down(pres, d, 64);
      if ((l & (2 * d - 1)) == 0 && op) {
        acc = pres ? F::combine(acc, o, e_err) : o;
        pres = 1;
      }
    }
    if (l == 0) {
      wpart[tid >> 6] = acc;
      wpres[tid >> 6] = pres;
    }
    __syncthreads();
    if (tid == 0) {
      A tot{};
      int tp = 0;
      for (int k = 0; k < (int)(kBlockThreads / 64); ++k)
        if (wpres[k]) {
          tot = tp ? F::combine(tot, wpart[k], e_err) : wpart[k];
          tp = 1;
        }
      // A split row's pieces each leave a partial; spmv_long folds them in
      // piece order once every piece has run (the kernel boundary orders
      // them: no cross-workgroup fence, which on gfx950 would write back
      // the XCD's L2).
      partials[b.pbase + b.piece] = tot;
    }
  }
  if (e_err) raise_error(err, b.row0, e_err);
}

template <class E, int ADD, int MUL, bool HW, class WT = E>
__global__ __launch_bounds__(kBlockThreads) void spmv_blocks(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tgt, const WT* __restrict__ w, const E* __restrict__ x,
    E* __restrict__ y, WriteArgs wa, typename Fold<E, ADD>::A* __restrict__ partials,
    unsigned long long* __restrict__ err, int variant) {
  __shared__ BlockLds<E, ADD> L;
  const Block b = blocks[block_index(variant & kVarXcd)];
  block_body<E, ADD, MUL, HW>(b, off, tgt, w, [&](uint32_t c) { return x[c]; }, y, wa, partials, err,
                              (variant & kVarNT) != 0, L);
}

// Hot-column cache: the persistent variant. Each workgroup first copies
// the hottest columns' x (xhot, ranks [0, nhot)) into LDS, then takes row
// blocks in turn; a tagged target (bit 31 set) names a hot rank read from
// LDS, any other target a column of x read from memory. The gathers the
// L2 serves (its request rate bounds this product) drop by the hot
// columns' share of the entries.
constexpr uint32_t kHotTag = 0x80000000u;
constexpr uint32_t kHotLdsBytes = 48 * 1024;

template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(kBlockThreads) void spmv_lds(
    const Block* __restrict__ blocks, uint32_t nblocks, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tgt, const E* __restrict__ w, const E* __restrict__ x,
    const E* __restrict__ xhot, uint32_t nhot, E* __restrict__ y, WriteArgs wa,
    typename Fold<E, ADD>::A* __restrict__ partials, unsigned long long* __restrict__ err, int variant) {
  __shared__ BlockLds<E, ADD> L;
  __shared__ E hot[kHotLdsBytes / sizeof(E)];
  for (uint32_t i = threadIdx.x; i < nhot; i += kBlockThreads) hot[i] = xhot[i];
  __syncthreads();
  const bool nt = (variant & kVarNT) != 0;
  auto gather = [&](uint32_t t) { return (t & kHotTag) ? hot[t & ~kHotTag] : x[t]; };
  for (uint32_t bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
    const Block b = blocks[bi];
    block_body<E, ADD, MUL, HW>(b, off, tgt, w, gather, y, wa, partials, err, nt, L);
    __syncthreads();
  }
}

// Folds each split row's piece partials in piece order and writes the row.
template <class E, int ADD>
__global__ __launch_bounds__(256) void spmv_long(const uint32_t* __restrict__ long_rows,
                                                 const uint32_t* __restrict__ long_pbase,
                                                 uint32_t nlong,
                                                 const typename Fold<E, ADD>::A* __restrict__ partials,
                                                 E* __restrict__ y, WriteArgs wa,
                                                 unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nlong) return;
  uint32_t e_err = 0;
  const uint32_t p0 = long_pbase[i], p1 = long_pbase[i + 1];
  typename F::A all = partials[p0];
  for (uint32_t p = p0 + 1; p < p1; ++p) all = F::combine(all, partials[p], e_err);
  const uint64_t row = long_rows[i];
  E res = F::finish(all, e_err);
  write_elem<E, ADD>(wa, row, &y[row], true, res, e_err);
  if (e_err) raise_error(err, row, e_err);
}

// The LDS-free variant for monoids whose result does not depend on the
// grouping (all but i64 plus and float min/max): each row's group of T lanes
// reads the row's entries directly, lane l taking entries l, l + T, ...
// (coalesced across the wave, since neighbouring groups hold neighbouring
// rows), four loads in flight per lane; the group then combines its lanes in
// lane order, so `any` still keeps the row's first entry. Long-row pieces
// stride the whole workgroup over the piece the same way.
template <class E, int ADD>
constexpr bool grouping_free() {
  return !(std::is_same<E, int64_t>::value && ADD == PLUS) && !(is_float_v<E> && (ADD == MIN || ADD == MAX));
}

template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(kBlockThreads) void spmv_direct(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tgt, const E* __restrict__ w, const E* __restrict__ x,
    E* __restrict__ y, WriteArgs wa, typename Fold<E, ADD>::A* __restrict__ partials,
    unsigned long long* __restrict__ err, int variant) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  __shared__ A wpart[kBlockThreads / 64];
  __shared__ int wpres[kBlockThreads / 64];
  const bool nt = variant & kVarNT;
  const Block b = blocks[block_index(variant & kVarXcd)];
  const uint32_t tid = threadIdx.x;
  uint32_t e_err = 0;
  // Folds entries i, i + step, ... below end into (acc, pres), four at a time.
  auto fold_strided = [&](uint64_t i, uint64_t end, uint32_t step, A& acc, int& pres) {
    for (; i < end; i += 4 * (uint64_t)step) {
      uint32_t c[4];
      E wk[4], xk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint64_t j = i + (uint64_t)k * step;
        c[k] = 0;
        wk[k] = E(1);
        if (j < end) {
          if constexpr (reads_x<MUL>()) c[k] = ld_stream(tgt + j, nt);
          if constexpr (HW) wk[k] = ld_stream(w + j, nt);
        }
      }
      // The four gathers unconditional (column 0 past the end), so they
      // are in flight together.
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        xk[k] = E();
        if constexpr (reads_x<MUL>()) xk[k] = x[c[k]];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint64_t j = i + (uint64_t)k * step;
        if (j < end) {
          const E wv = wk[k], xv = xk[k];
          A v = F::lift(multiply<E, ADD, MUL>(wv, xv, e_err));
          acc = pres ? F::combine(acc, v, e_err) : v;
          pres = 1;
        }
      }
    }
  };
  if (b.nrows) {
    const uint32_t nr = b.nrows;
    uint32_t T = 1;
    if (nr < kBlockThreads) {
      uint32_t q = kBlockThreads / nr;
      T = 1u << (31 - __builtin_clz(q));
      if (T > 64) T = 64;
    }
    const uint32_t groups = kBlockThreads / T, g = tid / T, l = tid % T;
    for (uint32_t r = g; r < nr; r += groups) {
      const uint64_t row = (uint64_t)b.row0 + r;
      A acc{};
      int pres = 0;
      fold_strided(off[row] + l, off[row + 1], T, acc, pres);
      for (uint32_t d = 1; d < T; d <<= 1) {
        A o = shfl_down(acc, d, T);
        int op = __shfl_down(pres, d, T);
        if ((l & (2 * d - 1)) == 0 && op) {
          acc = pres ? F::combine(acc, o, e_err) : o;
          pres = 1;
        }
      }
      if (l == 0) {
        E res = pres ? F::finish(acc, e_err) : E();
        write_elem<E, ADD>(wa, row, &y[row], pres != 0, res, e_err);
      }
    }
  } else {
    const uint64_t rb = off[b.row0], re = off[b.row0 + 1];
    const uint64_t e0 = rb + (uint64_t)b.piece * kBlockNnz;
    const uint64_t e1 = e0 + kBlockNnz < re ? e0 + kBlockNnz : re;
    A acc{};
    int pres = 0;
    fold_strided(e0 + tid, e1, kBlockThreads, acc, pres);
    const uint32_t l = tid & 63;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      A o = shfl_down(acc, d, 64);
      int op = __shfl_down(pres, d, 64);
      if ((l & (2 * d - 1)) == 0 && op) {
        acc = pres ? F::combine(acc, o, e_err) : o;
        pres = 1;
      }
    }
    if (l == 0) {
      wpart[tid >> 6] = acc;
      wpres[tid >> 6] = pres;
    }
    __syncthreads();
    if (tid == 0) {
      A tot{};
      int tp = 0;
      for (int k = 0; k < (int)(kBlockThreads / 64); ++k)
        if (wpres[k]) {
          tot = tp ? F::combine(tot, wpart[k], e_err) : wpart[k];
          tp = 1;
        }
      partials[b.pbase + b.piece] = tot;
    }
  }
  if (e_err) raise_error(err, b.row0, e_err);
}

// C consecutive elements of a row of x, read as one 16-byte access when
// C * sizeof(E) == 16 (the caller has checked the alignment).
template <class E, int C>
struct alignas(C * sizeof(E)) Span {
  E v[C];
};

// SpMM: LPR lanes per task (a short row or a piece of a split row), each
// lane folding the column groups lane, lane + LPR, ... (C adjacent columns
// a group, one load an entry) of its row over the task's entries; each
// column folds the entries in order.
template <class E, int ADD, int MUL, bool HW, int C>
__global__ __launch_bounds__(256) void spmm_rows(
    const uint32_t* __restrict__ tgt, const E* __restrict__ w, const E* __restrict__ x,
    E* __restrict__ y, uint64_t width, const SpmmTask* __restrict__ tasks, uint64_t ntasks, uint32_t lpr,
    typename Fold<E, ADD>::A* __restrict__ partials, WriteArgs wa, unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  const uint32_t tasks_per_block = 256 / lpr;
  const uint64_t task = (uint64_t)blockIdx.x * tasks_per_block + threadIdx.x / lpr;
  const uint32_t lane = threadIdx.x % lpr;
  if (task >= ntasks) return;
  const SpmmTask tk = tasks[task];
  const uint64_t row = tk.row;
  const uint64_t e0 = tk.ec & ((1ull << kSpmmCountShift) - 1);
  const uint64_t e1 = e0 + (tk.ec >> kSpmmCountShift);
  uint32_t e_err = 0;
  for (uint64_t c = (uint64_t)lane * C; c < width; c += (uint64_t)lpr * C) {
    A acc[C];
    int pres = 0;
    // Eight entries at a time, their loads in flight together (a slot past
    // the end reads row 0's columns and drops them); folded in order.
    constexpr int kU = 8;
    for (uint64_t e = e0; e < e1; e += kU) {
      uint32_t t4[kU];
      E wk[kU];
      Span<E, C> xv[kU];
#pragma unroll
      for (int k = 0; k < kU; ++k) {
        t4[k] = 0;
        wk[k] = E(1);
        if (e + k < e1) {
          if constexpr (reads_x<MUL>()) t4[k] = tgt[e + k];
          if constexpr (HW) wk[k] = w[e + k];
        }
      }
#pragma unroll
      for (int k = 0; k < kU; ++k) {
        if constexpr (reads_x<MUL>()) {
          xv[k] = *reinterpret_cast<const Span<E, C>*>(x + (uint64_t)t4[k] * width + c);
        } else {
#pragma unroll
          for (int q = 0; q < C; ++q) xv[k].v[q] = E();
        }
      }
#pragma unroll
      for (int k = 0; k < kU; ++k)
        if (e + k < e1) {
#pragma unroll
          for (int q = 0; q < C; ++q) {
            A v = F::lift(multiply<E, ADD, MUL>(wk[k], xv[k].v[q], e_err));
            acc[q] = pres ? F::combine(acc[q], v, e_err) : v;
          }
          pres = 1;
        }
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (tk.slot == ~0u) {
        E res = pres ? F::finish(acc[q], e_err) : E();
        write_elem<E, ADD>(wa, row, &y[row * width + c + q], pres != 0, res, e_err);
      } else {
        partials[(uint64_t)tk.slot * width + c + q] = acc[q];
      }
    }
  }
  if (e_err) raise_error(err, row, e_err);
}

// Folds each split row's piece partials in piece order, column by column.
template <class E, int ADD>
__global__ __launch_bounds__(256) void spmm_long(const uint32_t* __restrict__ long_rows,
                                                 const uint32_t* __restrict__ pbase, uint32_t nlong,
                                                 uint64_t width,
                                                 const typename Fold<E, ADD>::A* __restrict__ partials,
                                                 E* __restrict__ y, WriteArgs wa,
                                                 unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint64_t)nlong * width) return;
  const uint32_t t = (uint32_t)(i / width);
  const uint64_t c = i % width;
  uint32_t e_err = 0;
  const uint32_t p0 = pbase[t], p1 = pbase[t + 1];
  typename F::A all = partials[(uint64_t)p0 * width + c];
  for (uint32_t p = p0 + 1; p < p1; ++p) all = F::combine(all, partials[(uint64_t)p * width + c], e_err);
  const uint64_t row = long_rows[t];
  E res = F::finish(all, e_err);
  write_elem<E, ADD>(wa, row, &y[row * width + c], true, res, e_err);
  if (e_err) raise_error(err, row, e_err);
}

// ---- the tiled kernel's layout builders (spmv_tiles.hpp) ----------------
// tagged[slot of p] = p's target, tagged hot or warm, for every entry;
// padding 0. warm: the warm run's columns, ascending.
__global__ __launch_bounds__(256) void tile_tag(const uint32_t* __restrict__ tgt, const uint32_t* __restrict__ rank,
                                                uint32_t nhot, const uint32_t* __restrict__ warm, uint32_t nwarm,
                                                uint32_t ncols, uint64_t nnz, uint64_t padded,
                                                uint32_t* __restrict__ tagged) {
  for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < padded; p += (uint64_t)gridDim.x * 256) {
    uint32_t v = 0;
    if (p < nnz) {
      const uint32_t c = tgt[p], k = rank[c];
      if (k < nhot) {
        v = kTileHot | k;
      } else if (k - nhot < nwarm) {
        // The column's slot in the ascending warm run.
        uint32_t lo = 0, hi = nwarm;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) >> 1;
          if (warm[mid] < c)
            lo = mid + 1;
          else
            hi = mid;
        }
        v = ncols + lo;
      } else {
        v = c;
      }
    }
    const uint32_t li = (uint32_t)(p % kTileEntries);
    tagged[p - li + tile_slot(li >> 3, li & 7)] = v;
  }
}

// Marks each non-empty row's first entry.
__global__ __launch_bounds__(256) void tile_heads(const uint64_t* __restrict__ off, uint64_t n,
                                                  uint32_t* __restrict__ tagged) {
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (uint64_t)gridDim.x * 256) {
    const uint64_t p = off[r];
    if (off[r + 1] > p) {
      const uint32_t li = (uint32_t)(p % kTileEntries);
      tagged[p - li + tile_slot(li >> 3, li & 7)] |= kTileHead;
    }
  }
}

// Weights of one element type in the tile order; padding 0.
template <class E>
__global__ __launch_bounds__(256) void tile_weights(const E* __restrict__ w, uint64_t nnz, uint64_t padded,
                                                    E* __restrict__ out) {
  for (uint64_t p = (uint64_t)blockIdx.x * 256 + threadIdx.x; p < padded; p += (uint64_t)gridDim.x * 256) {
    const uint32_t li = (uint32_t)(p % kTileEntries);
    out[p - li + tile_slot(li >> 3, li & 7)] = p < nnz ? w[p] : E();
  }
}

__global__ void tile_clear(unsigned long long* p) { *p = ~0ull; }

// The tile layout of c for elements of esize bytes (and the part all sizes
// share), built once. WG_UNSUPPORTED when the layout does not apply (an id
// space of 2^30 or more, no entries, or no hot-column ranking).
int prepare_tiles(wg_context* ctx, DevCsr& c, size_t esize, uint32_t hotcap, uint32_t warmcap) {
  if (c.tile_state < 0) return WG_UNSUPPORTED;
  if (c.n >= (1ull << 30) || c.nnz == 0 || c.nnz >= (1ull << 40)) {
    c.tile_state = -1;
    return WG_UNSUPPORTED;
  }
  int s;
  if (!c.hot_state && (s = prepare_hot_columns(ctx, c))) return s;
  if (c.hot_state != 1) return WG_UNSUPPORTED;
  hipStream_t st = ctx->stream;
  if (c.tile_state == 0) {
    int cus = 0;
    WG_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const uint32_t W = (uint32_t)std::max(cus, 1) * kTileWaves;
    const uint64_t ntiles = (c.nnz + kTileEntries - 1) / kTileEntries;
    const uint64_t words = (c.n + 63) / 64;
    std::vector<uint64_t> bits(words + 1, 0);
    std::vector<uint32_t> pre(words + 1, 0);
    std::vector<uint32_t> wave(W + 1), heads(W);
    const std::vector<uint64_t>& off = c.h_offsets;
    for (uint64_t r = 0; r < c.n; ++r)
      if (off[r + 1] > off[r]) bits[r >> 6] |= 1ull << (r & 63);
    uint64_t acc = 0;
    for (uint64_t i = 0; i < words; ++i) {
      pre[i] = (uint32_t)acc;
      acc += __builtin_popcountll(bits[i]);
    }
    // Each wave's tiles, and the row starts before its first entry.
    uint64_t r = 0, starts = 0;
    for (uint32_t g = 0; g <= W; ++g) {
      wave[g] = (uint32_t)(ntiles * g / W);
      if (g == W) break;
      const uint64_t pos = (uint64_t)wave[g] * kTileEntries;
      for (; r < c.n && off[r] < pos; ++r)
        if (off[r + 1] > off[r]) ++starts;
      heads[g] = (uint32_t)starts;
    }
    // The rows a wave leaves open at its end (its last row start, when it
    // holds one): their values are joined from the wave records.
    std::vector<uint64_t> fix(words + 1, 0);
    {
      std::vector<uint64_t> last;  // ranks among non-empty rows, ascending
      for (uint32_t g = 0; g < W; ++g) {
        const uint64_t h1 = g + 1 < W ? heads[g + 1] : acc;
        if (h1 > heads[g]) last.push_back(h1 - 1);
      }
      uint64_t j = 0;
      size_t k = 0;
      for (uint64_t r = 0; r < c.n && k < last.size(); ++r) {
        if (off[r + 1] == off[r]) continue;
        if (j == last[k]) {
          fix[r >> 6] |= 1ull << (r & 63);
          ++k;
        }
        ++j;
      }
    }
    uint32_t *dw = nullptr, *dh = nullptr, *dp = nullptr;
    uint64_t *db = nullptr, *df = nullptr;
    if ((s = dalloc(ctx, (void**)&dw, (W + 1) * 4)) || (s = dalloc(ctx, (void**)&dh, W * 4)) ||
        (s = dalloc(ctx, (void**)&db, (words + 1) * 8)) || (s = dalloc(ctx, (void**)&dp, (words + 1) * 4)) ||
        (s = dalloc(ctx, (void**)&df, (words + 1) * 8))) {
      dfree(ctx, dw, (W + 1) * 4);
      dfree(ctx, dh, W * 4);
      dfree(ctx, db, (words + 1) * 8);
      dfree(ctx, dp, (words + 1) * 4);
      return s;
    }
    WG_HIP(ctx, hipMemcpyAsync(dw, wave.data(), (W + 1) * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(dh, heads.data(), W * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(db, bits.data(), (words + 1) * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(dp, pre.data(), (words + 1) * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(df, fix.data(), (words + 1) * 8, hipMemcpyHostToDevice, st));
    // The copies read host vectors that go out of scope here.
    WG_HIP(ctx, hipStreamSynchronize(st));
    c.tile_wave = dw;
    c.tile_heads = dh;
    c.nz_bits = db;
    c.nz_pre = dp;
    c.fix_bits = df;
    c.nz_rows = acc;
    c.tile_nwaves = W;
    c.tile_padded = ntiles * kTileEntries;
    c.tile_state = 1;
  }
  if (!c.tiles.count((int)esize)) {
    DevCsr::Tiles t;
    t.nhot = (uint32_t)std::min<uint64_t>(c.hot_n, hotcap);
    // The warm run: the next ranks, while their slots (n + slot) fit the
    // tag's column bits.
    uint64_t nw = std::min<uint64_t>(c.hot_n - t.nhot, warmcap);
    if (c.n + nw >= (1ull << 30)) nw = 0;
    t.nwarm = (uint32_t)nw;
    if (nw) {
      std::vector<uint32_t> ids(nw);
      WG_HIP(ctx, hipMemcpyAsync(ids.data(), c.hot_inv + t.nhot, nw * 4, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      std::sort(ids.begin(), ids.end());
      if ((s = dalloc(ctx, (void**)&t.warm_ids, nw * 4))) return s;
      if ((s = dalloc(ctx, &t.warm_x, nw * esize))) {
        dfree(ctx, t.warm_ids, nw * 4);
        return s;
      }
      WG_HIP(ctx, hipMemcpyAsync(t.warm_ids, ids.data(), nw * 4, hipMemcpyHostToDevice, st));
      // The copy reads a host vector that goes out of scope here.
      WG_HIP(ctx, hipStreamSynchronize(st));
    }
    if ((s = dalloc(ctx, (void**)&t.tagged, c.tile_padded * 4))) {
      dfree(ctx, t.warm_ids, nw * 4);
      dfree(ctx, t.warm_x, nw * esize);
      return s;
    }
    const unsigned grid = (unsigned)std::min<uint64_t>((c.tile_padded + 255) / 256, 16384);
    tile_tag<<<grid, 256, 0, st>>>(c.targets, c.hot_rank, t.nhot, t.warm_ids, t.nwarm, (uint32_t)c.n, c.nnz,
                                   c.tile_padded, t.tagged);
    tile_heads<<<(unsigned)std::min<uint64_t>((c.n + 255) / 256, 16384), 256, 0, st>>>(c.offsets, c.n, t.tagged);
    WG_HIP(ctx, hipGetLastError());
    c.tiles[(int)esize] = t;
  }
  return WG_OK;
}

// The weights of one element type in the tile order (copied once from wv,
// the CSR-order weights of that type, and kept with the layout).
template <class E>
int tile_weight_view(wg_context* ctx, DevCsr& c, const E* wv, const E** out) {
  DevCsr::Tiles& t = c.tiles[(int)sizeof(E)];
  const int d = DtypeOf<E>::v;
  if (!t.w[d]) {
    void* p = nullptr;
    if (int s = dalloc(ctx, &p, c.tile_padded * sizeof(E))) return s;
    tile_weights<E><<<(unsigned)std::min<uint64_t>((c.tile_padded + 255) / 256, 16384), 256, 0, ctx->stream>>>(
        wv, c.nnz, c.tile_padded, (E*)p);
    WG_HIP(ctx, hipGetLastError());
    t.w[d] = p;
  }
  *out = (const E*)t.w[d];
  return WG_OK;
}
template int tile_weight_view<double>(wg_context*, DevCsr&, const double*, const double**);

// y = A (+.x) x through the tiled kernel; wv: the weights of this element
// type in CSR order, or null (every weight 1).
template <class E, int ADD, int MUL>
static int launch_tiles(wg_context* ctx, DevCsr& c, const E* wv, const E* x, E* y, const WriteArgs& wa) {
  using A = typename Fold<E, ADD>::A;
  constexpr bool can_fail = std::is_same<E, int64_t>::value || std::is_same<E, uint32_t>::value;
  hipStream_t st = ctx->stream;
  DevCsr::Tiles& t = c.tiles[(int)sizeof(E)];
  int s;
  const E* wt = nullptr;
  if (wv && (s = tile_weight_view(ctx, c, wv, &wt))) return s;
  const size_t ybytes = ((c.nz_rows * sizeof(A)) + 255) & ~(size_t)255;
  const size_t need = ybytes + (size_t)c.tile_nwaves * sizeof(TileRec<A>);
  if (need > c.tile_out_bytes) {
    if (c.tile_out) {
      WG_HIP(ctx, hipStreamSynchronize(st));
      dfree(ctx, c.tile_out, c.tile_out_bytes);
      c.tile_out = nullptr;
      c.tile_out_bytes = 0;
    }
    if ((s = dalloc(ctx, &c.tile_out, need))) return s;
    c.tile_out_bytes = need;
  }
  A* ycomp = (A*)c.tile_out;
  TileRec<A>* recs = (TileRec<A>*)((char*)c.tile_out + ybytes);
  constexpr bool rx = reads_x<MUL>();
  if (rx && (t.nhot || t.nwarm))
    gather_tiers<E><<<gather_tiers_grid(t.nhot, t.nwarm), 256, 0, st>>>(
        x, c.hot_inv, t.nhot, (E*)c.hot_x, t.warm_ids, t.nwarm, (E*)t.warm_x, can_fail ? ctx->d_jerr : nullptr);
  else if (can_fail)
    tile_clear<<<1, 1, 0, st>>>(ctx->d_jerr);
  const unsigned grid = c.tile_nwaves / kTileWaves;
  const uint32_t ncols = (uint32_t)c.n;
  if (wt)
    spmv_tiles<E, ADD, MUL, true><<<grid, 1024, 0, st>>>(t.tagged, wt, x, (const E*)c.hot_x, t.nhot,
                                                         (const E*)t.warm_x, ncols, c.tile_wave, c.tile_heads,
                                                         c.nnz, ycomp, recs, ctx->d_jerr, nullptr);
  else
    spmv_tiles<E, ADD, MUL, false><<<grid, 1024, 0, st>>>(t.tagged, nullptr, x, (const E*)c.hot_x, t.nhot,
                                                          (const E*)t.warm_x, ncols, c.tile_wave, c.tile_heads,
                                                          c.nnz, ycomp, recs, ctx->d_jerr, nullptr);
  const TileGeom geo{c.nz_bits, c.nz_pre, c.fix_bits, c.offsets, c.n, c.tile_padded / kTileEntries, c.tile_nwaves};
  constexpr uint64_t wrows = 256 * kTileWriteRows;
  tile_write<E, ADD><<<(unsigned)((c.n + wrows - 1) / wrows), 256, 0, st>>>(geo, ycomp, recs, y, wa, ctx->d_jerr,
                                                                         ctx->d_err);
  WG_HIP(ctx, hipGetLastError());
  return WG_OK;
}

template <class E, int ADD, int MUL>
int launch_product(wg_context* ctx, DevCsr& csr, bool weighted, const E* x, E* y,
                   uint64_t width, const WriteArgs& wa) {
  if constexpr (std::is_same<E, uint32_t>::value && ADD == MAX) {
    return fail(ctx, WG_UNSUPPORTED, "u32 has no image of the max identity (i64::MIN)");
  } else {
    using A = typename Fold<E, ADD>::A;
    constexpr bool reads_w = MUL != SECOND && MUL != PAIR;
    const E* wv = nullptr;
    if (reads_w && weighted && csr.weights) {
      int s = weight_view(ctx, csr, DtypeOf<E>::v, (const void**)&wv);
      if (s) return s;
    }
    if (csr.targets64) return launch_wide<E, ADD, MUL>(ctx, csr, wv, x, y, width, wa);
    int s;
    if (width == 1) {
      if ((s = ensure_scratch(ctx, (size_t)std::max<uint32_t>(csr.npartials, 1) * sizeof(A)))) return s;
      if (csr.nblocks == 0) return WG_OK;
      A* part = (A*)ctx->scratch;
      constexpr bool direct_ok = grouping_free<E, ADD>();
      // The direct kernel groups a row's float sum by lane stride, the
      // blocked one by contiguous run: both within the rounding bound, but
      // not the same bits, so the tuner (whose choice follows timing) only
      // takes it where the bits are the same.
      constexpr bool direct_tunable = direct_ok && !(is_float_v<E> && ADD == PLUS);
      // The hot-first gathers change only where x is read from: the same
      // values, folded in the same order, so the result is the same bits.
      auto hot_ok = [&]() {
        if (!reads_x<MUL>()) return false;
        if (!csr.hot_state && prepare_hot_columns(ctx, csr)) return false;
        return csr.hot_state == 1;
      };
      // The LDS variant's tagged targets for this element size (built once),
      // and the resident workgroups it launches.
      constexpr uint32_t kHotCap = kHotLdsBytes / sizeof(E);
      auto lds_ok = [&]() -> bool {
        if (!hot_ok() || csr.n >= (1ull << 31)) return false;
        if (csr.lds_tgt.count(sizeof(E))) return true;
        uint32_t* t = nullptr;
        if (dalloc(ctx, (void**)&t, std::max<uint64_t>(csr.nnz, 1) * 4)) {
          (void)ok(ctx);
          return false;
        }
        const uint32_t nhot = (uint32_t)std::min<uint64_t>(csr.hot_n, kHotCap);
        tag_targets<<<2048, 256, 0, ctx->stream>>>(csr.targets, csr.hot_rank, nhot, csr.nnz, t);
        csr.lds_tgt[sizeof(E)] = {t, nhot};
        return hipGetLastError() == hipSuccess;
      };
      auto lds_grid = [&]() -> unsigned {
        int per_cu = 0, cus = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, wv ? (const void*)spmv_lds<E, ADD, MUL, true> : (const void*)spmv_lds<E, ADD, MUL, false>,
            kBlockThreads, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
        return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(csr.nblocks, (uint64_t)std::max(per_cu, 1) * std::max(cus, 1)));
      };
      // The tiled kernel's layout for this element size (built once); a
      // budget that cannot hold it leaves the other kernels.
      auto tile_ok = [&]() -> bool {
        const int st = prepare_tiles(ctx, csr, sizeof(E), tile_hot_cap<E>(), tile_warm_cap<E>());
        if (st == WG_OK) return true;
        (void)ok(ctx);
        return false;
      };
      int tile_status = WG_OK;
      auto launch = [&](int variant, E* out) {
        if (variant & kVarTile) {
          tile_status = launch_tiles<E, ADD, MUL>(ctx, csr, wv, x, out, wa);
          return;
        }
        const uint32_t* tg = csr.targets;
        const E* xs = x;
        if (variant & kVarLds) {
          const auto& lt = csr.lds_tgt[sizeof(E)];
          gather_hot<E><<<(unsigned)std::max<uint32_t>(1, (lt.second + 255) / 256), 256, 0, ctx->stream>>>(
              x, csr.hot_inv, lt.second, (E*)csr.hot_x);
          const unsigned grid = lds_grid();
          if (wv)
            spmv_lds<E, ADD, MUL, true><<<grid, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.nblocks, csr.offsets, lt.first, wv, x, (const E*)csr.hot_x, lt.second, out, wa,
                part, ctx->d_err, variant);
          else
            spmv_lds<E, ADD, MUL, false><<<grid, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.nblocks, csr.offsets, lt.first, nullptr, x, (const E*)csr.hot_x, lt.second, out,
                wa, part, ctx->d_err, variant);
          return;
        }
        if (variant & kVarHot) {
          gather_hot<E><<<(unsigned)std::max<uint64_t>(1, std::min<uint64_t>((csr.hot_n + 255) / 256, 16384)),
                          256, 0, ctx->stream>>>(x, csr.hot_inv, csr.hot_n, (E*)csr.hot_x);
          tg = csr.hot_tgt;
          xs = (const E*)csr.hot_x;
        }
        if (direct_ok && (variant & kVarDirect)) {
          if (wv)
            spmv_direct<E, ADD, MUL, true><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.offsets, tg, wv, xs, out, wa, part, ctx->d_err, variant);
          else
            spmv_direct<E, ADD, MUL, false><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.offsets, tg, nullptr, xs, out, wa, part, ctx->d_err, variant);
        } else {
          if (wv)
            spmv_blocks<E, ADD, MUL, true><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.offsets, tg, wv, xs, out, wa, part, ctx->d_err, variant);
          else
            spmv_blocks<E, ADD, MUL, false><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
                csr.blocks, csr.offsets, tg, (const E*)nullptr, xs, out, wa, part, ctx->d_err, variant);
        }
      };
      int variant = 0;
      const int key = ((DtypeOf<E>::v * 16 + ADD) * 16 + MUL) * 2 + (wv ? 1 : 0);
      if (ctx->spmv_variant >= 0) {
        variant = ctx->spmv_variant;
        if (variant & kVarTile) variant = tile_ok() ? kVarTile : (variant & ~kVarTile);
        if ((variant & kVarHot) && !hot_ok()) variant &= ~kVarHot;
        if ((variant & kVarLds) && !lds_ok()) variant &= ~kVarLds;
      } else if (csr.tuned.count(key)) {
        variant = csr.tuned[key];
      } else if (is_float_v<E> && ADD == PLUS && csr.nnz >= kTuneMinNnz) {
        // Float plus: the tiled kernel on large graphs, chosen by size, not
        // by timing (its sums group differently from the blocked kernel's).
        variant = tile_ok() ? kVarTile : 0;
        csr.tuned[key] = variant;
      } else if (csr.nnz >= kTuneMinNnz) {
        // Autotune on this graph: time every eligible variant into a scratch
        // y (the caller's y may be accumulated into), keep the fastest.
        void* ys = nullptr;
        if ((s = dalloc(ctx, &ys, csr.n * sizeof(E)))) return s;
        hipEvent_t e0, e1;
        WG_HIP(ctx, hipEventCreate(&e0));
        WG_HIP(ctx, hipEventCreate(&e1));
        float best = 1e30f;
        const bool hot = hot_ok();
        const bool lds = hot && lds_ok();
        const bool tile = tile_ok();
        for (int v = 0; v <= kVarTile; ++v) {
          if (v == kVarTile && !tile) continue;
          if ((v & kVarDirect) && !direct_tunable) continue;
          if ((v & kVarHot) && !hot) continue;
          // The LDS kernel is persistent, LDS-staged and reads x as given.
          if ((v & kVarLds) && (!lds || (v & (kVarDirect | kVarXcd | kVarHot)))) continue;
          launch(v, (E*)ys);
          if (tile_status) return tile_status;
          WG_HIP(ctx, hipEventRecord(e0, ctx->stream));
          for (int r = 0; r < 3; ++r) launch(v, (E*)ys);
          WG_HIP(ctx, hipEventRecord(e1, ctx->stream));
          WG_HIP(ctx, hipEventSynchronize(e1));
          float ms = 0;
          WG_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            variant = v;
          }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        dfree(ctx, ys, csr.n * sizeof(E));
        csr.tuned[key] = variant;
        if ((s = clear_error(ctx))) return s;
      }
      launch(variant, y);
      if (tile_status) return tile_status;
      if (csr.nlong && !(variant & kVarTile))
        spmv_long<E, ADD><<<(csr.nlong + 255) / 256, 256, 0, ctx->stream>>>(
            csr.long_rows, csr.long_pbase, csr.nlong, part, y, wa, ctx->d_err);
    } else {
      if ((s = prepare_spmm(ctx, csr))) return s;
      if ((s = ensure_scratch(ctx, (size_t)std::max<uint32_t>(csr.spmm_npartials, 1) * width * sizeof(A))))
        return s;
      // A lane takes 16 bytes of columns an entry when the rows of x split
      // into aligned 16-byte groups; else one column.
      constexpr int kC = (sizeof(E) == 4 || sizeof(E) == 8) ? (int)(16 / sizeof(E)) : 1;
      const bool groupable = kC > 1 && width % kC == 0 && (uintptr_t)x % 16 == 0;
      if (csr.spmm_ntasks == 0) return WG_OK;
      A* part = (A*)ctx->scratch;
      // A launch shape: lanes per row (a power of two, each lane taking
      // every lpr-th column group) and whether a group is 16 bytes of
      // columns or one. Every shape folds each column over the same entries
      // in the same order, so which one runs never changes a result.
      auto lanes_for = [&](bool grouped) {
        const uint64_t groups = grouped ? width / kC : width;
        uint32_t l = 1;
        while (l < groups && l < 64) l <<= 1;
        return l;
      };
      auto launch_mm = [&](int shape, E* out) {
        const uint32_t lpr = (uint32_t)(shape & 255);
        const bool grouped = (shape & 256) != 0;
        const uint64_t grid = (csr.spmm_ntasks + (256 / lpr) - 1) / (256 / lpr);
        auto run = [&](auto hw, auto cols) {
          spmm_rows<E, ADD, MUL, decltype(hw)::value, decltype(cols)::value>
              <<<(unsigned)grid, 256, 0, ctx->stream>>>(csr.targets, wv, x, out, width, csr.spmm_tasks,
                                                        csr.spmm_ntasks, lpr, part, wa, ctx->d_err);
        };
        if (wv) {
          if (grouped) run(std::true_type{}, std::integral_constant<int, kC>{});
          else run(std::true_type{}, std::integral_constant<int, 1>{});
        } else {
          if (grouped) run(std::false_type{}, std::integral_constant<int, kC>{});
          else run(std::false_type{}, std::integral_constant<int, 1>{});
        }
        if (csr.spmm_nlong) {
          const uint64_t items = (uint64_t)csr.spmm_nlong * width;
          spmm_long<E, ADD><<<(unsigned)((items + 255) / 256), 256, 0, ctx->stream>>>(
              csr.spmm_long_rows, csr.spmm_pbase, csr.spmm_nlong, width, part, out, wa, ctx->d_err);
        }
      };
      int shape = (int)lanes_for(groupable) | (groupable ? 256 : 0);
      const uint64_t mkey = ((uint64_t)(((DtypeOf<E>::v * 16 + ADD) * 16 + MUL) * 2 + (wv ? 1 : 0)) << 32) | width;
      if (ctx->spmm_shape > 0) {
        shape = ctx->spmm_shape;
        if ((shape & 256) && !groupable) shape &= 255;
      } else if (csr.spmm_tuned.count(mkey)) {
        shape = csr.spmm_tuned[mkey];
      } else if (csr.nnz >= kTuneMinNnz) {
        // Autotune the shape on this graph and width: each candidate timed
        // into a scratch y; the fastest is kept with the CSR.
        std::vector<int> cands;
        for (int grouped = groupable ? 1 : 0; grouped >= 0; --grouped)
          for (uint32_t l = lanes_for(grouped != 0), k = 0; k < 3 && l >= 1; l >>= 1, ++k)
            cands.push_back((int)l | (grouped ? 256 : 0));
        void* ys = nullptr;
        if ((s = dalloc(ctx, &ys, csr.n * width * sizeof(E)))) return s;
        hipEvent_t e0, e1;
        WG_HIP(ctx, hipEventCreate(&e0));
        WG_HIP(ctx, hipEventCreate(&e1));
        float best = 1e30f;
        for (int c : cands) {
          launch_mm(c, (E*)ys);
          WG_HIP(ctx, hipEventRecord(e0, ctx->stream));
          for (int r = 0; r < 3; ++r) launch_mm(c, (E*)ys);
          WG_HIP(ctx, hipEventRecord(e1, ctx->stream));
          WG_HIP(ctx, hipEventSynchronize(e1));
          float ms = 0;
          WG_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
          if (ms < best) {
            best = ms;
            shape = c;
          }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        dfree(ctx, ys, csr.n * width * sizeof(E));
        csr.spmm_tuned[mkey] = shape;
        if ((s = clear_error(ctx))) return s;
      }
      launch_mm(shape, y);
    }
    WG_HIP(ctx, hipGetLastError());
    return WG_OK;
  }
}

// Semiring::ALL order (compute/semiring.rs:149-163) to (+, x).
template <class E>
int dispatch(wg_context* ctx, DevCsr& csr, int semiring, bool weighted, const E* x, E* y,
             uint64_t width, const WriteArgs& wa) {
  switch (semiring) {
    case WG_PLUS_TIMES: return launch_product<E, PLUS, TIMES>(ctx, csr, weighted, x, y, width, wa);
    case WG_MIN_PLUS: return launch_product<E, MIN, PLUS>(ctx, csr, weighted, x, y, width, wa);
    case WG_MAX_PLUS: return launch_product<E, MAX, PLUS>(ctx, csr, weighted, x, y, width, wa);
    case WG_MAX_TIMES: return launch_product<E, MAX, TIMES>(ctx, csr, weighted, x, y, width, wa);
    case WG_MIN_TIMES: return launch_product<E, MIN, TIMES>(ctx, csr, weighted, x, y, width, wa);
    case WG_LOR_LAND: return launch_product<E, LOR, LAND>(ctx, csr, weighted, x, y, width, wa);
    case WG_ANY_PAIR: return launch_product<E, ANY, PAIR>(ctx, csr, weighted, x, y, width, wa);
    case WG_PLUS_PAIR: return launch_product<E, PLUS, PAIR>(ctx, csr, weighted, x, y, width, wa);
    case WG_PLUS_FIRST: return launch_product<E, PLUS, FIRST>(ctx, csr, weighted, x, y, width, wa);
    case WG_PLUS_SECOND: return launch_product<E, PLUS, SECOND>(ctx, csr, weighted, x, y, width, wa);
    case WG_MIN_FIRST: return launch_product<E, MIN, FIRST>(ctx, csr, weighted, x, y, width, wa);
    case WG_MIN_SECOND: return launch_product<E, MIN, SECOND>(ctx, csr, weighted, x, y, width, wa);
    case WG_MAX_SECOND: return launch_product<E, MAX, SECOND>(ctx, csr, weighted, x, y, width, wa);
  }
  return fail(ctx, WG_INVALID_ARGUMENT, "unknown semiring");
}

// PageRank's out-weights: PlusFirst over the forward CSR in f64
// (compute/pagerank/mod.rs:115-117).
int launch_out_weights(wg_context* ctx, wg_graph* g, bool weighted, const double* x, double* out) {
  WriteArgs wa{nullptr, 0, 0, 0, -1};
  return launch_product<double, PLUS, FIRST>(ctx, g->fwd, weighted, x, out, 1, wa);
}

// y = A (+.x) x (or A^T with transpose) in f64; unweighted entries read 1.
int launch_plus_times_f64(wg_context* ctx, wg_graph* g, bool transpose, bool weighted, const double* x,
                          double* y) {
  WriteArgs wa{nullptr, 0, 0, 0, -1};
  return launch_product<double, PLUS, TIMES>(ctx, transpose ? g->rev : g->fwd, weighted, x, y, 1, wa);
}

// The checks of spmv / spmm / multiply (compute/product/mod.rs:77-153).
static int product_checks(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                          wg_vector* y, bool spmv) {
  if (!ctx || !g || !p || !x || !y) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  int s;
  if ((s = check_graph(ctx, g, true)) || (s = check_vector(ctx, x)) || (s = check_vector(ctx, y)) ||
      (p->mask && (s = check_vector(ctx, p->mask))))
    return s;
  if (spmv) {
    for (uint64_t wd : {x->width, y->width})
      if (wd != 1) return fail(ctx, WG_WIDTH_MISMATCH, "width mismatch: expected 1, found " + std::to_string(wd));
  } else if (x->width != y->width) {
    return fail(ctx, WG_WIDTH_MISMATCH, "width mismatch: expected " + std::to_string(y->width) +
                                            ", found " + std::to_string(x->width));
  }
  if (x->dtype != y->dtype) return fail(ctx, WG_DTYPE_MISMATCH, "dtype mismatch");
  for (const wg_vector* v : {x, (const wg_vector*)y})
    if (v->id_space != g->n)
      return fail(ctx, WG_ID_SPACE_MISMATCH, "id space mismatch: expected " + std::to_string(g->n) +
                                                 ", found " + std::to_string(v->id_space));
  if (p->mask) {
    if (p->mask->id_space != g->n) return fail(ctx, WG_ID_SPACE_MISMATCH, "mask id space mismatch");
    if (p->mask->width != 1) return fail(ctx, WG_WIDTH_MISMATCH, "mask width must be 1");
  }
  if (p->semiring < 0 || p->semiring > 12) return fail(ctx, WG_INVALID_ARGUMENT, "unknown semiring");
  if (p->accumulate < -1 || p->accumulate > 9) return fail(ctx, WG_INVALID_ARGUMENT, "unknown binary op");
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  if (p->nrelations && !p->relations) return fail(ctx, WG_INVALID_ARGUMENT, "null relation list");
  return WG_OK;
}

static int product(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                   wg_vector* y, bool spmv) {
  int s = product_checks(ctx, g, p, x, y, spmv);
  if (s) return s;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_graph* gm = const_cast<wg_graph*>(g);
  if ((p->relations || p->per_relation) && (s = check_narrow(ctx, g))) return s;
  if ((p->relations || p->per_relation) &&
      (s = select_graph(ctx, gm, p->relations, p->nrelations, p->per_relation, &gm)))
    return s;
  DevCsr& csr = p->transpose ? gm->rev : gm->fwd;
  // Product::weighted None means "if the graph has weights"; a graph without
  // them reads every weight as 1 either way (matrix.rs:140-141).
  const bool weighted = (p->weighted < 0 || p->weighted != 0) && gm->weighted;
  WriteArgs wa{p->mask ? p->mask->data : nullptr, p->mask ? p->mask->dtype : 0,
               p->mask_complement, p->replace, p->accumulate};
  const bool can_fail = x->dtype == WG_I64 || x->dtype == WG_U32;
  if (can_fail && (s = clear_error(ctx))) return s;
  switch (x->dtype) {
    case WG_F32: s = dispatch<float>(ctx, csr, p->semiring, weighted, (const float*)x->data, (float*)y->data, x->width, wa); break;
    case WG_F64: s = dispatch<double>(ctx, csr, p->semiring, weighted, (const double*)x->data, (double*)y->data, x->width, wa); break;
    case WG_I64: s = dispatch<int64_t>(ctx, csr, p->semiring, weighted, (const int64_t*)x->data, (int64_t*)y->data, x->width, wa); break;
    case WG_BOOL: s = dispatch<uint8_t>(ctx, csr, p->semiring, weighted, (const uint8_t*)x->data, (uint8_t*)y->data, x->width, wa); break;
    case WG_U32: s = dispatch<uint32_t>(ctx, csr, p->semiring, weighted, (const uint32_t*)x->data, (uint32_t*)y->data, x->width, wa); break;
    default: return fail(ctx, WG_INVALID_ARGUMENT, "unknown dtype");
  }
  if (s) return s;
  if (can_fail) return collect_error(ctx, spmv ? "spmv" : "spmm");
  if (ctx->blocking) WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ok(ctx);
}

// ---- the product without allocation (wg_spmv_noalloc) -------------------
// The split rows' partials of wg_spmv_noalloc live in this array: one copy
// per device, held until the process ends, shared by every such call on
// that device. Host code reaches it only through the address
// hipGetSymbolAddress gives for the device (static_scratch below).
constexpr size_t kStaticScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

// The device address of `scratch` on `device` (current), looked up once per
// device; called when a context is created.
int static_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return WG_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return WG_DEVICE_ERROR;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return WG_OK;
}

// The blocked kernel (launch variant 0: wg_spmv's default below 2^22
// entries) over the CSR's own buffers; partials in the static scratch;
// weights as cached for this element type, or the f64 weights cast per
// entry. Sets *collect when the call can raise an error on the device.
template <class E, int ADD, int MUL>
static int launch_noalloc(wg_context* ctx, DevCsr& csr, bool weighted, const E* x, E* y, const WriteArgs& wa,
                          bool* collect) {
  if constexpr (std::is_same<E, uint32_t>::value && ADD == MAX) {
    return fail(ctx, WG_UNSUPPORTED, "u32 has no image of the max identity (i64::MIN)");
  } else {
    using A = typename Fold<E, ADD>::A;
    constexpr bool reads_w = MUL != SECOND && MUL != PAIR;
    *collect = std::is_same<E, int64_t>::value || std::is_same<E, uint32_t>::value;
    if (csr.nblocks == 0) return WG_OK;
    const size_t need = (size_t)csr.npartials * sizeof(A);
    if (need > kStaticScratchBytes)
      return fail(ctx, WG_MEMORY_LIMIT,
                  "wg_spmv_noalloc: the graph's split rows need " + std::to_string(need) +
                      " bytes of partials, above its fixed scratch of " + std::to_string(kStaticScratchBytes));
    A* part = (A*)ctx->static_scratch;
    hipStream_t st = ctx->stream;
    const int dt = DtypeOf<E>::v;
    const bool hw = reads_w && weighted && csr.weights;
    if (!hw) {
      spmv_blocks<E, ADD, MUL, false><<<csr.nblocks, kBlockThreads, 0, st>>>(
          csr.blocks, csr.offsets, csr.targets, (const E*)nullptr, x, y, wa, part, ctx->d_err, 0);
    } else if (std::is_same<E, double>::value || (csr.wstatus[dt] == WG_OK && csr.wview[dt])) {
      const E* w = std::is_same<E, double>::value ? (const E*)csr.weights : (const E*)csr.wview[dt];
      spmv_blocks<E, ADD, MUL, true><<<csr.nblocks, kBlockThreads, 0, st>>>(
          csr.blocks, csr.offsets, csr.targets, w, x, y, wa, part, ctx->d_err, 0);
    } else if (csr.wstatus[dt] > 0) {
      return fail(ctx, csr.wstatus[dt], "edge weight cast failed");
    } else {
      if constexpr (!std::is_same<E, double>::value) {
        spmv_blocks<E, ADD, MUL, true, double><<<csr.nblocks, kBlockThreads, 0, st>>>(
            csr.blocks, csr.offsets, csr.targets, csr.weights, x, y, wa, part, ctx->d_err, 0);
        *collect = true;  // the cast can fail (f32 range, i64 range)
      }
    }
    if (csr.nlong)
      spmv_long<E, ADD><<<(csr.nlong + 255) / 256, 256, 0, st>>>(csr.long_rows, csr.long_pbase, csr.nlong, part,
                                                                 y, wa, ctx->d_err);
    WG_HIP(ctx, hipGetLastError());
    return WG_OK;
  }
}

template <class E>
static int dispatch_noalloc(wg_context* ctx, DevCsr& csr, int semiring, bool weighted, const E* x, E* y,
                            const WriteArgs& wa, bool* collect) {
  switch (semiring) {
    case WG_PLUS_TIMES: return launch_noalloc<E, PLUS, TIMES>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MIN_PLUS: return launch_noalloc<E, MIN, PLUS>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MAX_PLUS: return launch_noalloc<E, MAX, PLUS>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MAX_TIMES: return launch_noalloc<E, MAX, TIMES>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MIN_TIMES: return launch_noalloc<E, MIN, TIMES>(ctx, csr, weighted, x, y, wa, collect);
    case WG_LOR_LAND: return launch_noalloc<E, LOR, LAND>(ctx, csr, weighted, x, y, wa, collect);
    case WG_ANY_PAIR: return launch_noalloc<E, ANY, PAIR>(ctx, csr, weighted, x, y, wa, collect);
    case WG_PLUS_PAIR: return launch_noalloc<E, PLUS, PAIR>(ctx, csr, weighted, x, y, wa, collect);
    case WG_PLUS_FIRST: return launch_noalloc<E, PLUS, FIRST>(ctx, csr, weighted, x, y, wa, collect);
    case WG_PLUS_SECOND: return launch_noalloc<E, PLUS, SECOND>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MIN_FIRST: return launch_noalloc<E, MIN, FIRST>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MIN_SECOND: return launch_noalloc<E, MIN, SECOND>(ctx, csr, weighted, x, y, wa, collect);
    case WG_MAX_SECOND: return launch_noalloc<E, MAX, SECOND>(ctx, csr, weighted, x, y, wa, collect);
  }
  return fail(ctx, WG_INVALID_ARGUMENT, "unknown semiring");
}

static int product_noalloc(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                           wg_vector* y) {
  if (int cs_ = check_narrow(ctx, g)) return cs_;
  int s = product_checks(ctx, g, p, x, y, true);
  if (s) return s;
  if (p->relations || p->nrelations || p->per_relation)
    return fail(ctx, WG_UNSUPPORTED, "wg_spmv_noalloc reads the graph's own CSR; a relation selection is built "
                                     "(allocated) by wg_spmv");
  if (!ctx->static_scratch) return fail(ctx, WG_DEVICE_ERROR, "no static scratch on this device");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  const DevCsr& csr = p->transpose ? g->rev : g->fwd;
  const bool weighted = (p->weighted < 0 || p->weighted != 0) && g->weighted;
  WriteArgs wa{p->mask ? p->mask->data : nullptr, p->mask ? p->mask->dtype : 0, p->mask_complement, p->replace,
               p->accumulate};
  DevCsr& c = const_cast<DevCsr&>(csr);
  bool collect = false;
  if ((s = clear_error(ctx))) return s;
  switch (x->dtype) {
    case WG_F32: s = dispatch_noalloc<float>(ctx, c, p->semiring, weighted, (const float*)x->data, (float*)y->data, wa, &collect); break;
    case WG_F64: s = dispatch_noalloc<double>(ctx, c, p->semiring, weighted, (const double*)x->data, (double*)y->data, wa, &collect); break;
    case WG_I64: s = dispatch_noalloc<int64_t>(ctx, c, p->semiring, weighted, (const int64_t*)x->data, (int64_t*)y->data, wa, &collect); break;
    case WG_BOOL: s = dispatch_noalloc<uint8_t>(ctx, c, p->semiring, weighted, (const uint8_t*)x->data, (uint8_t*)y->data, wa, &collect); break;
    case WG_U32: s = dispatch_noalloc<uint32_t>(ctx, c, p->semiring, weighted, (const uint32_t*)x->data, (uint32_t*)y->data, wa, &collect); break;
    default: return fail(ctx, WG_INVALID_ARGUMENT, "unknown dtype");
  }
  if (s) return s;
  if (collect) return collect_error(ctx, "spmv");
  if (ctx->blocking) WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ok(ctx);
}

}  // namespace wg

extern "C" {

int wg_spmv_noalloc(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                    wg_vector* y) {
  return wg::product_noalloc(ctx, g, p, x, y);
}

void wg_set_spmv_variant(wg_context* ctx, int variant) {
  if (ctx) ctx->spmv_variant = variant < 0 ? -1 : (variant & 63);
}

int wg_spmv_variant(const wg_graph* g, int transpose, int dtype, int semiring, int weighted) {
  static const int add_of[13] = {wg::PLUS, wg::MIN, wg::MAX, wg::MAX, wg::MIN, wg::LOR, wg::ANY,
                                 wg::PLUS, wg::PLUS, wg::PLUS, wg::MIN, wg::MIN, wg::MAX};
  static const int mul_of[13] = {wg::TIMES, wg::PLUS, wg::PLUS, wg::TIMES, wg::TIMES, wg::LAND,
                                 wg::PAIR, wg::PAIR, wg::FIRST, wg::SECOND, wg::FIRST, wg::SECOND,
                                 wg::SECOND};
  if (!g || semiring < 0 || semiring > 12) return -1;
  const wg::DevCsr& c = transpose ? g->rev : g->fwd;
  const int mul = mul_of[semiring];
  const bool hw = weighted && c.weights && mul != wg::SECOND && mul != wg::PAIR;
  const int key = ((dtype * 16 + add_of[semiring]) * 16 + mul) * 2 + (hw ? 1 : 0);
  auto it = c.tuned.find(key);
  return it == c.tuned.end() ? -1 : it->second;
}

int wg_spmm_shape(const wg_graph* g, int transpose, int dtype, int semiring, int weighted, uint64_t width) {
  static const int add_of[13] = {wg::PLUS, wg::MIN, wg::MAX, wg::MAX, wg::MIN, wg::LOR, wg::ANY,
                                 wg::PLUS, wg::PLUS, wg::PLUS, wg::MIN, wg::MIN, wg::MAX};
  static const int mul_of[13] = {wg::TIMES, wg::PLUS, wg::PLUS, wg::TIMES, wg::TIMES, wg::LAND,
                                 wg::PAIR, wg::PAIR, wg::FIRST, wg::SECOND, wg::FIRST, wg::SECOND,
                                 wg::SECOND};
  if (!g || semiring < 0 || semiring > 12) return -1;
  const wg::DevCsr& c = transpose ? g->rev : g->fwd;
  const int mul = mul_of[semiring];
  const bool hw = weighted && c.weights && mul != wg::SECOND && mul != wg::PAIR;
  const int key = ((dtype * 16 + add_of[semiring]) * 16 + mul) * 2 + (hw ? 1 : 0);
  auto it = c.spmm_tuned.find(((uint64_t)key << 32) | width);
  return it == c.spmm_tuned.end() ? -1 : it->second;
}

void wg_set_spmm_shape(wg_context* ctx, int shape) {
  if (ctx) ctx->spmm_shape = shape;
}

int wg_spmv(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y) {
  return wg::product(ctx, g, p, x, y, true);
}

int wg_spmm(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y) {
  return wg::product(ctx, g, p, x, y, false);
}

}  // extern "C"
