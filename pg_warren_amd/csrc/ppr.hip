// Batched personalised PageRank: R seed sets as the columns of one n x R
// power iteration (compute/pagerank/mod.rs:96-287 per column; the reference
// runs one seed vector per call). Each column is exactly the plain run with
// its seed vector: the same set-up (live count, out-weights, shares
// seed / sum(seed) with SQL float checks), the same update in f64, and its own
// stop test (L1 change < live nodes * tol); a column that stops is frozen at
// that iteration while the others go on.
//
// Layout: the ranks X are n x R row-major (the caller's vector). The
// messages msg = x / out live in column tiles of kTw columns, tile-major
// (tile t holds a 128-byte line per message row), holding the active
// columns only: when columns stop, the survivors are repacked, so an
// iteration reads ceil(active / kTw) tiles. Message rows are the nodes with
// out-edges, hottest (largest out-degree) first, so the most gathered lines
// sit together and stay in L2 and the infinity cache.
//
// A tile pass walks the reverse CSR in work items of whole rows (at most
// kItemNnz entries and kItemRows rows) or pieces of a longer row. A
// persistent workgroup takes items in turn: it stages the item's offsets
// and sources in LDS; its kG lane groups (kL lanes, kC columns a lane, one
// 16-byte load a lane per message line) each walk an equal run of entries
// with kGatherBatch lines in flight; rows that cross runs are summed in run
// order; then the rows advance x' = d (y + dangling_c p) + (1 - d) p and
// spread msg' = x' / out, several rows' loads in flight a group. Pieces of
// long rows go to a buffer that ppr_long sums in piece order. Per-column L1
// changes and dangling ranks go to per-workgroup partial slots that
// ppr_finish sums in a fixed order. No atomics and no cross-workgroup fences.
//
// The first iteration starts every column at x = 1 / live, so every column's
// messages, and each row's sum, are the same: tile 0's pass records the sums
// and ppr_first_rest advances the other tiles from them. From the second
// iteration on, zero rows (no in-edges and no seed share, or dead), which
// then hold 0 in every column, are left alone (PprRowIn).
#include <algorithm>
#include <functional>
#include <cmath>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

int launch_out_weights(wg_context* ctx, wg_graph* g, bool weighted, const double* x, double* out);

constexpr uint32_t kTw = 16;             // columns per tile: a 128-byte message line
constexpr uint32_t kC = 2;               // columns per lane in the gather kernels
constexpr uint32_t kL = kTw / kC;        // lanes per row group
constexpr uint32_t kG = 256 / kL;        // row groups per workgroup
constexpr uint32_t kRowStep = 256 / kTw; // rows per workgroup step, a lane per column
constexpr uint32_t kItemNnz = 2048;      // entries per work item
constexpr uint32_t kItemRows = 128;      // rows per work item
constexpr uint32_t kGatherBatch = 12;    // message lines in flight a lane
constexpr uint32_t kPprMaxCols = 256;

// Streaming data (the CSR, ranks, per-row inputs, new messages) bypasses
// the caches' keep policy so that the gathered message lines stay resident.
template <class T>
__device__ __forceinline__ T ld_nt(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <class T>
__device__ __forceinline__ void st_nt(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

// A lane's kC columns of a message line (one 16-byte access).
typedef double dvec2 __attribute__((ext_vector_type(2)));
struct Cols {
  double v[kC];
};
__device__ __forceinline__ Cols load_cols(const double* p) {
  const dvec2 q = *reinterpret_cast<const dvec2*>(p);
  return Cols{{q.x, q.y}};
}

// A work item: entries [b0, b0 + nent) of the reverse CSR, being the rows
// [row0, row0 + nrows), or (nrows == 0) a piece of row row0 whose partial
// sum goes to piece slot `slot`.
struct PprItem {
  uint64_t b0;
  uint32_t row0, nrows, slot, nent;
};

struct PprArgs {
  const uint64_t* off;    // reverse CSR
  const uint32_t* src;    // its sources as message rows
  const double* w;        // reverse weights or null
  uint64_t n;
  const uint32_t* hot;    // n: message row of each node (~0: no out-edges)
  uint64_t nm;            // message rows
  uint32_t R;             // row stride of X
  uint32_t nact;          // active columns (positions 0..nact-1)
  const uint32_t* cols;   // column of each position
  const double* out;      // out-weights
  const uint8_t* tomb;
  const uint32_t* seed_row;  // n: index into P, or ~0
  const double* P;           // seed rows x R shares
  double damping, teleport;
  const double* dangling;    // R: this iteration's dangling rank per column
  double* X;
  const double* Min;         // tiles x nm x kTw, positions of the active list
  double* Mout;
  double* part;              // per tile: nparts x 2 x kTw (change, dangling)
  double* piece;             // per tile: pieces x kTw
  uint32_t nparts;           // slots: ppr_block then ppr_long workgroups
  unsigned long long* err;
  const unsigned long long* cancel;  // the context's host-mapped cancel word
  unsigned* stop;                    // device word: the cancel word, forwarded
  double start;   // every live row's rank before the first iteration
  uint32_t mode;  // 0: the first iteration (the old ranks are `start`, not read);
                  // 1: the second (zero rows' ranks left alone); 2: later
                  // (zero rows' messages left alone too)
  double* y0;     // first iteration, tile 0: each row's gathered sum, or null
};

// A lane's columns: position t * kTw + kC * k + q is column c[q] when on[q].
struct LaneCols {
  uint32_t c[kC];
  bool on[kC];
  bool pair;  // both on, adjacent and 16-byte aligned in every row: one access
};
__device__ __forceinline__ LaneCols lane_cols(const PprArgs& a, uint32_t t, uint32_t k) {
  LaneCols lc;
#pragma unroll
  for (uint32_t q = 0; q < kC; ++q) {
    const uint32_t pos = t * kTw + kC * k + q;
    lc.on[q] = pos < a.nact;
    lc.c[q] = lc.on[q] ? a.cols[pos] : 0;
  }
  lc.pair = lc.on[0] && lc.on[1] && lc.c[1] == lc.c[0] + 1 && ((lc.c[0] | a.R) & 1) == 0;
  return lc;
}

// What a row's update reads besides its gathered sums.
//
// A zero row (dead, or with no in-edges and no seed share) ranks 0 in every
// column once the first iteration has run: its sum is empty and its seed
// share 0, so x' = d * (0 + dangling * 0) + (1 - d) * 0 = +0. From the second
// iteration on its ranks are neither read nor written (they hold that 0),
// its change and dangling terms are +0 and left out of the sums, and its
// message is 0; from the third on, when both message buffers hold that 0 in
// every position, the message is not written either.
struct PprRowIn {
  double xo[kC];
  double o;
  uint32_t h, sr;
  bool dead, zero;
};

__device__ __forceinline__ PprRowIn ppr_load(const PprArgs& a, uint64_t v, const LaneCols& lc, uint32_t sr,
                                             bool dead, bool empty) {
  PprRowIn in;
  in.sr = sr;
  in.dead = dead;
  in.zero = a.mode != 0 && (dead || (empty && sr == ~0u));
  in.o = 0.0;
  in.h = ~0u;
#pragma unroll
  for (uint32_t q = 0; q < kC; ++q) in.xo[q] = 0.0;
  if (in.zero) {
    if (a.mode == 1) in.h = ld_nt(a.hot + v);
    return in;
  }
  if (a.mode == 0) {
#pragma unroll
    for (uint32_t q = 0; q < kC; ++q)
      if (lc.on[q]) in.xo[q] = dead ? 0.0 : a.start;
  } else if (lc.pair) {
    const dvec2 x2 = ld_nt(reinterpret_cast<const dvec2*>(a.X + v * a.R + lc.c[0]));
    in.xo[0] = x2.x;
    in.xo[1] = x2.y;
  } else {
#pragma unroll
    for (uint32_t q = 0; q < kC; ++q)
      if (lc.on[q]) in.xo[q] = ld_nt(a.X + v * a.R + lc.c[q]);
  }
  in.o = ld_nt(a.out + v);
  in.h = ld_nt(a.hot + v);
  return in;
}

// The same, the row's seed index and tombstone read from memory.
__device__ __forceinline__ PprRowIn ppr_load_g(const PprArgs& a, uint64_t v, const LaneCols& lc, bool empty) {
  return ppr_load(a, v, lc, ld_nt(a.seed_row + v), a.tomb && ld_nt(a.tomb + v), empty);
}

__device__ __forceinline__ void ppr_advance(const PprArgs& a, uint64_t v, const LaneCols& lc, uint32_t t,
                                            uint32_t k, const double* y, const PprRowIn& in, double* change,
                                            double* dang, uint32_t& err) {
  if (in.zero) {
    if (in.h != ~0u) {
      const dvec2 mm = {0.0, 0.0};
      st_nt(reinterpret_cast<dvec2*>(a.Mout + ((uint64_t)t * a.nm + in.h) * kTw + kC * k), mm);
    }
    return;
  }
  double m[kC], nx[kC];
#pragma unroll
  for (uint32_t q = 0; q < kC; ++q) {
    m[q] = nx[q] = 0.0;
    if (!lc.on[q]) continue;
    const uint32_t c = lc.c[q];
    double next = 0.0;
    if (!in.dead) {
      const double p = in.sr == ~0u ? 0.0 : a.P[(uint64_t)in.sr * a.R + c];
      const double flowed = sqlf<ADD_>(y[q], sqlf<MUL_>(a.dangling[c], p, err), err);
      next = sqlf<ADD_>(sqlf<MUL_>(a.damping, flowed, err), sqlf<MUL_>(a.teleport, p, err), err);
    }
    nx[q] = next;
    if (sql_lt(0.0, in.o)) m[q] = sqlf<DIV_>(next, in.o, err);
    else dang[q] += next;
    change[q] += fabs(next - in.xo[q]);
  }
  if (lc.pair) {
    const dvec2 x2 = {nx[0], nx[1]};
    st_nt(reinterpret_cast<dvec2*>(a.X + v * a.R + lc.c[0]), x2);
  } else {
#pragma unroll
    for (uint32_t q = 0; q < kC; ++q)
      if (lc.on[q]) st_nt(a.X + v * a.R + lc.c[q], nx[q]);
  }
  if (in.h != ~0u) {
    const dvec2 mm = {m[0], m[1]};
    st_nt(reinterpret_cast<dvec2*>(a.Mout + ((uint64_t)t * a.nm + in.h) * kTw + kC * k), mm);
  }
}

// Sums (change, dangling) of the kG row groups of a workgroup (lane k's
// columns of each) into the workgroup's partial slot; `red` holds
// 2 x kG x kTw doubles.
__device__ __forceinline__ void ppr_block_partial(const PprArgs& a, uint32_t t, uint32_t slot,
                                                  const double* change, const double* dang, double* red) {
  const uint32_t tid = threadIdx.x, k = tid % kL, g = tid / kL;
#pragma unroll
  for (uint32_t q = 0; q < kC; ++q) {
    red[(0 * kG + g) * kTw + kC * k + q] = change[q];
    red[(1 * kG + g) * kTw + kC * k + q] = dang[q];
  }
  __syncthreads();
  if (tid < 2 * kTw) {
    const uint32_t which = tid / kTw, col = tid % kTw;
    double s = 0.0;
    for (uint32_t q = 0; q < kG; ++q) s += red[(which * kG + q) * kTw + col];
    a.part[((uint64_t)t * a.nparts + slot) * 2 * kTw + tid] = s;
  }
}

// Four workgroups a CU unweighted (LDS and registers sized for it), three
// weighted (the weights take another 16 KB of LDS).
template <bool HW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(HW ? 3 : 4))) void ppr_block(PprArgs a, uint32_t t, const PprItem* __restrict__ items,
                                                 uint32_t nitems, uint32_t npieces) {
  __shared__ uint32_t ss[kItemNnz];
  __shared__ double sw[HW ? kItemNnz : 1];
  __shared__ uint64_t so[kItemRows + 1];
  __shared__ uint32_t ssr[kItemRows];  // the rows' seed indices
  __shared__ uint8_t sdead[kItemRows]; // and tombstones
  __shared__ double Y[kItemRows * kTw];          // rows summed within one run
  __shared__ double segf[kG * kTw], segl[kG * kTw];  // a run's first / last row, when cut
  const uint32_t tid = threadIdx.x, k = tid % kL, g = tid / kL;
  const LaneCols lc = lane_cols(a, t, k);
  const bool any_on = lc.on[0];
  const double* tile = a.Min + (uint64_t)t * a.nm * kTw + kC * k;
  auto gather = [&](uint32_t s) -> Cols { return load_cols(tile + (uint64_t)s * kTw); };
  double change[kC] = {}, dang[kC] = {};
  uint32_t e_err = 0;
  __shared__ unsigned stop;
  uint32_t polls = 0;
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x, ++polls) {
    // Cancellation, checked once a work item (the load overlaps the item's
    // staging): one workgroup in 64 reads the host-mapped cancel word every
    // eighth item and forwards a raised word to a device word, which every
    // workgroup reads (host reads from every workgroup at every item held
    // the pass up on the bus). A raised word ends the pass, and the host
    // reports the call cancelled.
    if (tid == 0) {
      if ((blockIdx.x & 63) == 0 && (polls & 7) == 0 &&
          __hip_atomic_load(a.cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM))
        __hip_atomic_store(a.stop, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      stop = __hip_atomic_load(a.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const PprItem item = items[it];
    const bool piece = item.nrows == 0;
    const uint32_t nr = piece ? 1 : item.nrows;
    const uint64_t b0 = item.b0;
    const uint32_t nent = item.nent;
    if (piece) {
      if (tid < 2) so[tid] = b0 + (tid ? nent : 0);
    } else {
      for (uint32_t i = tid; i <= nr; i += 256) {
        so[i] = ld_nt(a.off + (uint64_t)item.row0 + i);
        if (i < nr) {
          ssr[i] = ld_nt(a.seed_row + (uint64_t)item.row0 + i);
          sdead[i] = a.tomb ? ld_nt(a.tomb + (uint64_t)item.row0 + i) : 0;
        }
      }
    }
    // The item's entries, a thread's eight loads in flight before its LDS
    // stores (addresses clamped to the item's last entry, stores guarded).
    if (nent) {
      constexpr uint32_t kS = kItemNnz / 256;
      uint32_t sv[kS];
      double wv[kS];
#pragma unroll
      for (uint32_t k = 0; k < kS; ++k) {
        const uint64_t e = b0 + min(tid + k * 256, nent - 1);
        sv[k] = ld_nt(a.src + e);
        if (HW) wv[k] = ld_nt(a.w + e);
      }
#pragma unroll
      for (uint32_t k = 0; k < kS; ++k) {
        const uint32_t i = tid + k * 256;
        if (i < nent) {
          ss[i] = sv[k];
          if (HW) sw[i] = wv[k];
        }
      }
    }
    __syncthreads();
    if (stop) break;
    // Group g's run of entries [c0, c1) (item-relative).
    const uint32_t chunk = (nent + kG - 1) / kG;
    const uint32_t c0 = min(g * chunk, nent), c1 = min(c0 + chunk, nent);
    if (c0 < c1) {
      uint32_t lo = 0, hi = nr;  // so[lo] <= b0 + c0 < so[hi]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (so[mid] - b0 <= c0) lo = mid;
        else hi = mid;
      }
      uint32_t r = lo, rend = (uint32_t)(so[r + 1] - b0);
      double acc[kC] = {};
      auto flush = [&]() {
        double* dst = so[r] - b0 < c0 ? &segf[g * kTw] : rend > c1 ? &segl[g * kTw] : &Y[r * kTw];
#pragma unroll
        for (uint32_t q = 0; q < kC; ++q) dst[kC * k + q] = acc[q];
      };
      for (uint32_t e = c0; e < c1; e += kGatherBatch) {
        const uint32_t nb = min(kGatherBatch, c1 - e);
        Cols m[kGatherBatch];
        double ws[kGatherBatch];
        if (nb == kGatherBatch) {
#pragma unroll
          for (uint32_t j = 0; j < kGatherBatch; ++j) {
            m[j] = gather(ss[e + j]);
            ws[j] = HW ? sw[e + j] : 1.0;
          }
        } else {
#pragma unroll
          for (uint32_t j = 0; j < kGatherBatch; ++j) {
            m[j] = Cols{};
            ws[j] = 1.0;
            if (j < nb) {
              m[j] = gather(ss[e + j]);
              if (HW) ws[j] = sw[e + j];
            }
          }
        }

        if (e + nb <= rend) {
          // No row ends before the batch does.
#pragma unroll
          for (uint32_t j = 0; j < kGatherBatch; ++j)
            if (j < nb)
#pragma unroll
              for (uint32_t q = 0; q < kC; ++q) acc[q] += HW ? ws[j] * m[j].v[q] : m[j].v[q];
        } else {
#pragma unroll
          for (uint32_t j = 0; j < kGatherBatch; ++j) {
            if (j < nb) {
              while (e + j >= rend) {
                flush();
#pragma unroll
                for (uint32_t q = 0; q < kC; ++q) acc[q] = 0.0;
                ++r;
                rend = (uint32_t)(so[r + 1] - b0);
              }
#pragma unroll
              for (uint32_t q = 0; q < kC; ++q) acc[q] += HW ? ws[j] * m[j].v[q] : m[j].v[q];
            }
          }
        }
      }
      flush();
    }
    __syncthreads();
    // A row's sums: within one run, or its runs' parts in run order.
    auto row_sum = [&](uint32_t r, double* y) {
      const uint32_t rb = (uint32_t)(so[r] - b0), re = (uint32_t)(so[r + 1] - b0);
#pragma unroll
      for (uint32_t q = 0; q < kC; ++q) y[q] = 0.0;
      if (rb == re) return;
      const uint32_t ga = rb / chunk, gb = (re - 1) / chunk;
      if (ga == gb) {
#pragma unroll
        for (uint32_t q = 0; q < kC; ++q) y[q] = Y[r * kTw + kC * k + q];
        return;
      }
#pragma unroll
      for (uint32_t q = 0; q < kC; ++q) y[q] = segl[ga * kTw + kC * k + q];
      for (uint32_t s = ga + 1; s <= gb; ++s)
#pragma unroll
        for (uint32_t q = 0; q < kC; ++q) y[q] += segf[s * kTw + kC * k + q];
    };
    if (piece) {
      if (g == 0) {
        double y[kC];
        row_sum(0, y);
#pragma unroll
        for (uint32_t q = 0; q < kC; ++q) a.piece[((uint64_t)t * npieces + item.slot) * kTw + kC * k + q] = y[q];
      }
    } else if (any_on) {
      constexpr uint32_t kBatch = 2;
      for (uint32_t r0 = g; r0 < nr; r0 += kBatch * kG) {
        PprRowIn in[kBatch];
        double y[kBatch][kC];
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
          const uint32_t r = r0 + b * kG;
          if (r < nr) {
            in[b] = ppr_load(a, (uint64_t)item.row0 + r, lc, ssr[r], sdead[r] != 0, so[r] == so[r + 1]);
            row_sum(r, y[b]);
            if (a.y0 && k == 0) a.y0[item.row0 + r] = y[b][0];
          }
        }
#pragma unroll
        for (uint32_t b = 0; b < kBatch; ++b) {
          const uint32_t r = r0 + b * kG;
          if (r < nr) ppr_advance(a, (uint64_t)item.row0 + r, lc, t, k, y[b], in[b], change, dang, e_err);
        }
      }
    }
    __syncthreads();
  }
  ppr_block_partial(a, t, blockIdx.x, change, dang, Y);
  if (e_err) raise_error(a.err, blockIdx.x, e_err);
}

// The long rows of a tile, kG a workgroup: pieces summed in piece order,
// then the update.
__global__ __launch_bounds__(256) void ppr_long(PprArgs a, uint32_t t, const uint32_t* __restrict__ rows,
                                                const uint32_t* __restrict__ pbase, uint32_t nlong,
                                                uint32_t npieces, uint32_t slot0) {
  __shared__ double red[2 * kG * kTw];
  const uint32_t i = blockIdx.x * kG + threadIdx.x / kL, k = threadIdx.x % kL;
  const LaneCols lc = lane_cols(a, t, k);
  uint32_t e_err = 0;
  double change[kC] = {}, dang[kC] = {};
  if (i < nlong && lc.on[0]) {
    const uint64_t v = rows[i];
    double y[kC] = {};
    for (uint32_t p = pbase[i]; p < pbase[i + 1]; ++p)
#pragma unroll
      for (uint32_t q = 0; q < kC; ++q) y[q] += a.piece[((uint64_t)t * npieces + p) * kTw + kC * k + q];
    if (a.y0 && k == 0) a.y0[v] = y[0];
    ppr_advance(a, v, lc, t, k, y, ppr_load_g(a, v, lc, false), change, dang, e_err);
  }
  ppr_block_partial(a, t, slot0 + blockIdx.x, change, dang, red);
  if (e_err) raise_error(a.err, blockIdx.x, e_err);
}

// A workgroup per active position, in a fixed order: the iteration's L1
// change and the next dangling rank; a column whose change falls below the
// bound stops (when count is set; the set-up pass only takes the dangling).
__global__ __launch_bounds__(256) void ppr_finish(const double* __restrict__ part, uint32_t nparts,
                                                  const uint32_t* __restrict__ cols,
                                                  double* __restrict__ dangling, uint32_t* __restrict__ iterations,
                                                  uint8_t* __restrict__ active, double bound, int has_tol,
                                                  int count) {
  __shared__ double sc[256], sd[256];
  const uint32_t pos = blockIdx.x, t = pos / kTw, k = pos % kTw, tid = threadIdx.x;
  const double* p = part + (uint64_t)t * nparts * 2 * kTw;
  double c = 0.0, d = 0.0;
  for (uint32_t b = tid; b < nparts; b += 256) {
    c += p[(uint64_t)b * 2 * kTw + k];
    d += p[(uint64_t)b * 2 * kTw + kTw + k];
  }
  sc[tid] = c;
  sd[tid] = d;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (tid < s) {
      sc[tid] += sc[tid + s];
      sd[tid] += sd[tid + s];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const uint32_t col = cols[pos];
    dangling[col] = sd[0];
    if (count) {
      iterations[col] += 1;
      if (has_tol && sql_lt(sc[0], bound)) active[col] = 0;
    }
  }
}

// Before the first iteration every column holds x = start at live rows (0
// at dead ones), so its messages are the same in every column: tile 0's
// messages x / out (every position) and each workgroup's dangling sum,
// which every tile's partial slot takes.
__global__ __launch_bounds__(256) void ppr_start(uint64_t n, uint32_t ntiles, double start,
                                                 const uint8_t* __restrict__ tomb,
                                                 const uint32_t* __restrict__ hot, const double* __restrict__ out,
                                                 double* __restrict__ M, double* __restrict__ part, uint32_t nparts,
                                                 unsigned long long* err) {
  __shared__ double red[256];
  const uint32_t tid = threadIdx.x;
  uint32_t e_err = 0;
  double dang = 0.0;
  for (uint64_t v = (uint64_t)blockIdx.x * 256 + tid; v < n; v += (uint64_t)gridDim.x * 256) {
    const double x = (tomb && tomb[v]) ? 0.0 : start;
    const double o = out[v];
    double m = 0.0;
    if (sql_lt(0.0, o)) m = sqlf<DIV_>(x, o, e_err);
    else dang += x;
    const uint32_t h = hot[v];
    if (h != ~0u) {
      const dvec2 mm = {m, m};
      dvec2* dst = reinterpret_cast<dvec2*>(M + (uint64_t)h * kTw);
#pragma unroll
      for (uint32_t q = 0; q < kTw / 2; ++q) dst[q] = mm;
    }
  }
  red[tid] = dang;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  if (tid < ntiles) {
    double* p = part + ((uint64_t)tid * nparts + blockIdx.x) * 2 * kTw;
#pragma unroll
    for (uint32_t k = 0; k < kTw; ++k) {
      p[k] = 0.0;
      p[kTw + k] = red[0];
    }
  }
  if (e_err) raise_error(err, blockIdx.x, e_err);
}

// The first iteration of tiles 1 .. ntiles - 1 (tile blockIdx.y + 1), whose
// gathered sums equal tile 0's column 0 (y0) bit for bit: every position's
// messages are the same, and each is folded in the same order. Workgroup
// blockIdx.x takes every gridDim.x-th run of kG rows and fills partial slot
// blockIdx.x (gridDim.x is the partial slot count).
__global__ __launch_bounds__(256) void ppr_first_rest(PprArgs a) {
  __shared__ double red[2 * kG * kTw];
  const uint32_t t = blockIdx.y + 1, tid = threadIdx.x, k = tid % kL, g = tid / kL;
  const LaneCols lc = lane_cols(a, t, k);
  uint32_t e_err = 0;
  double change[kC] = {}, dang[kC] = {};
  if (lc.on[0]) {
    for (uint64_t v = (uint64_t)blockIdx.x * kG + g; v < a.n; v += (uint64_t)gridDim.x * kG) {
      const double yv = ld_nt(a.y0 + v);
      double y[kC];
#pragma unroll
      for (uint32_t q = 0; q < kC; ++q) y[q] = yv;
      ppr_advance(a, v, lc, t, k, y, ppr_load_g(a, v, lc, false), change, dang, e_err);
    }
  }
  ppr_block_partial(a, t, blockIdx.x, change, dang, red);
  if (e_err) raise_error(a.err, blockIdx.x, e_err);
}

// The ranks of a call that runs no iteration: start at live rows.
__global__ __launch_bounds__(256) void ppr_fill_start(uint64_t n, uint32_t R, double start,
                                                      const uint8_t* __restrict__ tomb, double* __restrict__ X) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n * R; i += (uint64_t)gridDim.x * 256)
    X[i] = (tomb && tomb[i / R]) ? 0.0 : start;
}

// Moves the surviving positions' messages into a packed layout: position
// newpos takes old position from[newpos].
__global__ __launch_bounds__(256) void ppr_repack(uint64_t n, const double* __restrict__ Mold,
                                                  double* __restrict__ Mnew, const uint32_t* __restrict__ from,
                                                  uint32_t nact) {
  const uint32_t k = threadIdx.x % kTw;
  const uint32_t ntiles = (nact + kTw - 1) / kTw;
  for (uint64_t i = (uint64_t)blockIdx.x * kRowStep + threadIdx.x / kTw; i < n * ntiles;
       i += (uint64_t)gridDim.x * kRowStep) {
    const uint64_t t = i / n, v = i % n;
    const uint32_t pos = (uint32_t)(t * kTw + k);
    double m = 0.0;
    if (pos < nact) {
      const uint32_t f = from[pos];
      m = Mold[((uint64_t)(f / kTw) * n + v) * kTw + f % kTw];
    }
    Mnew[(t * n + v) * kTw + k] = m;
  }
}

__global__ void ppr_count_live(uint64_t n, const uint8_t* __restrict__ tomb, unsigned long long* live) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    c += !(tomb && tomb[i]);
  for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(live, c);
}

// Work items over the reverse CSR, in row order: runs of whole rows (at
// most kItemNnz entries and kItemRows rows) and, for a row longer than
// kItemNnz, one item per kItemNnz-entry piece. Built once per graph.
int prepare_ppr_plan(wg_context* ctx, wg_graph* g) {
  if (g->ppr_items) return WG_OK;
  const std::vector<uint64_t>& off = g->rev.h_offsets;
  const uint64_t n = g->n;
  std::vector<PprItem> items;
  std::vector<uint32_t> lrows, lbase;
  uint32_t npieces = 0;
  for (uint64_t v = 0; v < n;) {
    const uint64_t deg = off[v + 1] - off[v];
    if (deg > kItemNnz) {
      const uint32_t np = (uint32_t)((deg + kItemNnz - 1) / kItemNnz);
      lrows.push_back((uint32_t)v);
      lbase.push_back(npieces);
      for (uint32_t p = 0; p < np; ++p) {
        const uint64_t p0 = off[v] + (uint64_t)p * kItemNnz;
        items.push_back({p0, (uint32_t)v, 0, npieces + p, (uint32_t)std::min<uint64_t>(kItemNnz, off[v + 1] - p0)});
      }
      npieces += np;
      ++v;
      continue;
    }
    const uint64_t r0 = v;
    uint64_t ent = 0;
    while (v < n && v - r0 < kItemRows) {
      const uint64_t d = off[v + 1] - off[v];
      if (d > kItemNnz || ent + d > kItemNnz) break;
      ent += d;
      ++v;
    }
    items.push_back({off[r0], (uint32_t)r0, (uint32_t)(v - r0), 0, (uint32_t)ent});
  }
  lbase.push_back(npieces);
  void *pi, *pl, *pb;
  int s;
  if ((s = graph_buffer(ctx, g, kSlotPprItems, std::max<size_t>(items.size(), 1) * sizeof(PprItem), &pi)) ||
      (s = graph_buffer(ctx, g, kSlotPprLrows, std::max<size_t>(lrows.size(), 1) * 4, &pl)) ||
      (s = graph_buffer(ctx, g, kSlotPprLbase, lbase.size() * 4, &pb)))
    return s;
  if (!items.empty())
    WG_HIP(ctx, hipMemcpyAsync(pi, items.data(), items.size() * sizeof(PprItem), hipMemcpyHostToDevice, ctx->stream));
  if (!lrows.empty())
    WG_HIP(ctx, hipMemcpyAsync(pl, lrows.data(), lrows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  WG_HIP(ctx, hipMemcpyAsync(pb, lbase.data(), lbase.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  g->ppr_items = pi;
  g->ppr_lrows = (uint32_t*)pl;
  g->ppr_lbase = (uint32_t*)pb;
  g->ppr_nitems = (uint32_t)items.size();
  g->ppr_npieces = npieces;
  g->ppr_nlong = (uint32_t)lrows.size();
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_pagerank_batch(wg_context* ctx, const wg_graph* gc, const wg_vector* tombstones,
                                 const wg_pagerank_opts* o, const uint64_t* seed_offsets,
                                 const uint64_t* seed_nodes, const double* seed_weights, uint32_t R,
                                 wg_vector* ranks, uint32_t* iterations) {
  if (!ctx || !gc || !o || !seed_offsets || !ranks) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n;
  if (R == 0 || R > kPprMaxCols) return fail(ctx, WG_INVALID_ARGUMENT, "between 1 and 256 seed sets");
  if (ranks->dtype != WG_F64 || ranks->id_space != n || ranks->width != R)
    return fail(ctx, WG_INVALID_ARGUMENT, "ranks must be an f64 vector over the graph, one column per seed set");
  if (tombstones && (tombstones->dtype != WG_BOOL || tombstones->id_space != n || tombstones->width != 1))
    return fail(ctx, WG_INVALID_ARGUMENT, "tombstones must be a bool vector over the graph");
  for (uint64_t i = 0; i < seed_offsets[R]; ++i)
    if (seed_nodes[i] >= n)
      return fail(ctx, WG_NODE_OUTSIDE_ID_SPACE, "seed node outside the id space");
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint8_t* tomb = tombstones ? (const uint8_t*)tombstones->data : nullptr;
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
    int s = dalloc(ctx, p, std::max<size_t>(bytes, 8));
    if (!s) held.push_back({*p, std::max<size_t>(bytes, 8)});
    return s;
  };
  int s;
  // Live rows and 1 / live (mod.rs:104-107).
  unsigned long long* live_d;
  if ((s = get((void**)&live_d, 8))) return s;
  WG_HIP(ctx, hipMemsetAsync(live_d, 0, 8, st));
  ppr_count_live<<<1024, 256, 0, st>>>(n, tomb, live_d);
  unsigned long long live = 0;
  WG_HIP(ctx, hipMemcpyAsync(&live, live_d, 8, hipMemcpyDeviceToHost, st));
  std::vector<uint8_t> htomb(tombstones ? n : 0);
  if (tombstones && n) WG_HIP(ctx, hipMemcpyAsync(htomb.data(), tomb, n, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  uint32_t herr = 0;
  const double nodes = (double)live;
  const double start = sqlf<DIV_>(1.0, nodes, herr);
  if (herr) return fail(ctx, (int)herr, "pagerank: 1 / live nodes: division by zero");
  // Each column's shares seed / sum(seed) over live rows (mod.rs:155-199):
  // the dense seed vector is the weights summed per node (in input order),
  // its sum taken in ascending rows by 65536-row chunks as reduce_live does.
  std::vector<uint32_t> seed_row(n, ~0u);
  std::vector<uint32_t> rows_of;  // seed row index -> node
  std::vector<std::vector<std::pair<uint64_t, double>>> cols(R);
  for (uint32_t c = 0; c < R; ++c) {
    std::vector<std::pair<uint64_t, double>> e;
    for (uint64_t i = seed_offsets[c]; i < seed_offsets[c + 1]; ++i)
      e.push_back({seed_nodes[i], seed_weights ? seed_weights[i] : 1.0});
    std::stable_sort(e.begin(), e.end(), [](auto& x, auto& y) { return x.first < y.first; });
    std::vector<std::pair<uint64_t, double>> dense;
    for (auto& p : e) {
      if (!dense.empty() && dense.back().first == p.first) dense.back().second += p.second;
      else dense.push_back(p);
    }
    double sum = 0.0, chunk = 0.0;
    uint64_t cur = ~0ull;
    for (auto& p : dense) {
      if (tombstones && htomb[p.first]) continue;
      const uint64_t ch = p.first >> 16;
      if (ch != cur) {
        sum += chunk;
        chunk = 0.0;
        cur = ch;
      }
      chunk += p.second;
    }
    sum += chunk;
    const double total = sqlf<DIV_>(1.0, sum, herr);
    if (herr) return fail(ctx, (int)herr, "pagerank: 1 / sum(seed) of seed set " + std::to_string(c) + ": division by zero");
    for (auto& p : dense) {
      if (tombstones && htomb[p.first]) continue;
      const double share = sqlf<MUL_>(p.second, total, herr);
      if (herr) return fail(ctx, (int)herr, "pagerank: seed share out of range");
      if (seed_row[p.first] == ~0u) {
        seed_row[p.first] = (uint32_t)rows_of.size();
        rows_of.push_back((uint32_t)p.first);
      }
      cols[c].push_back({seed_row[p.first], share});
    }
  }
  std::vector<double> P(std::max<size_t>(rows_of.size(), 1) * R, 0.0);
  for (uint32_t c = 0; c < R; ++c)
    for (auto& p : cols[c]) P[p.first * R + c] = p.second;
  const double teleport = sqlf<SUB_>(1.0, o->damping, herr);
  if (herr) return fail(ctx, (int)herr, "pagerank: 1 - damping");
  uint32_t bound_err = 0;
  const double bound = o->has_tol ? sqlf<MUL_>(nodes, o->tol, bound_err) : 0.0;
  const bool weighted = o->weighted < 0 ? g->weighted : o->weighted != 0;
  DevCsr& rev = g->rev;
  if ((s = prepare_ppr_plan(ctx, g)) || (s = prepare_hot_order(ctx, g))) return s;
  const uint32_t nitems = g->ppr_nitems, npieces = g->ppr_npieces, nlong = g->ppr_nlong;
  const PprItem* items = (const PprItem*)g->ppr_items;
  // One persistent workgroup per resident slot.
  int per_cu = 0, cus = 0;
  WG_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                  &per_cu, weighted ? (const void*)ppr_block<true> : (const void*)ppr_block<false>, 256, 0));
  WG_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(std::max(per_cu, 1) * std::max(cus, 1),
                                                                 std::max<uint32_t>(nitems, 1)));
  const uint32_t ntiles = (R + kTw - 1) / kTw;
  const uint32_t long_blocks = (nlong + kG - 1) / kG;
  const uint32_t nparts = grid + long_blocks;
  const uint64_t nm = g->nhot;
  const size_t tiles_bytes = (size_t)ntiles * std::max<uint64_t>(nm, 1) * kTw * 8;
  double *X = (double*)ranks->data, *M0, *M1, *out, *part, *piece, *Pd, *dangling;
  uint32_t *seed_row_d, *cols_d, *from_d, *iters_d;
  unsigned* stop_d;
  uint8_t* active_d;
  double* y0 = nullptr;
  if ((s = get((void**)&M0, tiles_bytes)) || (s = get((void**)&M1, tiles_bytes)) ||
      (s = get((void**)&out, std::max<uint64_t>(n, 1) * 8)) ||
      (s = get((void**)&part, (size_t)ntiles * nparts * 2 * kTw * 8)) ||
      (s = get((void**)&piece, (size_t)ntiles * std::max<uint32_t>(npieces, 1) * kTw * 8)) ||
      (s = get((void**)&Pd, P.size() * 8)) || (s = get((void**)&dangling, R * 8)) ||
      (s = get((void**)&seed_row_d, std::max<uint64_t>(n, 1) * 4)) ||
      (s = get((void**)&cols_d, R * 4)) || (s = get((void**)&from_d, R * 4)) ||
      (s = get((void**)&iters_d, R * 4)) || (s = get((void**)&active_d, R)) ||
      (s = get((void**)&stop_d, 4)))
    return s;
  // Each row's first-iteration sum (tile 0's), which the other tiles take.
  if (ntiles > 1 && (s = get((void**)&y0, std::max<uint64_t>(n, 1) * 8))) return s;
  WG_HIP(ctx, hipMemsetAsync(stop_d, 0, 4, st));
  WG_HIP(ctx, hipMemcpyAsync(Pd, P.data(), P.size() * 8, hipMemcpyHostToDevice, st));
  if (n) WG_HIP(ctx, hipMemcpyAsync(seed_row_d, seed_row.data(), n * 4, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemsetAsync(iters_d, 0, R * 4, st));
  WG_HIP(ctx, hipMemsetAsync(active_d, 1, R, st));
  if ((s = clear_error(ctx))) return s;
  // out = PlusFirst(A) . 1 (the values do not matter to (x) first).
  if ((s = launch_out_weights(ctx, g, weighted, out, out))) return s;
  // Slots other than ppr_start's hold zeros for the set-up sums.
  WG_HIP(ctx, hipMemsetAsync(part, 0, (size_t)ntiles * nparts * 2 * kTw * 8, st));
  ppr_start<<<grid, 256, 0, st>>>(n, ntiles, start, tomb, g->hot_index, out, M0, part, nparts, ctx->d_err);
  std::vector<uint32_t> active_cols(R);
  for (uint32_t c = 0; c < R; ++c) active_cols[c] = c;
  WG_HIP(ctx, hipMemcpyAsync(cols_d, active_cols.data(), R * 4, hipMemcpyHostToDevice, st));
  // The first dangling ranks, from the start partials.
  ppr_finish<<<R, 256, 0, st>>>(part, nparts, cols_d, dangling, iters_d, active_d, 0.0, 0, 0);
  if ((s = collect_error(ctx, "pagerank batch"))) return s;
  const double* wv = weighted && rev.weights ? rev.weights : nullptr;
  const uint32_t max_iter = bound_err ? 1 : o->max_iter;
  if (max_iter == 0 && n) ppr_fill_start<<<2048, 256, 0, st>>>(n, R, start, tomb, X);
  std::vector<uint8_t> act(R, 1);
  double *Min = M0, *Mout = M1;
  // An iteration's launches (every tile's persistent pass and split rows,
  // then the columns' finish) are captured once as a HIP graph per message
  // parity and replayed while the active columns stay the same; a repack
  // captures them again. Kernel launches only: nothing in them is a memset.
  // The first iteration (tile 0's pass and the other tiles from its sums)
  // and the second run eagerly; from the third on the launches do not
  // change while the active columns stay the same.
  struct Captured {
    hipGraphExec_t exec = nullptr;
    const double* min = nullptr;
    uint32_t nact = 0;
  } cap[2];
  auto drop = [&]() {
    for (auto& c : cap)
      if (c.exec) {
        (void)hipGraphExecDestroy(c.exec);
        c = Captured{};
      }
  };
  struct Dropper {
    std::function<void()> f;
    ~Dropper() { f(); }
  } dropper{drop};
  for (uint32_t it = 0; it < max_iter && !active_cols.empty(); ++it) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    const uint32_t nact = (uint32_t)active_cols.size();
    const uint32_t tiles = (nact + kTw - 1) / kTw;
    PprArgs a{rev.offsets, g->hot_src, wv, n, g->hot_index, nm, R, nact, cols_d, out, tomb, seed_row_d, Pd,
              o->damping, teleport, dangling, X, Min, Mout, part, piece, nparts, ctx->d_err,
              ctx->d_mail + kMailCancel, stop_d, start, std::min<uint32_t>(it, 2), it == 0 ? y0 : nullptr};
    auto enqueue = [&]() {
      for (uint32_t t = 0; t < (it == 0 ? 1 : tiles); ++t) {
        if (wv) ppr_block<true><<<grid, 256, 0, st>>>(a, t, items, nitems, npieces);
        else ppr_block<false><<<grid, 256, 0, st>>>(a, t, items, nitems, npieces);
        if (long_blocks)
          ppr_long<<<long_blocks, 256, 0, st>>>(a, t, g->ppr_lrows, g->ppr_lbase, nlong, npieces, grid);
      }
      if (it == 0 && tiles > 1) ppr_first_rest<<<dim3(nparts, tiles - 1), 256, 0, st>>>(a);
      ppr_finish<<<nact, 256, 0, st>>>(part, nparts, cols_d, dangling, iters_d, active_d, bound,
                                       o->has_tol && !bound_err, 1);
    };
    if (ctx->graph_launch && it >= 2) {
      Captured& c = cap[Min == M0 ? 0 : 1];
      if (!c.exec || c.min != Min || c.nact != nact) {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        c = Captured{};
        hipGraph_t graph = nullptr;
        WG_HIP(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue();
        const hipError_t ce = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess) return hip_fail(ctx, ce, "hipStreamEndCapture");
        const hipError_t ie = hipGraphInstantiate(&c.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return hip_fail(ctx, ie, "hipGraphInstantiate");
        c.min = Min;
        c.nact = nact;
      }
      WG_HIP(ctx, hipGraphLaunch(c.exec, st));
    } else {
      enqueue();
    }
    WG_HIP(ctx, hipGetLastError());
    std::swap(Min, Mout);
    WG_HIP(ctx, hipMemcpyAsync(act.data(), active_d, R, hipMemcpyDeviceToHost, st));
    if ((s = collect_error(ctx, "pagerank batch"))) return s;
    // A pass the cancel word cut short is never taken as an iteration.
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    // Survivors keep their relative order; the messages are repacked.
    std::vector<uint32_t> next_cols, from;
    for (uint32_t p = 0; p < nact; ++p)
      if (act[active_cols[p]]) {
        next_cols.push_back(active_cols[p]);
        from.push_back(p);
      }
    if (next_cols.size() != nact && !next_cols.empty()) {
      WG_HIP(ctx, hipMemcpyAsync(from_d, from.data(), from.size() * 4, hipMemcpyHostToDevice, st));
      ppr_repack<<<2048, 256, 0, st>>>(nm, Min, Mout, from_d, (uint32_t)next_cols.size());
      std::swap(Min, Mout);
      WG_HIP(ctx, hipMemcpyAsync(cols_d, next_cols.data(), next_cols.size() * 4, hipMemcpyHostToDevice, st));
    }
    active_cols.swap(next_cols);
  }
  if (iterations) WG_HIP(ctx, hipMemcpy(iterations, iters_d, R * 4, hipMemcpyDeviceToHost));
  if (bound_err) return fail(ctx, (int)bound_err, "pagerank: live nodes * tol out of range");
  return ok(ctx);
}
