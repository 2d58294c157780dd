// Semiring SpMV and SpMM on gfx950 (compute/product/{mod,kernel}.rs).
//
// SpMV: one 256-thread workgroup per row block. A block stages the products
// of its (at most 2048) entries in LDS after coalesced loads of the target
// and weight streams and a gather of x, then folds each row: short blocks
// give each row a power-of-two group of lanes, each lane folding a
// contiguous run of the row and the group combining the runs in order with
// DPP/bpermute shuffles; rows longer than a block are split into pieces whose
// partials the last piece to finish (an arrival ticket) folds in piece order.
// Folding contiguous runs in order keeps the reference's left-to-right
// result exactly for every monoid but float plus.
//
// SpMM: a group of lanes per row, lanes over the columns of the row-major X,
// each lane folding its columns over the row's entries in ascending order
// (the reference's fold order); long rows split as above.
#include <algorithm>

#include "semiring.hpp"
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

template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(kBlockThreads) void spmv_blocks(
    const Block* __restrict__ blocks, const uint64_t* __restrict__ off,
    const uint32_t* __restrict__ tgt, const E* __restrict__ w, const E* __restrict__ x,
    E* __restrict__ y, WriteArgs wa, typename Fold<E, ADD>::A* __restrict__ partials,
    uint32_t* __restrict__ tickets, unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  __shared__ E prod[kBlockNnz];
  __shared__ uint32_t rs[kBlockRows + 1];
  __shared__ A wpart[kBlockThreads / 64];
  __shared__ int wpres[kBlockThreads / 64];

  const Block b = blocks[blockIdx.x];
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

  // Stage the products: every load of the block issued before any use.
  {
    uint32_t col[kBlockNnz / kBlockThreads];
    E wv[kBlockNnz / kBlockThreads];
#pragma unroll
    for (int k = 0; k < (int)(kBlockNnz / kBlockThreads); ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      if (i < cnt) {
        if constexpr (reads_x<MUL>()) col[k] = tgt[e0 + i];
        if constexpr (HW) wv[k] = w[e0 + i];
      }
    }
#pragma unroll
    for (int k = 0; k < (int)(kBlockNnz / kBlockThreads); ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      if (i < cnt) {
        E wk = E(1);
        if constexpr (HW) wk = wv[k];
        E xv = E();
        if constexpr (reads_x<MUL>()) xv = x[col[k]];
        prod[i] = multiply<E, ADD, MUL>(wk, xv, e_err);
      }
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
      const uint64_t row = b.row0;
      if (b.npieces == 1) {
        E res = tp ? F::finish(tot, e_err) : E();
        write_elem<E, ADD>(wa, row, &y[row], tp != 0, res, e_err);
      } else {
        partials[b.pbase + b.piece] = tot;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t prev = atomicAdd(&tickets[b.ticket], 1u);
        if (prev == b.npieces - 1) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          A all = load_coherent(&partials[b.pbase]);
          for (uint32_t p = 1; p < b.npieces; ++p)
            all = F::combine(all, load_coherent(&partials[b.pbase + p]), e_err);
          tickets[b.ticket] = 0;
          E res = F::finish(all, e_err);
          write_elem<E, ADD>(wa, row, &y[row], true, res, e_err);
        }
      }
    }
  }
  if (e_err) raise_error(err, b.row0, e_err);
}

// SpMM: LPR lanes per row task, each lane folding columns lane, lane + LPR,
// ... of its row over the task's entries.
template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(256) void spmm_rows(
    const uint64_t* __restrict__ off, const uint32_t* __restrict__ tgt, const E* __restrict__ w,
    const E* __restrict__ x, E* __restrict__ y, uint64_t width, uint64_t n, uint32_t lpr,
    const uint32_t* __restrict__ extra, uint32_t nextra, const uint32_t* __restrict__ ticket_of_row,
    const uint32_t* __restrict__ pbase, uint32_t* __restrict__ tickets,
    typename Fold<E, ADD>::A* __restrict__ partials, WriteArgs wa,
    unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  const uint32_t tasks_per_block = 256 / lpr;
  const uint64_t task = (uint64_t)blockIdx.x * tasks_per_block + threadIdx.x / lpr;
  const uint32_t lane = threadIdx.x % lpr;
  if (task >= n + nextra) return;
  uint64_t row;
  uint32_t piece;
  if (task < n) {
    row = task;
    piece = 0;
  } else {
    row = extra[2 * (task - n)];
    piece = extra[2 * (task - n) + 1];
  }
  const uint64_t rb = off[row], re = off[row + 1];
  const uint64_t deg = re - rb;
  const uint32_t np = deg > kSpmmPiece ? (uint32_t)((deg + kSpmmPiece - 1) / kSpmmPiece) : 1;
  const uint64_t e0 = rb + (uint64_t)piece * kSpmmPiece;
  const uint64_t e1 = e0 + kSpmmPiece < re ? e0 + kSpmmPiece : re;
  uint32_t e_err = 0;
  const uint32_t tk = np > 1 ? ticket_of_row[row] : 0;
  for (uint64_t c = lane; c < width; c += lpr) {
    A acc{};
    int pres = 0;
    for (uint64_t e = e0; e < e1; ++e) {
      E wk = E(1);
      if constexpr (HW) wk = w[e];
      E xv = E();
      if constexpr (reads_x<MUL>()) xv = x[(uint64_t)tgt[e] * width + c];
      A v = F::lift(multiply<E, ADD, MUL>(wk, xv, e_err));
      acc = pres ? F::combine(acc, v, e_err) : v;
      pres = 1;
    }
    if (np == 1) {
      E res = pres ? F::finish(acc, e_err) : E();
      write_elem<E, ADD>(wa, row, &y[row * width + c], pres != 0, res, e_err);
    } else {
      partials[(uint64_t)(pbase[tk] + piece) * width + c] = acc;
    }
  }
  if (np > 1) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    // Every lane of the task has stored its columns; lane 0 takes the ticket
    // once they all have (the group's lanes are in one wave).
    __builtin_amdgcn_wave_barrier();
    uint32_t prev = 0;
    if (lane == 0) prev = atomicAdd(&tickets[tk], 1u);
    prev = __shfl(prev, (int)(threadIdx.x & 63) - (int)lane, 64);
    if (prev == np - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      const uint64_t base = pbase[tk];
      for (uint64_t c = lane; c < width; c += lpr) {
        A all = load_coherent(&partials[base * width + c]);
        for (uint32_t p = 1; p < np; ++p)
          all = F::combine(all, load_coherent(&partials[(base + p) * width + c]), e_err);
        E res = F::finish(all, e_err);
        write_elem<E, ADD>(wa, row, &y[row * width + c], true, res, e_err);
      }
      if (lane == 0) tickets[tk] = 0;
    }
  }
  if (e_err) raise_error(err, row, e_err);
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
    int s;
    if (width == 1) {
      if ((s = ensure_scratch(ctx, (size_t)std::max<uint32_t>(csr.npartials, 1) * sizeof(A)))) return s;
      if (csr.nblocks == 0) return WG_OK;
      A* part = (A*)ctx->scratch;
      if (wv)
        spmv_blocks<E, ADD, MUL, true><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
            csr.blocks, csr.offsets, csr.targets, wv, x, y, wa, part, csr.tickets, ctx->d_err);
      else
        spmv_blocks<E, ADD, MUL, false><<<csr.nblocks, kBlockThreads, 0, ctx->stream>>>(
            csr.blocks, csr.offsets, csr.targets, nullptr, x, y, wa, part, csr.tickets, ctx->d_err);
    } else {
      if ((s = prepare_spmm(ctx, csr))) return s;
      if ((s = ensure_scratch(ctx, (size_t)std::max<uint32_t>(csr.spmm_npartials, 1) * width * sizeof(A))))
        return s;
      uint32_t lpr = 1;
      while (lpr < width && lpr < 64) lpr <<= 1;
      const uint64_t tasks = csr.n + csr.spmm_nextra;
      const uint64_t grid = (tasks + (256 / lpr) - 1) / (256 / lpr);
      if (grid == 0) return WG_OK;
      A* part = (A*)ctx->scratch;
      if (wv)
        spmm_rows<E, ADD, MUL, true><<<(unsigned)grid, 256, 0, ctx->stream>>>(
            csr.offsets, csr.targets, wv, x, y, width, csr.n, lpr, csr.spmm_extra, csr.spmm_nextra,
            csr.spmm_ticket_of_row, csr.spmm_pbase, csr.spmm_tickets, part, wa, ctx->d_err);
      else
        spmm_rows<E, ADD, MUL, false><<<(unsigned)grid, 256, 0, ctx->stream>>>(
            csr.offsets, csr.targets, nullptr, x, y, width, csr.n, lpr, csr.spmm_extra,
            csr.spmm_nextra, csr.spmm_ticket_of_row, csr.spmm_pbase, csr.spmm_tickets, part, wa,
            ctx->d_err);
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

// The checks of spmv / spmm / multiply (compute/product/mod.rs:77-153).
static int product(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                   wg_vector* y, bool spmv) {
  if (!ctx || !g || !p || !x || !y) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
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
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_graph* gm = const_cast<wg_graph*>(g);
  DevCsr& csr = p->transpose ? gm->rev : gm->fwd;
  const bool weighted = p->weighted < 0 ? g->weighted : p->weighted != 0;
  WriteArgs wa{p->mask ? p->mask->data : nullptr, p->mask ? p->mask->dtype : 0,
               p->mask_complement, p->replace, p->accumulate};
  const bool can_fail = x->dtype == WG_I64 || x->dtype == WG_U32;
  int s;
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

}  // namespace wg

extern "C" {

int wg_spmv(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y) {
  return wg::product(ctx, g, p, x, y, true);
}

int wg_spmm(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y) {
  return wg::product(ctx, g, p, x, y, false);
}

}  // extern "C"
