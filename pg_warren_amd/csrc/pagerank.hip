// PageRank by power iteration on gfx950 (compute/pagerank/mod.rs:96-287).
//
// One fused kernel per iteration over the reverse CSR's row blocks:
//   gather   y_i = sum_j w_ji * msg_j       (f64, the SpMV block scheme)
//   advance  x_i' = d (y_i + dangling p_i) + (1 - d) p_i, SQL float8 checks
//   spread   msg_i' = x_i' / out_i, or 0 and x_i' joins the next dangling sum
// Block partials of the L1 change and of the next dangling rank are summed
// in a fixed order by a one-workgroup pr_finish launch, which also takes the
// stop decision (L1 change < live nodes * tol, in SQL order) into a device
// flag; later launches of the batch see the flag and return at once, so the
// host enqueues iterations in batches and reads the flag once per batch.
// The file ends with the adjoint (wg_pagerank_backward).
#include <algorithm>
#include <cmath>

#include "semiring.hpp"
#include "spmv_tiles.hpp"
#include "wg_internal.hpp"

namespace wg {

struct PrState {
  double dangling[2];  // dangling rank read by the iteration of each parity
  double delta;
  uint32_t stop;
  uint32_t iterations;
  uint32_t arrivals;
  uint32_t pad;
};

struct PrArgs {
  const Block* blocks;
  const uint64_t* off;
  const uint32_t* src;
  const double* w;
  const double* out;    // weighted out-degree (PlusFirst over the forward CSR)
  const double* p;      // personalisation, or null: uniform p_uniform
  double p_uniform;
  const uint8_t* tomb;  // tombstoned rows, or null
  double damping, teleport, bound;
  int has_tol;
  uint32_t nblocks, nlong;
  double* x;
  double* part;        // 2 * (nblocks + nlong): change, then dangling
  double* piece;       // split rows' gather partials
  PrState* st;
  unsigned long long* err;
};

__device__ __forceinline__ bool pr_halted(const PrArgs& a) {
  return __hip_atomic_load(&a.st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
         __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != ~0ull;
}

// advance + spread of one row; returns its L1 change and adds its dangling
// rank to `dang`.
__device__ __forceinline__ double pr_row(const PrArgs& a, uint64_t row, double y, double dangling,
                                         double* msg_out, double& dang, uint32_t& err) {
  double xo = a.x[row], next = 0.0;
  if (!(a.tomb && a.tomb[row])) {
    const double p = a.p ? a.p[row] : a.p_uniform;
    double flowed = sqlf<ADD_>(y, sqlf<MUL_>(dangling, p, err), err);
    next = sqlf<ADD_>(sqlf<MUL_>(a.damping, flowed, err), sqlf<MUL_>(a.teleport, p, err), err);
  }
  a.x[row] = next;
  const double o = a.out[row];
  if (sql_lt(0.0, o)) {
    msg_out[row] = sqlf<DIV_>(next, o, err);
  } else {
    dang += next;
    msg_out[row] = 0.0;
  }
  return fabs(next - xo);
}

// One workgroup after each iteration: fixed-order sums of the partials
// (the kernel boundary makes every block's stores visible, with no
// cross-workgroup fence), and the stop decision.
__global__ __launch_bounds__(256) void pr_finish(PrArgs a, int parity) {
  __shared__ double sc[256], sd[256];
  if (pr_halted(a)) return;
  const uint32_t tid = threadIdx.x, total = a.nblocks + a.nlong;
  double c = 0.0, d = 0.0;
  for (uint32_t i = tid; i < total; i += 256) {
    c += a.part[i];
    d += a.part[total + i];
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
    a.st->delta = sc[0];
    a.st->dangling[parity ^ 1] = sd[0];
    a.st->iterations += 1;
    if (a.has_tol && sql_lt(sc[0], a.bound)) a.st->stop = 1;
  }
}

// The split rows of an iteration: each folds its pieces' gather partials in
// order, then advances; its L1 change and dangling rank take a slot of
// their own, so the fixed-order total does not depend on scheduling.
__global__ __launch_bounds__(256) void pr_long(PrArgs a, const uint32_t* __restrict__ long_rows,
                                               const uint32_t* __restrict__ long_pbase,
                                               double* __restrict__ msg_out, int parity) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= a.nlong || pr_halted(a)) return;
  uint32_t e_err = 0;
  double y = a.piece[long_pbase[i]];
  for (uint32_t p = long_pbase[i] + 1; p < long_pbase[i + 1]; ++p) y += a.piece[p];
  double dl = 0.0;
  const double ch = pr_row(a, long_rows[i], y, a.st->dangling[parity], msg_out, dl, e_err);
  a.part[a.nblocks + i] = ch;
  a.part[2 * a.nblocks + a.nlong + i] = dl;
  if (e_err) raise_error(a.err, long_rows[i], e_err);
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const uint32_t tid = threadIdx.x;
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (tid == 0)
    for (int k = 0; k < 4; ++k) t += sh[k];
  __syncthreads();
  return t;
}

template <bool HW>
__global__ __launch_bounds__(kBlockThreads) void pr_iterate(PrArgs a, const double* __restrict__ msg_in,
                                                            double* __restrict__ msg_out, int parity) {
  __shared__ double prod[kBlockNnz];
  __shared__ uint32_t rs[kBlockRows + 1];
  __shared__ double wsum[8];
  if (pr_halted(a)) return;
  const double dangling = a.st->dangling[parity];
  const Block b = a.blocks[blockIdx.x];
  const uint32_t tid = threadIdx.x;
  uint32_t e_err = 0;
  uint64_t e0, e1;
  if (b.nrows) {
    e0 = a.off[b.row0];
    e1 = a.off[b.row0 + b.nrows];
  } else {
    const uint64_t rb = a.off[b.row0], re = a.off[b.row0 + 1];
    e0 = rb + (uint64_t)b.piece * kBlockNnz;
    e1 = e0 + kBlockNnz < re ? e0 + kBlockNnz : re;
  }
  const uint32_t cnt = (uint32_t)(e1 - e0);
  {
    // All the block's message gathers issued before the first use: a lane
    // past the entries reads message 0 (which exists) and drops it.
    constexpr int kPer = (int)(kBlockNnz / kBlockThreads);
    uint32_t col[kPer];
    double wv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      col[k] = 0;
      wv[k] = 1.0;
      if (i < cnt) {
        col[k] = a.src[e0 + i];
        if (HW) wv[k] = a.w[e0 + i];
      }
    }
    double mv[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) mv[k] = msg_in[col[k]];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const uint32_t i = tid + k * kBlockThreads;
      if (i < cnt) prod[i] = HW ? wv[k] * mv[k] : mv[k];
    }
  }
  if (b.nrows)
    for (uint32_t k = tid; k <= b.nrows; k += kBlockThreads) rs[k] = (uint32_t)(a.off[b.row0 + k] - e0);
  __syncthreads();

  double change = 0.0, dang = 0.0;
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
      const uint32_t s0 = rs[r], s1 = rs[r + 1];
      const uint32_t len = s1 - s0, chunk = (len + T - 1) / T;
      uint32_t s = s0 + l * chunk, t = s + chunk < s1 ? s + chunk : s1;
      double acc = 0.0;
      for (uint32_t i = s; i < t; ++i) acc += prod[i];
      for (uint32_t d = 1; d < T; d <<= 1) {
        double o = __shfl_down(acc, d, T);
        if ((l & (2 * d - 1)) == 0) acc += o;
      }
      if (l == 0) change += pr_row(a, (uint64_t)b.row0 + r, acc, dangling, msg_out, dang, e_err);
    }
    double c = block_sum(change, wsum);
    double dd = block_sum(dang, wsum + 4);
    if (tid == 0) {
      a.part[blockIdx.x] = c;
      a.part[a.nblocks + a.nlong + blockIdx.x] = dd;
    }
  } else {
    const uint32_t chunk = (cnt + kBlockThreads - 1) / kBlockThreads;
    uint32_t s = tid * chunk, t = s + chunk < cnt ? s + chunk : cnt;
    double acc = 0.0;
    for (uint32_t i = s; i < t; ++i) acc += prod[i];
    acc = block_sum(acc, wsum);
    // A piece of a split row: its gather partial waits for pr_long.
    if (tid == 0) {
      a.part[blockIdx.x] = 0.0;
      a.part[a.nblocks + a.nlong + blockIdx.x] = 0.0;
      a.piece[b.pbase + b.piece] = acc;
    }
  }
  if (e_err) raise_error(a.err, b.row0, e_err);
}

// ---- the gather through the tiled SpMV (large graphs) ----------------------
// On graphs of kPrTileMinNnz entries or more an iteration is: the hot and
// warm message copies (pr_tiers), the tiled fold of each row's messages
// (spmv_tiles.hpp, the stop flag as its halt word), then pr_tile_rows:
// advance and spread of every row from its folded value (the rows a wave
// leaves open joined from the wave records), with fixed-order block partials of the L1 change and of the
// dangling rank, summed by pr_reduce (a block per span of partials) and
// pr_finish_sum (one workgroup, the stop decision).
constexpr uint64_t kPrTileMinNnz = 1ull << 22;
constexpr uint32_t kPrReduceSpan = 256;

__global__ __launch_bounds__(256) void pr_tiers(const double* __restrict__ msg, const uint32_t* __restrict__ inv,
                                                uint32_t nh, double* __restrict__ xh,
                                                const uint32_t* __restrict__ warm, uint32_t nw,
                                                double* __restrict__ xw, const PrState* __restrict__ st) {
  if (__hip_atomic_load(&st->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // Four messages a thread, a grid apart, their gathers in flight together.
  const uint64_t q = (uint64_t)blockIdx.x * 256 + threadIdx.x, T = (uint64_t)gridDim.x * 256;
  const uint64_t total = (uint64_t)nh + nw;
  uint32_t id[4];
  double v[4];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint64_t i = q + k * T;
    id[k] = i < nh ? inv[i] : i < total ? warm[i - nh] : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) v[k] = msg[id[k]];
#pragma unroll
  for (uint32_t k = 0; k < 4; ++k) {
    const uint64_t i = q + k * T;
    if (i < nh)
      xh[i] = v[k];
    else if (i < total)
      xw[i - nh] = v[k];
  }
}

// Every row: its folded in-messages from the row's slot (0 for a row with no
// in-edges, the empty sum), then advance and spread.
__global__ __launch_bounds__(256) void pr_tile_rows(PrArgs a, const TileGeom g, const double* __restrict__ ycomp,
                                                    const TileRec<double>* __restrict__ recs, uint64_t n,
                                                    double* __restrict__ msg_out, double* __restrict__ bpart,
                                                    uint32_t nbp, int parity) {
  __shared__ double wsum[8];
  if (pr_halted(a)) return;
  const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  double change = 0.0, dang = 0.0;
  uint32_t e_err = 0;
  if (r < n) {
    const uint64_t wd = g.nz_bits[r >> 6];
    const uint32_t b = (uint32_t)(r & 63);
    double y = 0.0;
    if ((wd >> b) & 1)
      y = tile_row_value<double, PLUS>(g, r, g.nz_pre[r >> 6] + __popcll(wd & ((1ull << b) - 1)),
                                       (g.fix_bits[r >> 6] >> b) & 1, ycomp, recs);
    change = pr_row(a, r, y, a.st->dangling[parity], msg_out, dang, e_err);
  }
  const double c = block_sum(change, wsum);
  const double dd = block_sum(dang, wsum + 4);
  if (threadIdx.x == 0) {
    bpart[blockIdx.x] = c;
    bpart[nbp + blockIdx.x] = dd;
  }
  if (e_err) raise_error(a.err, r, e_err);
}

// Block k sums partials [k * span, (k + 1) * span) of both halves, in a
// fixed order (thread-strided, then a tree).
__global__ __launch_bounds__(256) void pr_reduce(const PrArgs a, const double* __restrict__ bpart, uint32_t nbp,
                                                 uint32_t span, double* __restrict__ out, uint32_t nout) {
  __shared__ double sc[256], sd[256];
  if (pr_halted(a)) return;
  const uint32_t tid = threadIdx.x, lo = blockIdx.x * span, hi = min(lo + span, nbp);
  double c = 0.0, d = 0.0;
  for (uint32_t i = lo + tid; i < hi; i += 256) {
    c += bpart[i];
    d += bpart[nbp + i];
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
    out[blockIdx.x] = sc[0];
    out[nout + blockIdx.x] = sd[0];
  }
}

// The totals of pr_reduce's outputs (fixed order) and the stop decision.
__global__ __launch_bounds__(256) void pr_finish_sum(PrArgs a, const double* __restrict__ red, uint32_t nred,
                                                     int parity) {
  __shared__ double sc[256], sd[256];
  if (pr_halted(a)) return;
  const uint32_t tid = threadIdx.x;
  double c = 0.0, d = 0.0;
  for (uint32_t i = tid; i < nred; i += 256) {
    c += red[i];
    d += red[nred + i];
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
    a.st->delta = sc[0];
    a.st->dangling[parity ^ 1] = sd[0];
    a.st->iterations += 1;
    if (a.has_tol && sql_lt(sc[0], a.bound)) a.st->stop = 1;
  }
}

// msg = x / out (or 0, the rank joining the dangling sum) for the first
// iteration, and the live-row count / seed sums of the set-up.
__global__ __launch_bounds__(256) void pr_spread0(const double* __restrict__ x,
                                                  const double* __restrict__ out,
                                                  double* __restrict__ msg, uint64_t n,
                                                  double* __restrict__ part,
                                                  unsigned long long* err) {
  __shared__ double wsum[4];
  uint32_t e_err = 0;
  double dang = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double o = out[i];
    if (sql_lt(0.0, o)) {
      msg[i] = sqlf<DIV_>(x[i], o, e_err);
    } else {
      dang += x[i];
      msg[i] = 0.0;
    }
    if (e_err) {
      raise_error(err, i, e_err);
      e_err = 0;
    }
  }
  double t = block_sum(dang, wsum);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Fills x with `start` at live rows and 0 at tombstoned ones, and counts the
// live rows.
__global__ __launch_bounds__(256) void pr_init(double* __restrict__ x, const uint8_t* __restrict__ tomb,
                                               uint64_t n, double start, unsigned long long* live) {
  uint64_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const bool dead = tomb && tomb[i];
    x[i] = dead ? 0.0 : start;
    c += !dead;
  }
  for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(live, (unsigned long long)c);
}

template <class S>
__global__ __launch_bounds__(256) void pr_seed_sum(const S* __restrict__ seed, const uint8_t* __restrict__ tomb,
                                                   uint64_t n, double* __restrict__ part) {
  __shared__ double wsum[4];
  double s = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    if (!(tomb && tomb[i])) s += (double)seed[i];
  double t = block_sum(s, wsum);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <class S>
__global__ __launch_bounds__(256) void pr_shares(const S* __restrict__ seed, const uint8_t* __restrict__ tomb,
                                                 uint64_t n, double total, double* __restrict__ p,
                                                 unsigned long long* err) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    uint32_t e = 0;
    p[i] = (tomb && tomb[i]) ? 0.0 : sqlf<MUL_>((double)seed[i], total, e);
    if (e) raise_error(err, i, e);
  }
}

__global__ void pr_set_state(PrState* st, double dangling0) {
  st->dangling[0] = dangling0;
  st->dangling[1] = 0.0;
  st->delta = 0.0;
  st->stop = 0;
  st->iterations = 0;
  st->arrivals = 0;
}

constexpr unsigned kReduceGrid = 1024;

// PageRank's buffers, kept with the graph between calls (one slot each), so
// a run allocates nothing once the graph has been ranked before.
struct PrBuffers {
  wg_context* ctx;
  wg_graph* g;
  int next_slot = 100;
  PrBuffers(wg_context* c, wg_graph* gr) : ctx(c), g(gr) {}
  template <class T>
  int get(T** p, size_t count) {
    return get_at(p, count, next_slot++);
  }
  template <class T>
  int get_at(T** p, size_t count, int slot) {
    void* q = nullptr;
    int s = graph_buffer(ctx, g, slot, count * sizeof(T), &q);
    if (s) return s;
    *p = (T*)q;
    return WG_OK;
  }
};

int launch_out_weights(wg_context* ctx, wg_graph* g, bool weighted, const double* x, double* out);

static double host_sum(const std::vector<double>& v) {
  double s = 0.0;
  for (double d : v) s += d;
  return s;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_pagerank(wg_context* ctx, const wg_graph* gc, const wg_vector* tombstones,
                           const wg_pagerank_opts* o, const wg_vector* seed, wg_vector* result,
                           uint32_t* iterations) {
  if (!ctx || !gc || !o || !result) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n;
  if (result->dtype != WG_F64 || result->width != 1 || result->id_space != n)
    return fail(ctx, WG_INVALID_ARGUMENT, "the ranks vector must be f64, width 1, over the graph");
  if (tombstones && (tombstones->dtype != WG_BOOL || tombstones->id_space != n || tombstones->width != 1))
    return fail(ctx, WG_INVALID_ARGUMENT, "tombstones must be a bool vector over the graph");
  if (iterations) *iterations = 0;
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint8_t* tomb = tombstones ? (const uint8_t*)tombstones->data : nullptr;
  double* x = (double*)result->data;
  PrBuffers buf(ctx, g);
  int s;
  unsigned long long* live_d;
  if ((s = buf.get(&live_d, 1))) return s;
  WG_HIP(ctx, hipMemsetAsync(live_d, 0, 8, st));
  if ((s = clear_error(ctx))) return s;
  // x = 1 / live at live rows, the live count first (mod.rs:105-114).
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, kReduceGrid));
  pr_init<<<grid, 256, 0, st>>>(x, tomb, n, 0.0, live_d);
  unsigned long long live = 0;
  WG_HIP(ctx, hipMemcpyAsync(&live, live_d, 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  uint32_t herr = 0;
  const double nodes = (double)live;
  const double start = sqlf<DIV_>(1.0, nodes, herr);
  if (herr) return fail(ctx, (int)herr, "pagerank: 1 / live nodes: division by zero");
  pr_init<<<grid, 256, 0, st>>>(x, tomb, n, start, live_d);
  // out = PlusFirst(A) . x over the forward CSR (mod.rs:115-117).
  const bool weighted = o->weighted < 0 ? g->weighted : o->weighted != 0;
  double *out, *msg0, *msg1, *p = nullptr, *part, *piece;
  if ((s = buf.get(&out, std::max<uint64_t>(n, 1)))) return s;
  if ((s = launch_out_weights(ctx, g, weighted, x, out))) return s;
  // Personalisation (mod.rs:118-121, 155-199).
  if (seed) {
    if (seed->dtype == WG_BOOL)
      return fail(ctx, WG_DTYPE_MISMATCH, "dtype mismatch: expected f64, found bool");
    if (seed->width != 1) {
      if (o->max_iter != 0)
        return fail(ctx, WG_WIDTH_MISMATCH, "width mismatch: expected " + std::to_string(seed->width) + ", found 1");
    } else {
      if (seed->id_space != n) return fail(ctx, WG_ID_SPACE_MISMATCH, "seed id space mismatch");
      if ((s = buf.get_at(&part, kReduceGrid, 150))) return s;
      switch (seed->dtype) {
        case WG_F64: pr_seed_sum<double><<<grid, 256, 0, st>>>((const double*)seed->data, tomb, n, part); break;
        case WG_F32: pr_seed_sum<float><<<grid, 256, 0, st>>>((const float*)seed->data, tomb, n, part); break;
        case WG_I64: pr_seed_sum<int64_t><<<grid, 256, 0, st>>>((const int64_t*)seed->data, tomb, n, part); break;
        case WG_U32: pr_seed_sum<uint32_t><<<grid, 256, 0, st>>>((const uint32_t*)seed->data, tomb, n, part); break;
      }
      std::vector<double> hp(grid);
      WG_HIP(ctx, hipMemcpyAsync(hp.data(), part, grid * 8, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      const double total = sqlf<DIV_>(1.0, host_sum(hp), herr);
      if (herr) return fail(ctx, (int)herr, "pagerank: 1 / sum(seed): division by zero");
      if ((s = buf.get_at(&p, std::max<uint64_t>(n, 1), 151))) return s;
      switch (seed->dtype) {
        case WG_F64: pr_shares<double><<<grid, 256, 0, st>>>((const double*)seed->data, tomb, n, total, p, ctx->d_err); break;
        case WG_F32: pr_shares<float><<<grid, 256, 0, st>>>((const float*)seed->data, tomb, n, total, p, ctx->d_err); break;
        case WG_I64: pr_shares<int64_t><<<grid, 256, 0, st>>>((const int64_t*)seed->data, tomb, n, total, p, ctx->d_err); break;
        case WG_U32: pr_shares<uint32_t><<<grid, 256, 0, st>>>((const uint32_t*)seed->data, tomb, n, total, p, ctx->d_err); break;
      }
      if ((s = collect_error(ctx, "pagerank seed shares"))) return s;
    }
  }
  const double teleport = sqlf<SUB_>(1.0, o->damping, herr);
  if (herr) return fail(ctx, (int)herr, "pagerank: 1 - damping");
  uint32_t bound_err = 0;
  const double bound = o->has_tol ? sqlf<MUL_>(nodes, o->tol, bound_err) : 0.0;
  if (o->max_iter == 0) {
    WG_HIP(ctx, hipStreamSynchronize(st));
    return ok(ctx);
  }
  // First messages and dangling rank.
  DevCsr& rev = g->rev;
  if ((s = buf.get(&msg0, std::max<uint64_t>(n, 1)))) return s;
  if ((s = buf.get(&msg1, std::max<uint64_t>(n, 1)))) return s;
  const uint32_t nparts = 2 * (rev.nblocks + rev.nlong);
  if ((s = buf.get(&part, std::max<uint32_t>(std::max<uint32_t>(nparts, kReduceGrid), 1)))) return s;
  if ((s = buf.get(&piece, std::max<uint32_t>(rev.npartials, 1)))) return s;
  PrState* state;
  if ((s = buf.get(&state, 1))) return s;
  pr_spread0<<<grid, 256, 0, st>>>(x, out, msg0, n, part, ctx->d_err);
  {
    std::vector<double> hp(grid);
    WG_HIP(ctx, hipMemcpyAsync(hp.data(), part, grid * 8, hipMemcpyDeviceToHost, st));
    if ((s = collect_error(ctx, "pagerank"))) return s;
    pr_set_state<<<1, 1, 0, st>>>(state, host_sum(hp));
  }
  const double* wv = nullptr;
  if (weighted && rev.weights) wv = rev.weights;
  PrArgs a{rev.blocks, rev.offsets, rev.targets, wv, out, p, start, tomb, o->damping, teleport,
           bound, o->has_tol && !bound_err, rev.nblocks, rev.nlong, x, part, piece, state,
           ctx->d_err};
  // With a tolerance whose bound fails, the reference raises that error
  // after the first iteration (mod.rs:144-147).
  const uint32_t max_iter = bound_err ? 1 : o->max_iter;
  // The tiled gather on large graphs (its layout built once per graph).
  bool tiled = false;
  const double* wt = nullptr;
  double *ycomp = nullptr, *bpart = nullptr, *red = nullptr;
  TileRec<double>* recs = nullptr;
  const uint32_t nbp = (uint32_t)((n + 255) / 256);
  const uint32_t nred = (nbp + kPrReduceSpan - 1) / kPrReduceSpan;
  if (rev.nnz >= kPrTileMinNnz) {
    tiled = prepare_tiles(ctx, rev, sizeof(double), tile_hot_cap<double>(), tile_warm_cap<double>()) == WG_OK;
    if (!tiled) (void)ok(ctx);
  }
  if (tiled) {
    if (wv && (s = tile_weight_view<double>(ctx, rev, wv, &wt))) return s;
    if ((s = buf.get_at(&ycomp, std::max<uint64_t>(rev.nz_rows, 1), 160))) return s;
    if ((s = buf.get_at(&recs, rev.tile_nwaves, 161))) return s;
    if ((s = buf.get_at(&bpart, 2 * (size_t)nbp, 162))) return s;
    if ((s = buf.get_at(&red, 2 * (size_t)nred, 163))) return s;
  }
  uint32_t launched = 0;
  PrState hs{};
  const uint32_t batch = 8;
  while (launched < max_iter) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    const uint32_t k = std::min(batch, max_iter - launched);
    for (uint32_t i = 0; i < k; ++i, ++launched) {
      const int parity = launched & 1;
      const double* mi = parity ? msg1 : msg0;
      double* mo = parity ? msg0 : msg1;
      if (rev.nblocks == 0) break;
      if (tiled) {
        const DevCsr::Tiles& t = rev.tiles[(int)sizeof(double)];
        if (t.nhot || t.nwarm)
          pr_tiers<<<(unsigned)(((uint64_t)t.nhot + t.nwarm + 4 * 256 - 1) / (4 * 256)), 256, 0, st>>>(
              mi, rev.hot_inv, t.nhot, (double*)rev.hot_x, t.warm_ids, t.nwarm, (double*)t.warm_x, state);
        const unsigned tg = rev.tile_nwaves / kTileWaves;
        if (wt)
          spmv_tiles<double, PLUS, TIMES, true><<<tg, 1024, 0, st>>>(
              t.tagged, wt, mi, (const double*)rev.hot_x, t.nhot, (const double*)t.warm_x, (uint32_t)n,
              rev.tile_wave, rev.tile_heads, rev.nnz, ycomp, recs, ctx->d_jerr, &state->stop);
        else
          spmv_tiles<double, PLUS, SECOND, false><<<tg, 1024, 0, st>>>(
              t.tagged, nullptr, mi, (const double*)rev.hot_x, t.nhot, (const double*)t.warm_x, (uint32_t)n,
              rev.tile_wave, rev.tile_heads, rev.nnz, ycomp, recs, ctx->d_jerr, &state->stop);
        const TileGeom geo{rev.nz_bits, rev.nz_pre, rev.fix_bits, rev.offsets, rev.n, rev.tile_padded / kTileEntries,
                           rev.tile_nwaves};
        pr_tile_rows<<<nbp, 256, 0, st>>>(a, geo, ycomp, recs, n, mo, bpart, nbp, parity);
        pr_reduce<<<nred, 256, 0, st>>>(a, bpart, nbp, kPrReduceSpan, red, nred);
        pr_finish_sum<<<1, 256, 0, st>>>(a, red, nred, parity);
        continue;
      }
      if (wv)
        pr_iterate<true><<<rev.nblocks, kBlockThreads, 0, st>>>(a, mi, mo, parity);
      else
        pr_iterate<false><<<rev.nblocks, kBlockThreads, 0, st>>>(a, mi, mo, parity);
      if (rev.nlong)
        pr_long<<<(rev.nlong + 255) / 256, 256, 0, st>>>(a, rev.long_rows, rev.long_pbase, mo, parity);
      pr_finish<<<1, 256, 0, st>>>(a, parity);
    }
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipMemcpyAsync(&hs, state, sizeof hs, hipMemcpyDeviceToHost, st));
    if ((s = collect_error(ctx, "pagerank"))) return s;
    if (hs.stop || rev.nblocks == 0) break;
  }
  if (iterations) *iterations = rev.nblocks ? hs.iterations : max_iter;
  if (bound_err) return fail(ctx, (int)bound_err, "pagerank: live nodes * tol out of range");
  return ok(ctx);
}

// ---------------------------------------------------------------------------
// Adjoint of PageRank. The forward fixed point is, at live rows j,
//   x_j = d (sum_i w_ij x_i / s_i [s_i > 0] + delta p_j) + (1 - d) p_j,
//   delta = sum_{i: s_i <= 0} x_i,  s_i = sum_j w_ij (the out-weight),
// and x_j = 0 at tombstoned rows (compute/pagerank/mod.rs:96-287). For a
// loss L(x) with gradient g = dL/dx, the adjoint lambda solves
//   lambda_i = d (sum_j w_ij lt_j / s_i) + g_i      (s_i > 0)
//   lambda_i = d (sum_j p_j lt_j) + g_i             (s_i <= 0)
// with lt = lambda at live rows and 0 at tombstoned ones: a pull over the
// forward CSR (the forward pass pulls over the reverse), iterated from
// lambda = g with the forward's stop rule (L1 change below live * tol).
// Then, for an edge a -> b with s_a > 0,
//   dL/dw_ab = (d x_a / s_a) (lt_b - sum_j w_aj lt_j / s_a),
// and 0 where s_a <= 0 (the dangling indicator has no derivative).
namespace wg {

// p_i: the seed share, or start at live rows without a seed.
__device__ __forceinline__ double adj_share(const double* p, double start, const uint8_t* tomb, uint64_t i) {
  if (tomb && tomb[i]) return 0.0;
  return p ? p[i] : start;
}

__global__ __launch_bounds__(256) void adj_start(const double* __restrict__ g, const uint8_t* __restrict__ tomb,
                                                 const double* __restrict__ p, double start, uint64_t n,
                                                 double* __restrict__ lambda, double* __restrict__ lt,
                                                 double* __restrict__ part) {
  __shared__ double wsum[4];
  double pi = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double l = g[i];
    const double t = (tomb && tomb[i]) ? 0.0 : l;
    lambda[i] = l;
    lt[i] = t;
    pi += adj_share(p, start, tomb, i) * t;
  }
  const double s = block_sum(pi, wsum);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = 0.0;
    part[gridDim.x + blockIdx.x] = s;
  }
}

// One adjoint step from u = A lt and pi = sum p lt of the previous step.
__global__ __launch_bounds__(256) void adj_advance(const double* __restrict__ u, const double* __restrict__ out,
                                                   const double* __restrict__ g, const uint8_t* __restrict__ tomb,
                                                   const double* __restrict__ p, double start, uint64_t n,
                                                   double damping, const double* __restrict__ scal,
                                                   double* __restrict__ lambda, double* __restrict__ lt,
                                                   double* __restrict__ part) {
  __shared__ double wsum[8];
  const double pi = scal[0];
  double change = 0.0, pin = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    const double s = out[i];
    const double z = sql_lt(0.0, s) ? u[i] / s : pi;
    const double l = damping * z + g[i];
    change += fabs(l - lambda[i]);
    lambda[i] = l;
    const double t = (tomb && tomb[i]) ? 0.0 : l;
    lt[i] = t;
    pin += adj_share(p, start, tomb, i) * t;
  }
  const double c = block_sum(change, wsum);
  const double q = block_sum(pin, wsum + 4);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = c;
    part[gridDim.x + blockIdx.x] = q;
  }
}

// The partials in a fixed order: scal[0] = pi, scal[1] = the L1 change.
__global__ __launch_bounds__(256) void adj_reduce(const double* __restrict__ part, uint32_t grid,
                                                  double* __restrict__ scal) {
  __shared__ double sc[256], sq[256];
  const uint32_t tid = threadIdx.x;
  double c = 0.0, q = 0.0;
  for (uint32_t b = tid; b < grid; b += 256) {
    c += part[b];
    q += part[grid + b];
  }
  sc[tid] = c;
  sq[tid] = q;
  __syncthreads();
  for (uint32_t d = 128; d > 0; d >>= 1) {
    if (tid < d) {
      sc[tid] += sc[tid + d];
      sq[tid] += sq[tid + d];
    }
    __syncthreads();
  }
  if (tid == 0) {
    scal[0] = sq[0];
    scal[1] = sc[0];
  }
}

// dL/dw for every edge of the forward CSR, a wavefront a row.
__global__ __launch_bounds__(256) void adj_gradient(const uint64_t* __restrict__ off, const uint32_t* __restrict__ tgt,
                                                    uint64_t n, const double* __restrict__ x,
                                                    const double* __restrict__ out, const double* __restrict__ u,
                                                    const double* __restrict__ lt, double damping,
                                                    double* __restrict__ dldw) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t waves = (uint64_t)gridDim.x * 4;
  for (uint64_t a = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); a < n; a += waves) {
    const uint64_t b = off[a], e = off[a + 1];
    const double s = out[a];
    const bool flows = sql_lt(0.0, s);
    const double coef = flows ? damping * x[a] / s : 0.0;
    const double pa = flows ? u[a] / s : 0.0;
    for (uint64_t k = b + lane; k < e; k += 64) dldw[k] = flows ? coef * (lt[tgt[k]] - pa) : 0.0;
  }
}

int launch_plus_times_f64(wg_context* ctx, wg_graph* g, bool transpose, bool weighted, const double* x,
                          double* y);

}  // namespace wg

extern "C" int wg_pagerank_backward(wg_context* ctx, const wg_graph* gc, const wg_vector* tombstones,
                                    const wg_pagerank_opts* o, const wg_vector* seed, const wg_vector* ranks,
                                    const wg_vector* dl_dranks, wg_vector* dl_dweights, wg_vector* adjoint,
                                    uint32_t* iterations) {
  if (!ctx || !gc || !o || !ranks || !dl_dranks || !dl_dweights)
    return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n, m = g->nnz;
  if (iterations) *iterations = 0;
  for (const wg_vector* v : {ranks, dl_dranks, (const wg_vector*)adjoint})
    if (v && (v->dtype != WG_F64 || v->width != 1 || v->id_space != n))
      return fail(ctx, WG_INVALID_ARGUMENT, "ranks, their gradient and the adjoint must be f64, width 1, over the graph");
  if (dl_dweights->dtype != WG_F64 || dl_dweights->width != 1 || dl_dweights->id_space != m)
    return fail(ctx, WG_INVALID_ARGUMENT, "the weight gradient must be f64, width 1, one row per edge");
  if (tombstones && (tombstones->dtype != WG_BOOL || tombstones->id_space != n || tombstones->width != 1))
    return fail(ctx, WG_INVALID_ARGUMENT, "tombstones must be a bool vector over the graph");
  const bool weighted = o->weighted < 0 ? g->weighted : o->weighted != 0;
  if (!weighted || !g->fwd.weights)
    return fail(ctx, WG_INVALID_ARGUMENT, "pagerank backward differentiates the edge weights: the run must be weighted");
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint8_t* tomb = tombstones ? (const uint8_t*)tombstones->data : nullptr;
  PrBuffers buf(ctx, g);
  buf.next_slot = 170;
  int s;
  unsigned long long* live_d;
  double *scratch, *out, *p = nullptr, *lambda, *lt, *u, *part, *scal;
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, kReduceGrid));
  if ((s = buf.get(&live_d, 1)) || (s = buf.get(&scratch, std::max<uint64_t>(n, 1))) ||
      (s = buf.get(&out, std::max<uint64_t>(n, 1))) || (s = buf.get(&lt, std::max<uint64_t>(n, 1))) ||
      (s = buf.get(&u, std::max<uint64_t>(n, 1))) || (s = buf.get(&part, 2 * kReduceGrid)) ||
      (s = buf.get(&scal, 2)))
    return s;
  if (adjoint) {
    lambda = (double*)adjoint->data;
  } else if ((s = buf.get(&lambda, std::max<uint64_t>(n, 1)))) {
    return s;
  }
  WG_HIP(ctx, hipMemsetAsync(live_d, 0, 8, st));
  if ((s = clear_error(ctx))) return s;
  // The live count and start, as the forward run takes them.
  pr_init<<<grid, 256, 0, st>>>(scratch, tomb, n, 0.0, live_d);
  unsigned long long live = 0;
  WG_HIP(ctx, hipMemcpyAsync(&live, live_d, 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  uint32_t herr = 0;
  const double nodes = (double)live;
  const double start = sqlf<DIV_>(1.0, nodes, herr);
  if (herr) return fail(ctx, (int)herr, "pagerank: 1 / live nodes: division by zero");
  if ((s = launch_out_weights(ctx, g, true, scratch, out))) return s;
  if (seed) {
    if (seed->dtype == WG_BOOL) return fail(ctx, WG_DTYPE_MISMATCH, "dtype mismatch: expected f64, found bool");
    if (seed->width != 1) return fail(ctx, WG_WIDTH_MISMATCH, "width mismatch: expected " + std::to_string(seed->width) + ", found 1");
    if (seed->id_space != n) return fail(ctx, WG_ID_SPACE_MISMATCH, "seed id space mismatch");
    switch (seed->dtype) {
      case WG_F64: pr_seed_sum<double><<<grid, 256, 0, st>>>((const double*)seed->data, tomb, n, part); break;
      case WG_F32: pr_seed_sum<float><<<grid, 256, 0, st>>>((const float*)seed->data, tomb, n, part); break;
      case WG_I64: pr_seed_sum<int64_t><<<grid, 256, 0, st>>>((const int64_t*)seed->data, tomb, n, part); break;
      case WG_U32: pr_seed_sum<uint32_t><<<grid, 256, 0, st>>>((const uint32_t*)seed->data, tomb, n, part); break;
    }
    std::vector<double> hp(grid);
    WG_HIP(ctx, hipMemcpyAsync(hp.data(), part, grid * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    const double total = sqlf<DIV_>(1.0, host_sum(hp), herr);
    if (herr) return fail(ctx, (int)herr, "pagerank: 1 / sum(seed): division by zero");
    if ((s = buf.get(&p, std::max<uint64_t>(n, 1)))) return s;
    switch (seed->dtype) {
      case WG_F64: pr_shares<double><<<grid, 256, 0, st>>>((const double*)seed->data, tomb, n, total, p, ctx->d_err); break;
      case WG_F32: pr_shares<float><<<grid, 256, 0, st>>>((const float*)seed->data, tomb, n, total, p, ctx->d_err); break;
      case WG_I64: pr_shares<int64_t><<<grid, 256, 0, st>>>((const int64_t*)seed->data, tomb, n, total, p, ctx->d_err); break;
      case WG_U32: pr_shares<uint32_t><<<grid, 256, 0, st>>>((const uint32_t*)seed->data, tomb, n, total, p, ctx->d_err); break;
    }
    if ((s = collect_error(ctx, "pagerank seed shares"))) return s;
  }
  uint32_t bound_err = 0;
  const double bound = o->has_tol ? sqlf<MUL_>(nodes, o->tol, bound_err) : 0.0;
  if (bound_err) return fail(ctx, (int)bound_err, "pagerank: live nodes * tol out of range");
  const double* g_in = (const double*)dl_dranks->data;
  adj_start<<<grid, 256, 0, st>>>(g_in, tomb, p, start, n, lambda, lt, part);
  adj_reduce<<<1, 256, 0, st>>>(part, grid, scal);
  uint32_t it = 0;
  double hs[2] = {0.0, 0.0};
  while (it < o->max_iter) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    ++it;
    if ((s = launch_plus_times_f64(ctx, g, false, true, lt, u))) return s;
    adj_advance<<<grid, 256, 0, st>>>(u, out, g_in, tomb, p, start, n, o->damping, scal, lambda, lt, part);
    adj_reduce<<<1, 256, 0, st>>>(part, grid, scal);
    WG_HIP(ctx, hipGetLastError());
    if (o->has_tol) {
      WG_HIP(ctx, hipMemcpyAsync(hs, scal, 16, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      if (sql_lt(hs[1], bound)) break;
    }
  }
  // The gradient from the final adjoint.
  if ((s = launch_plus_times_f64(ctx, g, false, true, lt, u))) return s;
  if (m) {
    const unsigned ggrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 3) / 4, 65536));
    adj_gradient<<<ggrid, 256, 0, st>>>(g->fwd.offsets, g->fwd.targets, n, (const double*)ranks->data, out, u,
                                        lt, o->damping, (double*)dl_dweights->data);
  }
  WG_HIP(ctx, hipGetLastError());
  if ((s = collect_error(ctx, "pagerank backward"))) return s;
  if (iterations) *iterations = it;
  return ok(ctx);
}
