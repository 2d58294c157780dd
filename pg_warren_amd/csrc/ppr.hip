// Batched personalised PageRank: R seed sets as the columns of one n x R
// power iteration (compute/pagerank/mod.rs:96-287 per column; the reference
// runs one seed vector per call). Each column is exactly the plain run with
// its seed vector: the same set-up (live count, out-weights, shares
// seed / sum(seed) with SQL float checks), the same update in f64, and its own
// stop test (L1 change < live nodes * tol); a column that stops is frozen at
// that iteration while the others go on.
//
// One fused kernel per iteration (ppr_iterate): persistent workgroups whose
// wavefronts take rows in turn; the lanes of a wavefront hold the active
// columns, so a neighbour's row of messages is one coalesced read of
// 8 x (active columns) bytes; the gather y = A^T msg is followed in registers
// by the advance x' = d (y + dangling_c p) + (1 - d) p and the next spread
// msg' = x' / out. Per-column L1 changes and dangling ranks go to per-block
// partials, which ppr_finish (one workgroup, a thread per column) sums in a
// fixed order, deciding which columns stop. Rows longer than kPprPiece
// entries are gathered in pieces by whole workgroups and advanced by
// ppr_long. No atomics and no cross-workgroup fences.
#include <algorithm>
#include <cmath>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

int launch_out_weights(wg_context* ctx, wg_graph* g, bool weighted, const double* x, double* out);

constexpr uint32_t kPprPiece = 4096;   // rows longer than this are split
constexpr uint32_t kPprMaxCols = 256;  // columns a launch handles (4 per lane)
constexpr uint32_t kPprGrid = 2048;    // persistent workgroups of ppr_iterate

struct PprArgs {
  const uint64_t* off;    // reverse CSR
  const uint32_t* src;
  const double* w;        // reverse weights or null
  uint64_t n;
  uint32_t R;             // row stride of X, M (columns of the batch)
  uint32_t nact;          // active columns this iteration
  const uint32_t* cols;   // their indices
  const double* out;      // out-weights
  const uint8_t* tomb;
  const uint32_t* seed_row;  // n: index into P, or ~0
  const double* P;           // seed rows x R shares
  double damping, teleport;
  const double* dangling;    // R: this iteration's dangling rank per column
  double* X;
  const double* Min;
  double* Mout;
  double* part;              // (grid + nlong) x 2 x R: change, dangling
  double* piece;             // split-row pieces x R
  unsigned long long* err;
};

// Rows handled by ppr_iterate: all but the split ones.
__device__ __forceinline__ double ppr_update(const PprArgs& a, uint64_t v, uint32_t c, double y,
                                             double& change, double& dang, uint32_t& err) {
  const uint64_t idx = v * a.R + c;
  const double xo = a.X[idx];
  double next = 0.0;
  if (!(a.tomb && a.tomb[v])) {
    const uint32_t sr = a.seed_row[v];
    const double p = sr == ~0u ? 0.0 : a.P[(uint64_t)sr * a.R + c];
    const double flowed = sqlf<ADD_>(y, sqlf<MUL_>(a.dangling[c], p, err), err);
    next = sqlf<ADD_>(sqlf<MUL_>(a.damping, flowed, err), sqlf<MUL_>(a.teleport, p, err), err);
  }
  a.X[idx] = next;
  const double o = a.out[v];
  if (sql_lt(0.0, o)) {
    a.Mout[idx] = sqlf<DIV_>(next, o, err);
  } else {
    dang += next;
    a.Mout[idx] = 0.0;
  }
  change += fabs(next - xo);
  return next;
}

template <bool HW>
__global__ __launch_bounds__(256) void ppr_iterate(PprArgs a, uint32_t nlong_pieces) {
  __shared__ double red[2][4][kPprMaxCols];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int J = kPprMaxCols / 64;
  uint32_t col[J];
  bool on[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const uint32_t k = lane + 64 * j;
    on[j] = k < a.nact;
    col[j] = on[j] ? a.cols[k] : 0;
  }
  double change[J], dang[J];
#pragma unroll
  for (int j = 0; j < J; ++j) change[j] = dang[j] = 0.0;
  uint32_t e_err = 0;
  const uint64_t nwaves = (uint64_t)gridDim.x * 4;
  for (uint64_t v = (uint64_t)blockIdx.x * 4 + wv; v < a.n; v += nwaves) {
    const uint64_t b = a.off[v], e = a.off[v + 1];
    if (e - b > kPprPiece) continue;  // split rows: ppr_piece / ppr_long
    double acc[J];
#pragma unroll
    for (int j = 0; j < J; ++j) acc[j] = 0.0;
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t u = a.src[i];
      const double wgt = HW ? a.w[i] : 1.0;
      const double* mrow = a.Min + u * a.R;
#pragma unroll
      for (int j = 0; j < J; ++j)
        if (on[j]) acc[j] += HW ? wgt * mrow[col[j]] : mrow[col[j]];
    }
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (on[j]) ppr_update(a, v, col[j], acc[j], change[j], dang[j], e_err);
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    red[0][wv][lane + 64 * j] = change[j];
    red[1][wv][lane + 64 * j] = dang[j];
  }
  __syncthreads();
  for (uint32_t k = tid; k < a.nact; k += 256) {
    double c = 0.0, d = 0.0;
    for (int q = 0; q < 4; ++q) {
      c += red[0][q][k];
      d += red[1][q][k];
    }
    a.part[(uint64_t)blockIdx.x * 2 * a.R + k] = c;
    a.part[(uint64_t)blockIdx.x * 2 * a.R + a.R + k] = d;
  }
  if (e_err) raise_error(a.err, blockIdx.x, e_err);
}

// One piece of a split row: four wavefronts take its entries in turn and
// their column sums meet in LDS.
template <bool HW>
__global__ __launch_bounds__(256) void ppr_piece(PprArgs a, const uint32_t* __restrict__ pieces) {
  __shared__ double red[4][kPprMaxCols];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int J = kPprMaxCols / 64;
  const uint32_t v = pieces[3 * blockIdx.x], p = pieces[3 * blockIdx.x + 1];
  const uint32_t slot = pieces[3 * blockIdx.x + 2];
  const uint64_t b = a.off[v], e = a.off[v + 1];
  const uint64_t p0 = b + (uint64_t)p * kPprPiece, p1 = p0 + kPprPiece < e ? p0 + kPprPiece : e;
  double acc[J];
  uint32_t col[J];
  bool on[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    acc[j] = 0.0;
    const uint32_t k = lane + 64 * j;
    on[j] = k < a.nact;
    col[j] = on[j] ? a.cols[k] : 0;
  }
  for (uint64_t i = p0 + wv; i < p1; i += 4) {
    const uint64_t u = a.src[i];
    const double wgt = HW ? a.w[i] : 1.0;
    const double* mrow = a.Min + u * a.R;
#pragma unroll
    for (int j = 0; j < J; ++j)
      if (on[j]) acc[j] += HW ? wgt * mrow[col[j]] : mrow[col[j]];
  }
#pragma unroll
  for (int j = 0; j < J; ++j) red[wv][lane + 64 * j] = acc[j];
  __syncthreads();
  for (uint32_t k = tid; k < a.nact; k += 256)
    a.piece[(uint64_t)slot * a.R + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

// A split row: its pieces' sums in piece order, then the update.
__global__ __launch_bounds__(256) void ppr_long(PprArgs a, const uint32_t* __restrict__ rows,
                                                const uint32_t* __restrict__ pbase, uint32_t nlong,
                                                uint32_t grid_parts) {
  const uint32_t i = blockIdx.x;
  const uint64_t v = rows[i];
  uint32_t e_err = 0;
  for (uint32_t k = threadIdx.x; k < a.nact; k += 256) {
    const uint32_t c = a.cols[k];
    double y = 0.0;
    for (uint32_t p = pbase[i]; p < pbase[i + 1]; ++p) y += a.piece[(uint64_t)p * a.R + k];
    double change = 0.0, dang = 0.0;
    ppr_update(a, v, c, y, change, dang, e_err);
    a.part[(uint64_t)(grid_parts + i) * 2 * a.R + k] = change;
    a.part[(uint64_t)(grid_parts + i) * 2 * a.R + a.R + k] = dang;
  }
  if (e_err) raise_error(a.err, v, e_err);
}

// Per column, in a fixed order: the iteration's L1 change and the next
// dangling rank; a column whose change falls below the bound stops.
__global__ __launch_bounds__(256) void ppr_finish(const double* __restrict__ part, uint32_t nparts,
                                                  uint32_t R, uint32_t nact, const uint32_t* __restrict__ cols,
                                                  double* __restrict__ dangling, double* __restrict__ delta,
                                                  uint32_t* __restrict__ iterations, uint8_t* __restrict__ active,
                                                  double bound, int has_tol) {
  for (uint32_t k = threadIdx.x; k < nact; k += blockDim.x) {
    double c = 0.0, d = 0.0;
    for (uint32_t b = 0; b < nparts; ++b) {
      c += part[(uint64_t)b * 2 * R + k];
      d += part[(uint64_t)b * 2 * R + R + k];
    }
    const uint32_t col = cols[k];
    delta[col] = c;
    dangling[col] = d;
    iterations[col] += 1;
    if (has_tol && sql_lt(c, bound)) active[col] = 0;
  }
}

// x = start at live rows, msg = x / out, and per-block dangling partials.
__global__ __launch_bounds__(256) void ppr_start(uint64_t n, uint32_t R, double start,
                                                 const uint8_t* __restrict__ tomb,
                                                 const double* __restrict__ out, double* __restrict__ X,
                                                 double* __restrict__ M, double* __restrict__ part,
                                                 unsigned long long* err) {
  // Thread k of a block owns column k (R <= 256); blocks take rows in turn.
  const uint32_t k = threadIdx.x;
  uint32_t e_err = 0;
  double dang = 0.0;
  if (k < R) {
    for (uint64_t v = blockIdx.x; v < n; v += gridDim.x) {
      const double x = (tomb && tomb[v]) ? 0.0 : start;
      X[v * R + k] = x;
      const double o = out[v];
      if (sql_lt(0.0, o)) {
        M[v * R + k] = sqlf<DIV_>(x, o, e_err);
      } else {
        dang += x;
        M[v * R + k] = 0.0;
      }
    }
    part[(uint64_t)blockIdx.x * 2 * R + k] = 0.0;
    part[(uint64_t)blockIdx.x * 2 * R + R + k] = dang;
  }
  if (e_err) raise_error(err, blockIdx.x, e_err);
}

__global__ void ppr_count_live(uint64_t n, const uint8_t* __restrict__ tomb, unsigned long long* live) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    c += !(tomb && tomb[i]);
  for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(live, c);
}

}  // namespace wg

using namespace wg;

extern "C" int wg_pagerank_batch(wg_context* ctx, const wg_graph* gc, const wg_vector* tombstones,
                                 const wg_pagerank_opts* o, const uint64_t* seed_offsets,
                                 const uint64_t* seed_nodes, const double* seed_weights, uint32_t R,
                                 wg_vector* ranks, uint32_t* iterations) {
  if (!ctx || !gc || !o || !seed_offsets || !ranks) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
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
  // Split rows and their pieces.
  std::vector<uint32_t> lrows, lbase, pieces;
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t deg = rev.h_offsets[v + 1] - rev.h_offsets[v];
    if (deg <= kPprPiece) continue;
    lbase.push_back((uint32_t)(pieces.size() / 3));
    const uint32_t np = (uint32_t)((deg + kPprPiece - 1) / kPprPiece);
    for (uint32_t p = 0; p < np; ++p) {
      pieces.push_back((uint32_t)v);
      pieces.push_back(p);
      pieces.push_back((uint32_t)(pieces.size() / 3));
    }
    lrows.push_back((uint32_t)v);
  }
  const uint32_t npieces = (uint32_t)(pieces.size() / 3), nlong = (uint32_t)lrows.size();
  lbase.push_back(npieces);
  const size_t nR = (size_t)std::max<uint64_t>(n, 1) * R;
  double *X = (double*)ranks->data, *M0, *M1, *out, *part, *piece, *Pd, *dangling, *delta;
  uint32_t *seed_row_d, *cols_d, *iters_d, *lrows_d, *lbase_d, *pieces_d;
  uint8_t* active_d;
  const uint32_t nparts = kPprGrid + nlong;
  if ((s = get((void**)&M0, nR * 8)) || (s = get((void**)&M1, nR * 8)) ||
      (s = get((void**)&out, std::max<uint64_t>(n, 1) * 8)) ||
      (s = get((void**)&part, (size_t)nparts * 2 * R * 8)) ||
      (s = get((void**)&piece, (size_t)std::max<uint32_t>(npieces, 1) * R * 8)) ||
      (s = get((void**)&Pd, P.size() * 8)) || (s = get((void**)&dangling, R * 8)) ||
      (s = get((void**)&delta, R * 8)) || (s = get((void**)&seed_row_d, std::max<uint64_t>(n, 1) * 4)) ||
      (s = get((void**)&cols_d, R * 4)) || (s = get((void**)&iters_d, R * 4)) ||
      (s = get((void**)&active_d, R)) || (s = get((void**)&lrows_d, std::max<uint32_t>(nlong, 1) * 4)) ||
      (s = get((void**)&lbase_d, lbase.size() * 4)) ||
      (s = get((void**)&pieces_d, std::max<size_t>(pieces.size(), 1) * 4)))
    return s;
  WG_HIP(ctx, hipMemcpyAsync(Pd, P.data(), P.size() * 8, hipMemcpyHostToDevice, st));
  if (n) WG_HIP(ctx, hipMemcpyAsync(seed_row_d, seed_row.data(), n * 4, hipMemcpyHostToDevice, st));
  if (nlong) WG_HIP(ctx, hipMemcpyAsync(lrows_d, lrows.data(), nlong * 4, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemcpyAsync(lbase_d, lbase.data(), lbase.size() * 4, hipMemcpyHostToDevice, st));
  if (npieces) WG_HIP(ctx, hipMemcpyAsync(pieces_d, pieces.data(), pieces.size() * 4, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemsetAsync(iters_d, 0, R * 4, st));
  WG_HIP(ctx, hipMemsetAsync(active_d, 1, R, st));
  if ((s = clear_error(ctx))) return s;
  // out = PlusFirst(A) . 1 (the values do not matter to (x) first).
  if ((s = launch_out_weights(ctx, g, weighted, out, out))) return s;
  ppr_start<<<kPprGrid, 256, 0, st>>>(n, R, start, tomb, out, X, M0, part, ctx->d_err);
  std::vector<uint32_t> all(R);
  for (uint32_t c = 0; c < R; ++c) all[c] = c;
  WG_HIP(ctx, hipMemcpyAsync(cols_d, all.data(), R * 4, hipMemcpyHostToDevice, st));
  // The first dangling ranks: ppr_finish over the start partials, leaving
  // iterations and activity as they were.
  ppr_finish<<<1, 256, 0, st>>>(part, kPprGrid, R, R, cols_d, dangling, delta, iters_d, active_d, 0.0, 0);
  WG_HIP(ctx, hipMemsetAsync(iters_d, 0, R * 4, st));
  if ((s = collect_error(ctx, "pagerank batch"))) return s;
  const double* wv = weighted && rev.weights ? rev.weights : nullptr;
  const uint32_t max_iter = bound_err ? 1 : o->max_iter;
  std::vector<uint8_t> act(R, 1);
  std::vector<uint32_t> active_cols = all;
  double *Min = M0, *Mout = M1;
  for (uint32_t it = 0; it < max_iter && !active_cols.empty(); ++it) {
    if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
    const uint32_t nact = (uint32_t)active_cols.size();
    WG_HIP(ctx, hipMemcpyAsync(cols_d, active_cols.data(), nact * 4, hipMemcpyHostToDevice, st));
    PprArgs a{rev.offsets, rev.targets, wv, n, R, nact, cols_d, out, tomb, seed_row_d, Pd,
              o->damping, teleport, dangling, X, Min, Mout, part, piece, ctx->d_err};
    if (wv)
      ppr_iterate<true><<<kPprGrid, 256, 0, st>>>(a, npieces);
    else
      ppr_iterate<false><<<kPprGrid, 256, 0, st>>>(a, npieces);
    if (npieces) {
      if (wv)
        ppr_piece<true><<<npieces, 256, 0, st>>>(a, pieces_d);
      else
        ppr_piece<false><<<npieces, 256, 0, st>>>(a, pieces_d);
      ppr_long<<<nlong, 256, 0, st>>>(a, lrows_d, lbase_d, nlong, kPprGrid);
    }
    ppr_finish<<<1, 256, 0, st>>>(part, nparts, R, nact, cols_d, dangling, delta, iters_d, active_d, bound,
                                  o->has_tol && !bound_err);
    WG_HIP(ctx, hipGetLastError());
    std::swap(Min, Mout);
    WG_HIP(ctx, hipMemcpyAsync(act.data(), active_d, R, hipMemcpyDeviceToHost, st));
    if ((s = collect_error(ctx, "pagerank batch"))) return s;
    active_cols.clear();
    for (uint32_t c = 0; c < R; ++c)
      if (act[c]) active_cols.push_back(c);
  }
  if (iterations) WG_HIP(ctx, hipMemcpy(iterations, iters_d, R * 4, hipMemcpyDeviceToHost));
  if (bound_err) return fail(ctx, (int)bound_err, "pagerank: live nodes * tol out of range");
  return ok(ctx);
}
