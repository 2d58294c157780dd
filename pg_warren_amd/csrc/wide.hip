// Graphs with u64 targets (TargetWidth::U64, graph/csr/target.rs:16-31):
// upload with Csr::from_parts' checks on the device
// (graph/csr/mod.rs:78-137), and the reverse CSR by a stable radix sort of
// the entries by target (graph/csr/transpose.rs:24-44). Such a graph is
// held with u64 targets only; products read it through wide.hpp, and the
// other entry points refuse it (WG_UNSUPPORTED).
#include <algorithm>
#include <string>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr unsigned kT = 256;

unsigned grid_of(uint64_t items) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kT - 1) / kT, 16384));
}

__device__ __forceinline__ uint64_t gid() { return (uint64_t)blockIdx.x * kT + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * kT; }

// The row holding entry e: the last r in [0, n) with off[r] <= e.
__device__ __forceinline__ uint64_t row_of_entry(const uint64_t* __restrict__ off, uint64_t n, uint64_t e) {
  uint64_t lo = 0, hi = n;  // off[lo] <= e < off[hi]
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (off[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// The first row whose offset decreases (OffsetsNotMonotone).
__global__ __launch_bounds__(kT) void offsets_descent(const uint64_t* __restrict__ off, uint64_t n,
                                                      unsigned long long* __restrict__ bad) {
  for (uint64_t r = gid(); r < n; r += gstride())
    if (off[r + 1] < off[r]) {
      atomicMin(bad, (unsigned long long)r);
      return;
    }
}

// The first rows (ascending) holding a target outside the id space and a
// pair of targets not strictly ascending: bad[0], bad[1].
template <class T>
__global__ __launch_bounds__(kT) void csr_entries(const uint64_t* __restrict__ off, const T* __restrict__ tgt,
                                                  uint64_t m, uint64_t n, unsigned long long* __restrict__ bad) {
  for (uint64_t e = gid(); e < m; e += gstride()) {
    const uint64_t t = tgt[e];
    const bool out = t >= n, desc = e > 0 && (uint64_t)tgt[e - 1] >= t;
    if (!out && !desc) continue;
    const uint64_t r = row_of_entry(off, n, e);
    if (out) atomicMin(&bad[0], (unsigned long long)r);
    if (desc && e > off[r]) atomicMin(&bad[1], (unsigned long long)r);
  }
}

__global__ __launch_bounds__(kT) void iota64(uint64_t* __restrict__ v, uint64_t m) {
  for (uint64_t i = gid(); i < m; i += gstride()) v[i] = i;
}

// Reverse entry j: the source row of forward entry idx[j] and its weight.
__global__ __launch_bounds__(kT) void reverse_place(const uint64_t* __restrict__ fo, uint64_t n,
                                                    const uint64_t* __restrict__ idx, uint64_t m,
                                                    const double* __restrict__ w, uint64_t* __restrict__ rt,
                                                    double* __restrict__ rw) {
  for (uint64_t j = gid(); j < m; j += gstride()) {
    const uint64_t e = idx[j];
    rt[j] = row_of_entry(fo, n, e);
    if (w) rw[j] = w[e];
  }
}

// Reverse offsets: ro[r] = the first sorted key at or above r, r in [0, n].
__global__ __launch_bounds__(kT) void reverse_offsets(const uint64_t* __restrict__ keys, uint64_t m, uint64_t n,
                                                      uint64_t* __restrict__ ro) {
  for (uint64_t r = gid(); r <= n; r += gstride()) {
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (keys[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    ro[r] = lo;
  }
}

unsigned key_bits(uint64_t n) {
  unsigned b = 1;
  while (b < 64 && (1ull << b) < n) ++b;
  return b;
}

}  // namespace

// Csr::from_parts' per-row checks (graph/csr/mod.rs:117-126) over a CSR
// already on the device whose offsets are known to be monotone: the first
// row holding a target outside the id space or targets not strictly
// ascending, the former named first within a row.
template <class T>
int check_rows_device(wg_context* ctx, uint64_t n, const uint64_t* off, const T* tgt, uint64_t m) {
  if (!m) return WG_OK;
  hipStream_t st = ctx->stream;
  unsigned long long* bad = nullptr;
  int s;
  if ((s = dalloc(ctx, (void**)&bad, 16))) return s;
  unsigned long long hb[2] = {~0ull, ~0ull};
  hipError_t e = hipMemsetAsync(bad, 0xff, 16, st);
  if (e == hipSuccess) {
    csr_entries<T><<<grid_of(m), kT, 0, st>>>(off, tgt, m, n, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hb, bad, 16, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dfree(ctx, bad, 16);
  if (e != hipSuccess) return hip_fail(ctx, e, "row checks");
  if (hb[0] != ~0ull && hb[0] <= hb[1])
    return fail(ctx, WG_INVALID_CSR, "target outside id space in row " + std::to_string(hb[0]));
  if (hb[1] != ~0ull) return fail(ctx, WG_INVALID_CSR, "row " + std::to_string(hb[1]) + " not sorted");
  return WG_OK;
}
template int check_rows_device<uint32_t>(wg_context*, uint64_t, const uint64_t*, const uint32_t*, uint64_t);

int check_narrow(wg_context* ctx, const wg_graph* g) {
  if (g && g->fwd.targets64)
    return fail(ctx, WG_UNSUPPORTED, "a graph with u64 targets is read only by products (wg_spmv, wg_spmm)");
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_graph_upload_u64(wg_context* ctx, uint64_t n, const uint64_t* off, const uint64_t* tgt,
                                   const double* w, wg_graph** out) {
  if (!ctx || !out) return WG_INVALID_ARGUMENT;
  *out = nullptr;
  if (!off) return fail(ctx, WG_INVALID_CSR, "no offsets");
  if (n == UINT64_MAX) return fail(ctx, WG_INVALID_CSR, "id space too large");
  // Csr::from_parts' checks, in its order (graph/csr/mod.rs:78-137).
  if (off[0] != 0) return fail(ctx, WG_INVALID_CSR, "first offset is not zero");
  const uint64_t m = off[n];
  if (m && !tgt) return fail(ctx, WG_INVALID_CSR, "no targets");
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  wg_graph* g = new (std::nothrow) wg_graph();
  if (!g) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  g->ctx = ctx;
  g->n = n;
  g->nnz = m;
  g->weighted = w != nullptr;
  DevCsr& f = g->fwd;
  DevCsr& r = g->rev;
  f.n = r.n = n;
  f.nnz = r.nnz = m;
  int s = WG_OK;
  unsigned long long* bad = nullptr;
  uint64_t *keys2 = nullptr, *idx = nullptr, *idx2 = nullptr;
  auto cleanup = [&](int status) {
    dfree(ctx, bad, 24);
    dfree(ctx, keys2, std::max<uint64_t>(m, 1) * 8);
    dfree(ctx, idx, std::max<uint64_t>(m, 1) * 8);
    dfree(ctx, idx2, std::max<uint64_t>(m, 1) * 8);
    if (status) destroy_graph(g);
    return status;
  };
  if ((s = dalloc(ctx, (void**)&f.offsets, (n + 1) * 8)) ||
      (s = dalloc(ctx, (void**)&f.targets64, std::max<uint64_t>(m, 1) * 8)) ||
      (w && (s = dalloc(ctx, (void**)&f.weights, std::max<uint64_t>(m, 1) * 8))) ||
      (s = dalloc(ctx, (void**)&bad, 24)))
    return cleanup(s);
  WG_HIP(ctx, hipMemcpyAsync(f.offsets, off, (n + 1) * 8, hipMemcpyHostToDevice, st));
  if (m) WG_HIP(ctx, hipMemcpyAsync(f.targets64, tgt, m * 8, hipMemcpyHostToDevice, st));
  if (m && w) WG_HIP(ctx, hipMemcpyAsync(f.weights, w, m * 8, hipMemcpyHostToDevice, st));
  WG_HIP(ctx, hipMemsetAsync(bad, 0xff, 24, st));
  offsets_descent<<<grid_of(n), kT, 0, st>>>(f.offsets, n, bad);
  WG_HIP(ctx, hipGetLastError());
  unsigned long long hb[3];
  WG_HIP(ctx, hipMemcpyAsync(hb, bad, 24, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  if (hb[0] != ~0ull) return cleanup(fail(ctx, WG_INVALID_CSR, "offsets not monotone at row " + std::to_string(hb[0])));
  if ((s = check_rows_device<uint64_t>(ctx, n, f.offsets, f.targets64, m))) return cleanup(s);
  // The reverse CSR: entries stably sorted by target, each the source row
  // of its forward entry (ascending within a target, as transpose.rs
  // places them).
  if ((s = dalloc(ctx, (void**)&r.offsets, (n + 1) * 8)) ||
      (s = dalloc(ctx, (void**)&r.targets64, std::max<uint64_t>(m, 1) * 8)) ||
      (w && (s = dalloc(ctx, (void**)&r.weights, std::max<uint64_t>(m, 1) * 8))) ||
      (s = dalloc(ctx, (void**)&keys2, std::max<uint64_t>(m, 1) * 8)) ||
      (s = dalloc(ctx, (void**)&idx, std::max<uint64_t>(m, 1) * 8)) ||
      (s = dalloc(ctx, (void**)&idx2, std::max<uint64_t>(m, 1) * 8)))
    return cleanup(s);
  if (m) {
    iota64<<<grid_of(m), kT, 0, st>>>(idx, m);
    WG_HIP(ctx, hipGetLastError());
    if ((s = sort_u64_keyed(ctx, f.targets64, keys2, idx, idx2, m, key_bits(n)))) return cleanup(s);
    reverse_place<<<grid_of(m), kT, 0, st>>>(f.offsets, n, idx2, m, w ? f.weights : nullptr, r.targets64,
                                             r.weights);
    WG_HIP(ctx, hipGetLastError());
  }
  reverse_offsets<<<grid_of(n + 1), kT, 0, st>>>(keys2, m, n, r.offsets);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(st));
  s = cleanup(WG_OK);
  ctx->graphs.insert(g);
  *out = g;
  return ok(ctx);
}

extern "C" int wg_graph_read_u64(wg_context* ctx, const wg_graph* g, int reverse, uint64_t* off, uint64_t* tgt,
                                 double* w) {
  if (!ctx || !g || !off) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int s = check_graph(ctx, g, true)) return s;
  if (!g->fwd.targets64) return fail(ctx, WG_INVALID_ARGUMENT, "the graph has u32 targets");
  const DevCsr& c = reverse ? g->rev : g->fwd;
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  WG_HIP(ctx, hipMemcpy(off, c.offsets, (g->n + 1) * 8, hipMemcpyDeviceToHost));
  if (g->nnz) {
    if (tgt) WG_HIP(ctx, hipMemcpy(tgt, c.targets64, g->nnz * 8, hipMemcpyDeviceToHost));
    if (w && c.weights) WG_HIP(ctx, hipMemcpy(w, c.weights, g->nnz * 8, hipMemcpyDeviceToHost));
  }
  return ok(ctx);
}
