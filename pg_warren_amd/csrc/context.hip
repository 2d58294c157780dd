// Context, memory budget, cancellation, device vectors and the device CSR
// pair of libwarren_gpu.
#include <algorithm>
#include <cstdio>
#include <new>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

size_t dtype_size(int dtype) {
  switch (dtype) {
    case WG_F32: return 4;
    case WG_F64: return 8;
    case WG_I64: return 8;
    case WG_BOOL: return 1;
    case WG_U32: return 4;
  }
  return 0;
}

int fail(wg_context* ctx, int status, const std::string& msg) {
  if (ctx) {
    ctx->last_status = status;
    ctx->last_msg = msg;
  }
  return status;
}

int ok(wg_context* ctx) {
  if (ctx) {
    ctx->last_status = WG_OK;
    ctx->last_msg.clear();
  }
  return WG_OK;
}

int hip_fail(wg_context* ctx, hipError_t e, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  return fail(ctx, e == hipErrorOutOfMemory ? WG_OUT_OF_MEMORY : WG_DEVICE_ERROR, buf);
}

// The block cache: blocks up to kCacheBlock bytes, kCacheTotal in all.
constexpr size_t kCacheBlock = size_t(512) << 20;
constexpr size_t kCacheTotal = size_t(4) << 30;

static void flush_cache(wg_context* ctx) {
  for (auto& b : ctx->cache) (void)hipFree(b.second);
  ctx->cache.clear();
  ctx->cache_bytes = 0;
}

int dalloc(wg_context* ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 8;
  uint64_t used = ctx->in_use.load();
  if (ctx->limit && used + bytes > ctx->limit) {
    char buf[160];
    snprintf(buf, sizeof buf, "memory limit: %llu bytes in use, %zu more asked, limit %llu",
             (unsigned long long)used, bytes, (unsigned long long)ctx->limit);
    return fail(ctx, WG_MEMORY_LIMIT, buf);
  }
  auto hit = ctx->cache.find(bytes);
  if (hit != ctx->cache.end()) {
    *p = hit->second;
    ctx->cache_bytes -= bytes;
    ctx->cache.erase(hit);
    ctx->in_use += bytes;
    return WG_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && !ctx->cache.empty()) {
    // Give the cached blocks back and try once more.
    (void)hipGetLastError();
    flush_cache(ctx);
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    char buf[160];
    snprintf(buf, sizeof buf, "out of device memory allocating %zu bytes", bytes);
    return fail(ctx, WG_OUT_OF_MEMORY, buf);
  }
  ctx->in_use += bytes;
  return WG_OK;
}

// A freed block is cached once the context's stream has drained (no kernel
// or copy of it still uses the block; the next owner may write it from the
// host), else, or when too large, given back.
void dfree(wg_context* ctx, void* p, size_t bytes) {
  if (!p) return;
  if (bytes == 0) bytes = 8;
  ctx->in_use -= bytes;
  if (bytes <= kCacheBlock && hipStreamSynchronize(ctx->stream) == hipSuccess) {
    while (ctx->cache_bytes + bytes > kCacheTotal && !ctx->cache.empty()) {
      auto largest = std::prev(ctx->cache.end());
      (void)hipFree(largest->second);
      ctx->cache_bytes -= largest->first;
      ctx->cache.erase(largest);
    }
    ctx->cache.emplace(bytes, p);
    ctx->cache_bytes += bytes;
    return;
  }
  (void)hipFree(p);
}

int ensure_scratch(wg_context* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return WG_OK;
  if (ctx->scratch) {
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, ctx->scratch, ctx->scratch_bytes);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t want = std::max<size_t>(bytes, 1 << 20);
  int s = dalloc(ctx, &ctx->scratch, want);
  if (s) return s;
  ctx->scratch_bytes = want;
  return WG_OK;
}

__global__ void set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

int clear_error(wg_context* ctx) {
  set_u64<<<1, 1, 0, ctx->stream>>>(ctx->d_err, ~0ull);
  WG_HIP(ctx, hipGetLastError());
  return WG_OK;
}

static const char* status_text(int s) {
  switch (s) {
    case WG_FLOAT_OVERFLOW: return "value out of range: overflow";
    case WG_FLOAT_UNDERFLOW: return "value out of range: underflow";
    case WG_BIGINT_OUT_OF_RANGE: return "bigint out of range";
    case WG_DIVISION_BY_ZERO: return "division by zero";
  }
  return "kernel error";
}

int collect_error(wg_context* ctx, const char* what) {
  WG_HIP(ctx, hipMemcpyAsync(ctx->h_err, ctx->d_err, 8, hipMemcpyDeviceToHost, ctx->stream));
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  unsigned long long w = *ctx->h_err;
  if (w != ~0ull) {
    int status = (int)(w & 0xff);
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s (at index %llu)", what, status_text(status), w >> 8);
    return fail(ctx, status, buf);
  }
  return ok(ctx);
}

int wait_mail(wg_context* ctx, unsigned long long seq, unsigned long long* out, int k) {
  unsigned long long* m = ctx->h_mail;
  for (uint32_t spin = 1; __atomic_load_n(&m[0], __ATOMIC_ACQUIRE) != seq; ++spin) {
    if (spin % 64) continue;
    const hipError_t e = hipStreamQuery(ctx->stream);
    if (e == hipErrorNotReady) continue;
    if (e != hipSuccess) return hip_fail(ctx, e, "waiting for a kernel's results");
    // The stream drained: a post it held is visible by now.
    if (__atomic_load_n(&m[0], __ATOMIC_ACQUIRE) != seq)
      return fail(ctx, WG_DEVICE_ERROR, "a kernel ended without posting its results");
  }
  for (int i = 0; i < k; ++i) out[i] = __atomic_load_n(&m[1 + i], __ATOMIC_RELAXED);
  return WG_OK;
}

bool is_cancelled(wg_context* ctx) { return ctx->cancelled.load() != 0; }

template <class E>
__global__ void cast_weights(const double* __restrict__ w, E* __restrict__ out, uint64_t m,
                             unsigned long long* err) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < m; i += stride) {
    uint32_t e = 0;
    E v = weight_cast<E>(w[i], e);
    if (e) raise_error(err, i, e);
    out[i] = v;
  }
}

int weight_view(wg_context* ctx, DevCsr& csr, int dtype, const void** out) {
  *out = nullptr;
  if (!csr.weights) return WG_OK;
  if (dtype == WG_F64) {
    *out = csr.weights;
    return WG_OK;
  }
  if (csr.wstatus[dtype] >= 0) {
    *out = csr.wview[dtype];
    if (csr.wstatus[dtype] != WG_OK)
      return fail(ctx, csr.wstatus[dtype], std::string("edge weight cast: ") +
                                               status_text(csr.wstatus[dtype]));
    return WG_OK;
  }
  size_t es = dtype_size(dtype);
  void* p = nullptr;
  int s = dalloc(ctx, &p, csr.nnz * es);
  if (s) return s;
  if ((s = clear_error(ctx))) {
    dfree(ctx, p, csr.nnz * es);
    return s;
  }
  uint64_t m = csr.nnz;
  unsigned grid = (unsigned)std::min<uint64_t>((m + 255) / 256, 65536);
  if (grid == 0) grid = 1;
  switch (dtype) {
    case WG_F32: cast_weights<float><<<grid, 256, 0, ctx->stream>>>(csr.weights, (float*)p, m, ctx->d_err); break;
    case WG_I64: cast_weights<int64_t><<<grid, 256, 0, ctx->stream>>>(csr.weights, (int64_t*)p, m, ctx->d_err); break;
    case WG_BOOL: cast_weights<uint8_t><<<grid, 256, 0, ctx->stream>>>(csr.weights, (uint8_t*)p, m, ctx->d_err); break;
    case WG_U32: cast_weights<uint32_t><<<grid, 256, 0, ctx->stream>>>(csr.weights, (uint32_t*)p, m, ctx->d_err); break;
  }
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    dfree(ctx, p, csr.nnz * es);
    return hip_fail(ctx, le, "cast_weights launch");
  }
  s = collect_error(ctx, "edge weight cast");
  if (s == WG_DEVICE_ERROR || s == WG_OUT_OF_MEMORY) {
    // A runtime failure, not the cast's own verdict: nothing is cached.
    dfree(ctx, p, csr.nnz * es);
    return s;
  }
  csr.wview[dtype] = p;
  csr.wstatus[dtype] = s;
  *out = p;
  return s;
}

// Host partition of a CSR's rows into SpMV blocks.
static void partition_rows(const std::vector<uint64_t>& off, std::vector<Block>& blocks,
                           uint32_t& nlong, uint32_t& npartials, uint64_t& maxdeg) {
  uint64_t n = off.size() - 1;
  blocks.clear();
  nlong = npartials = 0;
  maxdeg = 0;
  uint64_t r = 0;
  while (r < n) {
    uint64_t deg = off[r + 1] - off[r];
    maxdeg = std::max(maxdeg, deg);
    if (deg > kBlockNnz) {
      uint32_t np = (uint32_t)((deg + kBlockNnz - 1) / kBlockNnz);
      for (uint32_t p = 0; p < np; ++p)
        blocks.push_back(Block{(uint32_t)r, 0, p, np, nlong, npartials, 0, 0});
      ++nlong;
      npartials += np;
      ++r;
      continue;
    }
    uint64_t r0 = r, nnz = 0;
    while (r < n && r - r0 < kBlockRows) {
      uint64_t d = off[r + 1] - off[r];
      if (d > kBlockNnz || nnz + d > kBlockNnz) break;
      maxdeg = std::max(maxdeg, d);
      nnz += d;
      ++r;
    }
    blocks.push_back(Block{(uint32_t)r0, (uint32_t)(r - r0), 0, 1, 0, 0, 0, 0});
  }
}

// The SpMV partition of a CSR whose offsets, targets and weights are on the
// device and whose h_offsets are set: its blocks, split rows and their
// partial slots.
static int partition_csr(wg_context* ctx, DevCsr& c) {
  int s;
  std::vector<Block> blocks;
  partition_rows(c.h_offsets, blocks, c.nlong, c.npartials, c.max_degree);
  c.nblocks = (uint32_t)blocks.size();
  if ((s = dalloc(ctx, (void**)&c.blocks, blocks.size() * sizeof(Block)))) return s;
  if (!blocks.empty())
    WG_HIP(ctx, hipMemcpyAsync(c.blocks, blocks.data(), blocks.size() * sizeof(Block),
                               hipMemcpyHostToDevice, ctx->stream));
  // The split rows and their partial slots, in block order.
  std::vector<uint32_t> lrows, lbase;
  std::vector<uint2> pieces;
  for (const Block& b : blocks)
    if (b.nrows == 0) {
      if (b.piece == 0) {
        lrows.push_back(b.row0);
        lbase.push_back(b.pbase);
      }
      pieces.push_back(make_uint2(b.row0, b.piece));
    }
  lbase.push_back(c.npartials);
  if ((s = dalloc(ctx, (void**)&c.pieces, std::max<size_t>(pieces.size(), 1) * sizeof(uint2)))) return s;
  if (!pieces.empty())
    WG_HIP(ctx, hipMemcpyAsync(c.pieces, pieces.data(), pieces.size() * sizeof(uint2), hipMemcpyHostToDevice,
                               ctx->stream));
  if ((s = dalloc(ctx, (void**)&c.long_rows, std::max<size_t>(lrows.size(), 1) * 4))) return s;
  if ((s = dalloc(ctx, (void**)&c.long_pbase, lbase.size() * 4))) return s;
  if (!lrows.empty())
    WG_HIP(ctx, hipMemcpyAsync(c.long_rows, lrows.data(), lrows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  WG_HIP(ctx, hipMemcpyAsync(c.long_pbase, lbase.data(), lbase.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  // The copies read host memory the caller may free on return.
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return WG_OK;
}

static int upload_csr(wg_context* ctx, DevCsr& c, uint64_t n, const uint64_t* off,
                      const uint32_t* tgt, const double* w) {
  c.n = n;
  c.nnz = off[n];
  c.h_offsets.assign(off, off + n + 1);
  int s;
  if ((s = dalloc(ctx, (void**)&c.offsets, (n + 1) * 8))) return s;
  if ((s = dalloc(ctx, (void**)&c.targets, c.nnz * 4))) return s;
  WG_HIP(ctx, hipMemcpyAsync(c.offsets, off, (n + 1) * 8, hipMemcpyHostToDevice, ctx->stream));
  if (c.nnz)
    WG_HIP(ctx, hipMemcpyAsync(c.targets, tgt, c.nnz * 4, hipMemcpyHostToDevice, ctx->stream));
  if (w) {
    if ((s = dalloc(ctx, (void**)&c.weights, c.nnz * 8))) return s;
    if (c.nnz)
      WG_HIP(ctx, hipMemcpyAsync(c.weights, w, c.nnz * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  return partition_csr(ctx, c);
}

int adopt_device_csr(wg_context* ctx, DevCsr& c) {
  c.h_offsets.resize(c.n + 1);
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  WG_HIP(ctx, hipMemcpy(c.h_offsets.data(), c.offsets, (c.n + 1) * 8, hipMemcpyDeviceToHost));
  return partition_csr(ctx, c);
}

void release_graph(wg_graph* g) {
  if (!g->ctx) return;
  wg_context* ctx = g->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  for (wg_graph* r : g->relations) release_graph(r);
  for (auto& v : g->selections) release_graph(v.second);
  free_csr(ctx, g->fwd);
  free_csr(ctx, g->rev);
  for (auto& w : g->workspace) dfree(ctx, w.first, w.second);
  g->workspace.clear();
  for (auto& c : g->cached) dfree(ctx, c.second.first, c.second.second);
  g->cached.clear();
  g->ctx = nullptr;
}

void destroy_graph(wg_graph* g) {
  release_graph(g);
  for (wg_graph* r : g->relations) destroy_graph(r);
  for (auto& v : g->selections) destroy_graph(v.second);
  delete g;
}

int check_graph(wg_context* ctx, const wg_graph* g, bool wide_ok) {
  if (!g) return fail(ctx, WG_INVALID_ARGUMENT, "null graph");
  if (g->ctx != ctx)
    return fail(ctx, WG_INVALID_ARGUMENT,
                g->ctx ? "the graph belongs to another context" : "the graph's context was destroyed");
  if (!wide_ok) return check_narrow(ctx, g);
  return WG_OK;
}

int check_vector(wg_context* ctx, const wg_vector* v) {
  if (!v) return fail(ctx, WG_INVALID_ARGUMENT, "null vector");
  if (v->ctx != ctx)
    return fail(ctx, WG_INVALID_ARGUMENT,
                v->ctx ? "the vector belongs to another context" : "the vector's context was destroyed");
  return WG_OK;
}

void free_csr(wg_context* ctx, DevCsr& c) {
  dfree(ctx, c.offsets, (c.n + 1) * 8);
  dfree(ctx, c.targets, c.nnz * 4);
  dfree(ctx, c.targets64, std::max<uint64_t>(c.nnz, 1) * 8);
  if (c.weights) dfree(ctx, c.weights, c.targets64 ? std::max<uint64_t>(c.nnz, 1) * 8 : c.nnz * 8);
  dfree(ctx, c.blocks, c.nblocks * sizeof(Block));
  dfree(ctx, c.long_rows, std::max<uint32_t>(c.nlong, 1) * 4);
  dfree(ctx, c.long_pbase, (c.nlong + 1) * 4);
  dfree(ctx, c.pieces, std::max<uint32_t>(c.npartials, 1) * sizeof(uint2));
  for (int d = 0; d < 5; ++d)
    if (c.wview[d]) dfree(ctx, c.wview[d], c.nnz * dtype_size(d));
  if (c.hot_state == 1) {
    dfree(ctx, c.hot_tgt, std::max<uint64_t>(c.nnz, 1) * 4);
    dfree(ctx, c.hot_inv, std::max<uint64_t>(c.hot_n, 1) * 4);
    dfree(ctx, c.hot_x, std::max<uint64_t>(c.hot_n, 1) * 8);
    dfree(ctx, c.hot_rank, std::max<uint64_t>(c.n, 1) * 4);
  }
  for (auto& t : c.lds_tgt) dfree(ctx, t.second.first, std::max<uint64_t>(c.nnz, 1) * 4);
  for (auto& t : c.tiles) {
    dfree(ctx, t.second.tagged, c.tile_padded * 4);
    dfree(ctx, t.second.warm_ids, (uint64_t)t.second.nwarm * 4);
    dfree(ctx, t.second.warm_x, (uint64_t)t.second.nwarm * t.first);
    for (int d = 0; d < 5; ++d)
      if (t.second.w[d]) dfree(ctx, t.second.w[d], c.tile_padded * dtype_size(d));
  }
  if (c.tile_state == 1) {
    dfree(ctx, c.tile_wave, (c.tile_nwaves + 1) * 4);
    dfree(ctx, c.tile_heads, std::max<uint32_t>(c.tile_nwaves, 1) * 4);
    dfree(ctx, c.nz_bits, ((c.n + 63) / 64 + 1) * 8);
    dfree(ctx, c.nz_pre, ((c.n + 63) / 64 + 1) * 4);
    dfree(ctx, c.fix_bits, ((c.n + 63) / 64 + 1) * 8);
  }
  if (c.tile_out) dfree(ctx, c.tile_out, c.tile_out_bytes);
  if (c.spmm_ready) {
    dfree(ctx, c.spmm_tasks, std::max<uint64_t>(c.spmm_ntasks, 1) * sizeof(SpmmTask));
    dfree(ctx, c.spmm_pbase, (c.spmm_nlong + 1) * 4);
    dfree(ctx, c.spmm_long_rows, std::max<uint32_t>(c.spmm_nlong, 1) * 4);
  }
  c = DevCsr();
}

int graph_buffer(wg_context* ctx, wg_graph* g, int slot, size_t bytes, void** out) {
  bytes = std::max<size_t>(bytes, 8);
  auto it = g->cached.find(slot);
  if (it != g->cached.end() && it->second.second >= bytes) {
    *out = it->second.first;
    return WG_OK;
  }
  if (it != g->cached.end()) {
    WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
    dfree(ctx, it->second.first, it->second.second);
    g->cached.erase(it);
  }
  void* p = nullptr;
  int s = dalloc(ctx, &p, bytes);
  if (s) return s;
  g->cached[slot] = {p, bytes};
  *out = p;
  return WG_OK;
}

int prepare_spmm(wg_context* ctx, DevCsr& c) {
  if (c.spmm_ready) return WG_OK;
  // Tasks by entry count: a counting pass, then each task placed in its
  // count's run, counts descending, rows and pieces ascending within one.
  std::vector<uint64_t> at(kSpmmPiece + 2, 0);
  std::vector<uint32_t> pbase, lrows;
  auto piece_count = [](uint64_t deg, uint32_t p) {
    return (uint32_t)std::min<uint64_t>(kSpmmPiece, deg - (uint64_t)p * kSpmmPiece);
  };
  uint64_t ntasks = 0;
  for (uint64_t r = 0; r < c.n; ++r) {
    const uint64_t deg = c.h_offsets[r + 1] - c.h_offsets[r];
    const uint32_t np = deg > kSpmmPiece ? (uint32_t)((deg + kSpmmPiece - 1) / kSpmmPiece) : 1;
    for (uint32_t p = 0; p < np; ++p) ++at[np > 1 ? piece_count(deg, p) : deg];
    ntasks += np;
  }
  for (uint64_t k = kSpmmPiece + 1, sum = 0; k-- > 0;) {
    const uint64_t cnt = at[k];
    at[k] = sum;
    sum += cnt;
  }
  std::vector<SpmmTask> tasks(ntasks);
  uint32_t np_total = 0;
  for (uint64_t r = 0; r < c.n; ++r) {
    const uint64_t rb = c.h_offsets[r], deg = c.h_offsets[r + 1] - rb;
    if (deg <= kSpmmPiece) {
      tasks[at[deg]++] = SpmmTask{rb | deg << kSpmmCountShift, (uint32_t)r, ~0u};
      continue;
    }
    const uint32_t np = (uint32_t)((deg + kSpmmPiece - 1) / kSpmmPiece);
    pbase.push_back(np_total);
    lrows.push_back((uint32_t)r);
    for (uint32_t p = 0; p < np; ++p) {
      const uint32_t cnt = piece_count(deg, p);
      tasks[at[cnt]++] = SpmmTask{(rb + (uint64_t)p * kSpmmPiece) | (uint64_t)cnt << kSpmmCountShift,
                                  (uint32_t)r, np_total + p};
    }
    np_total += np;
  }
  pbase.push_back(np_total);
  c.spmm_ntasks = ntasks;
  c.spmm_nlong = (uint32_t)lrows.size();
  c.spmm_npartials = np_total;
  int s;
  if ((s = dalloc(ctx, (void**)&c.spmm_tasks, std::max<uint64_t>(ntasks, 1) * sizeof(SpmmTask))) ||
      (s = dalloc(ctx, (void**)&c.spmm_pbase, (c.spmm_nlong + 1) * 4)) ||
      (s = dalloc(ctx, (void**)&c.spmm_long_rows, std::max<uint32_t>(c.spmm_nlong, 1) * 4))) {
    // Nothing stays charged from a failed preparation (dfree skips null).
    dfree(ctx, c.spmm_tasks, std::max<uint64_t>(ntasks, 1) * sizeof(SpmmTask));
    dfree(ctx, c.spmm_pbase, (c.spmm_nlong + 1) * 4);
    c.spmm_tasks = nullptr;
    c.spmm_pbase = nullptr;
    c.spmm_long_rows = nullptr;
    return s;
  }
  if (ntasks)
    WG_HIP(ctx, hipMemcpyAsync(c.spmm_tasks, tasks.data(), ntasks * sizeof(SpmmTask), hipMemcpyHostToDevice,
                               ctx->stream));
  WG_HIP(ctx, hipMemcpyAsync(c.spmm_pbase, pbase.data(), pbase.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  if (!lrows.empty())
    WG_HIP(ctx, hipMemcpyAsync(c.spmm_long_rows, lrows.data(), lrows.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  c.spmm_ready = true;
  return WG_OK;
}

__global__ __launch_bounds__(256) void set_weights(const uint64_t* __restrict__ edges,
                                                   const double* __restrict__ values,
                                                   const uint64_t* __restrict__ rev_pos, uint64_t count,
                                                   double* __restrict__ fwd_w, double* __restrict__ rev_w) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256) {
    const uint64_t e = edges[i];
    fwd_w[e] = values[i];
    rev_w[rev_pos[e]] = values[i];
  }
}

__global__ __launch_bounds__(256) void renumber_sources(const uint32_t* __restrict__ src,
                                                        const uint32_t* __restrict__ index,
                                                        uint32_t* __restrict__ out, uint64_t nnz) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * 256)
    out[i] = index[src[i]];
}

__global__ __launch_bounds__(256) void column_counts(const uint32_t* __restrict__ tgt, uint64_t nnz,
                                                     unsigned int* __restrict__ count) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nnz; i += (uint64_t)gridDim.x * 256)
    atomicAdd(&count[tgt[i]], 1u);
}

int prepare_hot_columns(wg_context* ctx, DevCsr& c) {
  if (c.hot_state) return WG_OK;
  const uint64_t n = c.n, m = c.nnz;
  hipStream_t st = ctx->stream;
  // How many entries gather each column (a one-time histogram).
  unsigned int* cnt_d;
  int s;
  if ((s = dalloc(ctx, (void**)&cnt_d, std::max<uint64_t>(n, 1) * 4))) {
    c.hot_state = -1;
    return ok(ctx);
  }
  WG_HIP(ctx, hipMemsetAsync(cnt_d, 0, std::max<uint64_t>(n, 1) * 4, st));
  if (m) column_counts<<<(unsigned)std::min<uint64_t>((m + 255) / 256, 16384), 256, 0, st>>>(c.targets, m, cnt_d);
  std::vector<uint32_t> cnt(n);
  if (n) WG_HIP(ctx, hipMemcpyAsync(cnt.data(), cnt_d, n * 4, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  // Ranks by descending count, ascending id within a count.
  uint32_t maxc = 0;
  for (uint32_t v : cnt) maxc = std::max(maxc, v);
  std::vector<uint64_t> next((size_t)maxc + 2, 0);
  for (uint32_t v : cnt) next[v]++;
  uint64_t pos = 0;
  for (uint64_t d = (uint64_t)maxc + 1; d-- > 1;) {
    const uint64_t k = next[d];
    next[d] = pos;
    pos += k;
  }
  std::vector<uint32_t> rank(n, ~0u), inv(pos);
  for (uint64_t v = 0; v < n; ++v)
    if (cnt[v]) {
      rank[v] = (uint32_t)next[cnt[v]]++;
      inv[rank[v]] = (uint32_t)v;
    }
  const uint64_t nh = pos;
  if ((s = dalloc(ctx, (void**)&c.hot_tgt, std::max<uint64_t>(m, 1) * 4))) {
    dfree(ctx, cnt_d, std::max<uint64_t>(n, 1) * 4);
    c.hot_tgt = nullptr;
    c.hot_state = -1;
    return ok(ctx);
  }
  if ((s = dalloc(ctx, (void**)&c.hot_inv, std::max<uint64_t>(nh, 1) * 4)) ||
      (s = dalloc(ctx, &c.hot_x, std::max<uint64_t>(nh, 1) * 8))) {
    dfree(ctx, cnt_d, std::max<uint64_t>(n, 1) * 4);
    dfree(ctx, c.hot_tgt, std::max<uint64_t>(m, 1) * 4);
    if (c.hot_inv) dfree(ctx, c.hot_inv, std::max<uint64_t>(nh, 1) * 4);
    c.hot_tgt = nullptr;
    c.hot_inv = nullptr;
    c.hot_x = nullptr;
    c.hot_state = -1;
    return ok(ctx);
  }
  // The rank table stays in the histogram's buffer.
  if (n) WG_HIP(ctx, hipMemcpyAsync(cnt_d, rank.data(), n * 4, hipMemcpyHostToDevice, st));
  if (nh) WG_HIP(ctx, hipMemcpyAsync(c.hot_inv, inv.data(), nh * 4, hipMemcpyHostToDevice, st));
  if (m)
    renumber_sources<<<2048, 256, 0, st>>>(c.targets, (const uint32_t*)cnt_d, c.hot_tgt, m);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(st));
  c.hot_rank = (uint32_t*)cnt_d;
  c.hot_n = nh;
  c.hot_state = 1;
  return WG_OK;
}

int prepare_hot_order(wg_context* ctx, wg_graph* g) {
  if (g->hot_src) return WG_OK;
  const uint64_t n = g->n;
  const std::vector<uint64_t>& fo = g->fwd.h_offsets;
  // Counting sort by out-degree, descending; ascending ids within a degree.
  const uint64_t maxd = g->fwd.max_degree;
  std::vector<uint64_t> next(maxd + 2, 0);
  for (uint64_t v = 0; v < n; ++v) next[fo[v + 1] - fo[v]]++;
  uint64_t pos = 0;
  for (uint64_t d = maxd + 1; d-- > 1;) {
    const uint64_t c = next[d];
    next[d] = pos;
    pos += c;
  }
  std::vector<uint32_t> index(n, ~0u);
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t d = fo[v + 1] - fo[v];
    if (d) index[v] = (uint32_t)next[d]++;
  }
  void *pi, *ps;
  int s;
  if ((s = graph_buffer(ctx, g, kSlotHotIndex, std::max<uint64_t>(n, 1) * 4, &pi))) return s;
  if ((s = graph_buffer(ctx, g, kSlotHotSrc, std::max<uint64_t>(g->nnz, 1) * 4, &ps))) return s;
  if (n) WG_HIP(ctx, hipMemcpyAsync(pi, index.data(), n * 4, hipMemcpyHostToDevice, ctx->stream));
  if (g->nnz)
    renumber_sources<<<2048, 256, 0, ctx->stream>>>(g->rev.targets, (const uint32_t*)pi, (uint32_t*)ps, g->nnz);
  WG_HIP(ctx, hipGetLastError());
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  g->hot_index = (uint32_t*)pi;
  g->hot_src = (uint32_t*)ps;
  g->nhot = pos;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" {

int wg_context_create(int device, uint64_t memory_limit, wg_context** out) {
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return WG_DEVICE_ERROR;
  if (device < 0 || device >= count) return WG_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return WG_DEVICE_ERROR;
  wg_context* ctx = new (std::nothrow) wg_context();
  if (!ctx) return WG_OUT_OF_MEMORY;
  ctx->device = device;
  ctx->limit = memory_limit;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void**)&ctx->d_err, 8) != hipSuccess || hipMalloc((void**)&ctx->d_jerr, 8) != hipSuccess ||
      hipMemset(ctx->d_jerr, 0xff, 8) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_err, 8, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_mail, kMailWords * 8, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void**)&ctx->d_mail, ctx->h_mail, 0) != hipSuccess ||
      static_scratch(device, &ctx->static_scratch) != WG_OK) {
    delete ctx;
    return WG_DEVICE_ERROR;
  }
  memset(ctx->h_mail, 0, kMailWords * 8);  // the cancel word starts lowered
  *out = ctx;
  return WG_OK;
}

void wg_context_destroy(wg_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  // Whatever the caller still holds loses its device memory now and keeps
  // a host record that its own destroy call frees without touching ctx.
  for (wg_bfs_cursor* c : ctx->cursors) release_cursor(c);
  for (wg_gathered* r : ctx->gathereds) release_gathered(r);
  for (wg_vector* v : ctx->vectors) {
    dfree(ctx, v->data, v->bytes);
    v->data = nullptr;
    v->ctx = nullptr;
  }
  for (wg_graph* g : ctx->graphs) release_graph(g);
  ctx->cursors.clear();
  ctx->gathereds.clear();
  ctx->vectors.clear();
  ctx->graphs.clear();
  flush_cache(ctx);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->d_err) (void)hipFree(ctx->d_err);
  if (ctx->d_jerr) (void)hipFree(ctx->d_jerr);
  if (ctx->h_err) (void)hipHostFree(ctx->h_err);
  if (ctx->h_mail) (void)hipHostFree(ctx->h_mail);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int wg_last_error(const wg_context* ctx, char* buf, size_t len) {
  if (!ctx) return WG_INVALID_ARGUMENT;
  if (buf && len) {
    size_t k = std::min(len - 1, ctx->last_msg.size());
    memcpy(buf, ctx->last_msg.data(), k);
    buf[k] = 0;
  }
  return ctx->last_status;
}

int wg_host_alloc(wg_context* ctx, uint64_t bytes, void** out) {
  if (!ctx || !out) return WG_INVALID_ARGUMENT;
  *out = nullptr;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
    *out = nullptr;
    (void)hipGetLastError();
    return fail(ctx, WG_OUT_OF_MEMORY, "page-locked host allocation of " + std::to_string(bytes) + " bytes");
  }
  return ok(ctx);
}

void wg_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

void wg_cancel(wg_context* ctx) {
  if (!ctx) return;
  ctx->cancelled = 1;
  if (ctx->h_mail) __atomic_store_n(&ctx->h_mail[kMailCancel], 1ull, __ATOMIC_RELEASE);
}
void wg_cancel_reset(wg_context* ctx) {
  if (!ctx) return;
  ctx->cancelled = 0;
  if (ctx->h_mail) __atomic_store_n(&ctx->h_mail[kMailCancel], 0ull, __ATOMIC_RELEASE);
}

void wg_set_progress(wg_context* ctx, const wg_progress* port) {
  if (!ctx) return;
  ctx->progress = port ? *port : wg_progress{};
}
uint64_t wg_memory_in_use(const wg_context* ctx) { return ctx ? ctx->in_use.load() : 0; }

int wg_set_memory_limit(wg_context* ctx, uint64_t limit) {
  if (!ctx) return WG_INVALID_ARGUMENT;
  ctx->limit = limit;
  return wg::ok(ctx);
}

int wg_synchronize(wg_context* ctx) {
  if (!ctx) return WG_INVALID_ARGUMENT;
  return collect_error(ctx, "deferred kernel");
}

void wg_set_graph_launch(wg_context* ctx, int on) {
  if (ctx) ctx->graph_launch = on != 0;
}

void wg_set_blocking(wg_context* ctx, int blocking) {
  if (ctx) ctx->blocking = blocking;
}

void* wg_stream(wg_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int wg_vector_create(wg_context* ctx, int dtype, uint64_t id_space, uint64_t width,
                     const void* host, wg_vector** out) {
  *out = nullptr;
  if (dtype < 0 || dtype > 4) return fail(ctx, WG_INVALID_ARGUMENT, "unknown dtype");
  if (width == 0) return fail(ctx, WG_INVALID_ARGUMENT, "zero width");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_vector* v = new (std::nothrow) wg_vector();
  if (!v) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  v->ctx = ctx;
  v->dtype = dtype;
  v->id_space = id_space;
  v->width = width;
  v->bytes = id_space * width * dtype_size(dtype);
  int s = dalloc(ctx, &v->data, v->bytes);
  if (s) {
    delete v;
    return s;
  }
  hipError_t e = host ? hipMemcpy(v->data, host, v->bytes, hipMemcpyHostToDevice)
                      : hipMemsetAsync(v->data, 0, v->bytes ? v->bytes : 8, ctx->stream);
  if (e == hipSuccess && !host) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    dfree(ctx, v->data, v->bytes);
    delete v;
    return hip_fail(ctx, e, "vector create");
  }
  ctx->vectors.insert(v);
  *out = v;
  return ok(ctx);
}

void wg_vector_destroy(wg_vector* v) {
  if (!v) return;
  if (v->ctx) {
    v->ctx->vectors.erase(v);
    (void)hipStreamSynchronize(v->ctx->stream);
    dfree(v->ctx, v->data, v->bytes);
  }
  delete v;
}

int wg_vector_write(wg_context* ctx, wg_vector* v, const void* host) {
  if (!ctx) return WG_INVALID_ARGUMENT;
  if (int s = check_vector(ctx, v)) return s;
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (v->bytes) WG_HIP(ctx, hipMemcpy(v->data, host, v->bytes, hipMemcpyHostToDevice));
  return ok(ctx);
}

int wg_vector_read(wg_context* ctx, const wg_vector* v, void* host) {
  if (!ctx) return WG_INVALID_ARGUMENT;
  if (int s = check_vector(ctx, v)) return s;
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (v->bytes) WG_HIP(ctx, hipMemcpy(host, v->data, v->bytes, hipMemcpyDeviceToHost));
  return ok(ctx);
}

int wg_vector_info(const wg_vector* v, int* dtype, uint64_t* id_space, uint64_t* width) {
  if (!v) return WG_INVALID_ARGUMENT;
  if (dtype) *dtype = v->dtype;
  if (id_space) *id_space = v->id_space;
  if (width) *width = v->width;
  return WG_OK;
}

void* wg_vector_device_ptr(wg_vector* v) { return v ? v->data : nullptr; }

// Csr::from_parts checks (graph/csr/mod.rs:78-137).
// Csr::from_parts' checks on the offsets (graph/csr/mod.rs:97-116); the
// per-row checks over the targets run on the device after the copy.
static int validate_offsets(wg_context* ctx, uint64_t n, const uint64_t* off, const uint32_t* tgt) {
  if (n >= (1ull << 32)) return fail(ctx, WG_INVALID_CSR, "id space exceeds u32 targets");
  if (off[0] != 0) return fail(ctx, WG_INVALID_CSR, "first offset is not zero");
  for (uint64_t r = 0; r < n; ++r)
    if (off[r + 1] < off[r]) return fail(ctx, WG_INVALID_CSR, "offsets not monotone at row " + std::to_string(r));
  if (off[n] && !tgt) return fail(ctx, WG_INVALID_CSR, "no targets");
  return WG_OK;
}

int wg_graph_upload(wg_context* ctx, uint64_t n, const uint64_t* off, const uint32_t* tgt,
                    const double* w, wg_graph** out) {
  *out = nullptr;
  if (!off) return fail(ctx, WG_INVALID_CSR, "no offsets");
  int s = validate_offsets(ctx, n, off, tgt);
  if (s) return s;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_graph* g = new (std::nothrow) wg_graph();
  if (!g) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  g->ctx = ctx;
  g->n = n;
  g->nnz = off[n];
  g->weighted = w != nullptr;
  s = upload_csr(ctx, g->fwd, n, off, tgt, w);
  if (!s) s = check_rows_device<uint32_t>(ctx, n, g->fwd.offsets, g->fwd.targets, g->nnz);
  // The reverse CSR, transposed on the device (graph/csr/transpose.rs:24-44).
  if (!s) s = build_reverse(ctx, g);
  if (s) {
    free_csr(ctx, g->fwd);
    free_csr(ctx, g->rev);
    delete g;
    return s;
  }
  ctx->graphs.insert(g);
  *out = g;
  return ok(ctx);
}

void wg_graph_destroy(wg_graph* g) {
  // A relation or selection goes with the graph that owns it.
  if (!g || g->owner) return;
  if (wg_context* ctx = g->ctx) {
    ctx->graphs.erase(g);
    // A cursor over this graph can read it no more: its device memory goes
    // too (it leaves the context's list, orphaned), and its next step
    // reports the graph gone.
    for (auto it = ctx->cursors.begin(); it != ctx->cursors.end();) {
      if (cursor_reads(*it, g)) {
        release_cursor(*it);
        it = ctx->cursors.erase(it);
      } else {
        ++it;
      }
    }
  }
  destroy_graph(g);
}

int wg_graph_info(const wg_graph* g, uint64_t* n, uint64_t* nnz, int* weighted) {
  if (!g) return WG_INVALID_ARGUMENT;
  if (n) *n = g->n;
  if (nnz) *nnz = g->nnz;
  if (weighted) *weighted = g->weighted;
  return WG_OK;
}

int wg_graph_read_forward(wg_context* ctx, const wg_graph* g, uint64_t* off, uint32_t* tgt, double* w) {
  if (!ctx || !g || !off) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int s = check_graph(ctx, g)) return s;
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  WG_HIP(ctx, hipMemcpy(off, g->fwd.offsets, (g->n + 1) * 8, hipMemcpyDeviceToHost));
  if (g->nnz) {
    if (tgt) WG_HIP(ctx, hipMemcpy(tgt, g->fwd.targets, g->nnz * 4, hipMemcpyDeviceToHost));
    if (w && g->fwd.weights)
      WG_HIP(ctx, hipMemcpy(w, g->fwd.weights, g->nnz * 8, hipMemcpyDeviceToHost));
  }
  return ok(ctx);
}

int wg_graph_read_reverse(wg_context* ctx, const wg_graph* g, uint64_t* off, uint32_t* tgt,
                          double* w) {
  if (!ctx || !g || !off || (g->nnz && !tgt)) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int s = check_graph(ctx, g)) return s;
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  WG_HIP(ctx, hipMemcpy(off, g->rev.offsets, (g->n + 1) * 8, hipMemcpyDeviceToHost));
  if (g->nnz) {
    WG_HIP(ctx, hipMemcpy(tgt, g->rev.targets, g->nnz * 4, hipMemcpyDeviceToHost));
    if (w && g->rev.weights)
      WG_HIP(ctx, hipMemcpy(w, g->rev.weights, g->nnz * 8, hipMemcpyDeviceToHost));
  }
  return ok(ctx);
}

int wg_graph_update_weights(wg_context* ctx, wg_graph* g, const uint64_t* edges, const double* values,
                            uint64_t count) {
  if (!ctx || !g || (count && (!edges || !values))) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int s = check_graph(ctx, g)) return s;
  if (!g->fwd.weights) return fail(ctx, WG_INVALID_ARGUMENT, "the graph has no edge weights");
  for (uint64_t i = 0; i < count; ++i)
    if (edges[i] >= g->nnz) return fail(ctx, WG_INVALID_ARGUMENT, "edge index outside the graph");
  if (count == 0) return ok(ctx);
  WG_HIP(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const uint64_t m = g->nnz;
  void* rp;
  int s;
  const bool built = g->cached.count(kSlotRevPos) != 0;
  if ((s = graph_buffer(ctx, g, kSlotRevPos, m * 8, &rp))) return s;
  // Each forward entry's place in the reverse CSR, computed once.
  if (!built && (s = reverse_position_map(ctx, g, (uint64_t*)rp))) return s;
  void *de, *dv;
  if ((s = dalloc(ctx, &de, count * 8))) return s;
  if ((s = dalloc(ctx, &dv, count * 8))) {
    dfree(ctx, de, count * 8);
    return s;
  }
  hipError_t e = hipMemcpyAsync(de, edges, count * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(dv, values, count * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const unsigned grid = (unsigned)std::min<uint64_t>((count + 255) / 256, 4096);
    set_weights<<<grid, 256, 0, st>>>((const uint64_t*)de, (const double*)dv, (const uint64_t*)rp, count,
                                      g->fwd.weights, g->rev.weights);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  dfree(ctx, de, count * 8);
  dfree(ctx, dv, count * 8);
  if (e != hipSuccess) return hip_fail(ctx, e, "update weights");
  // A weight changed in one direction breaks A = A^T.
  g->rev_is_fwd = false;
  // Weights cast to other element types, and every copy in the tiled
  // product's order, are made again on next use.
  for (DevCsr* c : {&g->fwd, &g->rev}) {
    for (int d = 0; d < 5; ++d) {
      if (d == WG_F64) continue;
      if (c->wview[d]) dfree(ctx, c->wview[d], c->nnz * dtype_size(d));
      c->wview[d] = nullptr;
      c->wstatus[d] = -1;
    }
    for (auto& t : c->tiles)
      for (int d = 0; d < 5; ++d) {
        if (t.second.w[d]) dfree(ctx, t.second.w[d], c->tile_padded * dtype_size(d));
        t.second.w[d] = nullptr;
      }
  }
  return ok(ctx);
}

}  // extern "C"

namespace wg {
// A streaming copy for measuring the achievable HBM rate: each thread keeps
// four 16-byte loads in flight, stores bypass the caches (non-temporal), and
// the grid is 32 workgroups of 256 threads per CU.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy16(u32x4* __restrict__ dst, const u32x4* __restrict__ src,
                                              uint64_t n16) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u32x4 a = __builtin_nontemporal_load(src + i), b = __builtin_nontemporal_load(src + i + stride),
                c = __builtin_nontemporal_load(src + i + 2 * stride),
                d = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + stride);
    __builtin_nontemporal_store(c, dst + i + 2 * stride);
    __builtin_nontemporal_store(d, dst + i + 3 * stride);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}
}  // namespace wg

extern "C" int wg_device_copy(wg_context* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (bytes & 15)) return wg::fail(ctx, WG_INVALID_ARGUMENT, "copy size must be a multiple of 16");
  uint64_t n16 = bytes / 16;
  int cus = 0;
  WG_HIP(ctx, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  unsigned grid = (unsigned)std::min<uint64_t>((n16 + 255) / 256, (uint64_t)std::max(cus, 1) * 32);
  if (grid == 0) return WG_OK;
  wg::copy16<<<grid, 256, 0, ctx->stream>>>((wg::u32x4*)dst, (const wg::u32x4*)src, n16);
  WG_HIP(ctx, hipGetLastError());
  if (ctx->blocking) WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return wg::ok(ctx);
}
