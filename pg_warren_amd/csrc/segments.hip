// The segment codec of a CSR (graph/segment/mod.rs, graph/segment/csr.rs):
// decoding WRSG segments straight into device memory, and encoding a
// device CSR back into them.
//
// A segment is a 16-byte little-endian header -- magic "WRSG", format (2
// for a CSR's segments), kind, element width, a reserved byte, the element
// count -- and its payload. Decoding follows AnyCsr::decode (csr.rs:84-127)
// and Csr::from_segments (graph/csr/mod.rs:78-137) check for check, and the
// first violation in their order is the error. The headers are read on
// the host; the payloads are copied to the device once, and the checks
// that read every element (stored rows ascending and non-empty, offsets
// monotone, each row's targets inside the id space and strictly ascending)
// run there as first-violation reductions. Sparse offsets (the rows
// segment) become dense offsets on the device, one binary search per row.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "wg_internal.hpp"

namespace wg {
namespace {

constexpr unsigned kThreads = 256;
constexpr uint64_t kHeader = 16;
// SegmentKind (graph/segment/mod.rs:47-80).
constexpr uint8_t kOffsets = 1, kTargets = 2, kWeights = 3, kRows = 11;
constexpr uint8_t kCsrFormat = 2;

unsigned grid_of(uint64_t items) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kThreads - 1) / kThreads, 8192));
}

std::string str(uint64_t v) { return std::to_string(v); }

int seg_fail(wg_context* ctx, const std::string& what) { return fail(ctx, WG_INVALID_SEGMENT, what); }
int csr_fail(wg_context* ctx, const std::string& what) { return fail(ctx, WG_INVALID_CSR, what); }

struct Header {
  uint8_t width = 0;
  uint64_t count = 0;
};

uint64_t le64(const uint8_t* p) {
  uint64_t v;
  std::memcpy(&v, p, 8);
  return v;
}

// read_header (mod.rs:322-343).
int read_header(wg_context* ctx, const uint8_t* b, uint64_t len, uint8_t kind, Header* h) {
  if (!b || len < kHeader) return seg_fail(ctx, "SegmentTruncated");
  if (std::memcmp(b, "WRSG", 4) != 0) return seg_fail(ctx, "SegmentMagic");
  if (b[4] != kCsrFormat) return seg_fail(ctx, "SegmentFormat: found " + str(b[4]));
  if (b[5] != kind) return seg_fail(ctx, "SegmentKind: expected " + str(kind) + ", found " + str(b[5]));
  if (b[7] != 0) return seg_fail(ctx, "SegmentReserved: found " + str(b[7]));
  h->width = b[6];
  h->count = le64(b + 8);
  return WG_OK;
}

// expect_width and expect_length (mod.rs:346-365).
int check_elements(wg_context* ctx, const Header& h, uint64_t payload, unsigned width) {
  if (h.width != width)
    return seg_fail(ctx, "SegmentWidth: expected " + str(width) + ", found " + str(h.width));
  const unsigned __int128 expected = (unsigned __int128)h.count * width;
  if (expected != payload) {
    const uint64_t e = expected > UINT64_MAX ? UINT64_MAX : (uint64_t)expected;
    return seg_fail(ctx, "SegmentLength: expected " + str(e) + ", found " + str(payload));
  }
  return WG_OK;
}

__device__ __forceinline__ uint64_t first_index() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t stride() { return (uint64_t)gridDim.x * blockDim.x; }

// *first = min i in [1, count) with a[i] <= a[i-1] (strict) or a[i] <
// a[i-1] (not strict). A thread's first hit is its lowest.
template <class T>
__global__ __launch_bounds__(kThreads) void first_descent(const T* __restrict__ a, uint64_t count, bool strict,
                                                          unsigned long long* __restrict__ first) {
  for (uint64_t i = first_index() + 1; i < count; i += stride()) {
    const bool bad = strict ? a[i] <= a[i - 1] : a[i] < a[i - 1];
    if (bad) {
      atomicMin(first, (unsigned long long)i);
      return;
    }
  }
}

// dense[r] = sparse[the first stored row at or after r], r in [0, n].
__global__ __launch_bounds__(kThreads) void densify(const uint32_t* __restrict__ rows, uint64_t R,
                                                    const uint64_t* __restrict__ sparse, uint64_t n,
                                                    uint64_t* __restrict__ dense) {
  for (uint64_t r = first_index(); r <= n; r += stride()) {
    uint64_t lo = 0, hi = R;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((uint64_t)rows[mid] < r)
        lo = mid + 1;
      else
        hi = mid;
    }
    dense[r] = sparse[lo];
  }
}

// The first rows (ascending) holding a target outside the id space and a
// pair not strictly ascending: bad[0], bad[1].
__global__ __launch_bounds__(kThreads) void check_entries(const uint32_t* __restrict__ row_of,
                                                          const uint64_t* __restrict__ off,
                                                          const uint32_t* __restrict__ tgt, uint64_t m,
                                                          uint64_t n, unsigned long long* __restrict__ bad) {
  for (uint64_t e = first_index(); e < m; e += stride()) {
    const uint32_t r = row_of[e];
    const uint32_t t = tgt[e];
    if (t >= n) atomicMin(&bad[0], (unsigned long long)r);
    if (e > off[r] && tgt[e - 1] >= t) atomicMin(&bad[1], (unsigned long long)r);
  }
}

// Segment encoding: a payload's little-endian copy after its header.
void put_header(uint8_t* out, uint8_t kind, uint8_t width, uint64_t count) {
  std::memcpy(out, "WRSG", 4);
  out[4] = kCsrFormat;
  out[5] = kind;
  out[6] = width;
  out[7] = 0;
  std::memcpy(out + 8, &count, 8);
}

}  // namespace

// One CSR's segments decoded and checked into c (device arrays, dense
// offsets), then partitioned.
static int decode_csr(wg_context* ctx, uint64_t n, const wg_csr_segments* sg, DevCsr& c, bool* weighted) {
  hipStream_t st = ctx->stream;
  int s;
  // AnyCsr::decode (csr.rs:90-127): offsets, weights, the rows' and the
  // targets' headers, then the elements at the id space's width (u32).
  Header ho, hw, hr, ht;
  if ((s = read_header(ctx, sg->offsets, sg->offsets_len, kOffsets, &ho))) return s;
  if ((s = check_elements(ctx, ho, sg->offsets_len - kHeader, 8))) return s;
  if (sg->weights) {
    if ((s = read_header(ctx, sg->weights, sg->weights_len, kWeights, &hw))) return s;
    if ((s = check_elements(ctx, hw, sg->weights_len - kHeader, 8))) return s;
  }
  if (sg->rows && (s = read_header(ctx, sg->rows, sg->rows_len, kRows, &hr))) return s;
  if ((s = read_header(ctx, sg->targets, sg->targets_len, kTargets, &ht))) return s;
  if ((s = check_elements(ctx, ht, sg->targets_len - kHeader, 4))) return s;
  if (sg->rows && (s = check_elements(ctx, hr, sg->rows_len - kHeader, 4))) return s;
  // Csr::from_segments (graph/csr/mod.rs:78-137).
  const uint64_t nrows = sg->rows ? hr.count : 0;
  const uint64_t expected = sg->rows ? nrows + 1 : n + 1;
  if (ho.count != expected)
    return csr_fail(ctx, "OffsetsLength: expected " + str(expected) + ", found " + str(ho.count));
  const uint8_t* op = sg->offsets + kHeader;
  if (le64(op) != 0) return csr_fail(ctx, "FirstOffsetNotZero");
  const uint64_t m = ht.count;
  c.n = n;
  c.nnz = m;
  if ((s = dalloc(ctx, (void**)&c.offsets, (n + 1) * 8))) return s;
  if ((s = dalloc(ctx, (void**)&c.targets, m * 4))) return s;
  if (m) WG_HIP(ctx, hipMemcpyAsync(c.targets, sg->targets + kHeader, m * 4, hipMemcpyHostToDevice, st));
  unsigned long long* flags = nullptr;
  if ((s = dalloc(ctx, (void**)&flags, 4 * 8))) return s;
  struct Flags {
    wg_context* ctx;
    unsigned long long* p;
    ~Flags() { dfree(ctx, p, 4 * 8); }
  } keep{ctx, flags};
  auto reset = [&]() { return hipMemsetAsync(flags, 0xff, 4 * 8, st); };
  auto read = [&](unsigned long long* h, int k) {
    hipError_t e = hipMemcpyAsync(h, flags, k * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
  };
  unsigned long long hf[4];
  if (sg->rows) {
    // check_stored_rows (graph/csr/mod.rs:239-258): ascending, inside the
    // id space (the last), each non-empty.
    const uint8_t* rp = sg->rows + kHeader;
    uint32_t* drows = nullptr;
    uint64_t* dsparse = nullptr;
    if ((s = dalloc(ctx, (void**)&drows, std::max<uint64_t>(nrows, 1) * 4))) return s;
    if ((s = dalloc(ctx, (void**)&dsparse, (nrows + 1) * 8))) {
      dfree(ctx, drows, std::max<uint64_t>(nrows, 1) * 4);
      return s;
    }
    struct Free2 {
      wg_context* ctx;
      void *a, *b;
      size_t na, nb;
      ~Free2() {
        dfree(ctx, a, na);
        dfree(ctx, b, nb);
      }
    } f2{ctx, drows, dsparse, std::max<uint64_t>(nrows, 1) * 4, (nrows + 1) * 8};
    if (nrows) WG_HIP(ctx, hipMemcpyAsync(drows, rp, nrows * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(dsparse, op, (nrows + 1) * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, reset());
    first_descent<uint32_t><<<grid_of(nrows), kThreads, 0, st>>>(drows, nrows, true, flags);
    first_descent<uint64_t><<<grid_of(nrows + 1), kThreads, 0, st>>>(dsparse, nrows + 1, true, flags + 1);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, read(hf, 2));
    if (hf[0] != ~0ull) return csr_fail(ctx, "StoredRowsNotAscending: index " + str(hf[0]));
    if (nrows) {
      uint32_t last;
      std::memcpy(&last, rp + (nrows - 1) * 4, 4);
      if (last >= n) return csr_fail(ctx, "StoredRowOutsideIdSpace: row " + str(last));
    }
    if (hf[1] != ~0ull) {
      uint32_t row;
      std::memcpy(&row, rp + (hf[1] - 1) * 4, 4);
      return csr_fail(ctx, "StoredRowEmpty: row " + str(row));
    }
    // Non-empty stored rows make the offsets strictly ascending.
    const uint64_t last_off = le64(op + nrows * 8);
    if (last_off != m) return csr_fail(ctx, "FinalOffset: expected " + str(m) + ", found " + str(last_off));
    densify<<<grid_of(n + 1), kThreads, 0, st>>>(drows, nrows, dsparse, n, c.offsets);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
  } else {
    WG_HIP(ctx, hipMemcpyAsync(c.offsets, op, (n + 1) * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, reset());
    first_descent<uint64_t><<<grid_of(n + 1), kThreads, 0, st>>>(c.offsets, n + 1, false, flags);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, read(hf, 1));
    if (hf[0] != ~0ull) return csr_fail(ctx, "OffsetsNotMonotone: row " + str(hf[0] - 1));
    const uint64_t last_off = le64(op + n * 8);
    if (last_off != m) return csr_fail(ctx, "FinalOffset: expected " + str(m) + ", found " + str(last_off));
  }
  // The rows' entries: the first row with a target outside the id space or
  // a pair out of order, the former first within a row.
  if ((s = adopt_device_csr(ctx, c))) return s;
  if (m) {
    uint32_t* row_of = nullptr;
    if ((s = dalloc(ctx, (void**)&row_of, m * 4))) return s;
    struct Free1 {
      wg_context* ctx;
      void* a;
      size_t na;
      ~Free1() { dfree(ctx, a, na); }
    } f1{ctx, row_of, m * 4};
    if ((s = expand_rows(ctx, c, row_of))) return s;
    WG_HIP(ctx, reset());
    check_entries<<<grid_of(m), kThreads, 0, st>>>(row_of, c.offsets, c.targets, m, n, flags);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, read(hf, 2));
    if (hf[0] != ~0ull && hf[0] <= hf[1]) {
      const uint64_t r = hf[0];
      const uint32_t* tp = (const uint32_t*)(sg->targets + kHeader);
      uint64_t target = 0;
      for (uint64_t p = c.h_offsets[r]; p < c.h_offsets[r + 1]; ++p) {
        uint32_t t;
        std::memcpy(&t, tp + p, 4);
        if (t >= n) {
          target = t;
          break;
        }
      }
      return csr_fail(ctx, "TargetOutsideIdSpace: row " + str(r) + ", target " + str(target));
    }
    if (hf[1] != ~0ull) return csr_fail(ctx, "RowNotSorted: row " + str(hf[1]));
  }
  if (sg->weights) {
    if (hw.count != m) return csr_fail(ctx, "WeightsLength: expected " + str(m) + ", found " + str(hw.count));
    if ((s = dalloc(ctx, (void**)&c.weights, m * 8))) return s;
    if (m) WG_HIP(ctx, hipMemcpyAsync(c.weights, sg->weights + kHeader, m * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
  }
  *weighted = sg->weights != nullptr;
  return WG_OK;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_graph_load_segments(wg_context* ctx, uint64_t id_space, const wg_csr_segments* forward,
                                      const wg_csr_segments* reverse, wg_graph** out) {
  if (!ctx || !forward || !out) return WG_INVALID_ARGUMENT;
  *out = nullptr;
  if (!forward->offsets || !forward->targets) return fail(ctx, WG_INVALID_ARGUMENT, "null segment");
  // TargetWidth::for_id_space (graph/csr/target.rs:25-31): u64 targets
  // beyond a 2^32 id space, which this build does not hold.
  if (id_space >= (1ull << 32)) return fail(ctx, WG_UNSUPPORTED, "id space beyond u32 targets");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  wg_graph* g = new (std::nothrow) wg_graph();
  if (!g) return fail(ctx, WG_OUT_OF_MEMORY, "host allocation");
  g->ctx = ctx;
  g->n = id_space;
  bool weighted = false, rweighted = false;
  int s = decode_csr(ctx, id_space, forward, g->fwd, &weighted);
  g->weighted = weighted;
  g->nnz = g->fwd.nnz;
  if (!s && reverse) {
    s = decode_csr(ctx, id_space, reverse, g->rev, &rweighted);
    // A stored reverse CSR is taken as the transpose, as the reference
    // takes it; only its edge count and weights must match.
    if (!s && (g->rev.nnz != g->nnz || rweighted != weighted))
      s = fail(ctx, WG_INVALID_CSR, "the reverse CSR does not hold the forward CSR's edges");
    bool same = false;
    if (!s) s = same_csr(ctx, g->fwd, g->rev, &same);
    g->rev_is_fwd = same;
  } else if (!s) {
    s = build_reverse(ctx, g);
  }
  if (s) {
    destroy_graph(g);
    return s;
  }
  ctx->graphs.insert(g);
  *out = g;
  return ok(ctx);
}

extern "C" int wg_csr_segment_lengths(uint64_t id_space, uint64_t nnz, uint64_t* offsets_len,
                                      uint64_t* targets_len, uint64_t* weights_len) {
  if (offsets_len) *offsets_len = kHeader + (id_space + 1) * 8;
  if (targets_len) *targets_len = kHeader + nnz * 4;
  if (weights_len) *weights_len = kHeader + nnz * 8;
  return WG_OK;
}

extern "C" int wg_graph_encode_segments(wg_context* ctx, const wg_graph* g, int reverse, uint8_t* offsets,
                                        uint8_t* targets, uint8_t* weights) {
  if (!ctx || !g) return WG_INVALID_ARGUMENT;
  if (int cs_ = check_graph(ctx, g)) return cs_;
  if (weights && !g->weighted) return fail(ctx, WG_INVALID_ARGUMENT, "the graph has no weights");
  const DevCsr& c = reverse ? g->rev : g->fwd;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // csr.rs:130-137, dense offsets: offsets, targets (u32) and weights.
  if (offsets) {
    put_header(offsets, kOffsets, 8, g->n + 1);
    WG_HIP(ctx, hipMemcpy(offsets + kHeader, c.offsets, (g->n + 1) * 8, hipMemcpyDeviceToHost));
  }
  if (targets) {
    put_header(targets, kTargets, 4, c.nnz);
    if (c.nnz) WG_HIP(ctx, hipMemcpy(targets + kHeader, c.targets, c.nnz * 4, hipMemcpyDeviceToHost));
  }
  if (weights) {
    put_header(weights, kWeights, 8, c.nnz);
    if (c.nnz) WG_HIP(ctx, hipMemcpy(weights + kHeader, c.weights, c.nnz * 8, hipMemcpyDeviceToHost));
  }
  return ok(ctx);
}

// ---- a relation's CSR in the reference's layout ---------------------------
// stores_sparse (graph/csr/sparse.rs:10-40): sparse offsets (a u32 row id
// and a u64 offset per stored row, and the end offset) when they take at
// most a quarter of the bytes of dense offsets; a graph's only relation is
// its all-relations CSR, whose offsets are always dense.
static bool stores_sparse(uint64_t stored, uint64_t id_space) {
  const uint64_t sparse = stored * (4 + 8) + 8, dense = (id_space + 1) * 8;
  return sparse * 4 <= dense;
}

static int relation_csr(wg_context* ctx, const wg_graph* g, uint32_t relation, int reverse, const DevCsr** out,
                        bool* sparse, uint64_t* stored) {
  if (int cs_ = check_graph(ctx, g)) return cs_;
  if (relation >= g->relation_count)
    return fail(ctx, WG_UNKNOWN_RELATION, "relation " + std::to_string(relation) + " outside the dictionary");
  const wg_graph* r = g->relation_count > 1 ? g->relations[relation] : g;
  const DevCsr& c = reverse ? r->rev : r->fwd;
  uint64_t k = 0;
  for (uint64_t v = 0; v < c.n; ++v) k += c.h_offsets[v + 1] > c.h_offsets[v];
  *out = &c;
  *stored = k;
  *sparse = g->relation_count > 1 && stores_sparse(k, g->n);
  return WG_OK;
}

extern "C" int wg_relation_segment_lengths(wg_context* ctx, const wg_graph* g, uint32_t relation, int reverse,
                                           uint64_t* rows_len, uint64_t* offsets_len, uint64_t* targets_len,
                                           uint64_t* weights_len) {
  if (!ctx || !g) return WG_INVALID_ARGUMENT;
  const DevCsr* c = nullptr;
  bool sparse = false;
  uint64_t k = 0;
  if (int s = relation_csr(ctx, g, relation, reverse, &c, &sparse, &k)) return s;
  if (rows_len) *rows_len = sparse ? kHeader + k * 4 : 0;
  if (offsets_len) *offsets_len = kHeader + (sparse ? k + 1 : g->n + 1) * 8;
  if (targets_len) *targets_len = kHeader + c->nnz * 4;
  if (weights_len) *weights_len = g->weighted ? kHeader + c->nnz * 8 : 0;
  return ok(ctx);
}

extern "C" int wg_relation_encode_segments(wg_context* ctx, const wg_graph* g, uint32_t relation, int reverse,
                                           uint8_t* rows, uint8_t* offsets, uint8_t* targets, uint8_t* weights) {
  if (!ctx || !g) return WG_INVALID_ARGUMENT;
  const DevCsr* cp = nullptr;
  bool sparse = false;
  uint64_t k = 0;
  if (int s = relation_csr(ctx, g, relation, reverse, &cp, &sparse, &k)) return s;
  if (weights && !g->weighted) return fail(ctx, WG_INVALID_ARGUMENT, "the graph has no weights");
  const DevCsr& c = *cp;
  WG_HIP(ctx, hipSetDevice(ctx->device));
  WG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const std::vector<uint64_t>& off = c.h_offsets;
  // encode_csr (csr.rs:123-130): rows (sparse only), offsets, targets, weights.
  if (sparse) {
    if (rows) {
      put_header(rows, kRows, 4, k);
      uint8_t* p = rows + kHeader;
      for (uint64_t v = 0; v < c.n; ++v)
        if (off[v + 1] > off[v]) {
          const uint32_t id = (uint32_t)v;
          memcpy(p, &id, 4);
          p += 4;
        }
    }
    if (offsets) {
      put_header(offsets, kOffsets, 8, k + 1);
      uint8_t* p = offsets + kHeader;
      for (uint64_t v = 0; v < c.n; ++v)
        if (off[v + 1] > off[v]) {
          memcpy(p, &off[v], 8);
          p += 8;
        }
      memcpy(p, &off[c.n], 8);
    }
  } else if (offsets) {
    put_header(offsets, kOffsets, 8, g->n + 1);
    memcpy(offsets + kHeader, off.data(), (g->n + 1) * 8);
  }
  if (targets) {
    put_header(targets, kTargets, 4, c.nnz);
    if (c.nnz) WG_HIP(ctx, hipMemcpy(targets + kHeader, c.targets, c.nnz * 4, hipMemcpyDeviceToHost));
  }
  if (weights) {
    put_header(weights, kWeights, 8, c.nnz);
    if (c.nnz) WG_HIP(ctx, hipMemcpy(weights + kHeader, c.weights, c.nnz * 8, hipMemcpyDeviceToHost));
  }
  return ok(ctx);
}
