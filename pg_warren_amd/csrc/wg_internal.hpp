// Host-side internals of libwarren_gpu: the context (stream, memory budget,
// cancellation, error word), device vectors and the device CSR pair.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <string>
#include <map>
#include <set>
#include <vector>

#include "../../include/warren_gpu.h"

namespace wg {

// A row block of the SpMV partition: either up to kBlockRows whole rows
// whose entries total at most kBlockNnz (nrows > 0), or one piece of a row
// too long for a block (nrows == 0, piece < npieces). A split row's pieces
// each write a partial to slot pbase + piece; spmv_long folds them in piece
// order.
struct Block {
  uint32_t row0;
  uint32_t nrows;
  uint32_t piece;
  uint32_t npieces;
  uint32_t lrow;    // long rows: index of the row among the split rows
  uint32_t pbase;   // long rows: first partial slot of the row
  uint32_t pad0, pad1;
};

constexpr uint32_t kBlockThreads = 256;
constexpr size_t kMailWords = 32;  // wg_context mailbox: seq + 30 values + the cancel word
// The mailbox's last word: 1 once wg_cancel is called, polled by persistent
// kernels (a host-mapped word, visible to a running kernel).
constexpr size_t kMailCancel = kMailWords - 1;
constexpr uint32_t kBlockNnz = 2048;   // entries staged in LDS per block
constexpr uint32_t kBlockRows = 512;   // rows a short block may hold

// A CSR resident in device memory, with its SpMV partition and the edge
// weights cast once per element type (SemiringElement::from_weight,
// compute/semiring.rs:243-246); a cast that fails keeps its status.
// A SpMM task: entries [e0, e0 + cnt) of `row`, packed as e0 | cnt << 55
// (cnt <= kSpmmPiece); slot ~0 for a whole row, else the partial slot of
// this piece of a split row.
struct SpmmTask {
  uint64_t ec;
  uint32_t row, slot;
};
constexpr int kSpmmCountShift = 55;

struct DevCsr {
  uint64_t n = 0, nnz = 0;
  uint64_t* offsets = nullptr;   // n + 1
  uint32_t* targets = nullptr;   // nnz
  uint64_t* targets64 = nullptr; // nnz, instead of targets: u64 targets (wg_graph_upload_u64)
  double* weights = nullptr;     // nnz or null
  std::vector<uint64_t> h_offsets;  // host copy of the offsets (partitioning)
  Block* blocks = nullptr;
  uint32_t nblocks = 0;
  uint32_t nlong = 0;        // rows split into pieces
  uint32_t npartials = 0;    // pieces of split rows
  uint32_t* long_rows = nullptr;   // nlong: the row of each split row
  uint32_t* long_pbase = nullptr;  // nlong + 1: its partial slots
  uint2* pieces = nullptr;         // npartials: (row, piece) of each split-row piece, in block order
  uint64_t max_degree = 0;
  void* wview[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int wstatus[5] = {-1, -1, -1, -1, -1};  // -1: not cast yet
  // SpMM: long rows split into pieces of at most kSpmmPiece entries; one
  // task per short row or piece, the tasks ordered by entry count (most
  // first, then by row and piece) so that the rows a wavefront folds
  // together are about as long.
  SpmmTask* spmm_tasks = nullptr;
  uint64_t spmm_ntasks = 0;
  uint32_t* spmm_pbase = nullptr;          // nlong + 1: partial slots per split row
  uint32_t* spmm_long_rows = nullptr;      // nlong: the row of each split row
  uint32_t spmm_nlong = 0, spmm_npartials = 0;
  bool spmm_ready = false;
  // Autotuned SpMV launch variant per (dtype, monoid, multiply, weighted).
  std::map<int, int> tuned;
  // Autotuned SpMM launch shape per (dtype, monoid, multiply, weighted,
  // width): lanes per row | 256 when a lane takes 16-byte column groups.
  std::map<uint64_t, int> spmm_tuned;
  // Hot-first column numbering for the gathers (built on first use): the
  // gathered columns ranked by how many entries read them (descending, ties
  // by id), the targets in that numbering, the original id of each rank,
  // and room for x permuted into it. hot_state: 0 not built, 1 ready,
  // -1 not available (the memory budget refused it).
  uint32_t* hot_tgt = nullptr;
  uint32_t* hot_inv = nullptr;
  uint32_t* hot_rank = nullptr;  // n: each column's rank (~0: never gathered)
  uint64_t hot_n = 0;
  void* hot_x = nullptr;
  int hot_state = 0;
  // Targets tagged for the LDS hot-column cache, per element size: bit 31
  // and the rank for the first `nhot` ranks, the column id otherwise.
  std::map<int, std::pair<uint32_t*, uint32_t>> lds_tgt;
  // The tiled SpMV's layout (spmv_tiles.hpp), built on first use: per
  // element size the tagged targets in tile order and the weights of each
  // element type in that order; per CSR the waves' tile runs and row starts
  // before them, the non-empty-row bitmap with a count per 64 rows, and room
  // for the finished rows and the waves' records. tile_state: 0 not built,
  // 1 ready, -1 not available (id space of 2^30 or more, or the budget).
  struct Tiles {
    uint32_t* tagged = nullptr;
    uint32_t nhot = 0;
    uint32_t nwarm = 0;           // columns in the warm run
    uint32_t* warm_ids = nullptr;  // nwarm columns, ascending
    void* warm_x = nullptr;        // their x, nwarm elements
    void* w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  };
  std::map<int, Tiles> tiles;
  uint64_t tile_padded = 0;        // entries including the last tile's padding
  uint32_t* tile_wave = nullptr;   // tile_nwaves + 1
  uint32_t* tile_heads = nullptr;  // tile_nwaves
  uint32_t tile_nwaves = 0;
  uint64_t* nz_bits = nullptr;     // (n + 63) / 64 words
  uint32_t* nz_pre = nullptr;      // (n + 63) / 64
  uint64_t* fix_bits = nullptr;    // (n + 63) / 64 words: rows joined from wave records
  uint64_t nz_rows = 0;
  void* tile_out = nullptr;        // finished rows and wave records
  size_t tile_out_bytes = 0;
  int tile_state = 0;
};

constexpr uint32_t kSpmmPiece = 256;

}  // namespace wg

struct wg_context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t limit = 0;
  std::atomic<uint64_t> in_use{0};
  std::atomic<int> cancelled{0};
  wg_progress progress{};  // where long work reports (all null: nowhere)
  int blocking = 1;
  int last_status = 0;
  std::string last_msg;
  unsigned long long* d_err = nullptr;  // (index << 8 | status), ~0 when clear
  unsigned long long* d_jerr = nullptr; // the tiled SpMV's error word, by row slot
  unsigned long long* h_err = nullptr;  // pinned mirror
  // Coherent pinned words a kernel posts a few results into: mail[0] the
  // sequence number, then the values; the host polls (wait_mail) instead of
  // copying and synchronising.
  unsigned long long* h_mail = nullptr;  // host view
  unsigned long long* d_mail = nullptr;  // device view
  unsigned long long mail_seq = 0;
  // Freed device blocks kept for reuse, by exact size (hipFree waits for
  // the whole device; a search frees some twenty buffers).
  std::multimap<size_t, void*> cache;
  size_t cache_bytes = 0;
  void* scratch = nullptr;              // partial results of split rows
  size_t scratch_bytes = 0;
  int spmv_variant = -1;  // -1: autotune; else the launch variant to force
  int spmm_shape = -1;    // -1: autotune; else the SpMM launch shape to force
  int graph_launch = 1;   // replay iterative launch sequences as captured HIP graphs
  void* static_scratch = nullptr;  // this device's copy of wg_spmv_noalloc's partials
  // The live objects handed out through this context. wg_context_destroy
  // releases their device memory and orphans them (their ctx becomes null),
  // so destroying one of them afterwards only frees its host record.
  std::set<wg_graph*> graphs;
  std::set<wg_vector*> vectors;
  std::set<wg_bfs_cursor*> cursors;
  std::set<wg_gathered*> gathereds;
};

struct wg_vector {
  wg_context* ctx = nullptr;
  int dtype = 0;
  uint64_t id_space = 0, width = 0;
  void* data = nullptr;
  size_t bytes = 0;
};

struct wg_graph {
  wg_context* ctx = nullptr;
  uint64_t n = 0, nnz = 0;
  bool weighted = false;
  bool rev_is_fwd = false;  // the reverse CSR equals the forward one (A = A^T)
  wg::DevCsr fwd, rev;
  // Search workspace sized by the graph, allocated on first use and kept
  // (charged to the context) until the graph is destroyed.
  std::vector<std::pair<void*, size_t>> workspace;
  void* bfs_ws[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t bfs_ws_bytes[7] = {};
  // Other kernels' per-graph buffers by slot: (pointer, bytes).
  std::map<int, std::pair<void*, size_t>> cached;
  // Hot-first numbering of message sources (built on first use): nodes with
  // out-edges ranked by descending out-degree (ties by id), hot_index[v]
  // their rank (~0 for nodes without out-edges), and the reverse CSR's
  // sources in that numbering. The most gathered messages then share the
  // fewest cache lines.
  uint32_t* hot_index = nullptr;
  uint32_t* hot_src = nullptr;
  uint64_t nhot = 0;
  // The batched PageRank's work items over the reverse CSR (built on first
  // use): runs of whole rows and pieces of long rows, the long rows and
  // their first piece slots.
  void* ppr_items = nullptr;
  uint32_t* ppr_lrows = nullptr;
  uint32_t* ppr_lbase = nullptr;
  uint32_t ppr_nitems = 0, ppr_npieces = 0, ppr_nlong = 0;
  // Relations (GraphCsr's relation CSRs, graph/csr/graph_csr.rs:15-153):
  // with two or more, each relation's CSR pair is a graph of its own and
  // fwd/rev above are the merged all-relations pair; with one, fwd/rev are
  // that relation's.
  uint32_t relation_count = 1;
  std::vector<wg_graph*> relations;
  // Relation selections built for products and searches, by key.
  std::map<std::string, wg_graph*> selections;
  // Relations and selections are owned by their graph and go with it.
  wg_graph* owner = nullptr;
};

namespace wg {
// A device buffer of at least `bytes` kept with the graph under `slot`
// (charged to the context, released with the graph).
int graph_buffer(wg_context* ctx, wg_graph* g, int slot, size_t bytes, void** out);
constexpr int kSlotHotIndex = 190, kSlotHotSrc = 191;
constexpr int kSlotPprItems = 192, kSlotPprLrows = 193, kSlotPprLbase = 194;
constexpr int kSlotRevPos = 195;  // forward entry -> reverse position
int prepare_hot_order(wg_context* ctx, wg_graph* g);
// A CSR whose offsets, targets and weights were built on the device: reads
// its offsets back and partitions it for the SpMV.
int adopt_device_csr(wg_context* ctx, DevCsr& c);
void free_csr(wg_context* ctx, DevCsr& c);
// Transposes g->fwd into g->rev on the device and sets g->rev_is_fwd.
int build_reverse(wg_context* ctx, wg_graph* g);
// pos[p] = the reverse-CSR position of forward entry p.
int reverse_position_map(wg_context* ctx, const wg_graph* g, uint64_t* pos);
// The graph a selection of relations reads (g itself, one of its
// relations, or a merged selection kept with g); rels NULL: every relation.
int select_graph(wg_context* ctx, wg_graph* g, const uint32_t* rels, uint32_t k, int per_relation,
                 wg_graph** out);
// Frees a graph, its relations and selections (an orphaned graph: only
// their host records).
void destroy_graph(wg_graph* g);
// Frees the device memory of a graph, its relations and selections and
// orphans them (ctx null); their host records stay for destroy_graph.
void release_graph(wg_graph* g);
// The same for a BFS cursor and a gathered result (bfs.hip, gather.hip).
void release_cursor(wg_bfs_cursor* c);
void release_gathered(wg_gathered* r);
// Whether a cursor reads graph g (or a relation or selection of g).
bool cursor_reads(const wg_bfs_cursor* c, const wg_graph* g);
// WG_INVALID_ARGUMENT unless g is a live graph of ctx (a graph of another
// context, or one orphaned by wg_context_destroy, is refused).
// wide_ok: a graph with u64 targets is accepted (products only).
int check_graph(wg_context* ctx, const wg_graph* g, bool wide_ok = false);
// WG_UNSUPPORTED for a graph with u64 targets.
int check_narrow(wg_context* ctx, const wg_graph* g);
// Csr::from_parts' per-row checks on the device (targets in the id space,
// rows strictly ascending) for offsets known to be monotone.
template <class T>
int check_rows_device(wg_context* ctx, uint64_t n, const uint64_t* off, const T* tgt, uint64_t m);
int check_vector(wg_context* ctx, const wg_vector* v);
// Whether two CSRs hold equal arrays (on the device).
int same_csr(wg_context* ctx, const DevCsr& a, const DevCsr& b, bool* same);
// row_of[p] = the row of entry p of a partitioned CSR.
int expand_rows(wg_context* ctx, const DevCsr& c, uint32_t* row_of);
// Stable ascending radix sort of m (key, value) pairs on key bits [0, bits),
// on the context stream.
int sort_u32_pairs(wg_context* ctx, const uint32_t* kin, uint32_t* kout, const uint32_t* vin, uint32_t* vout,
                   uint64_t m, unsigned bits);
int sort_u64_keyed(wg_context* ctx, const uint64_t* kin, uint64_t* kout, const uint64_t* vin, uint64_t* vout,
                   uint64_t m, unsigned bits);
int sort_u32_keyed(wg_context* ctx, const uint32_t* kin, uint32_t* kout, const uint64_t* vin, uint64_t* vout,
                   uint64_t m, unsigned bits);
// out[i] = in[0] + ... + in[i-1] (out[0] = 0), on the context stream.
// temp: device scratch of temp_bytes for the scan, or null (allocated and
// released inside); exclusive_scan_u64_bytes(m) is the size it needs.
int exclusive_scan_u64(wg_context* ctx, const uint64_t* in, uint64_t* out, uint64_t m, void* temp = nullptr,
                       size_t temp_bytes = 0);
size_t exclusive_scan_u64_bytes(uint64_t m);
// A CSR's entries as (row, column) pairs (nnz of each).
int csr_edges(wg_context* ctx, const DevCsr& c, uint32_t* src, uint32_t* dst);
// The unweighted, partitioned CSR of a device edge list: parallel edges
// folded, or kept (in input order) with keep.
int csr_of_edges(wg_context* ctx, uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst, bool keep,
                 DevCsr& out);
}  // namespace wg

namespace wg {

size_t dtype_size(int dtype);
int fail(wg_context* ctx, int status, const std::string& msg);
int ok(wg_context* ctx);
int hip_fail(wg_context* ctx, hipError_t e, const char* what);
// Waits until a kernel on the context's stream posts mailbox sequence
// number `seq`, then copies its first k values into out. The wait ends
// with an error if the stream drains or fails without the post.
int wait_mail(wg_context* ctx, unsigned long long seq, unsigned long long* out, int k);
// Device memory charged to the context's budget before it is allocated
// (ComputeMemory::charge, compute/memory.rs:97).
int dalloc(wg_context* ctx, void** p, size_t bytes);
void dfree(wg_context* ctx, void* p, size_t bytes);
int ensure_scratch(wg_context* ctx, size_t bytes);
// Clears the device error word (enqueued on the context stream).
int clear_error(wg_context* ctx);
// Waits for the stream and turns a raised error word into a status.
int collect_error(wg_context* ctx, const char* what);
bool is_cancelled(wg_context* ctx);
// The device address of wg_spmv_noalloc's static partials on `device` (the
// current device), looked up once per device (spmv.hip).
int static_scratch(int device, void** out);
// The weights of `csr` cast to `dtype` (cached); returns the cast status.
int weight_view(wg_context* ctx, DevCsr& csr, int dtype, const void** out);
int prepare_spmm(wg_context* ctx, DevCsr& csr);
// Builds csr's hot-first column numbering; returns WG_OK with hot_state -1
// when the budget cannot hold it.
int prepare_hot_columns(wg_context* ctx, DevCsr& csr);
// The tiled SpMV's layout of csr for elements of esize bytes (spmv.hip);
// WG_UNSUPPORTED when it does not apply.
int prepare_tiles(wg_context* ctx, DevCsr& c, size_t esize, uint32_t hotcap, uint32_t warmcap);
// The weights in the tiled layout's order, as elements of E (spmv.hip).
template <class E>
int tile_weight_view(wg_context* ctx, DevCsr& c, const E* wv, const E** out);

#define WG_HIP(ctx, expr)                                   \
  do {                                                      \
    hipError_t e_ = (expr);                                 \
    if (e_ != hipSuccess) return ::wg::hip_fail(ctx, e_, #expr); \
  } while (0)

}  // namespace wg
