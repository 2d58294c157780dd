/*
 * warren_gpu.h -- C ABI of the MI355X (gfx950) implementation of warren-core's
 * whole-graph compute path: device CSR, semiring SpMV/SpMM, PageRank, BFS and
 * Lanczos. Plain pointers and sizes only; every entry point returns a status
 * (0 = ok) whose numbering mirrors warren-core's error variants, and the
 * message of the last failure is read with wg_last_error().
 *
 * Reference interfaces (paths relative to crates/warren-core/src/ of
 * ExaDev/pg_warren) are cited at each declaration. The ABI is not thread-safe
 * per context (warren-core's callers drive a kernel from one thread,
 * compute/workers.rs:1-16); each context owns one HIP stream.
 */
#ifndef WARREN_GPU_H
#define WARREN_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: ComputeError / TraversalError / ArithmeticError
 * (compute/error.rs:10-210, traversal/error.rs:10-62) ---------------------- */
enum wg_status {
  WG_OK = 0,
  WG_CANCELLED = 1,            /* ComputeError::Cancelled */
  WG_DTYPE_MISMATCH = 2,       /* ComputeError::DtypeMismatch */
  WG_ID_SPACE_MISMATCH = 3,    /* ComputeError::IdSpaceMismatch */
  WG_WIDTH_MISMATCH = 4,       /* ComputeError::WidthMismatch */
  WG_MEMORY_LIMIT = 5,         /* ComputeError::MemoryLimit */
  WG_OUT_OF_MEMORY = 6,        /* ComputeError::OutOfMemory */
  WG_DIVISION_BY_ZERO = 7,     /* ArithmeticError::DivisionByZero */
  WG_FLOAT_OVERFLOW = 8,       /* ArithmeticError::FloatOverflow */
  WG_FLOAT_UNDERFLOW = 9,      /* ArithmeticError::FloatUnderflow */
  WG_BIGINT_OUT_OF_RANGE = 10, /* ArithmeticError::BigintOutOfRange */
  WG_INVALID_ARGUMENT = 11,    /* UnknownSemiring / UnknownBinaryOp / ZeroWidth */
  WG_NODE_OUTSIDE_ID_SPACE = 12, /* TraversalError::NodeOutsideIdSpace */
  WG_NODE_FILTER_ID_SPACE = 13,  /* TraversalError::NodeFilterIdSpace */
  WG_INVALID_CSR = 14,         /* GraphError from Csr::from_parts (graph/csr/mod.rs:78-137) */
  WG_UNSUPPORTED = 15,         /* a dtype / semiring pair with no meaning */
  WG_DEVICE_ERROR = 16,        /* a HIP runtime failure */
  WG_NOT_SYMMETRIC = 17,       /* SpectralError::NotSymmetric */
  WG_NON_POSITIVE_DEGREE = 18, /* SpectralError::NonPositiveDegree */
  WG_EIGENPAIR_COUNT = 19,     /* SpectralError::EigenpairCount */
  WG_NOT_FINITE = 20,          /* SpectralError::NotFinite */
  WG_EDGE_OUTSIDE_ID_SPACE = 21,       /* GraphError::EdgeOutsideIdSpace (graph/error.rs:91-98) */
  WG_RELATION_OUTSIDE_DICTIONARY = 22, /* GraphError::RelationOutsideDictionary (:99-104) */
  WG_UNKNOWN_RELATION = 23,            /* ComputeError::UnknownRelation (compute/error.rs:82-87) */
  WG_INVALID_SEGMENT = 24,             /* GraphError::Segment* (graph/error.rs): a segment's framing */
  WG_COMBINER_TYPE = 25,               /* ComputeError::CombinerType (compute/error.rs:109-116) */
  WG_SENDER_OUTSIDE_ID_SPACE = 26      /* ComputeError::SenderOutsideIdSpace (:117-124) */
};

/* ---- element types (compute/dtype.rs:12-21), plus u32: the image of i64 on
 * [0, 2^32-2] with 2^32-1 standing for i64::MAX (+infinity) ------------------ */
enum wg_dtype { WG_F32 = 0, WG_F64 = 1, WG_I64 = 2, WG_BOOL = 3, WG_U32 = 4 };

/* ---- semirings in Semiring::ALL order (compute/semiring.rs:149-163) ------ */
enum wg_semiring {
  WG_PLUS_TIMES = 0, WG_MIN_PLUS, WG_MAX_PLUS, WG_MAX_TIMES, WG_MIN_TIMES,
  WG_LOR_LAND, WG_ANY_PAIR, WG_PLUS_PAIR, WG_PLUS_FIRST, WG_PLUS_SECOND,
  WG_MIN_FIRST, WG_MIN_SECOND, WG_MAX_SECOND
};

/* ---- binary operators in BinaryOp::ALL order (compute/semiring.rs:42-53) - */
enum wg_binary_op {
  WG_OP_NONE = -1, WG_FIRST = 0, WG_SECOND, WG_PAIR, WG_PLUS, WG_TIMES, WG_MIN,
  WG_MAX, WG_LOR, WG_LAND, WG_ANY
};

/* ---- traversal direction (graph::Direction, traversal/adjacency.rs:20-60) */
enum wg_direction { WG_OUT = 0, WG_IN = 1, WG_BOTH = 2 /* traversals and gather */ };

typedef struct wg_context wg_context;
typedef struct wg_graph wg_graph;
typedef struct wg_vector wg_vector;

/* ---- context: Workers + ComputeMemory + Interrupts ------------------------
 * Replaces Workers::new (compute/workers.rs:65-80), ComputeMemory::new
 * (compute/memory.rs:31-60) and the Interrupts/Cancellation pair
 * (interrupts.rs:22-70). memory_limit = 0: no limit. */
int wg_context_create(int device, uint64_t memory_limit, wg_context** out);
/* Releases the context's device memory, including that of every graph,
 * vector, BFS cursor and gathered result made through it and not yet
 * destroyed: those are left orphaned, so a later wg_*_destroy of one only
 * frees its host record, and any other call given one returns
 * WG_INVALID_ARGUMENT. A graph, vector, cursor or result passed with a
 * context other than its own is refused the same way. */
void wg_context_destroy(wg_context* ctx);
/* Copies the last failure's message; returns its status. */
int wg_last_error(const wg_context* ctx, char* buf, size_t len);
/* Cancellation::cancel (interrupts.rs:40-70): the next check returns
 * WG_CANCELLED; wg_cancel_reset clears it. wg_cancel is the one call that
 * may be made from another thread while a call on the context runs (as the
 * reference's flag is raised by the thread waiting for the pool): it sets
 * an atomic word that calls check before each launch, search level, batch
 * of PageRank iterations and Lanczos step. Every other call on a context
 * comes from one thread at a time. */
void wg_cancel(wg_context* ctx);
void wg_cancel_reset(wg_context* ctx);

/* The progress port (progress.rs:9-16): long-running work reports the phase
 * it is in (total units, or -1 when unknown), the units done and named
 * counters, on the thread that made the call, where it checks for
 * cancellation. wg_eigen reports as spectral/lanczos.rs:502-629 does: phase
 * "lanczos" with total = the step limit, then after every step and every
 * Rayleigh-Ritz step done = steps taken and the details "converged" (pairs
 * converged, the components' exact pairs included), "restarts" and, once
 * estimated, "residual" (the largest wanted residual estimate over the
 * largest Ritz value's magnitude). Reporting never changes a result. A
 * null callback is skipped. */
typedef struct wg_progress {
  void (*phase)(void* user, const char* name, int64_t total);
  void (*done)(void* user, uint64_t done);
  void (*detail)(void* user, const char* name, double value);
  void* user;
} wg_progress;
/* Sets the context's progress port (copied); null reports nowhere. */
void wg_set_progress(wg_context* ctx, const wg_progress* port);
/* ComputeMemory::in_use (compute/memory.rs:97-120). */
uint64_t wg_memory_in_use(const wg_context* ctx);
/* A new budget for the calls that follow, as a fresh ComputeMemory::new
 * (compute/memory.rs:47-60) would give them: later allocations are charged
 * against it (0: no limit); bytes already held stay held, and a limit
 * below them refuses every new allocation with WG_MEMORY_LIMIT. */
int wg_set_memory_limit(wg_context* ctx, uint64_t limit);
/* Waits for the context's stream and reports a deferred kernel error. */
int wg_synchronize(wg_context* ctx);
/* 1 (default): every call waits and reports its own errors. 0: calls whose
 * element type cannot fail (float products) return without waiting, for
 * back-to-back launches; errors surface at wg_synchronize. */
void wg_set_blocking(wg_context* ctx, int blocking);
/* Whether iterative calls replay each iteration's launch sequence as a
 * captured HIP graph (1, the default: batched personalised PageRank captures
 * one per message parity and active-column set) or enqueue the launches one
 * by one (0). The results are the same either way. No counterpart in the
 * reference. */
void wg_set_graph_launch(wg_context* ctx, int on);
/* Page-locked host memory for the buffers results are copied into (visit
 * rows, gathered values, vectors read back): a copy into it is one DMA, where
 * a copy into ordinary (pageable) memory goes through a staging buffer at a
 * fraction of the link's rate. No counterpart in the reference, whose
 * results stay in host memory; the adapter's row buffers are its callers.
 * Host memory, not charged to the context's budget. */
int wg_host_alloc(wg_context* ctx, uint64_t bytes, void** out);
void wg_host_free(void* p);
/* The context's HIP stream (a hipStream_t), for timing with HIP events. */
void* wg_stream(wg_context* ctx);

/* ---- graphs: GraphCsr with a forward and a reverse CSR ---------------------
 * Replaces Csr::from_parts (graph/csr/mod.rs:56-137) for the merged
 * all-relations CSR (graph/csr/graph_csr.rs:15-153) and its reverse, built by
 * transpose (graph/csr/transpose.rs:16-44). offsets: n+1 u64, targets:
 * offsets[n] u32, each row strictly ascending; weights: offsets[n] f64 or
 * NULL (unweighted: every weight reads as 1). Host pointers; copied. */
int wg_graph_upload(wg_context* ctx, uint64_t n, const uint64_t* offsets,
                    const uint32_t* targets, const double* weights, wg_graph** out);
/* The same with u64 targets (TargetWidth::U64, graph/csr/target.rs:16-31:
 * the width warren-core uses for id spaces of 2^32 nodes or more; taken
 * here for any id space). Csr::from_parts' checks run on the device; the
 * reverse CSR is built by a stable radix sort by target. Such a graph is
 * read by wg_spmv and wg_spmm (plain, transposed, masked, accumulated; a
 * thread folds each output in the reference's order, so float plus too is
 * the reference's left fold) and by wg_graph_read_u64; every other entry
 * point, and a relation selection, returns WG_UNSUPPORTED for it. */
int wg_graph_upload_u64(wg_context* ctx, uint64_t n, const uint64_t* offsets,
                        const uint64_t* targets, const double* weights, wg_graph** out);
/* A u64-target graph's forward (reverse = 0) or reverse CSR: host buffers
 * of n+1, nnz and nnz (targets and weights may be NULL). */
int wg_graph_read_u64(wg_context* ctx, const wg_graph* g, int reverse, uint64_t* offsets,
                      uint64_t* targets, double* weights);
/* Also releases any BFS cursor still reading the graph (the cursor's next
 * step then returns WG_INVALID_ARGUMENT; wg_bfs_cursor_destroy still frees
 * it). */
void wg_graph_destroy(wg_graph* g);
/* n, nnz and whether the graph carries weights. */
int wg_graph_info(const wg_graph* g, uint64_t* n, uint64_t* nnz, int* weighted);
/* GraphCsr::build (graph/csr/graph_csr.rs:194-221) on the device: the
 * CSR pair of an edge list of m edges sources[i] -> targets[i] with
 * optional weights[i] and relations[i] (< relation_count; NULL: every edge
 * is relation 0). With relations and relation_count >= 2 each relation's
 * CSR pair is built too (for wg_graph_select) and the graph's own pair is
 * the merged all-relations CSR. Rows come out sorted by target; parallel edges of one
 * relation fold into one, weights summed in input order, and a target
 * reached through several relations appears once, the relations' sums
 * added in relation order (build/deduplication.rs:83-152); self loops stay.
 * The first edge outside the id space or the dictionary is the error
 * (build/mod.rs:329-352). *deduplicated (may be NULL) receives the edges
 * folded into another edge of the same relation. The arrays are host
 * memory, or device memory when device_pointers is non-zero. Unweighted
 * when weights is NULL. */
int wg_graph_build(wg_context* ctx, uint64_t n, uint64_t m, const uint32_t* sources,
                   const uint32_t* targets, const double* weights, const uint32_t* relations,
                   uint32_t relation_count, int device_pointers, wg_graph** out,
                   uint64_t* deduplicated);
/* Relation selection (GraphCsr::selected, graph/csr/graph_csr.rs:320-340,
 * and the matrix a product reads, compute/product/matrix.rs:126-157,
 * 247-294): the graph whose CSR pair holds the edges of `relations` (NULL:
 * every relation), for any entry point (wg_spmv, wg_pagerank, wg_bfs,
 * wg_eigen, ...). One relation: its CSR pair as built. Several: their rows
 * merged by column, an edge in several relations one entry whose weight is
 * their sum in relation order (1 for an unweighted relation's entry), or
 * with per_relation one entry per relation, in relation order. Built on
 * first use and kept with g (destroyed with it; wg_graph_destroy on a
 * selection does nothing). An index not below the relation count is
 * WG_UNKNOWN_RELATION (the smallest such). */
int wg_graph_select(wg_context* ctx, wg_graph* g, const uint32_t* relations, uint32_t nrelations,
                    int per_relation, wg_graph** out);
/* The relations a graph holds: relation_count of wg_graph_build when 2 or
 * more, else 1 (the graph is its only relation). */
int wg_graph_relation_count(const wg_graph* g, uint32_t* count);
/* ---- segments: the WRSG codec of a CSR (graph/segment/mod.rs:1-16,
 * graph/segment/csr.rs). Each segment is a 16-byte little-endian header
 * (magic "WRSG", format 2, kind, element width, reserved 0, u64 count)
 * and its payload: offsets (kind 1, u64), targets (kind 2, u32 below a
 * 2^32 id space), weights (kind 3, f64) and, with sparse offsets, rows
 * (kind 11, the stored rows' ids). */
typedef struct {
  const uint8_t* rows; uint64_t rows_len;       /* NULL: dense offsets */
  const uint8_t* offsets; uint64_t offsets_len;
  const uint8_t* targets; uint64_t targets_len;
  const uint8_t* weights; uint64_t weights_len; /* NULL: unweighted */
} wg_csr_segments;
/* AnyCsr::decode (graph/segment/csr.rs:84-127) straight to the device: a
 * graph from its forward CSR's segments (host memory), and its reverse
 * CSR's if given (taken as the transpose, as the reference takes it), else
 * transposed on the device. Every check of the decode and of
 * Csr::from_segments (graph/csr/mod.rs:78-137, 239-258) in their order;
 * the first violation is the error: WG_INVALID_SEGMENT for the framing
 * (SegmentTruncated, SegmentMagic, SegmentFormat, SegmentKind,
 * SegmentReserved, SegmentWidth, SegmentLength), WG_INVALID_CSR for the
 * CSR (OffsetsLength, FirstOffsetNotZero, StoredRowsNotAscending,
 * StoredRowOutsideIdSpace, StoredRowEmpty, OffsetsNotMonotone,
 * FinalOffset, TargetOutsideIdSpace, RowNotSorted, WeightsLength), the
 * variant and its fields in the message. */
int wg_graph_load_segments(wg_context* ctx, uint64_t id_space, const wg_csr_segments* forward,
                           const wg_csr_segments* reverse, wg_graph** out);
/* The byte lengths of a CSR's dense segments (csr.rs:130-137). */
int wg_csr_segment_lengths(uint64_t id_space, uint64_t nnz, uint64_t* offsets_len,
                           uint64_t* targets_len, uint64_t* weights_len);
/* AnyCsr::encode with dense offsets: the forward (reverse = 0) or reverse
 * CSR as its segments, into host buffers of wg_csr_segment_lengths bytes
 * (any may be NULL; weights only for a weighted graph). */
int wg_graph_encode_segments(wg_context* ctx, const wg_graph* g, int reverse, uint8_t* offsets,
                             uint8_t* targets, uint8_t* weights);
/* A relation's CSR (forward, or reverse = 1) as its segments in the
 * reference's layout: with sparse offsets -- the stored rows' ids (kind 11,
 * u32) and an offset per stored row plus the end -- where they take at most
 * a quarter of the bytes of dense offsets (graph/csr/sparse.rs:10-40,
 * graph_csr.rs:769-800), else dense; a graph of one relation holds it as
 * its all-relations CSR, always dense. The lengths first (rows_len 0 when
 * dense; weights_len 0 for an unweighted graph); any buffer may be NULL.
 * WG_UNKNOWN_RELATION for a relation outside the graph's dictionary. */
int wg_relation_segment_lengths(wg_context* ctx, const wg_graph* g, uint32_t relation, int reverse,
                                uint64_t* rows_len, uint64_t* offsets_len, uint64_t* targets_len,
                                uint64_t* weights_len);
int wg_relation_encode_segments(wg_context* ctx, const wg_graph* g, uint32_t relation, int reverse,
                                uint8_t* rows, uint8_t* offsets, uint8_t* targets, uint8_t* weights);
/* The forward CSR (host buffers of n+1, nnz, nnz; targets and weights may
 * be NULL). */
int wg_graph_read_forward(wg_context* ctx, const wg_graph* g, uint64_t* offsets,
                          uint32_t* targets, double* weights);
/* The reverse CSR as built on the device (host buffers of n+1, nnz, nnz;
 * weights may be NULL). */
int wg_graph_read_reverse(wg_context* ctx, const wg_graph* g, uint64_t* offsets,
                          uint32_t* targets, double* weights);
/* Sets the weights of `count` edges, named by their index in the forward CSR
 * (the order of the uploaded targets), in both orientations, without a
 * rebuild; weights cast to other element types are cast again on next use.
 * No counterpart in the reference (a warren-core graph is immutable, weights
 * change by a rebuild); it serves the finite-difference checks of the
 * PageRank adjoint. Host arrays; the graph must carry weights. */
int wg_graph_update_weights(wg_context* ctx, wg_graph* g, const uint64_t* edges,
                            const double* values, uint64_t count);

/* ---- vectors: AnyVector (compute/vector.rs:9-50), row-major id_space x
 * width, resident in device memory and charged to the context's budget --- */
int wg_vector_create(wg_context* ctx, int dtype, uint64_t id_space, uint64_t width,
                     const void* host_values, wg_vector** out);
void wg_vector_destroy(wg_vector* v);
int wg_vector_write(wg_context* ctx, wg_vector* v, const void* host_values);
int wg_vector_read(wg_context* ctx, const wg_vector* v, void* host_values);
int wg_vector_info(const wg_vector* v, int* dtype, uint64_t* id_space, uint64_t* width);
void* wg_vector_device_ptr(wg_vector* v);

/* ---- semiring products: Product (compute/product/mod.rs:33-56) ----------- */
typedef struct {
  int32_t semiring;        /* wg_semiring */
  int32_t transpose;       /* y = A^T (x) x: pull over the reverse CSR */
  int32_t weighted;        /* -1: if the graph has weights (Product::weighted None) */
  int32_t mask_complement;
  const wg_vector* mask;   /* width-1 value mask of any dtype, or NULL */
  int32_t replace;
  int32_t accumulate;      /* wg_binary_op, or WG_OP_NONE */
  const uint32_t* relations; /* the relations whose edges make up A; NULL: all
                                (RelationSelection::All), else nrelations indexes */
  uint32_t nrelations;
  int32_t per_relation;    /* an edge in several selected relations is an entry
                              in each rather than one entry (Product::per_relation) */
} wg_product;

/* spmv (compute/product/mod.rs:77-92): width-1 x and y of one dtype. */
int wg_spmv(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y);
/* spmm (compute/product/mod.rs:97-110): x and y of the same width. */
int wg_spmm(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
            wg_vector* y);

/* wg_spmv without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, checks, errors and results as wg_spmv's default launch (the
 * blocked kernel of graphs under 2^22 entries, whose bits it gives for every
 * semiring and dtype); it runs that launch on any graph, never tunes and
 * never builds a layout. Weights are read as cached for the vector's type by
 * an earlier wg_spmv, else the f64 weights are cast per entry. The partials
 * of rows longer than 2048 entries live in a fixed 1 MiB device array, one
 * per GPU for the process's life, shared by every wg_spmv_noalloc call on
 * that GPU: a graph whose split rows need more (over 131,072 row pieces
 * of f64 partials; fewer for wider accumulators, 21,845 for i64 plus) is
 * WG_MEMORY_LIMIT, and calls that overlap on one GPU (two
 * contexts' streams) race. A relation selection (relations, per_relation) is
 * WG_UNSUPPORTED: building one allocates; wg_spmv builds it. */
int wg_spmv_noalloc(wg_context* ctx, const wg_graph* g, const wg_product* p, const wg_vector* x,
                    wg_vector* y);

/* ---- SpMV autotuner: no counterpart in the reference (whose launch shape is
 * fixed at 4096-row tasks, compute/product/kernel.rs:15). By default (-1) the
 * first product over a graph orientation of at least 2^22 entries, for each
 * dtype and semiring, times the launch variants and keeps the fastest:
 * bit 0 = the LDS-free direct kernel (where the monoid's result does not
 * depend on grouping, and for float plus, where its grouping gives other bits
 * within the rounding bound: the tuner leaves it out there, so a tuned
 * result has the default launch's bits), bit 1 = non-temporal row streams,
 * bit 2 = XCD-contiguous
 * block order, bit 3 = gathers from x permuted into the hot-first column order
 * (columns ranked by how many entries read them; same values, same fold
 * order, same result bits), bit 4 = a persistent kernel that keeps the
 * hottest columns' x values in LDS and reads them there (bits 0, 2 and 3 are
 * then ignored). wg_set_spmv_variant forces one variant (0..31; a bit the
 * graph, dtype or monoid cannot take is dropped);
 * wg_spmv_variant reports the choice made (-1: not tuned yet). */
void wg_set_spmv_variant(wg_context* ctx, int variant);
int wg_spmv_variant(const wg_graph* g, int transpose, int dtype, int semiring, int weighted);
/* SpMM launch shapes (no counterpart in the reference; the autotuner the
 * batched workloads ask for): lanes per row (a power of two up to 64, each
 * lane taking every lanes-th column group) | 256 when a group is 16 bytes
 * of adjacent columns (4 f32 / u32, 2 f64 / i64) rather than one column.
 * Every shape folds each column over the same entries in the same order,
 * so the choice never changes a result. The first wg_spmm over a graph of
 * 2^22 entries or more times up to six shapes at its width and keeps the
 * fastest; wg_set_spmm_shape forces one (-1: autotune; a 16-byte grouping
 * the width or alignment cannot take is dropped), wg_spmm_shape reports
 * the choice kept for (orientation, dtype, semiring, weighted, width), -1
 * if none. */
void wg_set_spmm_shape(wg_context* ctx, int shape);
int wg_spmm_shape(const wg_graph* g, int transpose, int dtype, int semiring, int weighted, uint64_t width);

/* ---- PageRank: PageRank / pagerank (compute/pagerank/mod.rs:27-151) ----- */
typedef struct {
  double damping;     /* probability of following an edge (0.85) */
  double tol;         /* stop once the L1 change < live nodes * tol */
  int32_t has_tol;    /* 0: PageRank::tol None, run max_iter iterations */
  uint32_t max_iter;
  int32_t weighted;   /* -1: if the graph has weights */
  int32_t pad;
} wg_pagerank_opts;

/* NetworkX-style power iteration in f64 over the reverse CSR. tombstones: a
 * bool vector over the graph (non-zero = tombstoned) or NULL; seed: a width-1
 * numeric vector (personalisation seed / sum(seed) over live rows) or NULL
 * (uniform). ranks: an f64 width-1 vector over the graph, written.
 * Errors as the reference: division by zero (no live node, seed summing to
 * zero), float overflow/underflow in a step, bool seed (dtype mismatch). */
int wg_pagerank(wg_context* ctx, const wg_graph* g, const wg_vector* tombstones,
                const wg_pagerank_opts* opts, const wg_vector* seed, wg_vector* ranks,
                uint32_t* iterations);

/* Batched personalised PageRank: R (1..256) seed sets run as the columns of
 * one power iteration; column c equals wg_pagerank with the seed vector that
 * holds seed_weights[i] (1 when NULL) at seed_nodes[i] for i in
 * [seed_offsets[c], seed_offsets[c+1]) (repeats add up), and stops on its own
 * test (L1 change < live nodes * tol), frozen at that iteration while the
 * others go on. No counterpart in the reference, which takes one seed vector
 * per call (compute/pagerank/mod.rs:44-46). ranks: f64, id space n, width R;
 * iterations: R host counters. */
int wg_pagerank_batch(wg_context* ctx, const wg_graph* g, const wg_vector* tombstones,
                      const wg_pagerank_opts* opts, const uint64_t* seed_offsets,
                      const uint64_t* seed_nodes, const double* seed_weights, uint32_t R,
                      wg_vector* ranks, uint32_t* iterations);

/* Adjoint of wg_pagerank with respect to the edge weights: given the ranks x
 * of a weighted run (same graph, tombstones, options and seed) and
 * dl_dranks = dL/dx for a loss L, writes dl_dweights = dL/dw, one f64 per
 * edge in forward-CSR order (a vector with id space nnz). Solves the adjoint
 * fixed point lambda = d (P lambda) + d (p . lambda) [dangling] + dL/dx by
 * pulls over the forward CSR (the forward run pulls over the reverse), with
 * the forward's stop rule (L1 change < live nodes * tol, or max_iter), then
 * dL/dw_ab = (d x_a / s_a) (lambda_b - (P lambda)_a) for s_a > 0 (s = the
 * out-weight; 0 elsewhere), tombstoned rows taking no part. adjoint: an f64
 * vector over the graph that receives lambda, or NULL. No counterpart in the
 * reference (whose PageRank has no backward pass). */
int wg_pagerank_backward(wg_context* ctx, const wg_graph* g, const wg_vector* tombstones,
                         const wg_pagerank_opts* opts, const wg_vector* seed, const wg_vector* ranks,
                         const wg_vector* dl_dranks, wg_vector* dl_dweights, wg_vector* adjoint,
                         uint32_t* iterations);

/* ---- traversal: Visit / TraversalStats (traversal/bfs.rs:15-23,
 * traversal/stats.rs:12-20) ------------------------------------------------ */
typedef struct {
  uint64_t node;
  uint64_t via;    /* UINT64_MAX for a start (Visit::via None) */
  uint32_t depth;
  uint32_t pad;
} wg_visit;

typedef struct {
  uint64_t nodes_visited;
  uint64_t edges_traversed; /* row entries scanned, as the reference counts
                               them: a pushed level scans every frontier row
                               (up to the refused visit's entry when
                               max_results cuts it), a pulled level stops at
                               the first frontier hit */
  uint32_t truncated;
  uint32_t levels_pulled;
} wg_traversal_stats;

/* Traversals read the CSRs of Adjacency::new (traversal/adjacency.rs:68-152):
 * relations NULL: every relation, through the graph's own (all-relations)
 * CSR pair; else the subset relations[0..nrelations) (sorted, repeats
 * dropped; an index not below the relation count is WG_UNKNOWN_RELATION),
 * each relation's CSR pair in index order. direction WG_OUT reads the
 * forward rows, WG_IN the reverse ones, WG_BOTH the forward then the
 * reverse; a node's row entries are its rows in that order. At most 32 CSRs
 * (16 relations both ways). */

/* The device-resident search: depth and via (u32 vectors over the graph,
 * 0xffffffff where unreached) of every node reached from starts, levels
 * pushed or, when pulling != 0 and Beamer's switch (alpha 1, beta 24,
 * traversal/level.rs:19-29,184-211) says so, pulled. via is the one a push
 * gives (the smallest frontier node with an edge to the node) either way.
 * filter: n host bytes (non-zero admits) or NULL; max_depth < 0: unbounded.
 * Starts must lie in the id space (NodeOutsideIdSpace) and repeats count
 * once. The stats count the device's own work (pulled levels included). */
int wg_bfs_levels(wg_context* ctx, const wg_graph* g, int direction, const uint32_t* relations,
                  uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, const uint8_t* filter,
                  int64_t max_depth, int pulling, wg_vector* depth, wg_vector* via, wg_traversal_stats* stats);

/* bfs (traversal/bfs.rs:34-45): every node reachable from starts, once, in
 * BFS order -- by depth; the starts in the order given; within a depth in
 * the order the reference's push reaches them: by via, then by the first
 * CSR whose row of the via holds the node, then by node. filter prunes
 * admission (starts are always admitted); max_results < 0: unbounded, else
 * the output stops there and stats->truncated says whether a visit was cut.
 * The reference never pulls a bfs level, so the stats are a push's (no
 * level pulled) although the device pulls wide levels. out must hold n
 * visits. */
int wg_bfs(wg_context* ctx, const wg_graph* g, int direction, const uint32_t* relations, uint32_t nrelations,
           const uint64_t* starts, uint64_t nstarts, const uint8_t* filter, int64_t max_depth, int64_t max_results,
           wg_visit* out, uint64_t* count, wg_traversal_stats* stats);

/* reachable_set (traversal/reachable.rs:48-90): the direction-optimising
 * search's nodes by depth, pushed levels in the order reached, pulled
 * levels ascending by node (a pulled node's via is the first frontier node
 * its reversed rows name, in their concatenated order). out must hold n
 * ids. */
int wg_reachable_set(wg_context* ctx, const wg_graph* g, int direction, const uint32_t* relations,
                     uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, int64_t max_depth, uint64_t* out,
                     uint64_t* count, wg_traversal_stats* stats);

/* BfsCursor (traversal/bfs_cursor.rs:23-135): the search of wg_bfs (or,
 * with reachable, of wg_reachable_set: no filter or result limit, a pulled
 * level listed by node) handed out one level per wg_bfs_cursor_next call,
 * the starts first. Each level's visits are sorted and written on the
 * device and reach out (host, room for the graph's n visits) in one copy;
 * *more = 0 once no level follows (the search is exhausted, max_depth
 * levels are done, or max_results visits were handed out and another was
 * found: stats.truncated). The cursor keeps its own workspace, so several
 * may be open on one graph. Arguments are checked at creation. */
typedef struct wg_bfs_cursor wg_bfs_cursor;
int wg_bfs_cursor_create(wg_context* ctx, const wg_graph* g, int direction, const uint32_t* relations,
                         uint32_t nrelations, const uint64_t* starts, uint64_t nstarts, const uint8_t* filter,
                         int64_t max_depth, int64_t max_results, int reachable, wg_bfs_cursor** out);
int wg_bfs_cursor_next(wg_context* ctx, wg_bfs_cursor* cursor, wg_visit* out, uint64_t* count,
                       int* more);
int wg_bfs_cursor_stats(const wg_bfs_cursor* cursor, wg_traversal_stats* stats);
void wg_bfs_cursor_destroy(wg_bfs_cursor* cursor);

/* Multi-source BFS: SourceBatch::search (compute/centrality/multi_source.rs:
 * 34-193) widened from 64 sources to up to 512, one bit per source. Every
 * node's distance from every source along direction, level by level until no
 * node gains a source (or max_depth >= 0 levels). Outputs, each optional:
 * depths -- u32, id space n, width nsources (row v, column s = distance from
 * source s, 0xffffffff unreached); reach -- u32 n, the sources reaching each
 * node (the count SourceBatch reports per distance, summed); farness -- i64
 * n, the sum of those distances (closeness centrality's numerator). Sources
 * may repeat; each position is its own source. levels: levels run. */
int wg_msbfs(wg_context* ctx, const wg_graph* g, int direction, const uint64_t* sources,
             uint32_t nsources, int64_t max_depth, wg_vector* depths, wg_vector* reach,
             wg_vector* farness, uint32_t* levels);

/* ---- gather: compute/gather/mod.rs:1-154 ---------------------------------
 * Messages fanned out over the graph's edges and combined at each receiver
 * in collect order: ascending sender dense id, a sender's messages in push
 * order, the receiver's own first with include_self. A receiver's senders
 * are its distinct neighbours in the direction (WG_IN: in-neighbours, the
 * pull of a transposed product; WG_OUT; WG_BOTH: their union) over the
 * selected relations (NULL: all), or with per_relation a sender once per
 * relation linking it (both directions of one relation once). Combiners in
 * Combiner::ALL order (compute/gather/combiner.rs:13-57). */
enum wg_combiner {
  WG_COMBINE_COLLECT = 0, WG_COMBINE_COUNT, WG_COMBINE_SUM, WG_COMBINE_MIN, WG_COMBINE_MAX,
  WG_COMBINE_FIRST, WG_COMBINE_CONCAT, WG_COMBINE_BOOL_OR, WG_COMBINE_BOOL_AND, WG_COMBINE_MODE
};
typedef struct {
  int32_t combiner;         /* wg_combiner */
  int32_t direction;        /* wg_direction */
  const uint32_t* relations;
  uint32_t nrelations;
  int32_t per_relation;
  const wg_vector* mask;    /* receivers the mask selects (compute/mask.rs:35-45), or NULL */
  int32_t mask_complement;
  int32_t include_self;
} wg_gather_opts;
typedef struct wg_gathered wg_gathered;
/* gather (mod.rs:96-145) over `count` messages of element type dtype
 * (WG_F32, WG_F64, WG_I64 or WG_BOOL; byte messages go through
 * wg_gather_bytes; any other type is WG_UNSUPPORTED), senders[i] sending
 * messages[i] (host arrays,
 * push order). The result holds the receivers that got a message,
 * ascending, each with its value: count as i64; sum, min, max, first,
 * bool_or, bool_and and mode as one message (i64 sums that overflow are
 * WG_BIGINT_OUT_OF_RANGE, the lowest receiver's); collect as its messages
 * in collect order. A combiner the element type does not take is
 * WG_COMBINER_TYPE (concat always: it takes byte messages, wg_gather_bytes);
 * a sender outside the id space,
 * WG_SENDER_OUTSIDE_ID_SPACE. */
int wg_gather(wg_context* ctx, const wg_graph* g, const wg_gather_opts* opts, int dtype, uint64_t count,
              const uint32_t* senders, const void* messages, wg_gathered** out);
/* The receivers and the values held (collect: every collected message). */
int wg_gathered_info(const wg_gathered* r, uint64_t* receivers, uint64_t* values);
/* Host copies: receivers (u64 dense ids), values (i64 for count, else the
 * messages' type), offsets (collect: receivers + 1 positions into values).
 * Any may be NULL. */
int wg_gathered_read(wg_context* ctx, const wg_gathered* r, uint64_t* receivers, void* values,
                     uint64_t* offsets);
/* gather over byte-string messages (MessageKind::Bytes, combiner.rs:166-196):
 * message i is bytes[offsets[i] .. offsets[i + 1]) (host arrays, offsets
 * count + 1 non-decreasing from 0), sent by senders[i]. Combiners collect,
 * count, first and concat (the bytes of a receiver's messages back to back
 * in collect order); any other is WG_COMBINER_TYPE. The result's receivers
 * as wg_gather's; count gives i64 values (wg_gathered_read); collect gives
 * one string a collected message with offsets per receiver
 * (wg_gathered_read's offsets), first and concat one string a receiver.
 * The strings are copied out with wg_gathered_read_bytes. */
int wg_gather_bytes(wg_context* ctx, const wg_graph* g, const wg_gather_opts* opts, uint64_t count,
                    const uint32_t* senders, const uint8_t* bytes, const uint64_t* offsets, wg_gathered** out);
/* The total length of a byte result's strings. */
int wg_gathered_bytes_info(const wg_gathered* r, uint64_t* bytes);
/* Host copies of a byte result's strings: their bytes back to back and
 * values + 1 offsets into them (values as wg_gathered_info gives). */
int wg_gathered_read_bytes(wg_context* ctx, const wg_gathered* r, uint8_t* bytes, uint64_t* string_offsets);
void wg_gathered_destroy(wg_gathered* r);

/* ---- eigenpairs: eigen (spectral/lanczos.rs:69-224) ----------------------
 * The matrix (spectral/operator.rs:33-45): S = (A + A^T)/2 with symmetrize
 * (one product when the reverse CSR equals the forward one), else A, which
 * must be symmetric (Freivalds check, 1e-9); D = the row sums of S; over the
 * active nodes (those with an edge of S). */
enum wg_matrix {
  WG_ADJACENCY = 0,            /* S */
  WG_LAPLACIAN = 1,            /* D - S */
  WG_NORMALIZED_LAPLACIAN = 2, /* I - D^-1/2 S D^-1/2 */
  WG_RANDOM_WALK = 3           /* D^-1 S, solved through D^-1/2 S D^-1/2 */
};
enum wg_which { WG_LARGEST = 0, WG_SMALLEST = 1 };

/* EigenOptions (lanczos.rs:140-156) and GraphMatrix (operator.rs:79-93). */
typedef struct {
  int32_t matrix;       /* wg_matrix */
  int32_t which;        /* wg_which */
  int32_t weighted;     /* -1: if the graph has weights */
  int32_t symmetrize;   /* 1: (A + A^T)/2; 0: A, checked symmetric */
  uint32_t k;           /* eigenpairs */
  uint32_t basis_size;  /* Krylov basis vectors; 0: max(2k + 1, 20) (working_set.rs:79-88) */
  double tol;           /* converged when |beta s_last| <= tol * max|theta| (1e-6) */
  uint64_t max_iter;    /* operator applications; 0: 1000 x basis size, raised to k */
  int32_t precision;    /* WG_F32 or WG_F64: the basis and eigenvectors */
  int32_t pad;
  uint64_t seed;        /* start vectors: random::uniform(seed, stream, node) */
} wg_eigen_opts;

/* Eigen (lanczos.rs:174-198): steps taken, the connected components of the
 * active nodes, whether every pair met tol, how many pairs came exactly
 * from the components (deflation.rs), and the thick restarts made.
 * working_set (WorkingSet::bytes, working_set.rs:114-128): the bytes the
 * call holds at its peak beyond what was in use when it began (the
 * operator's and the components' buffers, the basis and its rotation
 * coefficients); also set when the call is refused WG_MEMORY_LIMIT at
 * admission. It leaves out the caches the first product over a graph
 * builds and keeps (the context's partials scratch, the graph's layouts):
 * once those exist, a budget of in-use + working_set bytes admits the same
 * call again and one byte less is refused before the first step. */
typedef struct {
  uint64_t iterations;
  uint64_t components;
  uint32_t converged;
  uint32_t exact;
  uint32_t restarts;
  uint32_t basis_size;
  uint64_t working_set;
} wg_eigen_info;

/* Thick-restart Lanczos with full reorthogonalisation (lanczos.rs:412-743).
 * The operator runs in f64 over the device CSR pair; the basis (n x size,
 * f32 or f64 per precision, every sum in f64) stays in device memory; the
 * restart rotation is an in-place FP64 block product on the matrix cores;
 * the projected problem is solved in FP64 on the host. Writes k eigenvalues
 * in `which` order, their residuals ||Mx - lambda x||_2 for the unit
 * eigenvectors as stored, and (if vectors is not NULL) the eigenvectors as
 * the columns of an id space n x width k vector of the precision's dtype
 * (zero at inactive nodes, the entry of largest magnitude positive).
 * Errors: WG_EIGENPAIR_COUNT (k = 0 or above the active nodes),
 * WG_NOT_SYMMETRIC, WG_NON_POSITIVE_DEGREE (normalised operators),
 * WG_NOT_FINITE, WG_MEMORY_LIMIT (the working set is checked before the
 * first step), WG_INVALID_ARGUMENT. */
int wg_eigen(wg_context* ctx, const wg_graph* g, const wg_eigen_opts* opts, double* values,
             double* residuals, wg_vector* vectors, wg_eigen_info* info);

/* The eigenpairs of a symmetric m x m matrix (m <= 256), solved on the
 * device in f64: the projected problem of each Rayleigh-Ritz step, where
 * the reference calls nalgebra's SymmetricEigen (spectral/ritz.rs:34-52).
 * One workgroup reduces a to tridiagonal form (Householder) and runs
 * implicit QR steps with the Wilkinson shift. a: row-major, host memory;
 * values[l] and vectors[i * m + l] (column l the unit eigenvector of
 * values[l]) are written to host memory, in the solver's order (unsorted).
 * WG_NOT_FINITE if a value is not finite or the steps do not settle. */
int wg_sym_eigen(wg_context* ctx, uint32_t m, const double* a, double* values, double* vectors);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference: a streaming copy kernel (16-byte
 * non-temporal loads and stores, four loads in flight per thread, 32
 * workgroups per CU, grid-stride) of `bytes` from src to dst, both device buffers of
 * at least that size, on the context stream. Calibrates the achievable HBM
 * bandwidth and the profiler's FETCH_SIZE / WRITE_SIZE counters. */
int wg_device_copy(wg_context* ctx, void* dst, const void* src, uint64_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* WARREN_GPU_H */
