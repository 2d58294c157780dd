/*
 * warren_oracle.h -- CPU restatement of warren-core's whole-graph compute
 * path, used ONLY as the correctness checker and the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (the pg_warren_amd package or
 * libwarren_gpu.so) links, loads or calls this library; only tests/, the
 * smoke() check and bench.py's cpu_baseline leg do.
 *
 * Every function restates the reference algorithm it cites (paths relative to
 * the reference tree, crates/warren-core/src/...). Element types are passed
 * as plain C types: f32 = float, f64 = double, i64 = int64_t, bool = uint8_t
 * (0/1), u32 = uint32_t (see or_dtype below).
 */
#ifndef WARREN_ORACLE_H
#define WARREN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the same numbering as the product's WG_* codes. */
enum {
  OR_OK = 0,
  OR_CANCELLED = 1,
  OR_DTYPE_MISMATCH = 2,
  OR_ID_SPACE_MISMATCH = 3,
  OR_WIDTH_MISMATCH = 4,
  OR_MEMORY_LIMIT = 5,
  OR_OUT_OF_MEMORY = 6,
  OR_DIVISION_BY_ZERO = 7,
  OR_FLOAT_OVERFLOW = 8,
  OR_FLOAT_UNDERFLOW = 9,
  OR_BIGINT_OUT_OF_RANGE = 10,
  OR_INVALID_ARGUMENT = 11,
  OR_NODE_OUTSIDE_ID_SPACE = 12,
  OR_NODE_FILTER_ID_SPACE = 13,
  OR_INVALID_CSR = 14,
  OR_UNSUPPORTED = 15,
  OR_EDGE_OUTSIDE_ID_SPACE = 21,
  OR_RELATION_OUTSIDE_DICTIONARY = 22
};

/* Element types (compute/dtype.rs:12-21) plus u32, the image of i64 on
 * [0, 2^32-2] with 2^32-1 standing for i64::MAX (+infinity). */
enum { OR_F32 = 0, OR_F64 = 1, OR_I64 = 2, OR_BOOL = 3, OR_U32 = 4 };

/* A CSR with dense u64 offsets (graph/csr/mod.rs:38-44). weights may be
 * NULL (an unweighted graph: every weight reads as 1). */
typedef struct {
  uint64_t n;
  const uint64_t* offsets; /* n + 1 */
  const uint32_t* targets; /* offsets[n] */
  const double* weights;   /* offsets[n] or NULL */
} or_csr;

/* Csr::from_parts checks (graph/csr/mod.rs:78-137). */
int or_csr_validate(const or_csr* csr);

/* Builds the CSR of an edge list as GraphCsr::build_as does for one relation
 * (graph/csr/build/mod.rs, build/deduplication.rs:15-80): rows sorted by
 * target, parallel edges folded with weights summed in input order, self
 * loops kept. offsets_out has n+1 slots, targets_out/weights_out m slots;
 * *nnz_out receives the folded edge count. weights / weights_out may be NULL.*/
int or_csr_build(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                 const double* weights, uint64_t* offsets_out, uint32_t* targets_out,
                 double* weights_out, uint64_t* nnz_out);

/* The merged all-relations CSR of an edge list as GraphCsr::build does
 * (graph/csr/build/mod.rs:248-288, 329-352; columns.rs:263-343;
 * deduplication.rs:83-152, 196-206): edges checked in order (the first
 * edge outside the id space, else outside the dictionary, is the error and
 * its index goes to *bad_edge_out), counted by source keeping input order,
 * each row sorted stably by (target, relation), weights summed per
 * relation in input order and the relation sums added in relation order.
 * relations may be NULL (every edge relation 0). *nnz_out = entries kept,
 * *folded_out = edges folded into an edge of the same relation.
 * keep_parallel: no folding, every edge an entry, rows in (target,
 * relation, input) order -- the per-relation multigraph a product reads
 * with Product::per_relation (compute/product/matrix.rs:276-279). */
int or_csr_build_merged(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* weights, const uint32_t* relations, uint32_t relation_count,
                        int keep_parallel, uint64_t* offsets_out, uint32_t* targets_out,
                        double* weights_out, uint64_t* nnz_out, uint64_t* folded_out,
                        uint64_t* bad_edge_out);

/* The reverse CSR by counting-sort transpose (graph/csr/transpose.rs:16-44).*/
int or_csr_transpose(const or_csr* fwd, uint64_t* offsets_out, uint32_t* targets_out,
                     double* weights_out);

/* Semiring product y = A (x) x or A^T (x) x (compute/product/mod.rs:77-205,
 * kernel.rs:84-239). Semiring index in Semiring::ALL order
 * (compute/semiring.rs:149-163); accumulate is a BinaryOp index
 * (semiring.rs:42-53) or -1 for none. rev may be NULL: a transposed product
 * then runs the sequential push scatter (kernel.rs:145-182). mask may be
 * NULL. threads <= 0 means hardware concurrency. */
int or_spmv(const or_csr* fwd, const or_csr* rev, int semiring, int dtype, int transpose,
            int weighted, const void* mask, int mask_dtype, int mask_complement, int replace,
            int accumulate, uint64_t width, const void* x, void* y, int threads);

/* PageRank (compute/pagerank/mod.rs:96-287). tombstones: n bytes (non-zero =
 * tombstoned) or NULL. seed: n doubles or NULL. */
int or_pagerank(const or_csr* fwd, const or_csr* rev, const uint8_t* tombstones,
                double damping, double tol, int has_tol, uint32_t max_iter,
                const double* seed, double* out, uint32_t* iterations, int threads);

/* Adjoint of or_pagerank with respect to the edge weights of fwd (which must
 * carry weights): from the forward ranks x and g = dL/dx, the adjoint
 * lambda (n) and dL/dw (one per forward entry). */
int or_pagerank_backward(const or_csr* fwd, const uint8_t* tombstones, double damping,
                         double tol, int has_tol, uint32_t max_iter, const double* seed,
                         const double* x, const double* g, double* lambda, double* dldw,
                         uint32_t* iterations, int threads);

/* A BFS visit (traversal/bfs.rs:15-23); via = UINT64_MAX for a start. */
typedef struct {
  uint64_t node;
  uint64_t via;
  uint32_t depth;
  uint32_t pad;
} or_visit;

typedef struct {
  uint64_t nodes_visited;
  uint64_t edges_traversed;
  uint32_t truncated;
  uint32_t levels_pulled;
} or_stats;

/* Level-synchronous BFS (traversal/level.rs:88-337, bfs.rs:34-111).
 * adj: the rows followed; pull_rev: the same edges read backwards, or NULL
 * for a push-only search (warren.bfs); given, levels may be pulled by
 * Beamer's switch with alpha 1, beta 24 (reachable_set, reachable.rs:79-90).
 * filter: n bytes or NULL. max_depth < 0: unbounded; max_results < 0:
 * unbounded. visits_out must hold n entries. */
int or_bfs(const or_csr* adj, const or_csr* pull_rev, const uint64_t* starts, uint64_t nstarts,
           const uint8_t* filter, int64_t max_depth, int64_t max_results, or_visit* visits_out,
           uint64_t* count_out, or_stats* stats_out);

/* or_bfs over an Adjacency's CSR lists (traversal/adjacency.rs:20-221):
 * rows[0..k) read in order for a node's row entries, rrows[0..kr) those of
 * the reversed adjacency (NULL: no pulled level); id space n. */
int or_bfs_rows(uint64_t n, const or_csr* const* rows, uint32_t k, const or_csr* const* rrows, uint32_t kr,
                const uint64_t* starts, uint64_t nstarts, const uint8_t* filter, int64_t max_depth,
                int64_t max_results, or_visit* visits_out, uint64_t* count_out, or_stats* stats_out);

/* Depth of every node from each source (UINT32_MAX: unreached): the
 * per-source ground truth a multi-source BFS must match
 * (compute/centrality/multi_source.rs:92-193 reports the same distances). */
int or_bfs_depths(const or_csr* adj, const uint64_t* sources, uint64_t nsources,
                  uint32_t* depths_out /* nsources x n */, int threads);

/* One BinaryOp on two elements of dtype (SemiringElement::apply,
 * semiring.rs:236-388), or with monoid >= 0 the semiring multiply
 * SemiringElement::multiply(monoid, op, a, b); a, b, out point at elements. */
int or_apply(int op, int monoid, int dtype, const void* a, const void* b, void* out);

/* An edge weight as an element (SemiringElement::from_weight, cast.rs:87-108).*/
int or_from_weight(int dtype, double weight, void* out);

/* The counter-based SplitMix64 uniform in [-1, 1) (compute/random.rs:11-26).*/
double or_uniform(uint64_t seed, uint64_t stream, uint64_t node);

#ifdef __cplusplus
}
#endif
#endif
