/*
 * warren_gpu_synth.h -- synthetic inputs for tests and benchmarks. These
 * entry points have no counterpart in the reference (which ships no R-MAT
 * generator; its fixtures are G(n,m), power-law and SBM graphs); they live in
 * libwarren_gpu so the GPU runs and the CPU checks read identical graphs.
 */
#ifndef WARREN_GPU_SYNTH_H
#define WARREN_GPU_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* R-MAT (Chakrabarti, Zhan, Faloutsos 2004): edge_factor * 2^scale edges,
 * each choosing a quadrant per level with probabilities a, b, c and
 * 1-a-b-c from a counter-based SplitMix64 stream of (seed, edge, level); no
 * vertex permutation. Self loops and duplicate edges are removed; symmetrize
 * adds the reverse of every edge first. Writes the CSR (rows ascending):
 * offsets[2^scale + 1], targets[capacity] with capacity >= the generated
 * edge count (times 2 when symmetrized); *nnz_out = edges kept. threads <= 0:
 * all hardware threads. */
int wg_rmat_csr(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                uint64_t seed, int symmetrize, uint64_t* offsets, uint32_t* targets,
                uint64_t capacity, uint64_t* nnz_out, int threads);

/* The raw edge list wg_rmat_csr starts from: edge_factor * 2^scale edges
 * in generation order, self loops and parallel edges included. */
int wg_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                  uint64_t seed, uint32_t* sources, uint32_t* targets, int threads);

/* Edge weights as a function of (seed, source, target) alone:
 * kind 0: uniform in (0, 1], k * 2^-24 for k in 1..2^24 (exact in f32);
 * kind 1: integers uniform in [p0, p1];
 * kind 2: log-normal exp(p0 + p1 * N(0,1)) (Box-Muller). */
int wg_edge_weights(uint64_t n, const uint64_t* offsets, const uint32_t* targets, uint64_t seed,
                    int kind, double p0, double p1, double* weights, int threads);

#ifdef __cplusplus
}
#endif
#endif
