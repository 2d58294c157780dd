// Lanczos eigenpairs of a graph operator on gfx950 (spectral/lanczos.rs,
// operator.rs, basis.rs, deflation.rs, components.rs, ritz.rs, dense.rs,
// working_set.rs).
//
// Device side, all in f64 arithmetic:
//   operator  y = M x through the f64 SpMV over the device CSR pair, plus
//             elementwise passes (scaling by D^-1/2, the Laplacian forms);
//   basis     n x size, column-major (each basis vector contiguous), f32 or
//             f64 elements; the three-term recurrence, classical
//             Gram-Schmidt projections Q^T w and subtractions w -= Q c run as
//             HBM-bound passes of 1024 rows a workgroup, each workgroup's
//             partial sums reduced in workgroup order (results do not
//             depend on scheduling);
//   restart   Q <- Q C in place, a workgroup owning a tile of rows, on the
//             FP64 matrix cores (16x16x4 f64 MFMA);
//   deflation projection onto the complement of the connected components'
//             exact eigenvectors, summed per component in node order.
// Host side: the thick-restart control flow of the reference, the FP64
// eigendecomposition of the small projected matrix H (cyclic Jacobi), the
// connected components (union-find with parity, once per call).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

int launch_plus_times_f64(wg_context* ctx, wg_graph* g, bool transpose, bool weighted, const double* x,
                          double* y);

constexpr uint32_t kRows = 1024;       // rows per workgroup of a dense pass (4 a thread)
constexpr uint32_t kMaxBasis = 256;    // basis vectors at most
constexpr uint32_t kRotRows = 32;      // rows per workgroup of the restart rotation
constexpr double kF32Breakdown = 1e-5, kF64Breakdown = 1e-10;  // lanczos.rs:63-64
constexpr double kF32Tie = 1e-6, kF64Tie = 1e-12;              // lanczos.rs:52-53
constexpr uint64_t kStepsPerBasisVector = 1000;                // lanczos.rs:58
constexpr uint32_t kMinBasis = 20;                             // working_set.rs:22
constexpr uint32_t kPiece = 4096;  // nodes per partial sum of a component's coefficients

// ---- device helpers --------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// random::uniform (compute/random.rs:21-26): a value in [-1, 1) of
// (seed, stream, node) alone.
__device__ __forceinline__ double uniform(uint64_t seed, uint64_t stream, uint64_t node) {
  const uint64_t golden = 0x9e3779b97f4a7c15ull;
  const uint64_t bits = mix64(mix64(seed ^ mix64(stream * golden)) ^ (node * golden));
  return (double)(bits >> 11) / (double)(1ull << 52) - 1.0;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  return v;
}

// The sum over a 256-thread workgroup, in a fixed order, to every thread.
__device__ __forceinline__ double block_total(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  return t;
}

template <class E>
__device__ __forceinline__ E round_to(double v) {
  return (E)v;
}

// ---- reductions --------------------------------------------------------------

// out[c] = sum over b of part[b * ncols + c], in a fixed order; a workgroup
// a column.
// A step's coefficients (2 kMaxBasis) and scalars (8) into coherent
// page-locked host memory, then the context mailbox's sequence number with
// release: the host polls one word instead of copying each value and
// synchronising the stream. One wavefront: every lane's stores are fenced
// at system scope before lane 0 posts.
constexpr uint32_t kPostWords = 2 * kMaxBasis + 8;
__global__ __launch_bounds__(64) void lz_post(const double* __restrict__ coef, const double* __restrict__ scal,
                                              double* post, unsigned long long* mail, unsigned long long seq) {
  for (uint32_t i = threadIdx.x; i < kPostWords; i += 64) {
    const double v = i < 2 * kMaxBasis ? coef[i] : scal[i - 2 * kMaxBasis];
    __hip_atomic_store(&post[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  __threadfence_system();
  if (threadIdx.x == 0) __hip_atomic_store(&mail[0], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void reduce_cols(const double* __restrict__ part, uint32_t nblocks,
                                                   uint32_t ncols, double* __restrict__ out) {
  __shared__ double sh[4];
  const uint32_t c = blockIdx.x;
  double s = 0.0;
  for (uint32_t b = threadIdx.x; b < nblocks; b += 256) s += part[(uint64_t)b * ncols + c];
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) out[c] = t;
}

__global__ __launch_bounds__(256) void dot_partial(const double* __restrict__ a, const double* __restrict__ b,
                                                   uint64_t n, double* __restrict__ part) {
  __shared__ double sh[4];
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double s = 0.0;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    if (i < n) s += a[i] * b[i];
  }
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// sum (y - value x)^2.
__global__ __launch_bounds__(256) void residual_partial(const double* __restrict__ y, const double* __restrict__ x,
                                                        double value, uint64_t n, double* __restrict__ part) {
  __shared__ double sh[4];
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double s = 0.0;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    if (i < n) {
      const double d = y[i] - value * x[i];
      s += d * d;
    }
  }
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Per workgroup: the largest |x| and the first row holding it.
__global__ __launch_bounds__(256) void top_partial(const double* __restrict__ x, uint64_t n,
                                                   double* __restrict__ val, uint64_t* __restrict__ idx) {
  __shared__ double sv[256];
  __shared__ uint64_t si[256];
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double best = -1.0;
  uint64_t at = ~0ull;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    if (i < n && fabs(x[i]) > best) {
      best = fabs(x[i]);
      at = i;
    }
  }
  sv[threadIdx.x] = best;
  si[threadIdx.x] = at;
  __syncthreads();
  for (uint32_t d = 128; d > 0; d >>= 1) {
    if (threadIdx.x < d) {
      const double ov = sv[threadIdx.x + d];
      const uint64_t oi = si[threadIdx.x + d];
      if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    val[blockIdx.x] = sv[0];
    idx[blockIdx.x] = si[0];
  }
}

// ---- operator ----------------------------------------------------------------

__global__ __launch_bounds__(256) void op_scale(const double* __restrict__ x, const double* __restrict__ s,
                                                double* __restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    out[i] = s[i] * x[i];
}

// out = the operator's form of t = y, or (y + yt) / 2 when symmetrising
// (operator.rs:237-259, 464-470).
__global__ __launch_bounds__(256) void op_combine(const double* __restrict__ y, const double* __restrict__ yt,
                                                  const double* __restrict__ x, const double* __restrict__ deg,
                                                  const double* __restrict__ scale, int matrix, uint64_t n,
                                                  double* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    double t = y[i];
    if (yt) t = (t + yt[i]) / 2.0;
    switch (matrix) {
      case WG_LAPLACIAN: t = deg[i] * x[i] - t; break;
      case WG_NORMALIZED_LAPLACIAN: t = x[i] - scale[i] * t; break;
      case WG_RANDOM_WALK: t = scale[i] * t; break;
      default: break;
    }
    out[i] = t;
  }
}

__global__ __launch_bounds__(256) void op_random(const uint8_t* __restrict__ active, uint64_t seed, uint64_t stream,
                                                 uint64_t n, double* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    out[i] = active[i] ? uniform(seed, stream, i) : 0.0;
}

__global__ __launch_bounds__(256) void fill_f64(double* __restrict__ out, double v, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) out[i] = v;
}

// Weights that are negative or NaN (a matrix with one has eigenvalues past
// the components' at the wanted end; operator.rs:206-223).
__global__ __launch_bounds__(256) void count_negative(const double* __restrict__ w, uint64_t m,
                                                      unsigned long long* __restrict__ count) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (uint64_t)gridDim.x * 256)
    c += (w[i] != w[i] || w[i] < 0.0);
  for (int d = 32; d > 0; d >>= 1) c += __shfl_down(c, d, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

// ---- basis -------------------------------------------------------------------

template <class E>
__global__ __launch_bounds__(256) void bas_column(const E* __restrict__ Q, uint64_t ld, uint32_t j, uint64_t n,
                                                  double* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    out[i] = (double)Q[j * ld + i];
}

// Q[:, col] = w / norm, rounded to E.
template <class E>
__global__ __launch_bounds__(256) void bas_push(E* __restrict__ Q, uint64_t ld, uint32_t col, uint64_t n,
                                                const double* __restrict__ w, double norm) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    Q[col * ld + i] = round_to<E>(w[i] / norm);
}

// w -= beta q_{j-1}; partial q_j . w (basis.rs:163-178).
template <class E>
__global__ __launch_bounds__(256) void bas_recur(const E* __restrict__ Q, uint64_t ld, uint64_t n, uint32_t j,
                                                 double beta, double* __restrict__ w, double* __restrict__ part) {
  __shared__ double sh[4];
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double s = 0.0;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    if (i < n) {
      const double x = w[i] - beta * (double)Q[(uint64_t)(j - 1) * ld + i];
      w[i] = x;
      s += x * (double)Q[(uint64_t)j * ld + i];
    }
  }
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Q^T w per workgroup (part: nblocks x len), after w -= c q_pj when
// pj >= 0 with c read from *pc (basis.rs:239-260).
template <class E>
__global__ __launch_bounds__(256) void bas_project(const E* __restrict__ Q, uint64_t ld, uint64_t n, uint32_t len,
                                                   double* __restrict__ w, int32_t pj, const double* __restrict__ pc,
                                                   double* __restrict__ part) {
  __shared__ double red[4][kMaxBasis];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double x[4];
  bool in[4];
  const double c = pj >= 0 ? *pc : 0.0;
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + tid + 256 * r;
    in[r] = i < n;
    x[r] = in[r] ? w[i] : 0.0;
    if (pj >= 0 && in[r]) {
      x[r] -= c * (double)Q[(uint64_t)pj * ld + i];
      w[i] = x[r];
    }
  }
#pragma unroll 4
  for (uint32_t j = 0; j < len; ++j) {
    double acc = 0.0;
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r)
      if (in[r]) acc += x[r] * (double)Q[(uint64_t)j * ld + base + tid + 256 * r];
    acc = wave_sum(acc);
    if (lane == 0) red[wv][j] = acc;
  }
  __syncthreads();
  for (uint32_t j = tid; j < len; j += 256)
    part[(uint64_t)blockIdx.x * len + j] = ((red[0][j] + red[1][j]) + red[2][j]) + red[3][j];
}

// w -= Q c, each row's terms in basis order; partial ||w||^2 (basis.rs:263-280).
template <class E>
__global__ __launch_bounds__(256) void bas_subtract(const E* __restrict__ Q, uint64_t ld, uint64_t n, uint32_t len,
                                                    const double* __restrict__ coef, double* __restrict__ w,
                                                    double* __restrict__ part) {
  __shared__ double c[kMaxBasis];
  __shared__ double sh[4];
  for (uint32_t j = threadIdx.x; j < len; j += 256) c[j] = coef[j];
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double x[4];
  bool in[4];
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    in[r] = i < n;
    x[r] = in[r] ? w[i] : 0.0;
  }
  for (uint32_t j = 0; j < len; ++j)
#pragma unroll
    for (uint32_t r = 0; r < 4; ++r)
      if (in[r]) x[r] -= c[j] * (double)Q[(uint64_t)j * ld + base + threadIdx.x + 256 * r];
  double s = 0.0;
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r)
    if (in[r]) {
      w[base + threadIdx.x + 256 * r] = x[r];
      s += x[r] * x[r];
    }
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out = Q coef (basis.rs:221-236).
template <class E>
__global__ __launch_bounds__(256) void bas_combine(const E* __restrict__ Q, uint64_t ld, uint64_t n, uint32_t len,
                                                   const double* __restrict__ coef, double* __restrict__ out) {
  __shared__ double c[kMaxBasis];
  for (uint32_t j = threadIdx.x; j < len; j += 256) c[j] = coef[j];
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    double s = 0.0;
    for (uint32_t j = 0; j < len; ++j) s += c[j] * (double)Q[(uint64_t)j * ld + i];
    out[i] = s;
  }
}

// Q[:, 0..kept) <- Q[:, 0..len) C in place (basis.rs:182-218): a workgroup
// owns kRotRows rows, reads them whole into LDS, then writes them. C is
// len x kept, column-major (column l at C + l * len). The 16x16 output
// blocks run on the FP64 matrix cores: v_mfma_f64_16x16x4 takes a 16x4 A
// block (lane l: row l % 16, k l / 16), a 4x16 B block (lane l: k l / 16,
// column l % 16) and accumulates a 16x16 block (lane l, register q: row
// 4 q + l / 16, column l % 16).
template <class E>
__global__ __launch_bounds__(256) void bas_rotate(E* __restrict__ Q, uint64_t ld, uint64_t n, uint32_t len,
                                                  const double* __restrict__ C, uint32_t kept) {
  __shared__ double tile[kMaxBasis][kRotRows];  // [j][row]
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint64_t row0 = (uint64_t)blockIdx.x * kRotRows;
  for (uint32_t e = tid; e < len * kRotRows; e += 256) {
    const uint32_t j = e / kRotRows, r = e % kRotRows;
    const uint64_t i = row0 + r;
    tile[j][r] = i < n ? (double)Q[(uint64_t)j * ld + i] : 0.0;
  }
  __syncthreads();
  // Output blocks: 2 row blocks x ceil(kept / 16) column blocks, one wave
  // each in turn.
  const uint32_t ncb = (kept + 15) / 16;
  const uint32_t kl = lane >> 4, cl = lane & 15;
  for (uint32_t blk = wv; blk < 2 * ncb; blk += 4) {
    const uint32_t rb = blk & 1, cb = blk >> 1;
    const uint32_t col = cb * 16 + cl;
    typedef double d4 __attribute__((ext_vector_type(4)));
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t k0 = 0; k0 < len; k0 += 4) {
      const uint32_t k = k0 + kl;
      const double a = k < len ? tile[k][rb * 16 + cl] : 0.0;
      const double b = (k < len && col < kept) ? C[(uint64_t)col * len + k] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    if (col < kept) {
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) {
        const uint64_t i = row0 + rb * 16 + 4 * q + kl;
        if (i < n) Q[(uint64_t)col * ld + i] = round_to<E>(acc[q]);
      }
    }
  }
}

// ---- deflation -----------------------------------------------------------------

// Per piece of a component's node list: sum u x and sum sign u x.
__global__ __launch_bounds__(256) void defl_pieces(const uint32_t* __restrict__ order,
                                                   const uint64_t* __restrict__ piece_start, const double* __restrict__ u,
                                                   const uint8_t* __restrict__ colour, const double* __restrict__ x,
                                                   double* __restrict__ part) {
  __shared__ double sh[4];
  const uint64_t s0 = piece_start[blockIdx.x], s1 = piece_start[blockIdx.x + 1];
  double a = 0.0, b = 0.0;
  for (uint64_t p = s0 + threadIdx.x; p < s1; p += 256) {
    const uint32_t v = order[p];
    const double t = u[v] * x[v];
    a += t;
    b += colour[v] ? -t : t;
  }
  const double ta = block_total(a, sh);
  const double tb = block_total(b, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
  }
}

// A workgroup a component: its pieces' sums in a fixed order.
__global__ __launch_bounds__(256) void defl_coef(const double* __restrict__ part, const uint32_t* __restrict__ first_piece,
                                                 uint32_t ncomp, double* __restrict__ coef) {
  __shared__ double sh[4];
  const uint32_t c = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (uint32_t p = first_piece[c] + threadIdx.x; p < first_piece[c + 1]; p += 256) {
    a += part[2 * p];
    b += part[2 * p + 1];
  }
  const double ta = block_total(a, sh);
  const double tb = block_total(b, sh);
  if (threadIdx.x == 0) {
    coef[2 * c] = ta;
    coef[2 * c + 1] = tb;
  }
}

// x -= u (connected coef0 + alternating sign coef1 [bipartite]); partial ||x||^2
// (deflation.rs:169-195).
__global__ __launch_bounds__(256) void defl_apply(double* __restrict__ x, const uint32_t* __restrict__ comp,
                                                  const uint8_t* __restrict__ colour, const double* __restrict__ u,
                                                  const uint8_t* __restrict__ bipartite, const double* __restrict__ coef,
                                                  double connected, double alternating, uint64_t n,
                                                  double* __restrict__ part) {
  __shared__ double sh[4];
  const uint64_t base = (uint64_t)blockIdx.x * kRows;
  double s = 0.0;
  for (uint32_t r = 0; r < 4; ++r) {
    const uint64_t i = base + threadIdx.x + 256 * r;
    if (i < n) {
      double v = x[i];
      if (u[i] != 0.0) {
        const uint32_t c = comp[i];
        const double along = connected * coef[2 * c];
        const double across = bipartite[c] ? alternating * (colour[i] ? -1.0 : 1.0) * coef[2 * c + 1] : 0.0;
        v -= u[i] * (along + across);
        x[i] = v;
      }
      s += v * v;
    }
  }
  const double t = block_total(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void defl_vector(const uint32_t* __restrict__ comp, const uint8_t* __restrict__ colour,
                                                   const double* __restrict__ u, uint32_t c, int alternating, uint64_t n,
                                                   double* __restrict__ x) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
    double v = 0.0;
    if (u[i] != 0.0 && comp[i] == c) v = (alternating && colour[i]) ? -u[i] : u[i];
    x[i] = v;
  }
}

// x = E(x * factor) read back as f64 (lanczos.rs:732-743).
template <class E>
__global__ __launch_bounds__(256) void settle_scale(double* __restrict__ x, double factor, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    x[i] = (double)round_to<E>(x[i] * factor);
}

template <class E>
__global__ __launch_bounds__(256) void store_column(const double* __restrict__ x, uint64_t n, uint32_t width,
                                                    uint32_t col, E* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    out[i * width + col] = round_to<E>(x[i]);
}

// ---- the projected eigenproblem, on the device ---------------------------------

// Eigenpairs of the symmetric m x m matrix held row-major in a (f64, m <=
// kMaxBasis), by one workgroup (SymmetricEigen of ritz.rs:34-52): Householder
// reduction to a tridiagonal T = Q^T A Q, then implicit QR steps with the
// Wilkinson shift on T (Golub & Van Loan, algorithms 8.3.1 and 8.3.3), the
// Givens rotations accumulated into Q. a is overwritten; qt receives Q^T
// (row l = the eigenvector of values[l]); *status 0 on success, 1 when a
// value is not finite or the QR steps do not settle. Reductions are in a
// fixed order, so the result does not vary from run to run. Up to
// kEigLdsM rows, Q^T lives in LDS while the rotations run (each sweep's
// rotations are a dependent chain down a column) and is copied out at the
// end.
constexpr int kEigThreads = 256;
constexpr int kEigLdsM = 128;

// Doubles of device workspace device_sym_eigen needs for an m x m matrix.
constexpr uint64_t eigen_workspace(uint64_t m) { return 2 * m * m + m + 1; }

__device__ double eig_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
  __syncthreads();
  if ((tid & 63) == 0) sh[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < kEigThreads / 64; ++k) t += sh[k];
  return t;
}

template <bool LDSQ>
__global__ __launch_bounds__(kEigThreads) void sym_eigen_dev(double* __restrict__ a, double* __restrict__ qt,
                                                             int m, double* __restrict__ values,
                                                             int* __restrict__ status) {
  __shared__ double v[kMaxBasis], p[kMaxBasis], w[kMaxBasis], t[kMaxBasis];
  __shared__ double d[kMaxBasis], e[kMaxBasis], cs[kMaxBasis], sn[kMaxBasis];
  __shared__ double sh[kEigThreads / 64];
  __shared__ double qs[LDSQ ? kEigLdsM * kEigLdsM : 1];
  __shared__ int ctl[4];  // l, h, done, failed
  __shared__ int lmax[kEigThreads / 64];
  const int tid = threadIdx.x;
  double* const q_ = LDSQ ? qs : qt;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * m + j]; };
  auto Qt = [&](int k, int r) -> double& { return q_[(size_t)k * m + r]; };
  int bad = 0;
  for (int idx = tid; idx < m * m; idx += kEigThreads) {
    Qt(idx / m, idx % m) = (idx / m == idx % m) ? 1.0 : 0.0;
    if (!isfinite(a[idx])) bad = 1;
  }
  if (tid == 0) ctl[3] = 0;
  __syncthreads();
  if (bad) ctl[3] = 1;
  __syncthreads();
  if (ctl[3]) {
    if (tid == 0) *status = 1;
    return;
  }
  for (int k = 0; k + 2 < m; ++k) {
    const int len = m - k - 1;
    double q = 0.0;
    for (int i = tid; i < len; i += kEigThreads) {
      v[i] = A(k + 1 + i, k);
      q += v[i] * v[i];
    }
    const double norm2 = eig_sum(q, sh);
    if (norm2 == 0.0) continue;
    const double alpha = v[0] > 0.0 ? -sqrt(norm2) : sqrt(norm2);
    __syncthreads();
    if (tid == 0) v[0] -= alpha;
    __syncthreads();
    q = 0.0;
    for (int i = tid; i < len; i += kEigThreads) q += v[i] * v[i];
    const double vv = eig_sum(q, sh);
    if (vv == 0.0) continue;
    const double beta = 2.0 / vv;
    // p = beta A22 v; A22 is symmetric, so thread i walks column i (each
    // row of A22 read contiguously across the threads).
    q = 0.0;
    for (int i = tid; i < len; i += kEigThreads) {
      double sum = 0.0;
      for (int j = 0; j < len; ++j) sum += A(k + 1 + j, k + 1 + i) * v[j];
      p[i] = beta * sum;
      q += p[i] * v[i];
    }
    const double pv = eig_sum(q, sh);
    const double kk = beta * pv / 2.0;
    for (int i = tid; i < len; i += kEigThreads) w[i] = p[i] - kk * v[i];
    __syncthreads();
    // A22 -= v w^T + w v^T (stays exactly symmetric).
    for (int idx = tid; idx < len * len; idx += kEigThreads) {
      const int i = idx / len, j = idx % len;
      A(k + 1 + i, k + 1 + j) -= v[i] * w[j] + w[i] * v[j];
    }
    for (int i = tid; i < len; i += kEigThreads) {
      const double val = i == 0 ? alpha : 0.0;
      A(k + 1 + i, k) = val;
      A(k, k + 1 + i) = val;
    }
    // Q[:, k+1:] (I - beta v v^T), on the rows of qt: t = rows^T v.
    for (int r = tid; r < m; r += kEigThreads) {
      double sum = 0.0;
      for (int j = 0; j < len; ++j) sum += Qt(k + 1 + j, r) * v[j];
      t[r] = sum;
    }
    __syncthreads();
    for (int idx = tid; idx < len * m; idx += kEigThreads) {
      const int j = idx / m, r = idx % m;
      Qt(k + 1 + j, r) -= (beta * v[j]) * t[r];
    }
    __syncthreads();
  }
  for (int i = tid; i < m; i += kEigThreads) {
    d[i] = A(i, i);
    e[i] = i + 1 < m ? A(i, i + 1) : 0.0;
  }
  if (tid == 0) {
    ctl[1] = m - 1;
    ctl[2] = 0;
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  int steps = 0;
  for (;;) {
    if (tid == 0) {
      int h = ctl[1];
      while (h > 0 && fabs(e[h - 1]) <= eps * (fabs(d[h - 1]) + fabs(d[h]))) {
        e[h - 1] = 0.0;
        --h;
      }
      ctl[1] = h;
      if (h == 0) {
        ctl[2] = 1;
      } else if (++steps > 60 * m) {
        ctl[2] = 1;
        ctl[3] = 1;
      }
    }
    __syncthreads();
    if (ctl[2]) break;
    {
      // The unreduced block ending at h starts at l, the largest i < h whose
      // e[i - 1] is negligible beside its diagonal neighbours (0 if none):
      // every thread tests its i, a block maximum picks l.
      const int h = ctl[1];
      int cand = 0;
      for (int i = tid + 1; i < h; i += kEigThreads)
        if (!(fabs(e[i - 1]) > eps * (fabs(d[i - 1]) + fabs(d[i])))) cand = i;
      for (int o = 32; o > 0; o >>= 1) cand = max(cand, __shfl_xor(cand, o, 64));
      if ((tid & 63) == 0) lmax[tid >> 6] = cand;
      __syncthreads();
      if (tid == 0) {
        int l = 0;
        for (int w = 0; w < kEigThreads / 64; ++w) l = max(l, lmax[w]);
        if (l > 0) e[l - 1] = 0.0;
        ctl[0] = l;
        const double delta = (d[h - 1] - d[h]) / 2.0;
        const double eh = e[h - 1];
        const double mu = d[h] - eh * eh / (delta + (delta >= 0.0 ? 1.0 : -1.0) * hypot(delta, eh));
        double x = d[l] - mu, z = e[l];
        // The chain's running d[k] and e[k] stay in registers (each step's
        // d[k + 1] and scaled e[k + 1] are read back by the next step);
        // every other element is read from LDS before it is written.
        double dk = d[l], ek = e[l];
        for (int k = l; k < h; ++k) {
          const double dk1 = d[k + 1];
          const double ek1 = k + 1 < h ? e[k + 1] : 0.0;
          const double r = hypot(x, z);
          const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : z / r;
          if (k > l) e[k - 1] = r;
          const double ndk = c * c * dk + 2.0 * c * s * ek + s * s * dk1;
          const double ndk1 = s * s * dk - 2.0 * c * s * ek + c * c * dk1;
          const double nek = c * s * (dk1 - dk) + (c * c - s * s) * ek;
          d[k] = ndk;
          e[k] = nek;
          cs[k] = c;
          sn[k] = s;
          if (k + 1 < h) {
            z = s * ek1;
            ek = c * ek1;
            x = nek;
          } else {
            d[k + 1] = ndk1;
          }
          dk = ndk1;
        }
      }
    }
    __syncthreads();
    // The sweep's rotations, a column of Q^T per thread, in order.
    const int l = ctl[0], h = ctl[1];
    for (int r = tid; r < m; r += kEigThreads) {
      // Element k + 2 is read before element k is written, a step ahead.
      double u = Qt(l, r), u1 = Qt(l + 1, r);
      for (int k = l; k < h; ++k) {
        const double nx = k + 2 <= h ? Qt(k + 2, r) : 0.0;
        const double c = cs[k], s = sn[k];
        Qt(k, r) = c * u + s * u1;
        u = -s * u + c * u1;
        u1 = nx;
      }
      Qt(h, r) = u;
    }
    __syncthreads();
  }
  if (LDSQ)
    for (int idx = tid; idx < m * m; idx += kEigThreads) qt[idx] = qs[idx];
  bad = 0;
  for (int i = tid; i < m; i += kEigThreads) {
    values[i] = d[i];
    if (!isfinite(d[i])) bad = 1;
  }
  if (bad) ctl[3] = 1;
  __syncthreads();
  if (tid == 0) *status = ctl[3];
}

// The device solve from the host: a (m x m row-major) in, values and
// vectors (vectors[i * m + l] = component i of eigenvector l) out; ws holds
// 2 m^2 + m + 1 doubles of device memory.
static int device_sym_eigen(wg_context* ctx, double* ws, const std::vector<double>& a, int m,
                            std::vector<double>& values, std::vector<double>& vectors, bool& ok) {
  hipStream_t st = ctx->stream;
  const size_t mm = (size_t)m * m;
  double* da = ws;
  double* dq = ws + mm;
  double* dv = ws + 2 * mm;
  int* dst = (int*)(ws + 2 * mm + m);
  WG_HIP(ctx, hipMemcpyAsync(da, a.data(), mm * 8, hipMemcpyHostToDevice, st));
  if (m <= kEigLdsM)
    sym_eigen_dev<true><<<1, kEigThreads, 0, st>>>(da, dq, m, dv, dst);
  else
    sym_eigen_dev<false><<<1, kEigThreads, 0, st>>>(da, dq, m, dv, dst);
  WG_HIP(ctx, hipGetLastError());
  std::vector<double> qt(mm);
  values.resize(m);
  int status = 1;
  WG_HIP(ctx, hipMemcpyAsync(qt.data(), dq, mm * 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipMemcpyAsync(values.data(), dv, (size_t)m * 8, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipMemcpyAsync(&status, dst, 4, hipMemcpyDeviceToHost, st));
  WG_HIP(ctx, hipStreamSynchronize(st));
  ok = status == 0;
  vectors.resize(mm);
  for (int i = 0; i < m; ++i)
    for (int l = 0; l < m; ++l) vectors[(size_t)i * m + l] = qt[(size_t)l * m + i];
  return WG_OK;
}

// ritz::order (ritz.rs:56-90): indexes of values at the `which` end, the
// earlier index first among equals; for largest, runs of magnitudes within
// tie * max|value| list their positive values first.
static std::vector<int> ritz_order(const std::vector<double>& values, int which, double tie) {
  std::vector<int> order(values.size());
  std::iota(order.begin(), order.end(), 0);
  auto cmp = [&](int a, int b) {
    const double x = values[a], y = values[b];
    if (which == WG_LARGEST) {
      if (std::fabs(x) != std::fabs(y)) return std::fabs(x) > std::fabs(y);
      if (x != y) return x > y;
    } else if (x != y) {
      return x < y;
    }
    return a < b;
  };
  std::sort(order.begin(), order.end(), cmp);
  if (which == WG_LARGEST) {
    double width = 0.0;
    for (double v : values) width = std::max(width, std::fabs(v));
    width *= tie;
    size_t start = 0;
    while (start < order.size()) {
      const double first = std::fabs(values[order[start]]);
      size_t end = start;
      while (end < order.size() && first - std::fabs(values[order[end]]) <= width) ++end;
      std::stable_sort(order.begin() + start, order.begin() + end,
                       [&](int a, int b) { return !(values[a] < 0.0) && values[b] < 0.0; });
      start = end;
    }
  }
  return order;
}

// ---- host: components (components.rs) --------------------------------------------

// ---- components on the device ------------------------------------------------
// The components of the active nodes over the forward CSR's edges (either
// direction connects), numbered by their lowest node, each node's colour in a
// two-colouring (the lowest node false) and whether the component is
// bipartite: labels start as the node ids and take the smallest label of a
// neighbour (the library's own u32 min-second products over A and A^T,
// accumulated by min) and then their label's label, until a pass changes
// nothing; a breadth-first search from the lowest nodes over both directions
// gives the depths, whose parity is the colour; an edge whose ends share a
// colour (or a self loop) makes its component odd. The colour matters only
// in bipartite components, where it is the unique two-colouring.

__global__ void cc_iota(uint64_t n, uint32_t* __restrict__ label) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
    label[v] = (uint32_t)v;
}

// y[v] = y[y[v]] (one jump; labels only fall, so racing reads are safe) and
// *changed = 1 where y differs from x.
__global__ void cc_jump(uint64_t n, const uint32_t* __restrict__ x, uint32_t* y, unsigned* __restrict__ changed) {
  unsigned c = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t l = y[v];
    const uint32_t ll = y[l];
    if (ll < l) {
      l = ll;
      y[v] = l;
    }
    c |= l != x[v];
  }
  if (__syncthreads_or(c) && threadIdx.x == 0) atomicOr(changed, 1u);
}

// Roots (active nodes that are their own label) flagged for the scan.
__global__ void cc_roots(uint64_t n, const uint8_t* __restrict__ active, const uint32_t* __restrict__ label,
                         uint64_t* __restrict__ flag) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
    flag[v] = active[v] && label[v] == v;
}

// comp[v] = the number of v's root; roots[number] = the root; colour bytes
// from the depths; p0 / p1 mark the active nodes of colour 0 / 1.
__global__ void cc_number(uint64_t n, const uint8_t* __restrict__ active, const uint32_t* __restrict__ label,
                          const uint64_t* __restrict__ number, const uint32_t* __restrict__ depth,
                          uint32_t* __restrict__ comp, uint32_t* __restrict__ roots, uint8_t* __restrict__ colour,
                          uint8_t* __restrict__ p0, uint8_t* __restrict__ p1) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const bool act = active[v] != 0;
    comp[v] = act ? (uint32_t)number[label[v]] : 0;
    if (act && label[v] == v) roots[number[v]] = (uint32_t)v;
    const uint8_t c = depth ? (uint8_t)((depth[v] != 0xffffffffu) && (depth[v] & 1u)) : 0;
    colour[v] = act ? c : 0;
    p0[v] = act && !c;
    p1[v] = act && c;
  }
}

// odd[comp[v]] = 1 where active v has a neighbour of its own colour.
__global__ void cc_odd(uint64_t n, const uint8_t* __restrict__ active, const uint32_t* __restrict__ comp,
                       const uint8_t* __restrict__ colour, const uint8_t* __restrict__ y0,
                       const uint8_t* __restrict__ y1, uint8_t* __restrict__ odd) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x)
    if (active[v] && (colour[v] ? y1[v] : y0[v])) odd[comp[v]] = 1;
}

struct Components {
  std::vector<uint32_t> comp;   // component of each active node (0 elsewhere)
  std::vector<uint8_t> colour;  // two-colouring, the lowest node false
  std::vector<uint8_t> bipartite;
  uint32_t count = 0;
};

// ---- host: the solver -----------------------------------------------------------------

struct Launch {
  unsigned grid(uint64_t n) const { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384)); }
  unsigned blocks(uint64_t n) const { return (unsigned)std::max<uint64_t>(1, (n + kRows - 1) / kRows); }
};

struct Eigen {
  wg_context* ctx;
  wg_graph* g;
  wg_eigen_opts o;
  hipStream_t st;
  uint64_t n = 0, dimension = 0;
  bool weighted = false, transpose = false, normalized = false;
  std::vector<std::pair<void*, size_t>> held;
  // operator buffers (f64)
  double *input = nullptr, *scaled = nullptr, *output = nullptr, *scratch = nullptr, *deg = nullptr, *scale = nullptr;
  uint8_t* active = nullptr;
  double *part = nullptr, *coef = nullptr, *scal = nullptr, *topv = nullptr;
  // Coherent page-locked host words a step's results are posted to (lz_post).
  double *h_post = nullptr, *d_post = nullptr;
  uint64_t* topi = nullptr;
  uint32_t nblocks = 0;
  std::vector<uint8_t> h_active;
  std::vector<double> h_deg;
  // deflation
  bool deflated = false, nonnegative = true;
  double known_connected = NAN, known_bipartite = NAN;
  Components comps;
  uint32_t* d_comp = nullptr;
  uint8_t *d_colour = nullptr, *d_bip = nullptr;
  double *d_unit = nullptr, *d_dcoef = nullptr, *d_dpart = nullptr;
  uint32_t* d_order = nullptr;
  uint64_t* d_pstart = nullptr;
  uint32_t* d_firstp = nullptr;
  uint32_t npieces = 0;
  uint64_t deflation_count = 0;
  Launch L;

  Eigen(wg_context* c, wg_graph* gr, const wg_eigen_opts& op) : ctx(c), g(gr), o(op), st(c->stream), n(gr->n) {}
  ~Eigen() {
    (void)hipStreamSynchronize(st);
    for (auto& h : held) dfree(ctx, h.first, h.second);
    if (h_post) (void)hipHostFree(h_post);
  }
  template <class T>
  int get(T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count * sizeof(T), 8);
    int s = dalloc(ctx, (void**)p, bytes);
    if (!s) held.push_back({(void*)*p, bytes});
    return s;
  }
  double f64_sum_blocks(uint32_t nb) {
    // part[0..nb) -> scal[0], read back.
    reduce_cols<<<1, 256, 0, st>>>(part, nb, 1, scal);
    double v = 0.0;
    (void)hipMemcpyAsync(&v, scal, 8, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return v;
  }

  // y = S x for the f64 vector x into `into` (operator.rs:458-472).
  int multiply(const double* x, double* into, bool w) {
    int s;
    if ((s = launch_plus_times_f64(ctx, g, false, w, x, into))) return s;
    if (transpose) {
      if ((s = launch_plus_times_f64(ctx, g, true, w, x, scratch))) return s;
    }
    return WG_OK;
  }

  // output = M input (operator.rs:237-259).
  // output = M src, src the input buffer unless given (an f64 basis
  // column is read in place).
  int apply(const double* src = nullptr) {
    if (!src) src = input;
    const double* x = src;
    if (normalized) {
      op_scale<<<L.grid(n), 256, 0, st>>>(src, scale, scaled, n);
      x = scaled;
    }
    int s;
    if ((s = multiply(x, output, weighted))) return s;
    op_combine<<<L.grid(n), 256, 0, st>>>(output, transpose ? scratch : nullptr, src, deg, scale, o.matrix, n,
                                          output);
    WG_HIP(ctx, hipGetLastError());
    return WG_OK;
  }

  int setup() {
    int s;
    nblocks = L.blocks(n);
    if ((s = get(&input, n)) || (s = get(&scaled, n)) || (s = get(&output, n)) || (s = get(&scratch, n)) ||
        (s = get(&deg, n)) || (s = get(&scale, n)) || (s = get(&active, n)) ||
        (s = get(&part, (size_t)nblocks * kMaxBasis)) || (s = get(&coef, 2 * kMaxBasis)) || (s = get(&scal, 8)) ||
        (s = get(&topv, nblocks)) || (s = get(&topi, nblocks)))
      return s;
    if (hipHostMalloc((void**)&h_post, kPostWords * 8, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer((void**)&d_post, h_post, 0) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, WG_OUT_OF_MEMORY, "page-locked host words for the step's results");
    }
    weighted = (o.weighted < 0 ? g->weighted : o.weighted != 0) && g->fwd.weights;
    normalized = o.matrix == WG_NORMALIZED_LAPLACIAN || o.matrix == WG_RANDOM_WALK;
    transpose = o.symmetrize && !g->rev_is_fwd;
    if (!o.symmetrize && (s = check_symmetric())) return s;
    // Active nodes and degrees (operator.rs:336-364).
    fill_f64<<<L.grid(n), 256, 0, st>>>(input, 1.0, n);
    if ((s = multiply(input, output, false))) return s;
    op_combine<<<L.grid(n), 256, 0, st>>>(output, transpose ? scratch : nullptr, input, nullptr, nullptr,
                                          WG_ADJACENCY, n, output);
    std::vector<double> h(n);
    WG_HIP(ctx, hipMemcpyAsync(h.data(), output, n * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    h_active.resize(n);
    dimension = 0;
    for (uint64_t i = 0; i < n; ++i) {
      h_active[i] = h[i] > 0.0;
      dimension += h_active[i];
    }
    if ((s = multiply(input, output, weighted))) return s;
    op_combine<<<L.grid(n), 256, 0, st>>>(output, transpose ? scratch : nullptr, input, nullptr, nullptr,
                                          WG_ADJACENCY, n, deg);
    h_deg.resize(n);
    WG_HIP(ctx, hipMemcpyAsync(h_deg.data(), deg, n * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    std::vector<double> hs(n, 0.0);
    if (normalized) {
      for (uint64_t i = 0; i < n; ++i) {
        if (h_active[i] && (std::isnan(h_deg[i]) || h_deg[i] <= 0.0))
          return fail(ctx, WG_NON_POSITIVE_DEGREE,
                      "node " + std::to_string(i) + " has non-positive degree " + std::to_string(h_deg[i]));
        hs[i] = h_active[i] ? 1.0 / std::sqrt(h_deg[i]) : 0.0;
      }
    }
    WG_HIP(ctx, hipMemcpyAsync(scale, hs.data(), n * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(active, h_active.data(), n, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    return WG_OK;
  }

  // Freivalds: A r = A^T r for a random r (operator.rs:368-396).
  int check_symmetric() {
    std::vector<uint8_t> ones(n, 1);
    WG_HIP(ctx, hipMemcpyAsync(active, ones.data(), n, hipMemcpyHostToDevice, st));
    op_random<<<L.grid(n), 256, 0, st>>>(active, 0x53594d4dull, 0, n, input);
    int s;
    if ((s = launch_plus_times_f64(ctx, g, false, weighted, input, output)) ||
        (s = launch_plus_times_f64(ctx, g, true, weighted, input, scratch)))
      return s;
    std::vector<double> a(n), b(n);
    WG_HIP(ctx, hipMemcpyAsync(a.data(), output, n * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(b.data(), scratch, n * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    double size = 0.0, diff = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      size = std::max(size, std::max(std::fabs(a[i]), std::fabs(b[i])));
      const double d = std::fabs(a[i] - b[i]);
      diff = (std::isnan(d) || std::isnan(diff)) ? NAN : std::max(diff, d);
    }
    if (std::isnan(diff) || diff > 1e-9 * size)
      return fail(ctx, WG_NOT_SYMMETRIC, "the matrix is not symmetric: use symmetrize, or an undirected graph");
    return WG_OK;
  }

  // Components (deflation.rs, components.rs) on the device: see cc_iota.
  int prepare_components() {
    comps = Components{};
    comps.comp.assign(n, 0);
    comps.colour.assign(n, 0);
    if (n == 0) return WG_OK;
    uint32_t *lx, *ly, *comp, *roots, *dep, *via;
    uint64_t *flag, *number;
    uint8_t *colour, *p0, *p1, *y0, *y1, *odd;
    unsigned* changed;
    int s;
    if ((s = get(&lx, n)) || (s = get(&ly, n)) || (s = get(&comp, n)) || (s = get(&roots, n)) ||
        (s = get(&dep, n)) || (s = get(&via, n)) || (s = get(&flag, n)) || (s = get(&number, n)) ||
        (s = get(&colour, n)) || (s = get(&p0, n)) || (s = get(&p1, n)) || (s = get(&y0, n)) ||
        (s = get(&y1, n)) || (s = get(&odd, n)) || (s = get(&changed, 1)))
      return s;
    cc_iota<<<L.grid(n), 256, 0, st>>>(n, lx);
    WG_HIP(ctx, hipGetLastError());
    wg_product pm{};
    pm.semiring = WG_MIN_SECOND;
    pm.weighted = 0;
    pm.accumulate = WG_MIN;
    // The plain launch: these products need no tuning.
    const int saved_variant = ctx->spmv_variant;
    ctx->spmv_variant = 0;
    for (;;) {
      wg_vector X{ctx, WG_U32, n, 1, lx, n * 4}, Y{ctx, WG_U32, n, 1, ly, n * 4};
      WG_HIP(ctx, hipMemcpyAsync(ly, lx, n * 4, hipMemcpyDeviceToDevice, st));
      pm.transpose = 0;
      if ((s = wg_spmv(ctx, g, &pm, &X, &Y))) break;
      if (!g->rev_is_fwd) {
        pm.transpose = 1;
        if ((s = wg_spmv(ctx, g, &pm, &X, &Y))) break;
      }
      WG_HIP(ctx, hipMemsetAsync(changed, 0, 4, st));
      cc_jump<<<L.grid(n), 256, 0, st>>>(n, lx, ly, changed);
      WG_HIP(ctx, hipGetLastError());
      unsigned h = 0;
      WG_HIP(ctx, hipMemcpyAsync(&h, changed, 4, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      std::swap(lx, ly);
      if (!h) break;
    }
    ctx->spmv_variant = saved_variant;
    if (s) return s;
    // Number the roots in node order; list them.
    cc_roots<<<L.grid(n), 256, 0, st>>>(n, active, lx, flag);
    WG_HIP(ctx, hipGetLastError());
    // The scan's scratch is held with the rest, so that the components pass
    // makes no allocation the working set does not count.
    const size_t scan_bytes = exclusive_scan_u64_bytes(n);
    char* scan_tmp = nullptr;
    if ((s = get(&scan_tmp, scan_bytes))) return s;
    if ((s = exclusive_scan_u64(ctx, flag, number, n, scan_tmp, scan_bytes))) return s;
    uint64_t last[2] = {0, 0};
    WG_HIP(ctx, hipMemcpyAsync(&last[0], number + n - 1, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(&last[1], flag + n - 1, 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    comps.count = (uint32_t)(last[0] + last[1]);
    cc_number<<<L.grid(n), 256, 0, st>>>(n, active, lx, number, nullptr, comp, roots, colour, p0, p1);
    WG_HIP(ctx, hipGetLastError());
    if (comps.count) {
      // Depths from the roots over both directions: the colours.
      std::vector<uint32_t> r32(comps.count);
      WG_HIP(ctx, hipMemcpyAsync(r32.data(), roots, comps.count * 4, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      std::vector<uint64_t> starts(r32.begin(), r32.end());
      wg_vector D{ctx, WG_U32, n, 1, dep, n * 4}, V{ctx, WG_U32, n, 1, via, n * 4};
      if ((s = wg_bfs_levels(ctx, g, WG_BOTH, nullptr, 0, starts.data(), starts.size(), nullptr, -1, 1, &D, &V,
                             nullptr)))
        return s;
      cc_number<<<L.grid(n), 256, 0, st>>>(n, active, lx, number, dep, comp, roots, colour, p0, p1);
      WG_HIP(ctx, hipGetLastError());
      // A neighbour of each colour, then odd components.
      wg_product po{};
      po.semiring = WG_LOR_LAND;
      po.weighted = 0;
      po.accumulate = WG_OP_NONE;
      wg_vector P0{ctx, WG_BOOL, n, 1, p0, n}, P1{ctx, WG_BOOL, n, 1, p1, n};
      wg_vector Y0{ctx, WG_BOOL, n, 1, y0, n}, Y1{ctx, WG_BOOL, n, 1, y1, n};
      ctx->spmv_variant = 0;
      s = wg_spmv(ctx, g, &po, &P0, &Y0);
      if (!s) s = wg_spmv(ctx, g, &po, &P1, &Y1);
      ctx->spmv_variant = saved_variant;
      if (s) return s;
      WG_HIP(ctx, hipMemsetAsync(odd, 0, comps.count, st));
      cc_odd<<<L.grid(n), 256, 0, st>>>(n, active, comp, colour, y0, y1, odd);
      WG_HIP(ctx, hipGetLastError());
    }
    comps.bipartite.assign(comps.count, 0);
    WG_HIP(ctx, hipMemcpyAsync(comps.comp.data(), comp, n * 4, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(comps.colour.data(), colour, n, hipMemcpyDeviceToHost, st));
    if (comps.count) WG_HIP(ctx, hipMemcpyAsync(comps.bipartite.data(), odd, comps.count, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    for (auto& b : comps.bipartite) b = !b;
    return WG_OK;
  }

  // Deflation::known and new (deflation.rs:73-121).
  int prepare_deflation() {
    const bool smallest = o.which == WG_SMALLEST;
    if ((o.matrix == WG_LAPLACIAN || o.matrix == WG_NORMALIZED_LAPLACIAN) && smallest) known_connected = 0.0;
    else if (o.matrix == WG_RANDOM_WALK && !smallest) {
      known_connected = 1.0;
      known_bipartite = -1.0;
    } else if (o.matrix == WG_NORMALIZED_LAPLACIAN && !smallest) {
      known_bipartite = 2.0;
    } else {
      return WG_OK;
    }
    uint64_t nbip = 0;
    for (uint8_t b : comps.bipartite) nbip += b;
    deflation_count = (std::isnan(known_connected) ? 0 : comps.count) + (std::isnan(known_bipartite) ? 0 : nbip);
    if (deflation_count == 0) return WG_OK;
    deflated = true;
    if (weighted && g->fwd.weights) {
      unsigned long long* cnt;
      int s;
      if ((s = get(&cnt, 1))) return s;
      WG_HIP(ctx, hipMemsetAsync(cnt, 0, 8, st));
      count_negative<<<L.grid(g->nnz), 256, 0, st>>>(g->fwd.weights, g->nnz, cnt);
      unsigned long long h = 0;
      WG_HIP(ctx, hipMemcpyAsync(&h, cnt, 8, hipMemcpyDeviceToHost, st));
      WG_HIP(ctx, hipStreamSynchronize(st));
      nonnegative = h == 0;
    }
    // Unit vectors: component_weight normalised per component.
    std::vector<double> unit(n, 0.0), norm2(comps.count, 0.0);
    for (uint64_t i = 0; i < n; ++i) {
      if (!h_active[i]) continue;
      unit[i] = normalized ? std::sqrt(h_deg[i]) : 1.0;
      norm2[comps.comp[i]] += unit[i] * unit[i];
    }
    for (uint64_t i = 0; i < n; ++i)
      if (unit[i] != 0.0) unit[i] /= std::sqrt(norm2[comps.comp[i]]);
    // Each component's nodes in order, in pieces of at most kPiece.
    std::vector<uint64_t> cstart(comps.count + 1, 0);
    for (uint64_t i = 0; i < n; ++i)
      if (h_active[i]) cstart[comps.comp[i] + 1]++;
    for (uint32_t c = 0; c < comps.count; ++c) cstart[c + 1] += cstart[c];
    std::vector<uint32_t> order(dimension);
    {
      std::vector<uint64_t> next(cstart.begin(), cstart.end() - 1);
      for (uint64_t i = 0; i < n; ++i)
        if (h_active[i]) order[next[comps.comp[i]]++] = (uint32_t)i;
    }
    std::vector<uint64_t> pstart;
    std::vector<uint32_t> firstp(comps.count + 1);
    for (uint32_t c = 0; c < comps.count; ++c) {
      firstp[c] = (uint32_t)pstart.size();
      for (uint64_t p = cstart[c]; p < cstart[c + 1]; p += kPiece) pstart.push_back(p);
    }
    firstp[comps.count] = (uint32_t)pstart.size();
    npieces = (uint32_t)pstart.size();
    pstart.push_back(dimension);
    int s;
    if ((s = get(&d_comp, n)) || (s = get(&d_colour, n)) || (s = get(&d_bip, comps.count)) ||
        (s = get(&d_unit, n)) || (s = get(&d_dcoef, 2 * (size_t)comps.count)) ||
        (s = get(&d_dpart, 2 * (size_t)std::max<uint32_t>(npieces, 1))) || (s = get(&d_order, dimension)) ||
        (s = get(&d_pstart, pstart.size())) || (s = get(&d_firstp, firstp.size())))
      return s;
    WG_HIP(ctx, hipMemcpyAsync(d_comp, comps.comp.data(), n * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(d_colour, comps.colour.data(), n, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(d_bip, comps.bipartite.data(), comps.count, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(d_unit, unit.data(), n * 8, hipMemcpyHostToDevice, st));
    if (dimension) WG_HIP(ctx, hipMemcpyAsync(d_order, order.data(), dimension * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(d_pstart, pstart.data(), pstart.size() * 8, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipMemcpyAsync(d_firstp, firstp.data(), firstp.size() * 4, hipMemcpyHostToDevice, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    return WG_OK;
  }

  // x onto the complement of the components' vectors; ||x||^2 into the
  // device scalar *sq (enqueued only).
  void deflate_enqueue(double* x, double* sq) {
    if (npieces) defl_pieces<<<npieces, 256, 0, st>>>(d_order, d_pstart, d_unit, d_colour, x, d_dpart);
    if (comps.count) defl_coef<<<comps.count, 256, 0, st>>>(d_dpart, d_firstp, comps.count, d_dcoef);
    defl_apply<<<nblocks, 256, 0, st>>>(x, d_comp, d_colour, d_unit, d_bip, d_dcoef,
                                        std::isnan(known_connected) ? 0.0 : 1.0,
                                        std::isnan(known_bipartite) ? 0.0 : 1.0, n, part);
    reduce_cols<<<1, 256, 0, st>>>(part, nblocks, 1, sq);
  }

  double norm(const double* x) {
    dot_partial<<<nblocks, 256, 0, st>>>(x, x, n, part);
    return std::sqrt(f64_sum_blocks(nblocks));
  }

  // settle (lanczos.rs:732-743) for the E of the basis.
  template <class E>
  int settle(double* x) {
    const double nrm = norm(x);
    top_partial<<<nblocks, 256, 0, st>>>(x, n, topv, topi);
    std::vector<double> tv(nblocks);
    std::vector<uint64_t> ti(nblocks);
    WG_HIP(ctx, hipMemcpyAsync(tv.data(), topv, nblocks * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipMemcpyAsync(ti.data(), topi, nblocks * 8, hipMemcpyDeviceToHost, st));
    WG_HIP(ctx, hipStreamSynchronize(st));
    double best = 0.0;
    uint64_t at = ~0ull;
    for (uint32_t b = 0; b < nblocks; ++b)
      if (tv[b] > best) {
        best = tv[b];
        at = ti[b];
      }
    double sign = 1.0;
    if (at != ~0ull) {
      double v = 0.0;
      WG_HIP(ctx, hipMemcpy(&v, x + at, 8, hipMemcpyDeviceToHost));
      sign = v < 0.0 ? -1.0 : 1.0;
    }
    settle_scale<E><<<L.grid(n), 256, 0, st>>>(x, sign / nrm, n);
    WG_HIP(ctx, hipGetLastError());
    return WG_OK;
  }

  // The eigenpair of `value` for the vector in `input` (lanczos.rs:697-725).
  template <class E>
  int eigenpair(double value, double& residual) {
    int s;
    if ((s = settle<E>(input))) return s;
    if ((s = apply())) return s;
    residual_partial<<<nblocks, 256, 0, st>>>(output, input, value, n, part);
    residual = std::sqrt(f64_sum_blocks(nblocks));
    if (o.matrix == WG_RANDOM_WALK) {
      op_scale<<<L.grid(n), 256, 0, st>>>(input, scale, input, n);
      if ((s = settle<E>(input))) return s;
    }
    return WG_OK;
  }
};

struct Pair {
  double value, residual;
  std::vector<char> vec;  // n elements of E, or empty
};

template <class E>
struct Solver {
  Eigen& op;
  wg_context* ctx;
  hipStream_t st;
  uint64_t n, ld;
  uint32_t size, wanted, len = 0;
  uint64_t dim, max_iter, iterations = 0, streams = 0;
  uint32_t restarts = 0;
  // What the Rayleigh-Ritz steps so far found, for progress (lanczos.rs:438-446).
  uint32_t est_converged = 0;
  bool has_residual = false;
  double est_residual = 0.0;
  E* Q = nullptr;
  double* C = nullptr;
  double* W = nullptr;  // the projected eigenproblem's device workspace
  std::vector<double> H;
  bool has_prev = false;
  double prev = 0.0;
  double breakdown;
  Solver(Eigen& e, uint32_t sz, uint32_t w, uint64_t d, uint64_t mi)
      : op(e), ctx(e.ctx), st(e.st), n(e.n), ld(e.n), size(sz), wanted(w), dim(d), max_iter(mi) {
    breakdown = sizeof(E) == 4 ? kF32Breakdown : kF64Breakdown;
  }
  double& h(uint32_t i, uint32_t j) { return H[(size_t)i * size + j]; }

  int init() {
    int s;
    if ((s = op.get(&Q, std::max<uint64_t>(ld, 1) * size)) || (s = op.get(&C, (size_t)size * size)) ||
        (s = op.get(&W, eigen_workspace(size))))
      return s;
    H.assign((size_t)size * size, 0.0);
    return WG_OK;
  }

  // Q^T w into coef[off..off+len) after w -= c q_pj (device c at pc).
  void project(double* w, int32_t pj, const double* pc, double* out) {
    if (len == 0) return;
    bas_project<E><<<op.nblocks, 256, 0, st>>>(Q, ld, n, len, w, pj, pc, op.part);
    reduce_cols<<<len, 256, 0, st>>>(op.part, op.nblocks, len, out);
  }
  // w -= Q coef; ||w|| into *out (device).
  void subtract(double* w, const double* c, double* out) {
    bas_subtract<E><<<op.nblocks, 256, 0, st>>>(Q, ld, n, len, c, w, op.part);
    reduce_cols<<<1, 256, 0, st>>>(op.part, op.nblocks, 1, out);
  }

  // basis.rs:130-156: the coefficients Q^T w and ||w|| after; with the
  // deflation, also ||w|| after projecting out the components' vectors
  // (read in the same round trip).
  int orthogonalize(double* w, bool use_prev, std::vector<double>& coefs, double& nrm, double* deflated = nullptr) {
    coefs.assign(len, 0.0);
    double* first = op.coef;             // len
    double* corr = op.coef + kMaxBasis;  // len
    double* alpha = op.scal + 1;
    double* nd = op.scal + 2;
    double* dd = op.scal + 3;
    int32_t pj = -1;
    const bool recurrence = use_prev && len >= 2;
    if (recurrence) {
      const uint32_t j = len - 1;
      bas_recur<E><<<op.nblocks, 256, 0, st>>>(Q, ld, n, j, prev, w, op.part);
      reduce_cols<<<1, 256, 0, st>>>(op.part, op.nblocks, 1, alpha);
      pj = (int32_t)j;
    } else if (len) {
      project(w, -1, nullptr, first);
      subtract(w, first, nd);
    }
    project(w, pj, alpha, corr);
    if (len) subtract(w, corr, nd);
    else {
      dot_partial<<<op.nblocks, 256, 0, st>>>(w, w, n, op.part);
      reduce_cols<<<1, 256, 0, st>>>(op.part, op.nblocks, 1, nd);
    }
    if (deflated) op.deflate_enqueue(w, dd);
    WG_HIP(ctx, hipGetLastError());
    // first, corr and the scalars, posted to host memory in one round trip.
    const unsigned long long seq = ++ctx->mail_seq;
    lz_post<<<1, 64, 0, st>>>(op.coef, op.scal, op.d_post, ctx->d_mail, seq);
    WG_HIP(ctx, hipGetLastError());
    int s = wait_mail(ctx, seq, nullptr, 0);
    if (s) return s;
    const double* hp = op.h_post;
    std::vector<double> hf(hp + (first - op.coef), hp + (first - op.coef) + len);
    std::vector<double> hc(hp + (corr - op.coef), hp + (corr - op.coef) + len);
    const double ha = hp[2 * kMaxBasis + (alpha - op.scal)], hn = hp[2 * kMaxBasis + (nd - op.scal)];
    const double hd = hp[2 * kMaxBasis + (dd - op.scal)];
    if (deflated) *deflated = std::sqrt(hd);
    if (recurrence) {
      coefs[len - 2] = prev;
      coefs[len - 1] = ha;
    } else {
      for (uint32_t i = 0; i < len; ++i) coefs[i] = hf[i];
    }
    for (uint32_t i = 0; i < len; ++i) coefs[i] += hc[i];
    nrm = std::sqrt(hn);
    return WG_OK;
  }

  void push(const double* w, double nrm) {
    bas_push<E><<<op.L.grid(n), 256, 0, st>>>(Q, ld, len, n, w, nrm);
    ++len;
  }

  // lanczos.rs:546-570. beta < 0: breakdown.
  int step(double& beta) {
    const uint32_t j = len - 1;
    int s;
    if constexpr (std::is_same<E, double>::value) {
      if ((s = op.apply(Q + (uint64_t)j * ld))) return s;
    } else {
      bas_column<E><<<op.L.grid(n), 256, 0, st>>>(Q, ld, j, n, op.input);
      if ((s = op.apply())) return s;
    }
    ++iterations;
    std::vector<double> coefs;
    double nrm, dn = 0.0;
    if ((s = orthogonalize(op.output, has_prev, coefs, nrm, op.deflated ? &dn : nullptr))) return s;
    double b = op.deflated ? dn : nrm;
    for (uint32_t i = 0; i < len; ++i) {
      h(i, j) = coefs[i];
      h(j, i) = coefs[i];
    }
    double before = b * b;
    for (double c : coefs) before += c * c;
    before = std::sqrt(before);
    if (!std::isfinite(b)) return fail(ctx, WG_NOT_FINITE, "the eigenproblem produced a non-finite value");
    beta = b > breakdown * before ? b : -1.0;
    return WG_OK;
  }

  int extend_randomly(bool& ok) {
    ++streams;
    op_random<<<op.L.grid(n), 256, 0, st>>>(op.active, op.o.seed, streams - 1, n, op.input);
    const double before = op.norm(op.input);
    std::vector<double> coefs;
    double nrm, dn = 0.0;
    int s;
    if ((s = orthogonalize(op.input, false, coefs, nrm, op.deflated ? &dn : nullptr))) return s;
    if (op.deflated) nrm = dn;
    ok = nrm > breakdown * before;
    if (!ok) return WG_OK;
    push(op.input, nrm);
    has_prev = false;
    return WG_OK;
  }

  int extend(double beta, bool& ok) {
    if (beta >= 0.0) {
      push(op.output, beta);
      has_prev = true;
      prev = beta;
      ok = true;
      return WG_OK;
    }
    return extend_randomly(ok);
  }

  struct Ritz {
    std::vector<double> values, vectors;  // vectors[i * m + l]
    uint32_t m = 0;
  };

  int ritz(Ritz& r) {
    r.m = len;
    std::vector<double> a((size_t)len * len);
    for (uint32_t i = 0; i < len; ++i)
      for (uint32_t j = 0; j < len; ++j) a[(size_t)i * len + j] = h(i, j);
    std::vector<double> vals, vecs;
    bool solved = false;
    int s;
    if ((s = device_sym_eigen(ctx, W, a, (int)len, vals, vecs, solved))) return s;
    if (!solved) return fail(ctx, WG_NOT_FINITE, "the eigenproblem produced a non-finite value");
    const double tie = sizeof(E) == 4 ? kF32Tie : kF64Tie;
    const std::vector<int> order = ritz_order(vals, op.o.which, tie);
    r.values.resize(len);
    r.vectors.resize((size_t)len * len);
    for (uint32_t l = 0; l < len; ++l) {
      r.values[l] = vals[order[l]];
      for (uint32_t i = 0; i < len; ++i) r.vectors[(size_t)i * len + l] = vecs[(size_t)i * len + order[l]];
    }
    return WG_OK;
  }

  // 0 converged, 1 out of steps, 2 continue (lanczos.rs:630-648).
  int outcome(const Ritz& r, double beta, uint32_t& converged) {
    double scale = 0.0;
    for (double v : r.values) scale = std::max(scale, std::fabs(v));
    converged = 0;
    double largest = 0.0;
    for (uint32_t l = 0; l < wanted; ++l) {
      const double est = std::fabs(beta * r.vectors[(size_t)(len - 1) * len + l]);
      if (est <= op.o.tol * scale) ++converged;
      largest = std::max(largest, est);
    }
    est_converged = converged;
    est_residual = scale > 0.0 ? largest / scale : largest;
    has_residual = true;
    const bool done = len == dim || converged == wanted;
    if (done) return 0;
    return iterations >= max_iter ? 1 : 2;
  }

  int restart(const Ritz& r, uint32_t converged) {
    const uint32_t free = size - wanted;
    const uint32_t kept = std::min(wanted + std::clamp(converged, free / 4, free / 2), size - 1);
    std::vector<double> c((size_t)len * kept);
    for (uint32_t l = 0; l < kept; ++l)
      for (uint32_t j = 0; j < len; ++j) c[(size_t)l * len + j] = r.vectors[(size_t)j * len + l];
    WG_HIP(ctx, hipMemcpyAsync(C, c.data(), c.size() * 8, hipMemcpyHostToDevice, st));
    const unsigned tiles = (unsigned)std::max<uint64_t>(1, (n + kRotRows - 1) / kRotRows);
    bas_rotate<E><<<tiles, 256, 0, st>>>(Q, ld, n, len, C, kept);
    WG_HIP(ctx, hipGetLastError());
    WG_HIP(ctx, hipStreamSynchronize(st));
    std::fill(H.begin(), H.end(), 0.0);
    for (uint32_t l = 0; l < kept; ++l) h(l, l) = r.values[l];
    len = kept;
    return WG_OK;
  }

  int finish(const Ritz& r, std::vector<Pair>& out, bool keep_vectors) {
    for (uint32_t l = 0; l < wanted; ++l) {
      std::vector<double> c(len);
      for (uint32_t j = 0; j < len; ++j) c[j] = r.vectors[(size_t)j * len + l];
      WG_HIP(ctx, hipMemcpyAsync(op.coef, c.data(), len * 8, hipMemcpyHostToDevice, st));
      bas_combine<E><<<op.L.grid(n), 256, 0, st>>>(Q, ld, n, len, op.coef, op.input);
      Pair p;
      p.value = r.values[l];
      int s;
      if ((s = op.eigenpair<E>(p.value, p.residual))) return s;
      if (keep_vectors) {
        p.vec.resize(n * sizeof(double));
        WG_HIP(ctx, hipMemcpyAsync(p.vec.data(), op.input, n * 8, hipMemcpyDeviceToHost, st));
        WG_HIP(ctx, hipStreamSynchronize(st));
      }
      out.push_back(std::move(p));
    }
    return WG_OK;
  }

  // lanczos.rs:502-542. converged_out: whether every pair met tol.
  // The steps taken and the last estimates, to the context's progress port
  // (lanczos.rs:618-629); the pairs converged count the components' exact
  // pairs, which come first.
  void report() {
    const wg_progress& p = ctx->progress;
    if (p.done) p.done(p.user, iterations);
    if (p.detail) {
      p.detail(p.user, "converged", (double)(op.o.k - wanted + est_converged));
      p.detail(p.user, "restarts", (double)restarts);
      if (has_residual) p.detail(p.user, "residual", est_residual);
    }
  }

  int run(std::vector<Pair>& out, bool& converged_out, bool keep_vectors) {
    int s;
    if ((s = init())) return s;
    if (ctx->progress.phase) ctx->progress.phase(ctx->progress.user, "lanczos", (int64_t)max_iter);
    bool ok;
    if ((s = extend_randomly(ok))) return s;
    if (!ok)
      return fail(ctx, WG_EIGENPAIR_COUNT, "cannot compute " + std::to_string(op.o.k) +
                                               " eigenpairs of a matrix of dimension " + std::to_string(dim));
    for (;;) {
      if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
      double beta;
      if ((s = step(beta))) return s;
      report();
      if (len < size && iterations < max_iter) {
        if ((s = extend(beta, ok))) return s;
        if (ok) continue;
        Ritz r;
        if ((s = ritz(r))) return s;
        converged_out = true;
        return finish(r, out, keep_vectors);
      }
      Ritz r;
      if ((s = ritz(r))) return s;
      uint32_t conv;
      const int oc = outcome(r, beta >= 0.0 ? beta : 0.0, conv);
      report();
      if (oc == 2) {
        ++restarts;
        if ((s = restart(r, conv))) return s;
        if ((s = extend(beta, ok))) return s;
        if (!ok) {
          Ritz r2;
          if ((s = ritz(r2))) return s;
          converged_out = true;
          return finish(r2, out, keep_vectors);
        }
        has_prev = false;
        continue;
      }
      converged_out = oc == 0;
      return finish(r, out, keep_vectors);
    }
  }
};

template <class E>
static int solve(Eigen& e, uint32_t size, uint32_t wanted, uint64_t dim, uint64_t max_iter,
                 std::vector<Pair>& found, bool& converged, uint64_t& iterations, uint32_t& restarts,
                 bool keep_vectors) {
  Solver<E> sv(e, size, wanted, dim, max_iter);
  int s = sv.run(found, converged, keep_vectors);
  iterations = sv.iterations;
  restarts = sv.restarts;
  return s;
}

}  // namespace wg

using namespace wg;

extern "C" int wg_eigen(wg_context* ctx, const wg_graph* gc, const wg_eigen_opts* o, double* values,
                        double* residuals, wg_vector* vectors, wg_eigen_info* info) {
  if (!ctx || !gc || !o || !values || !residuals) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (int cs_ = check_graph(ctx, gc)) return cs_;
  if (o->matrix < 0 || o->matrix > 3)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown matrix: expected adjacency, laplacian, normalized_laplacian or random_walk");
  if (o->which != WG_LARGEST && o->which != WG_SMALLEST)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown which: expected largest or smallest");
  if (o->precision != WG_F32 && o->precision != WG_F64)
    return fail(ctx, WG_INVALID_ARGUMENT, "unknown precision: expected f32 or f64");
  wg_graph* g = const_cast<wg_graph*>(gc);
  const uint64_t n = g->n;
  if (vectors && (vectors->dtype != o->precision || vectors->id_space != n || vectors->width != o->k))
    return fail(ctx, WG_INVALID_ARGUMENT, "vectors must be of the precision's dtype, over the graph, width k");
  if (info) memset(info, 0, sizeof *info);
  if (is_cancelled(ctx)) return fail(ctx, WG_CANCELLED, "cancelled");
  WG_HIP(ctx, hipSetDevice(ctx->device));
  const uint64_t in_use_at_start = ctx->in_use.load();
  uint64_t working_set = 0;
  Eigen e(ctx, g, *o);
  int s;
  if ((s = e.setup())) return s;
  if (o->k == 0 || o->k > e.dimension)
    return fail(ctx, WG_EIGENPAIR_COUNT, "cannot compute " + std::to_string(o->k) +
                                             " eigenpairs of a matrix of dimension " + std::to_string(e.dimension));
  if ((s = e.prepare_components()) || (s = e.prepare_deflation())) return s;
  const uint64_t count = e.deflated ? e.deflation_count : 0;
  uint32_t wanted = o->k;
  if (e.deflated)
    wanted = e.nonnegative ? (uint32_t)(o->k > count ? o->k - count : 0)
                           : (uint32_t)std::min<uint64_t>(o->k, e.dimension > count ? e.dimension - count : 0);
  // The basis (working_set.rs:79-88).
  const uint64_t dim = e.dimension - count;
  uint32_t size = 0;
  if (wanted) {
    const uint64_t want = o->basis_size ? o->basis_size : std::max<uint64_t>(2 * (uint64_t)o->k + 1, kMinBasis);
    size = (uint32_t)std::min<uint64_t>(dim, want);
    if (size > kMaxBasis) return fail(ctx, WG_INVALID_ARGUMENT, "basis size above 256");
    if (size < std::min<uint64_t>(dim, (uint64_t)wanted + 1))
      return fail(ctx, WG_INVALID_ARGUMENT, "basis size must exceed the pairs Lanczos computes");
    // Admission (working_set.rs:114-128): what the operator and the
    // components already hold plus the basis, its rotation coefficients and
    // the projected eigenproblem's workspace (as Solver::init allocates
    // them), before the first step.
    const uint64_t need = std::max<uint64_t>(std::max<uint64_t>(n, 1) * size * (o->precision == WG_F32 ? 4 : 8), 8) +
                          std::max<uint64_t>((uint64_t)size * size * 8, 8) + eigen_workspace(size) * 8;
    const uint64_t held = ctx->in_use.load();
    working_set = held - std::min(held, in_use_at_start) + need;
    if (info) info->working_set = working_set;
    if (ctx->limit && held + need > ctx->limit)
      return fail(ctx, WG_MEMORY_LIMIT, "compute memory limit exceeded: the eigenproblem needs " +
                                            std::to_string(need) + " more bytes");
  }
  uint64_t max_iter = o->max_iter ? o->max_iter : kStepsPerBasisVector * size;
  max_iter = std::max<uint64_t>(max_iter, o->k);
  std::vector<Pair> found;
  bool converged = true;
  uint64_t iterations = 0;
  uint32_t restarts = 0;
  const bool keep = vectors != nullptr;
  if (wanted) {
    s = o->precision == WG_F32 ? solve<float>(e, size, wanted, dim, max_iter, found, converged, iterations, restarts, keep)
                               : solve<double>(e, size, wanted, dim, max_iter, found, converged, iterations, restarts, keep);
    if (s) return s;
  }
  // The components' pairs (deflation.rs:150-163), then merge (lanczos.rs:338-365).
  std::vector<Pair> known;
  if (e.deflated) {
    uint64_t taken = 0;
    auto add = [&](double value, uint32_t c, bool alternating) -> int {
      if (taken >= o->k) return WG_OK;
      ++taken;
      defl_vector<<<e.L.grid(n), 256, 0, e.st>>>(e.d_comp, e.d_colour, e.d_unit, c, alternating, n, e.input);
      Pair p;
      p.value = value;
      int r = o->precision == WG_F32 ? e.eigenpair<float>(value, p.residual) : e.eigenpair<double>(value, p.residual);
      if (r) return r;
      if (keep) {
        p.vec.resize(n * 8);
        WG_HIP(ctx, hipMemcpyAsync(p.vec.data(), e.input, n * 8, hipMemcpyDeviceToHost, e.st));
        WG_HIP(ctx, hipStreamSynchronize(e.st));
      }
      known.push_back(std::move(p));
      return WG_OK;
    };
    if (!std::isnan(e.known_connected))
      for (uint32_t c = 0; c < e.comps.count; ++c)
        if ((s = add(e.known_connected, c, false))) return s;
    if (!std::isnan(e.known_bipartite))
      for (uint32_t c = 0; c < e.comps.count; ++c)
        if (e.comps.bipartite[c] && (s = add(e.known_bipartite, c, true))) return s;
  }
  const uint32_t kc = (uint32_t)known.size();
  std::vector<double> exact;
  for (auto& p : known) exact.push_back(p.value);
  exact.erase(std::unique(exact.begin(), exact.end()), exact.end());
  std::vector<Pair*> all;
  for (auto& p : known) all.push_back(&p);
  for (auto& p : found) all.push_back(&p);
  double width = 0.0;
  for (Pair* p : all) width = std::max(width, std::fabs(p->value));
  width *= o->precision == WG_F32 ? kF32Tie : kF64Tie;
  std::vector<double> keys;
  for (Pair* p : all) {
    double key = p->value;
    for (double ex : exact) {
      const bool beyond = o->which == WG_SMALLEST
                              ? p->value < ex
                              : (std::fabs(p->value) > std::fabs(ex) && std::signbit(p->value) == std::signbit(ex));
      if (e.nonnegative && beyond && std::fabs(p->value - ex) <= width) {
        key = ex;
        break;
      }
    }
    keys.push_back(key);
  }
  std::vector<int> order = ritz_order(keys, o->which, o->precision == WG_F32 ? kF32Tie : kF64Tie);
  if (order.size() > o->k) order.resize(o->k);
  uint32_t exact_count = 0;
  for (size_t l = 0; l < order.size(); ++l) {
    Pair* p = all[order[l]];
    exact_count += order[l] < (int)kc;
    values[l] = p->value;
    residuals[l] = p->residual;
    if (keep) {
      WG_HIP(ctx, hipMemcpyAsync(e.input, p->vec.data(), n * 8, hipMemcpyHostToDevice, e.st));
      if (o->precision == WG_F32)
        store_column<float><<<e.L.grid(n), 256, 0, e.st>>>(e.input, n, o->k, (uint32_t)l, (float*)vectors->data);
      else
        store_column<double><<<e.L.grid(n), 256, 0, e.st>>>(e.input, n, o->k, (uint32_t)l, (double*)vectors->data);
      WG_HIP(ctx, hipGetLastError());
    }
  }
  WG_HIP(ctx, hipStreamSynchronize(e.st));
  if (info) {
    info->iterations = iterations;
    info->components = e.comps.count;
    info->converged = converged;
    info->exact = exact_count;
    info->restarts = restarts;
    info->basis_size = size;
    info->working_set = working_set;
  }
  return ok(ctx);
}

extern "C" int wg_sym_eigen(wg_context* ctx, uint32_t m, const double* a, double* values, double* vectors) {
  if (!ctx || (m && (!a || !values || !vectors))) return fail(ctx, WG_INVALID_ARGUMENT, "null argument");
  if (m > kMaxBasis) return fail(ctx, WG_INVALID_ARGUMENT, "the matrix may have at most 256 rows");
  if (m == 0) return ok(ctx);
  WG_HIP(ctx, hipSetDevice(ctx->device));
  const size_t bytes = eigen_workspace(m) * 8;
  void* ws = nullptr;
  int s;
  if ((s = dalloc(ctx, &ws, bytes))) return s;
  std::vector<double> am(a, a + (size_t)m * m), vals, vecs;
  bool solved = false;
  s = device_sym_eigen(ctx, (double*)ws, am, (int)m, vals, vecs, solved);
  dfree(ctx, ws, bytes);
  if (s) return s;
  if (!solved) return fail(ctx, WG_NOT_FINITE, "the eigenproblem produced a non-finite value");
  std::copy(vals.begin(), vals.end(), values);
  std::copy(vecs.begin(), vecs.end(), vectors);
  return ok(ctx);
}
