// Products over graphs with u64 targets (TargetWidth::U64,
// graph/csr/target.rs:16-31: the width warren-core uses for id spaces of
// 2^32 nodes or more). A thread takes one (row, column) of y and folds the
// row's products left to right, in the reference's order
// (compute/product/kernel.rs:100-109), so every monoid, float plus
// included, gives the reference's result exactly; then the mask, replace
// and accumulate rules. At these id spaces y itself is 2^32 elements or
// more, so a thread per output is the natural grain.
#pragma once

#include "semiring.hpp"
#include "wg_internal.hpp"

namespace wg {

template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(256) void spmv_wide(const uint64_t* __restrict__ off, const uint64_t* __restrict__ tgt,
                                                 const E* __restrict__ w, const E* __restrict__ x,
                                                 E* __restrict__ y, uint64_t n, uint64_t width, WriteArgs wa,
                                                 unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  constexpr bool RX = MUL != FIRST && MUL != PAIR;
  const uint64_t items = n * width;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < items; i += (uint64_t)gridDim.x * 256) {
    const uint64_t r = i / width, c = i % width;
    uint32_t e_err = 0;
    typename F::A acc{};
    bool pres = false;
    for (uint64_t p = off[r], pe = off[r + 1]; p < pe; ++p) {
      const E xv = RX ? x[tgt[p] * width + c] : E();
      const typename F::A v = F::lift(multiply<E, ADD, MUL>(HW ? w[p] : E(1), xv, e_err));
      acc = pres ? F::combine(acc, v, e_err) : v;
      pres = true;
    }
    const E res = pres ? F::finish(acc, e_err) : E();
    write_elem<E, ADD>(wa, r, &y[i], pres, res, e_err);
    if (e_err) raise_error(err, r, e_err);
  }
}

template <class E, int ADD, int MUL>
int launch_wide(wg_context* ctx, const DevCsr& c, const E* wv, const E* x, E* y, uint64_t width,
                const WriteArgs& wa) {
  const uint64_t items = c.n * width;
  if (!items) return WG_OK;
  const unsigned grid = (unsigned)std::min<uint64_t>((items + 255) / 256, 65536);
  if (wv)
    spmv_wide<E, ADD, MUL, true><<<grid, 256, 0, ctx->stream>>>(c.offsets, c.targets64, wv, x, y, c.n, width, wa,
                                                                ctx->d_err);
  else
    spmv_wide<E, ADD, MUL, false><<<grid, 256, 0, ctx->stream>>>(c.offsets, c.targets64, nullptr, x, y, c.n,
                                                                 width, wa, ctx->d_err);
  WG_HIP(ctx, hipGetLastError());
  return WG_OK;
}

}  // namespace wg
