// The tiled SpMV (pull over one CSR, width 1) for large graphs: persistent
// 1024-thread workgroups, one per CU, each holding in 128 KB of LDS the x
// values of the most gathered columns.
//
// Layout (built once per CSR and element size, kept with the CSR):
//  * the entries cut into tiles of kTileEntries = 512; lane l of a wave owns
//    entries 8l .. 8l+7 of a tile, stored interleaved (entry q of lane l at
//    ((q / 4) * 64 + l) * 4 + q % 4) so that each 16-byte load of a wave
//    reads 1 KB contiguously;
//  * tagged targets: bit 31 set for a hot column (the low bits its LDS slot,
//    its rank by gather count), else the column (or its warm slot, below);
//    bit 30 marks the first entry
//    of a row. Edge weights in the same interleaved order, per element type.
//  * per wave a run of tiles and the number of row starts before it; per
//    64 rows a bitmap of the non-empty ones and their running count.
//
//  * columns ranked just below the hot ones (by gather count) make up the
//    warm run: their x copied, in ascending column order, into one buffer of
//    kTileWarmBytes that an XCD's L2 holds beside the streams, and their
//    entries tagged n + the column's slot in that run. The remaining columns
//    are read from x itself.
//
// A wave folds its tiles in order: each lane folds its eight entries left to
// right (a row start begins a new run), a segmented scan over the lanes in
// lane order carries the open run into the next lane, and the run open at
// the end of a tile is carried into the next tile. A finished row's value
// goes to slot j (the row's rank among non-empty rows) of a compact array;
// the run a wave begins inside a row and the run it leaves open at its end
// go to the wave's record. tile_write then writes every row (empty ones
// too) through the product's mask and accumulate rules, joining the rows
// that waves leave open from the records in wave order. Every join combines
// a run with the run that follows it, so every monoid but float plus gives
// the reference's left fold exactly
// (compute/product/kernel.rs:100-109); float plus is grouped by runs of up
// to eight entries and their lane-order scan.
#pragma once

#include "semiring.hpp"

namespace wg {

constexpr uint32_t kTileEntries = 512;
constexpr uint32_t kTileWaves = 16;
// LDS for the hot columns: 152 KB of the CU's 160 KB for 4- and 8-byte
// elements; 1-byte elements keep a power-of-two 128 K slots, read with a mask.
constexpr uint32_t kTileHotBytes = 152 * 1024, kTileHotBytesByte = 128 * 1024;
#ifndef WG_TILE_WARM_BYTES
#define WG_TILE_WARM_BYTES (2u << 20)
#endif
constexpr uint32_t kTileWarmBytes = WG_TILE_WARM_BYTES;
constexpr uint32_t kTileHot = 0x80000000u, kTileHead = 0x40000000u, kTileCol = 0x3fffffffu;

template <class E>
constexpr uint32_t tile_hot_cap() {
  return sizeof(E) == 1 ? kTileHotBytesByte : kTileHotBytes / sizeof(E);
}

template <class E>
constexpr uint32_t tile_warm_cap() {
  return kTileWarmBytes / sizeof(E);
}

template <class A>
struct TileRec {
  A P;  // the run before the wave's first row start (or the whole wave)
  A S;  // the run from the wave's last row start to its end
  uint32_t Sj, flags;  // flags: 1 P present, 2 the wave holds a row start
};

template <class T>
__device__ __forceinline__ T shfl_up_any(T v, unsigned delta) {
  constexpr int W = (sizeof(T) + 3) / 4;
  union U {
    T t;
    int w[W];
    __device__ U() {}
  } in, out;
  in.t = v;
#pragma unroll
  for (int i = 0; i < W; ++i) out.w[i] = __shfl_up(in.w[i], delta, 64);
  return out.t;
}

template <class T>
__device__ __forceinline__ T shfl_any(T v, int src) {
  constexpr int W = (sizeof(T) + 3) / 4;
  union U {
    T t;
    int w[W];
    __device__ U() {}
  } in, out;
  in.t = v;
#pragma unroll
  for (int i = 0; i < W; ++i) out.w[i] = __shfl(in.w[i], src, 64);
  return out.t;
}

// Slot of entry q of lane l within its tile.
__host__ __device__ __forceinline__ uint32_t tile_slot(uint32_t lane, uint32_t q) {
  return ((q >> 2) * 64 + lane) * 4 + (q & 3);
}

// A lane's eight weights (16-byte loads where the element allows).
template <class E>
__device__ __forceinline__ void load_weights(const E* __restrict__ w, uint64_t tb, uint32_t lane, E (&wq)[8]) {
  if constexpr (sizeof(E) == 4) {
    typedef uint32_t v4 __attribute__((ext_vector_type(4)));
    const v4 a = __builtin_nontemporal_load((const v4*)(w + tb + tile_slot(lane, 0)));
    const v4 b = __builtin_nontemporal_load((const v4*)(w + tb + tile_slot(lane, 4)));
    const uint32_t u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) __builtin_memcpy(&wq[q], &u[q], 4);
  } else if constexpr (sizeof(E) == 8) {
    typedef uint64_t v2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const v2 a = __builtin_nontemporal_load((const v2*)(w + tb + tile_slot(lane, 4 * h)));
      const v2 b = __builtin_nontemporal_load((const v2*)(w + tb + tile_slot(lane, 4 * h) + 2));
      const uint64_t u[4] = {a.x, a.y, b.x, b.y};
#pragma unroll
      for (int q = 0; q < 4; ++q) __builtin_memcpy(&wq[4 * h + q], &u[q], 8);
    }
  } else {
    typedef uint32_t v1;
    const v1 a = __builtin_nontemporal_load((const v1*)(w + tb + tile_slot(lane, 0)));
    const v1 b = __builtin_nontemporal_load((const v1*)(w + tb + tile_slot(lane, 4)));
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      wq[q] = (E)((a >> (8 * q)) & 0xff);
      wq[4 + q] = (E)((b >> (8 * q)) & 0xff);
    }
  }
}

template <class E, int ADD, int MUL, bool HW>
__global__ __launch_bounds__(1024) void spmv_tiles(
    const uint32_t* __restrict__ tagged, const E* __restrict__ w, const E* __restrict__ x,
    const E* __restrict__ xhot, uint32_t nhot, const E* __restrict__ xwarm, uint32_t ncols,
    const uint32_t* __restrict__ wave_tiles,
    const uint32_t* __restrict__ wave_heads, uint64_t nnz, typename Fold<E, ADD>::A* __restrict__ ycomp,
    TileRec<typename Fold<E, ADD>::A>* __restrict__ recs, unsigned long long* __restrict__ jerr,
    const uint32_t* __restrict__ halt) {
  using F = Fold<E, ADD>;
  using A = typename F::A;
  constexpr uint32_t HOTN = tile_hot_cap<E>();
  // A set halt word (an iterative caller's stop flag) makes the launch a no-op.
  if (halt && __hip_atomic_load(halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  constexpr bool RX = MUL != FIRST && MUL != PAIR;  // the product reads x
  __shared__ __attribute__((aligned(16))) E hot[HOTN];
  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  typedef uint32_t v4 __attribute__((ext_vector_type(4)));
  if constexpr (RX) {
    // The hot run staged in 16-byte units, a thread's loads all in flight
    // before its LDS stores (one load, wait and store at a time left the
    // workgroup waiting some 38 L2 round trips before its first tile).
    // Units wholly inside the run are vector loads from xhot (an allocation
    // of at least nhot elements, 256-byte aligned; a thread whose unit lies
    // past them loads the last one again and stores zeros); the unit the
    // run ends inside is then completed element by element by the thread
    // that stored it; slots past the run hold E().
    constexpr uint32_t PER = 16 / sizeof(E);
    constexpr uint32_t UNITS = HOTN / PER;
    static_assert(HOTN % PER == 0, "the hot run is a whole number of 16-byte units");
    constexpr uint32_t STEPS = (UNITS + 1023) / 1024;
    const uint32_t full = min(nhot, HOTN) / PER;
    const v4 zero = {0u, 0u, 0u, 0u};
    v4 buf[STEPS];
    if (full) {
#pragma unroll
      for (uint32_t s = 0; s < STEPS; ++s)
        buf[s] = *(const v4*)(xhot + (uint64_t)min(tid + s * 1024, full - 1) * PER);
    }
#pragma unroll
    for (uint32_t s = 0; s < STEPS; ++s) {
      const uint32_t u = tid + s * 1024;
      if (u < UNITS) *(v4*)(hot + (uint64_t)u * PER) = u < full ? buf[s] : zero;
    }
    if (full < UNITS && tid == full % 1024)
      for (uint32_t i = full * PER; i < nhot && i < (full + 1) * PER; ++i) hot[i] = xhot[i];
    __syncthreads();
  }
  const uint32_t gw = blockIdx.x * kTileWaves + wave;
  const uint32_t t0 = wave_tiles[gw], t1 = wave_tiles[gw + 1];
  const uint32_t Gw = wave_heads[gw];
  if (t0 >= t1) {
    if (lane == 0) {
      TileRec<A> r;
      r.P = A{};
      r.S = A{};
      r.Sj = 0;
      r.flags = 0;
      recs[gw] = r;
    }
    return;
  }
  const uint64_t lt_mask = lane ? ((1ull << lane) - 1) : 0;
  uint32_t G = Gw;  // row starts before the current tile
  A C{};            // the run open at the end of the previous tile
  bool Cp = false;
  A Pw{};
  bool Pwp = false;
  uint32_t jmin = ~0u, jcode = 0;  // the smallest row (by slot) whose product failed
  // The loads of U tiles are issued before the first of them is folded, so
  // that their latencies overlap (a wave with fewer tiles left loads its
  // last tile again and drops the copy).
  constexpr uint32_t U = sizeof(E) <= 4 ? 2 : 1;
  for (uint32_t t = t0; t < t1; t += U) {
    uint32_t TQ[U][8];
    E WQ[U][8];
    E GQ[U][8];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint64_t tb = (uint64_t)min(t + u, t1 - 1) * kTileEntries;
      const v4 ta = __builtin_nontemporal_load((const v4*)(tagged + tb + tile_slot(lane, 0)));
      const v4 tb4 = __builtin_nontemporal_load((const v4*)(tagged + tb + tile_slot(lane, 4)));
      TQ[u][0] = ta.x, TQ[u][1] = ta.y, TQ[u][2] = ta.z, TQ[u][3] = ta.w;
      TQ[u][4] = tb4.x, TQ[u][5] = tb4.y, TQ[u][6] = tb4.z, TQ[u][7] = tb4.w;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if constexpr (HW) {
        load_weights<E>(w, (uint64_t)min(t + u, t1 - 1) * kTileEntries, lane, WQ[u]);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) WQ[u][q] = E(1);
      }
    }
    // Gathers: a hot entry reads column 0 (and drops it) so that every lane
    // issues the same loads, all in flight together; a warm one reads its
    // slot of the warm run.
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        GQ[u][q] = E();
        if constexpr (RX) {
          const uint32_t tg = TQ[u][q], col = tg & kTileCol;
          const E* src = (tg & kTileHot) ? x : (col >= ncols ? xwarm + (col - ncols) : x + col);
          GQ[u][q] = *src;
        }
      }
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (u > 0 && t + u >= t1) break;
      const uint64_t tb = (uint64_t)(t + u) * kTileEntries;
      const uint32_t(&tq)[8] = TQ[u];
      const E(&wq)[8] = WQ[u];
      const E(&gq)[8] = GQ[u];
      const uint64_t e = tb + (uint64_t)lane * 8;
      uint32_t vm = 0xffu;
      if (e + 8 > nnz) vm = e >= nnz ? 0u : ((1u << (uint32_t)(nnz - e)) - 1);
      uint32_t hm = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) hm |= ((tq[q] >> 30) & 1u) << q;
      hm &= vm;
      const uint32_t nh = __popc(hm);
      uint32_t Gl = G, Gt = 0;
#pragma unroll
      for (int bit = 0; bit < 4; ++bit) {
        const uint64_t b = __ballot((nh >> bit) & 1);
        Gl += __popcll(b & lt_mask) << bit;
        Gt += __popcll(b) << bit;
      }
      // Products and the lane's runs, left to right.
      A sv[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        E xv = gq[q];
        if constexpr (RX) {
          // In bounds for every tag: cold entries read some slot and drop it.
          const uint32_t slot =
              (HOTN & (HOTN - 1)) == 0 ? (tq[q] & (HOTN - 1)) : min(tq[q] & kTileCol, HOTN - 1);
          const E hv = hot[slot];
          xv = (tq[q] & kTileHot) ? hv : xv;
        }
        uint32_t ee = 0;
        const A pv = F::lift(multiply<E, ADD, MUL>(wq[q], xv, ee));
        if (ee && ((vm >> q) & 1)) {
          // The failing entry's row: the last row start at or before it.
          const uint32_t upto = hm & ((2u << q) - 1);
          const uint32_t j = upto ? Gl + __popc(upto) - 1 : Gl - 1;
          if (j < jmin) {
            jmin = j;
            jcode = ee;
          }
        }
        uint32_t ce = 0;
        sv[q] = (q == 0 || ((hm >> q) & 1)) ? pv : F::combine(sv[q > 0 ? q - 1 : 0], pv, ce);
      }
      // The lane's open run: from its last row start (or its first entry) to
      // its last entry.
      const uint32_t nvl = vm ? 32 - __clz(vm) : 0;
      A T = sv[7];
#pragma unroll
      for (int q = 0; q < 7; ++q)
        if (nvl == (uint32_t)q + 1) T = sv[q];
      bool Tp = nvl != 0;
      const bool Fh = hm != 0;
      uint32_t ce = 0;
      if (lane == 0 && !Fh && Cp) {
        T = Tp ? F::combine(C, T, ce) : C;
        Tp = true;
      }
      // Segmented inclusive scan over lanes: flag = the lane holds a row start.
      A X = T;
      uint32_t fl = (Tp ? 1u : 0u) | (Fh ? 2u : 0u);
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const A ov = shfl_up_any(X, d);
        const uint32_t of = __shfl_up(fl, d, 64);
        if (lane >= (uint32_t)d && !(fl & 2)) {
          if (of & 1) {
            X = (fl & 1) ? F::combine(ov, X, ce) : ov;
            fl |= 1;
          }
          fl |= of & 2;
        }
      }
      A Xin = shfl_up_any(X, 1);
      bool Xinp = (__shfl_up(fl, 1, 64) & 1) != 0;
      if (lane == 0) {
        Xin = C;
        Xinp = Cp;
      }
      if (hm) {
        // The row ending just before the lane's first row start.
        const int f = __ffs(hm) - 1;
        A pre = sv[0];
#pragma unroll
        for (int q = 1; q < 7; ++q)
          if (q < f) pre = sv[q];
        A v = Xin;
        bool vp = Xinp;
        if (f > 0) {
          v = vp ? F::combine(v, pre, ce) : pre;
          vp = true;
        }
        if (Gl == Gw) {
          Pw = v;
          Pwp = vp;
        } else if (vp) {
          ycomp[Gl - 1] = v;
        }
        // Rows that start and end inside the lane.
        const uint32_t em = (hm & (hm - 1)) >> 1;
#pragma unroll
        for (int q = 0; q < 7; ++q)
          if ((em >> q) & 1) ycomp[Gl + __popc(hm & ((2u << q) - 1)) - 1] = sv[q];
      }
      C = shfl_any(X, 63);
      Cp = (__shfl(fl, 63, 64) & 1) != 0;
      G += Gt;
    }
  }
  if (jmin != ~0u) raise_error(jerr, jmin, jcode);
  const uint64_t pb = __ballot(Pwp);
  const bool has_head = G != Gw;
  A P{};
  bool Pp = false;
  if (has_head) {
    if (pb) {
      P = shfl_any(Pw, __ffsll((unsigned long long)pb) - 1);
      Pp = true;
    }
  } else {
    P = C;
    Pp = Cp;
  }
  if (lane == 0) {
    TileRec<A> r;
    r.P = P;
    r.S = has_head ? C : A{};
    r.Sj = has_head ? G - 1 : 0u;
    r.flags = (Pp ? 1u : 0u) | (has_head ? 2u : 0u);
    recs[gw] = r;
  }
}

// The rows a wave leaves open at its end, joined with the following waves'
// leading runs in wave order.
// Where the tiled layout puts a row's value: bit r of nz_bits marks a
// non-empty row, whose rank among those is nz_pre[r / 64] plus the marked
// rows before it in its word; bit r of fix_bits marks the last row start of
// a wave, whose value is joined from the wave records.
struct TileGeom {
  const uint64_t* nz_bits;
  const uint32_t* nz_pre;
  const uint64_t* fix_bits;
  const uint64_t* offsets;
  uint64_t n;
  uint64_t ntiles;
  uint32_t nwaves;
};

// The value of non-empty row r of rank j. A row a wave leaves open is the
// wave's open run joined, in wave order, with the runs that the following
// waves hold before their first row start.
template <class E, int ADD>
__device__ __forceinline__ typename Fold<E, ADD>::A tile_row_value(
    const TileGeom& g, uint64_t r, uint32_t j, bool open, const typename Fold<E, ADD>::A* __restrict__ ycomp,
    const TileRec<typename Fold<E, ADD>::A>* __restrict__ recs) {
  using F = Fold<E, ADD>;
  if (!open) return ycomp[j];
  // Wave g holds tiles [ntiles * g / nwaves, ntiles * (g + 1) / nwaves); the
  // row's start lies in the last wave whose first tile is at or before it.
  const uint64_t t = g.offsets[r] / kTileEntries;
  const uint32_t i = (uint32_t)min<uint64_t>(((t + 1) * g.nwaves + g.ntiles - 1) / g.ntiles - 1, g.nwaves - 1);
  typename F::A v = recs[i].S;
  uint32_t ce = 0;
  for (uint32_t k = i + 1; k < g.nwaves; ++k) {
    const uint32_t fk = recs[k].flags;
    if (fk & 1) v = F::combine(v, recs[k].P, ce);
    if (fk & 2) break;
  }
  return v;
}

// Rows a tile_write thread writes (four measured faster than one, eight or
// sixteen).
constexpr uint32_t kTileWriteRows = 4;

// Rows r0 .. r0 + R - 1 for a thread: each row's value (the identity for an
// empty row, the empty fold) written through the product's mask and
// accumulate rules; a plain write of the R rows is one vector store.
template <class E, int ADD>
__global__ __launch_bounds__(256) void tile_write(const TileGeom g,
                                                  const typename Fold<E, ADD>::A* __restrict__ ycomp,
                                                  const TileRec<typename Fold<E, ADD>::A>* __restrict__ recs,
                                                  E* __restrict__ y, WriteArgs wa,
                                                  const unsigned long long* __restrict__ jerr,
                                                  unsigned long long* __restrict__ err) {
  using F = Fold<E, ADD>;
  constexpr uint32_t R = kTileWriteRows;
  const uint64_t r0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * R;
  if (r0 >= g.n) return;
  const uint64_t wd = g.nz_bits[r0 >> 6], fw = g.fix_bits[r0 >> 6];
  const uint32_t b0 = (uint32_t)(r0 & 63);
  const uint32_t cnt = (uint32_t)min<uint64_t>(R, g.n - r0);
  uint32_t j = g.nz_pre[r0 >> 6] + __popcll(wd & ((1ull << b0) - 1));
  const unsigned long long je = *jerr;
  E res[R];
  uint32_t e_err[R];
  bool present[R];
#pragma unroll
  for (uint32_t k = 0; k < R; ++k) {
    present[k] = k < cnt && ((wd >> (b0 + k)) & 1);
    e_err[k] = 0;
    res[k] = E();
    if (present[k]) {
      res[k] = F::finish(tile_row_value<E, ADD>(g, r0 + k, j, (fw >> (b0 + k)) & 1, ycomp, recs), e_err[k]);
      if (je != ~0ull && (je >> 8) == j) e_err[k] = (uint32_t)(je & 0xff);
      ++j;
    }
  }
  struct alignas(R * sizeof(E)) Rows {
    E v[R];
  };
  if (!wa.mask && wa.accumulate < 0 && cnt == R && ((uintptr_t)y & (sizeof(Rows) - 1)) == 0) {
    Rows o;
#pragma unroll
    for (uint32_t k = 0; k < R; ++k) o.v[k] = present[k] ? res[k] : identity<E, ADD>();
    *(Rows*)(y + r0) = o;
  } else {
#pragma unroll
    for (uint32_t k = 0; k < R; ++k)
      if (k < cnt) write_elem<E, ADD>(wa, r0 + k, &y[r0 + k], present[k], res[k], e_err[k]);
  }
#pragma unroll
  for (uint32_t k = 0; k < R; ++k)
    if (e_err[k]) raise_error(err, r0 + k, e_err[k]);
}

}  // namespace wg
