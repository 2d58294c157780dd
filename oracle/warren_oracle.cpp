// warren_oracle.cpp -- CPU restatement of warren-core's whole-graph compute
// path. TEST INFRASTRUCTURE: the correctness checker and the timed CPU
// baseline, never part of the product. Citations are relative to the
// reference's crates/warren-core/src/.
//
// Parity status of each function (how it is pinned):
//   semiring ops / casts   golden tables of compute/semiring.rs:437-533 and
//                          compute/cast.rs:110-150 (tests/test_oracle.py)
//   or_spmv                known answers of compute/product/tests.rs:51-198,
//                          pull == push bit for bit, naive dense product
//   or_pagerank            bit-exact golden ranks of pagerank/tests.rs:78-123,
//                          ring closed form, NetworkX (golden fixture)
//   or_bfs                 visit lists of traversal/bfs.rs:143-235, naive BFS,
//                          pushed == pulled depths
//   or_bfs_depths          brute-force BFS on tiny graphs
//   or_uniform             SplitMix64 closed form (compute/random.rs:14-26)
//   or_csr_build/transpose golden CSR of the product tests' example graph
#include "warren_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// BinaryOp, Monoid and Semiring tables (compute/semiring.rs:15-211).
enum Op { FIRST = 0, SECOND, PAIR, PLUS, TIMES, MIN, MAX, LOR, LAND, ANY };
// Semiring::ALL order: plus_times min_plus max_plus max_times min_times
// lor_land any_pair plus_pair plus_first plus_second min_first min_second
// max_second.
const int kAdd[13] = {PLUS, MIN, MAX, MAX, MIN, LOR, ANY, PLUS, PLUS, PLUS, MIN, MIN, MAX};
const int kMul[13] = {TIMES, PLUS, PLUS, TIMES, TIMES, LAND, PAIR, PAIR, FIRST, SECOND,
                      FIRST, SECOND, SECOND};

struct Err {
  int code = OR_OK;
};

// SQL float order (compute/dtype.rs:142-150): NaN equal to NaN and above
// every number; -0 == 0.
template <class F>
int sql_cmp_f(F a, F b) {
  bool an = std::isnan(a), bn = std::isnan(b);
  if (an && bn) return 0;
  if (an) return 1;
  if (bn) return -1;
  return a < b ? -1 : (a > b ? 1 : 0);
}

// least / greatest keep the first operand on a tie (semiring.rs:264-278).
template <class F>
F least_f(F a, F b) { return sql_cmp_f(b, a) < 0 ? b : a; }
template <class F>
F greatest_f(F a, F b) { return sql_cmp_f(b, a) > 0 ? b : a; }

// float8 -> float4 (compute/cast.rs:87-97).
int float4_from_float8(double v, float* out) {
  float r = (float)v;
  if (std::isinf(r) && !std::isinf(v)) return OR_FLOAT_OVERFLOW;
  if (r == 0.0f && v != 0.0) return OR_FLOAT_UNDERFLOW;
  *out = r;
  return OR_OK;
}

// float8 -> int8, round half to even then range check (cast.rs:99-108).
int int8_from_float8(double v, int64_t* out) {
  double r = std::nearbyint(v);  // default rounding mode: ties to even
  if (!(r >= -9223372036854775808.0 && r < 9223372036854775808.0)) return OR_BIGINT_OUT_OF_RANGE;
  *out = (int64_t)r;
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Element types (compute/semiring.rs:236-388). Each has apply, identity,
// from_weight, multiply and is_nonzero.
template <class E>
struct Elem;

template <class F>
struct FloatElem {
  static F apply(int op, F a, F b, Err&) {
    switch (op) {
      case FIRST: case ANY: return a;
      case SECOND: return b;
      case PAIR: return (F)1;
      case PLUS: return a + b;
      case TIMES: return a * b;
      case MIN: return least_f(a, b);
      case MAX: return greatest_f(a, b);
      case LOR: return (a != 0 || b != 0) ? (F)1 : (F)0;
      case LAND: return (a != 0 && b != 0) ? (F)1 : (F)0;
    }
    return a;
  }
  static F identity(int monoid) {
    if (monoid == MIN) return std::numeric_limits<F>::infinity();
    if (monoid == MAX) return -std::numeric_limits<F>::infinity();
    return (F)0;
  }
  static F multiply(int, int mul, F w, F v, Err& e) { return apply(mul, w, v, e); }
  static bool nonzero(F v) { return v != 0; }  // NaN != 0
  static double to_f64(F v) { return (double)v; }
};

template <>
struct Elem<double> : FloatElem<double> {
  static double from_weight(double w, Err&) { return w; }
};
template <>
struct Elem<float> : FloatElem<float> {
  static float from_weight(double w, Err& e) {
    float r = 0;
    int c = float4_from_float8(w, &r);
    if (c) e.code = c;
    return r;
  }
};

const int64_t I64MAX = std::numeric_limits<int64_t>::max();
const int64_t I64MIN = std::numeric_limits<int64_t>::min();

template <>
struct Elem<int64_t> {
  static int64_t apply(int op, int64_t a, int64_t b, Err& e) {
    int64_t r;
    switch (op) {
      case FIRST: case ANY: return a;
      case SECOND: return b;
      case PAIR: return 1;
      case PLUS:
        if (__builtin_add_overflow(a, b, &r)) { e.code = OR_BIGINT_OUT_OF_RANGE; return 0; }
        return r;
      case TIMES:
        if (__builtin_mul_overflow(a, b, &r)) { e.code = OR_BIGINT_OUT_OF_RANGE; return 0; }
        return r;
      case MIN: return std::min(a, b);
      case MAX: return std::max(a, b);
      case LOR: return (a != 0 || b != 0) ? 1 : 0;
      case LAND: return (a != 0 && b != 0) ? 1 : 0;
    }
    return a;
  }
  static int64_t identity(int monoid) {
    if (monoid == MIN) return I64MAX;
    if (monoid == MAX) return I64MIN;
    return 0;
  }
  // The min/max identities stand for +-infinity and absorb plus, and times
  // by a weight above zero (semiring.rs:340-363).
  static int64_t multiply(int add, int mul, int64_t w, int64_t v, Err& e) {
    bool infinite = (add == MIN && v == I64MAX) || (add == MAX && v == I64MIN);
    bool absorbs = mul == PLUS || (mul == TIMES && w > 0);
    if (infinite && absorbs) return v;
    return apply(mul, w, v, e);
  }
  static int64_t from_weight(double w, Err& e) {
    int64_t r = 0;
    int c = int8_from_float8(w, &r);
    if (c) e.code = c;
    return r;
  }
  static bool nonzero(int64_t v) { return v != 0; }
  static double to_f64(int64_t v) { return (double)v; }
};

template <>
struct Elem<uint8_t> {  // bool (semiring.rs:370-388)
  static uint8_t apply(int op, uint8_t a, uint8_t b, Err&) {
    switch (op) {
      case FIRST: case ANY: return a;
      case SECOND: return b;
      case PAIR: return 1;
      case PLUS: case MAX: case LOR: return (a || b) ? 1 : 0;
      case TIMES: case MIN: case LAND: return (a && b) ? 1 : 0;
    }
    return a;
  }
  static uint8_t identity(int monoid) { return monoid == MIN ? 1 : 0; }
  static uint8_t multiply(int, int mul, uint8_t w, uint8_t v, Err& e) { return apply(mul, w, v, e); }
  static uint8_t from_weight(double w, Err&) { return w != 0.0 ? 1 : 0; }  // NaN -> true
  static bool nonzero(uint8_t v) { return v != 0; }
  static double to_f64(uint8_t v) { return v ? 1.0 : 0.0; }
};

// u32: i64 restricted to [0, 2^32-2], with 2^32-1 standing for i64::MAX.
// Every operation maps to i64, runs the i64 rule and maps back; a result
// outside the image is BigintOutOfRange. The max monoid's identity i64::MIN
// has no image, so max-semirings are refused for u32 before any work.
const uint32_t U32INF = 0xffffffffu;
inline int64_t u2i(uint32_t v) { return v == U32INF ? I64MAX : (int64_t)v; }
inline uint32_t i2u(int64_t v, Err& e) {
  if (v == I64MAX) return U32INF;
  if (v < 0 || v > (int64_t)0xfffffffe) { e.code = OR_BIGINT_OUT_OF_RANGE; return 0; }
  return (uint32_t)v;
}
template <>
struct Elem<uint32_t> {
  static uint32_t apply(int op, uint32_t a, uint32_t b, Err& e) {
    return i2u(Elem<int64_t>::apply(op, u2i(a), u2i(b), e), e);
  }
  static uint32_t identity(int monoid) { return monoid == MIN ? U32INF : 0; }
  static uint32_t multiply(int add, int mul, uint32_t w, uint32_t v, Err& e) {
    return i2u(Elem<int64_t>::multiply(add, mul, u2i(w), u2i(v), e), e);
  }
  static uint32_t from_weight(double w, Err& e) {
    int64_t r = Elem<int64_t>::from_weight(w, e);
    if (e.code) return 0;
    if (r == I64MAX) { e.code = OR_BIGINT_OUT_OF_RANGE; return 0; }  // not a weight
    return i2u(r, e);
  }
  static bool nonzero(uint32_t v) { return v != 0; }
  static double to_f64(uint32_t v) { return (double)u2i(v); }
};

// SQL float8 operators with their checks (compute/map/arithmetic.rs:49-79).
enum Arith { ADD_, SUB_, MUL_, DIV_ };
double sqlf(int op, double a, double b, Err& e) {
  if (op == DIV_ && b == 0.0 && !std::isnan(a)) { e.code = OR_DIVISION_BY_ZERO; return 0; }
  double r = op == ADD_ ? a + b : op == SUB_ ? a - b : op == MUL_ ? a * b : a / b;
  bool finite_operands = op == DIV_ ? !std::isinf(a) : (!std::isinf(a) && !std::isinf(b));
  if (std::isinf(r) && finite_operands) { e.code = OR_FLOAT_OVERFLOW; return 0; }
  bool under = op == MUL_ ? (a != 0.0 && b != 0.0)
               : op == DIV_ ? (a != 0.0 && !std::isinf(b)) : false;
  if (r == 0.0 && under) { e.code = OR_FLOAT_UNDERFLOW; return 0; }
  return r;
}

int hw_threads(int threads) {
  if (threads > 0) return threads;
  unsigned h = std::thread::hardware_concurrency();
  return h ? (int)h : 1;
}

// Runs body(task) for task in [0, tasks) on `threads` threads; returns the
// error of the lowest failing task (the reference returns some failing
// task's error; this restatement picks the lowest, deterministically).
template <class Body>
int parallel_tasks(uint64_t tasks, int threads, Body body) {
  std::vector<int> status(tasks, OR_OK);
  std::atomic<uint64_t> next(0);
  auto worker = [&]() {
    for (;;) {
      uint64_t t = next.fetch_add(1);
      if (t >= tasks) break;
      status[t] = body(t);
    }
  };
  int nt = (int)std::min<uint64_t>((uint64_t)hw_threads(threads), std::max<uint64_t>(tasks, 1));
  if (nt <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  for (uint64_t t = 0; t < tasks; ++t)
    if (status[t]) return status[t];
  return OR_OK;
}

// ---------------------------------------------------------------------------
// Products (compute/product/kernel.rs).
template <class E>
bool mask_selects(const void* mask, int mask_dtype, int complement, uint64_t row) {
  bool set;
  switch (mask_dtype) {
    case OR_F32: set = ((const float*)mask)[row] != 0.0f; break;
    case OR_F64: set = ((const double*)mask)[row] != 0.0; break;
    case OR_I64: set = ((const int64_t*)mask)[row] != 0; break;
    case OR_U32: set = ((const uint32_t*)mask)[row] != 0; break;
    default: set = ((const uint8_t*)mask)[row] != 0; break;
  }
  return set != (complement != 0);
}

struct ProductArgs {
  int add, mul;
  int weighted;
  const void* mask;
  int mask_dtype, complement, replace, accumulate;
  uint64_t width;
};

// Write::row (kernel.rs:219-239).
template <class E>
void write_row(const ProductArgs& a, uint64_t row, E* out, const E* t, Err& e) {
  if (a.mask && !mask_selects<E>(a.mask, a.mask_dtype, a.complement, row)) {
    if (a.replace)
      for (uint64_t k = 0; k < a.width; ++k) out[k] = Elem<E>::identity(a.add);
    return;
  }
  if (a.accumulate < 0) {
    if (t) for (uint64_t k = 0; k < a.width; ++k) out[k] = t[k];
    else for (uint64_t k = 0; k < a.width; ++k) out[k] = Elem<E>::identity(a.add);
  } else if (t) {
    for (uint64_t k = 0; k < a.width; ++k) out[k] = Elem<E>::apply(a.accumulate, out[k], t[k], e);
  }
}

// Fixed::multiply (kernel.rs:39-47): the weight is cast only when read.
template <class E>
inline E product_of(const ProductArgs& a, double weight, E value, Err& e) {
  E w = E();
  if (a.mul != SECOND && a.mul != PAIR) w = Elem<E>::from_weight(a.weighted ? weight : 1.0, e);
  return Elem<E>::multiply(a.add, a.mul, w, value, e);
}

const uint64_t ROWS_PER_TASK = 4096;  // kernel.rs:15

// pull_with (kernel.rs:84-121): each row folded in ascending column order,
// its first entry initialising the sum.
template <class E>
int pull(const or_csr* m, const ProductArgs& a, const E* x, E* y, int threads) {
  uint64_t n = m->n, w = a.width;
  uint64_t tasks = (n + ROWS_PER_TASK - 1) / ROWS_PER_TASK;
  return parallel_tasks(tasks, threads, [&](uint64_t task) -> int {
    Err e;
    std::vector<E> sum(w);
    uint64_t r1 = std::min(n, (task + 1) * ROWS_PER_TASK);
    for (uint64_t row = task * ROWS_PER_TASK; row < r1; ++row) {
      bool present = false;
      for (uint64_t p = m->offsets[row]; p < m->offsets[row + 1]; ++p) {
        uint64_t col = m->targets[p];
        double wt = m->weights ? m->weights[p] : 1.0;
        for (uint64_t k = 0; k < w; ++k) {
          E prod = product_of<E>(a, wt, x[col * w + k], e);
          sum[k] = present ? Elem<E>::apply(a.add, sum[k], prod, e) : prod;
          if (e.code) return e.code;
        }
        present = true;
      }
      write_row<E>(a, row, y + row * w, present ? sum.data() : nullptr, e);
      if (e.code) return e.code;
    }
    return OR_OK;
  });
}

// push_with (kernel.rs:145-182): y = A^T x by a sequential scatter over the
// forward rows in ascending order, then the write.
template <class E>
int push(const or_csr* m, const ProductArgs& a, const E* x, E* y) {
  uint64_t n = m->n, w = a.width;
  std::vector<E> sums(n * w);
  std::vector<uint8_t> present(n, 0);
  Err e;
  for (uint64_t row = 0; row < n; ++row) {
    for (uint64_t p = m->offsets[row]; p < m->offsets[row + 1]; ++p) {
      uint64_t col = m->targets[p];
      double wt = m->weights ? m->weights[p] : 1.0;
      for (uint64_t k = 0; k < w; ++k) {
        E prod = product_of<E>(a, wt, x[row * w + k], e);
        E& s = sums[col * w + k];
        s = present[col] ? Elem<E>::apply(a.add, s, prod, e) : prod;
        if (e.code) return e.code;
      }
      present[col] = 1;
    }
  }
  for (uint64_t row = 0; row < n; ++row) {
    write_row<E>(a, row, y + row * w, present[row] ? &sums[row * w] : nullptr, e);
    if (e.code) return e.code;
  }
  return OR_OK;
}

// pull_with's SpMV path (width 1, kernel.rs:100-109) with the semiring fixed
// at compile time, as warren-core monomorphises each semiring and element
// type (kernel.rs:51-70): the same operations in the same order as pull,
// the weight still cast per entry (kernel.rs:40-47), without a run-time
// dispatch or a heap buffer per element.
template <class E, int ADD, int MUL>
int pull_mono(const or_csr* m, const ProductArgs& a, const E* x, E* y, int threads) {
  const uint64_t n = m->n;
  const uint64_t tasks = (n + ROWS_PER_TASK - 1) / ROWS_PER_TASK;
  constexpr bool reads_w = MUL != SECOND && MUL != PAIR;
  const double* wts = a.weighted ? m->weights : nullptr;
  return parallel_tasks(tasks, threads, [&](uint64_t task) -> int {
    Err e;
    const uint64_t r1 = std::min(n, (task + 1) * ROWS_PER_TASK);
    for (uint64_t row = task * ROWS_PER_TASK; row < r1; ++row) {
      const uint64_t p0 = m->offsets[row], p1 = m->offsets[row + 1];
      E sum{};
      for (uint64_t p = p0; p < p1; ++p) {
        E w = E();
        if (reads_w) w = Elem<E>::from_weight(wts ? wts[p] : 1.0, e);
        const E prod = Elem<E>::multiply(ADD, MUL, w, x[m->targets[p]], e);
        sum = p == p0 ? prod : Elem<E>::apply(ADD, sum, prod, e);
        if (e.code) return e.code;
      }
      write_row<E>(a, row, y + row, p1 > p0 ? &sum : nullptr, e);
      if (e.code) return e.code;
    }
    return OR_OK;
  });
}

template <class E>
int pull_width1(const or_csr* m, const ProductArgs& a, const E* x, E* y, int threads) {
  if (a.width != 1) return pull<E>(m, a, x, y, threads);
#define OR_MONO(A_, M_) \
  if (a.add == A_ && a.mul == M_) return pull_mono<E, A_, M_>(m, a, x, y, threads);
  OR_MONO(PLUS, TIMES) OR_MONO(MIN, PLUS) OR_MONO(MAX, PLUS) OR_MONO(MAX, TIMES) OR_MONO(MIN, TIMES)
  OR_MONO(LOR, LAND) OR_MONO(ANY, PAIR) OR_MONO(PLUS, PAIR) OR_MONO(PLUS, FIRST) OR_MONO(PLUS, SECOND)
  OR_MONO(MIN, FIRST) OR_MONO(MIN, SECOND) OR_MONO(MAX, SECOND)
#undef OR_MONO
  return pull<E>(m, a, x, y, threads);
}

template <class E>
int spmv_typed(const or_csr* fwd, const or_csr* rev, int transpose, const ProductArgs& a,
               const void* x, void* y, int threads) {
  const E* xs = (const E*)x;
  E* ys = (E*)y;
  const or_csr* m = !transpose ? fwd : rev;
  if (m) return a.width == 1 ? pull_width1<E>(m, a, xs, ys, threads) : pull<E>(m, a, xs, ys, threads);
  return push<E>(fwd, a, xs, ys);
}

// The rows a product reads: offsets monotone from 0, targets in the id space
// and in column order. Unlike Csr::from_parts a column may repeat: the
// per-relation matrix holds an entry per relation (compute/product/
// matrix.rs:276-279).
bool valid_csr(const or_csr* c) {
  if (!c || !c->offsets || c->n >= (1ull << 32) || c->offsets[0] != 0) return false;
  for (uint64_t r = 0; r < c->n; ++r) {
    if (c->offsets[r + 1] < c->offsets[r]) return false;
    for (uint64_t p = c->offsets[r]; p < c->offsets[r + 1]; ++p)
      if (c->targets[p] >= c->n || (p > c->offsets[r] && c->targets[p - 1] > c->targets[p])) return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// BFS (traversal/level.rs).
struct Bitmap {
  std::vector<uint64_t> w;
  explicit Bitmap(uint64_t n) : w((n + 63) / 64, 0) {}
  bool get(uint64_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
  void set(uint64_t i) { w[i >> 6] |= 1ull << (i & 63); }
};

uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

template <class E>
int apply_typed(int op, int monoid, const void* a, const void* b, void* out) {
  Err e;
  E r = monoid >= 0 ? Elem<E>::multiply(monoid, op, *(const E*)a, *(const E*)b, e)
                    : Elem<E>::apply(op, *(const E*)a, *(const E*)b, e);
  if (e.code) return e.code;
  *(E*)out = r;
  return OR_OK;
}

}  // namespace

extern "C" {

int or_csr_validate(const or_csr* c) {
  if (!c || !c->offsets) return OR_INVALID_CSR;
  if (c->n >= (1ull << 32)) return OR_INVALID_CSR;  // u32 targets (target.rs:25-31)
  if (c->offsets[0] != 0) return OR_INVALID_CSR;
  for (uint64_t r = 0; r < c->n; ++r)
    if (c->offsets[r + 1] < c->offsets[r]) return OR_INVALID_CSR;
  for (uint64_t r = 0; r < c->n; ++r) {
    for (uint64_t p = c->offsets[r]; p < c->offsets[r + 1]; ++p) {
      if (c->targets[p] >= c->n) return OR_INVALID_CSR;
      if (p > c->offsets[r] && c->targets[p - 1] >= c->targets[p]) return OR_INVALID_CSR;
    }
  }
  return OR_OK;
}

int or_csr_build(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                 const double* weights, uint64_t* offsets_out, uint32_t* targets_out,
                 double* weights_out, uint64_t* nnz_out) {
  for (uint64_t i = 0; i < m; ++i)
    if (src[i] >= n || dst[i] >= n) return OR_NODE_OUTSIDE_ID_SPACE;
  // Counting sort by source, keeping input order within a source.
  std::vector<uint64_t> start(n + 1, 0);
  for (uint64_t i = 0; i < m; ++i) start[src[i] + 1]++;
  for (uint64_t r = 0; r < n; ++r) start[r + 1] += start[r];
  std::vector<uint64_t> order(m), cursor(start.begin(), start.end() - 1);
  for (uint64_t i = 0; i < m; ++i) order[cursor[src[i]]++] = i;
  uint64_t out = 0;
  offsets_out[0] = 0;
  std::vector<uint64_t> row;
  for (uint64_t r = 0; r < n; ++r) {
    row.assign(order.begin() + start[r], order.begin() + start[r + 1]);
    // Stable sort by target: equal targets stay in input order, so their
    // weights sum in input order (deduplication.rs:1-7).
    std::stable_sort(row.begin(), row.end(), [&](uint64_t a, uint64_t b) { return dst[a] < dst[b]; });
    for (size_t k = 0; k < row.size(); ++k) {
      uint64_t e = row[k];
      if (k > 0 && dst[row[k - 1]] == dst[e]) {
        if (weights && weights_out) weights_out[out - 1] += weights[e];
        continue;
      }
      targets_out[out] = dst[e];
      if (weights_out) weights_out[out] = weights ? weights[e] : 1.0;
      ++out;
    }
    offsets_out[r + 1] = out;
  }
  *nnz_out = out;
  return OR_OK;
}

int or_csr_build_merged(uint64_t n, uint64_t m, const uint32_t* src, const uint32_t* dst,
                        const double* weights, const uint32_t* rel, uint32_t nrel, int keep_parallel,
                        uint64_t* offsets_out, uint32_t* targets_out, double* weights_out,
                        uint64_t* nnz_out, uint64_t* folded_out, uint64_t* bad_edge_out) {
  // counted(), build/mod.rs:329-352: per edge, the id space, then the
  // relation.
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] >= n || dst[i] >= n) {
      *bad_edge_out = i;
      return OR_EDGE_OUTSIDE_ID_SPACE;
    }
    if (rel && rel[i] >= nrel) {
      *bad_edge_out = i;
      return OR_RELATION_OUTSIDE_DICTIONARY;
    }
  }
  auto relation = [&](uint64_t e) -> uint32_t { return rel ? rel[e] : 0u; };
  // dense_rows (columns.rs:263-303): counting sort by source, input order
  // kept within a row.
  std::vector<uint64_t> start(n + 1, 0);
  for (uint64_t i = 0; i < m; ++i) start[src[i] + 1]++;
  for (uint64_t r = 0; r < n; ++r) start[r + 1] += start[r];
  std::vector<uint64_t> order(m), cursor(start.begin(), start.end() - 1);
  for (uint64_t i = 0; i < m; ++i) order[cursor[src[i]]++] = i;
  uint64_t out = 0, folded = 0;
  offsets_out[0] = 0;
  std::vector<uint64_t> row;
  for (uint64_t r = 0; r < n; ++r) {
    row.assign(order.begin() + start[r], order.begin() + start[r + 1]);
    std::stable_sort(row.begin(), row.end(), [&](uint64_t a, uint64_t b) {
      return dst[a] != dst[b] ? dst[a] < dst[b] : relation(a) < relation(b);
    });
    size_t k = 0;
    if (keep_parallel) {
      for (uint64_t e : row) {
        targets_out[out] = dst[e];
        if (weights_out) weights_out[out] = weights ? weights[e] : 1.0;
        ++out;
      }
      k = row.size();
    }
    while (k < row.size()) {
      const uint32_t t = dst[row[k]];
      double total = 0.0;
      bool first = true;
      while (k < row.size() && dst[row[k]] == t) {
        // relation_sum (deduplication.rs:196-206).
        const uint32_t rr = relation(row[k]);
        double sum = weights ? weights[row[k]] : 1.0;
        ++k;
        while (k < row.size() && dst[row[k]] == t && relation(row[k]) == rr) {
          sum += weights ? weights[row[k]] : 1.0;
          ++folded;
          ++k;
        }
        total = first ? sum : total + sum;
        first = false;
      }
      targets_out[out] = t;
      if (weights_out) weights_out[out] = total;
      ++out;
    }
    offsets_out[r + 1] = out;
  }
  *nnz_out = out;
  *folded_out = folded;
  return OR_OK;
}

int or_csr_transpose(const or_csr* f, uint64_t* off, uint32_t* tgt, double* wts) {
  uint64_t n = f->n, m = f->offsets[n];
  // Count in-degrees into off[t + 2], prefix, then scatter through off[t + 1]
  // scanning sources in order (transpose.rs:24-44).
  std::vector<uint64_t> o(n + 2, 0);
  for (uint64_t p = 0; p < m; ++p)
    if ((uint64_t)f->targets[p] + 2 < n + 2) o[f->targets[p] + 2]++;
  for (uint64_t r = 1; r < n + 2; ++r) o[r] += o[r - 1];
  for (uint64_t s = 0; s < n; ++s) {
    for (uint64_t p = f->offsets[s]; p < f->offsets[s + 1]; ++p) {
      uint64_t pos = o[f->targets[p] + 1]++;
      tgt[pos] = (uint32_t)s;
      if (wts) wts[pos] = f->weights ? f->weights[p] : 1.0;
    }
  }
  for (uint64_t r = 0; r <= n; ++r) off[r] = o[r];
  (void)m;
  return OR_OK;
}

int or_spmv(const or_csr* fwd, const or_csr* rev, int semiring, int dtype, int transpose,
            int weighted, const void* mask, int mask_dtype, int complement, int replace,
            int accumulate, uint64_t width, const void* x, void* y, int threads) {
  if (semiring < 0 || semiring > 12) return OR_INVALID_ARGUMENT;
  if (accumulate < -1 || accumulate > 9) return OR_INVALID_ARGUMENT;
  if (width == 0) return OR_WIDTH_MISMATCH;
  if (!valid_csr(fwd) || (rev && !valid_csr(rev))) return OR_INVALID_CSR;
  ProductArgs a{kAdd[semiring], kMul[semiring], weighted, mask, mask_dtype, complement, replace,
                accumulate, width};
  switch (dtype) {
    case OR_F32: return spmv_typed<float>(fwd, rev, transpose, a, x, y, threads);
    case OR_F64: return spmv_typed<double>(fwd, rev, transpose, a, x, y, threads);
    case OR_I64: return spmv_typed<int64_t>(fwd, rev, transpose, a, x, y, threads);
    case OR_BOOL: return spmv_typed<uint8_t>(fwd, rev, transpose, a, x, y, threads);
    case OR_U32:
      if (a.add == MAX) return OR_UNSUPPORTED;
      return spmv_typed<uint32_t>(fwd, rev, transpose, a, x, y, threads);
  }
  return OR_INVALID_ARGUMENT;
}

// PageRank (pagerank/mod.rs:96-287). Sums run in CHUNK = 65536 partials
// (reduce.rs:20) combined in order; each chunk folds sequentially from 0.
int or_pagerank(const or_csr* fwd, const or_csr* rev, const uint8_t* tomb, double damping,
                double tol, int has_tol, uint32_t max_iter, const double* seed, double* x,
                uint32_t* iterations, int threads) {
  const uint64_t CHUNK = 1 << 16;
  uint64_t n = fwd->n;
  if (rev && rev->n != n) return OR_ID_SPACE_MISMATCH;
  auto dead = [&](uint64_t i) { return tomb && tomb[i]; };
  uint64_t live = 0;
  for (uint64_t i = 0; i < n; ++i) live += !dead(i);
  double nodes = (double)live;
  Err e;
  double start = sqlf(DIV_, 1.0, nodes, e);
  if (e.code) return e.code;
  for (uint64_t i = 0; i < n; ++i) x[i] = dead(i) ? 0.0 : start;
  // out = PlusFirst(A) over the forward CSR: the out-weight of each row.
  std::vector<double> out(n, 0.0), msg(n, 0.0), y(n, 0.0), shares;
  ProductArgs pf{PLUS, FIRST, 1, nullptr, 0, 0, 0, -1, 1};
  int c = pull_width1<double>(fwd, pf, x, out.data(), threads);
  if (c) return c;
  std::vector<double> tr;  // transposed matrix if no reverse given
  std::vector<uint64_t> roff;
  std::vector<uint32_t> rtgt;
  or_csr rcsr;
  if (!rev) {
    uint64_t m = fwd->offsets[n];
    roff.resize(n + 1);
    rtgt.resize(m);
    tr.resize(m);
    or_csr_transpose(fwd, roff.data(), rtgt.data(), tr.data());
    rcsr = {n, roff.data(), rtgt.data(), fwd->weights ? tr.data() : nullptr};
    rev = &rcsr;
  }
  bool seeded = seed != nullptr;
  if (seeded) {
    // reduce_live(seed, Sum) then 1 / sum, then seed * total (mod.rs:155-199).
    double sum = 0.0;
    for (uint64_t c0 = 0; c0 < n; c0 += CHUNK) {
      double part = 0.0;
      for (uint64_t i = c0; i < std::min(n, c0 + CHUNK); ++i)
        if (!dead(i)) part += seed[i];
      sum += part;
    }
    double total = sqlf(DIV_, 1.0, sum, e);
    if (e.code) return e.code;
    shares.assign(n, 0.0);
    for (uint64_t i = 0; i < n; ++i)
      if (!dead(i)) {
        shares[i] = sqlf(MUL_, seed[i], total, e);
        if (e.code) return e.code;
      }
  }
  double teleport = sqlf(SUB_, 1.0, damping, e);
  if (e.code) return e.code;
  ProductArgs pt{PLUS, TIMES, 1, nullptr, 0, 0, 0, -1, 1};
  uint32_t it = 0;
  uint64_t chunks = (n + CHUNK - 1) / CHUNK;
  std::vector<double> partial(chunks);
  std::vector<int> status(chunks);
  while (it < max_iter) {
    ++it;
    // spread (mod.rs:204-232)
    c = parallel_tasks(chunks, threads, [&](uint64_t ch) -> int {
      Err le;
      double d = 0.0;
      for (uint64_t i = ch * CHUNK; i < std::min(n, (ch + 1) * CHUNK); ++i) {
        if (sql_cmp_f(out[i], 0.0) > 0) {
          msg[i] = sqlf(DIV_, x[i], out[i], le);
          if (le.code) return le.code;
        } else {
          d += x[i];
          msg[i] = 0.0;
        }
      }
      partial[ch] = d;
      return OR_OK;
    });
    if (c) return c;
    double dangling = 0.0;
    for (uint64_t ch = 0; ch < chunks; ++ch) dangling += partial[ch];
    c = pull_width1<double>(rev, pt, msg.data(), y.data(), threads);
    if (c) return c;
    // advance (mod.rs:245-282)
    c = parallel_tasks(chunks, threads, [&](uint64_t ch) -> int {
      Err le;
      double change = 0.0;
      for (uint64_t i = ch * CHUNK; i < std::min(n, (ch + 1) * CHUNK); ++i) {
        double next = 0.0;
        if (!dead(i)) {
          double p = seeded ? shares[i] : start;
          double flowed = sqlf(ADD_, y[i], sqlf(MUL_, dangling, p, le), le);
          if (le.code) return le.code;
          double a1 = sqlf(MUL_, damping, flowed, le);
          if (le.code) return le.code;
          double a2 = sqlf(MUL_, teleport, p, le);
          if (le.code) return le.code;
          next = sqlf(ADD_, a1, a2, le);
          if (le.code) return le.code;
        }
        change += std::fabs(next - x[i]);
        x[i] = next;
      }
      partial[ch] = change;
      return OR_OK;
    });
    if (c) return c;
    double delta = 0.0;
    for (uint64_t ch = 0; ch < chunks; ++ch) delta += partial[ch];
    if (has_tol) {
      double bound = sqlf(MUL_, nodes, tol, e);
      if (e.code) return e.code;
      if (sql_cmp_f(delta, bound) < 0) break;
    }
  }
  *iterations = it;
  return OR_OK;
}

// LevelSearch (level.rs:88-337) with the Results admission of bfs.rs:84-111,
// over an Adjacency's rows (traversal/adjacency.rs:20-221): `rows` are the
// CSRs read in order (a node's row entries are its rows concatenated:
// Adjacency::row_entries), `rrows` those of the reversed adjacency, which
// a pulled level scans (NULL: every level pushed). Degrees are summed over
// the CSRs read (row_entry_bound).
int or_bfs_rows(uint64_t n, const or_csr* const* rows, uint32_t k, const or_csr* const* rrows, uint32_t kr,
                const uint64_t* starts, uint64_t nstarts, const uint8_t* filter, int64_t max_depth,
                int64_t max_results, or_visit* visits, uint64_t* count_out, or_stats* stats) {
  const uint64_t ALPHA = 1, BETA = 24;
  or_stats st{0, 0, 0, 0};
  uint64_t count = 0;
  for (uint64_t i = 0; i < nstarts; ++i)
    if (starts[i] >= n) return OR_NODE_OUTSIDE_ID_SPACE;
  Bitmap visited(n);
  std::vector<uint64_t> frontier, next;
  bool stop = false;
  // offer_into: admit (or refuse on max_results), then mark and list.
  auto offer = [&](uint64_t node, uint32_t depth, uint64_t via) -> bool {
    if (max_results >= 0 && (uint64_t)max_results == count) {
      st.truncated = 1;
      return false;
    }
    visits[count++] = or_visit{node, via, depth, 0};
    st.nodes_visited++;
    visited.set(node);
    next.push_back(node);
    return true;
  };
  for (uint64_t i = 0; i < nstarts && !stop; ++i)
    if (!visited.get(starts[i]))
      if (!offer(starts[i], 0, UINT64_MAX)) stop = true;
  std::swap(frontier, next);
  next.clear();
  uint32_t depth = 0;
  while (!stop && !frontier.empty() && !(max_depth >= 0 && (int64_t)depth >= max_depth)) {
    std::sort(frontier.begin(), frontier.end());
    bool pulls = false;
    if (rrows && frontier.size() * BETA > n) {
      uint64_t fe = 0;
      for (uint64_t f : frontier)
        for (uint32_t j = 0; j < k; ++j) fe += rows[j]->offsets[f + 1] - rows[j]->offsets[f];
      uint64_t limit = fe * ALPHA, ue = 0;
      pulls = true;
      for (uint64_t v = 0; v < n && pulls; ++v)
        if (!visited.get(v)) {
          for (uint32_t j = 0; j < kr; ++j) ue += rrows[j]->offsets[v + 1] - rrows[j]->offsets[v];
          if (ue >= limit) pulls = false;
        }
    }
    uint32_t d = depth + 1;
    if (!pulls) {
      for (uint64_t via : frontier) {
        for (uint32_t j = 0; j < k && !stop; ++j)
          for (uint64_t p = rows[j]->offsets[via]; p < rows[j]->offsets[via + 1] && !stop; ++p) {
            uint64_t node = rows[j]->targets[p];
            st.edges_traversed++;
            if (visited.get(node) || (filter && !filter[node])) continue;
            if (!offer(node, d, via)) stop = true;
          }
        if (stop) break;
      }
    } else {
      st.levels_pulled++;
      Bitmap fr(n);
      for (uint64_t f : frontier) fr.set(f);
      Bitmap seen = visited;  // the densified visited set at level start
      for (uint64_t node = 0; node < n && !stop; ++node) {
        if (seen.get(node)) continue;
        if (filter && !filter[node]) continue;
        bool hit = false;
        for (uint32_t j = 0; j < kr && !hit; ++j)
          for (uint64_t p = rrows[j]->offsets[node]; p < rrows[j]->offsets[node + 1]; ++p) {
            uint64_t via = rrows[j]->targets[p];
            st.edges_traversed++;
            if (!fr.get(via)) continue;
            if (!offer(node, d, via)) stop = true;
            hit = true;
            break;
          }
      }
    }
    std::swap(frontier, next);
    next.clear();
    depth = d;
  }
  *count_out = count;
  if (stats) *stats = st;
  return OR_OK;
}

int or_bfs(const or_csr* adj, const or_csr* rev, const uint64_t* starts, uint64_t nstarts,
           const uint8_t* filter, int64_t max_depth, int64_t max_results, or_visit* visits,
           uint64_t* count_out, or_stats* stats) {
  return or_bfs_rows(adj->n, &adj, 1, rev ? &rev : nullptr, rev ? 1 : 0, starts, nstarts, filter, max_depth,
                     max_results, visits, count_out, stats);
}

int or_bfs_depths(const or_csr* adj, const uint64_t* sources, uint64_t ns, uint32_t* depths,
                  int threads) {
  uint64_t n = adj->n;
  for (uint64_t s = 0; s < ns; ++s)
    if (sources[s] >= n) return OR_NODE_OUTSIDE_ID_SPACE;
  return parallel_tasks(ns, threads, [&](uint64_t s) -> int {
    uint32_t* d = depths + s * n;
    for (uint64_t i = 0; i < n; ++i) d[i] = UINT32_MAX;
    std::vector<uint32_t> q;
    q.reserve(1024);
    d[sources[s]] = 0;
    q.push_back((uint32_t)sources[s]);
    for (size_t h = 0; h < q.size(); ++h) {
      uint32_t u = q[h];
      for (uint64_t p = adj->offsets[u]; p < adj->offsets[u + 1]; ++p) {
        uint32_t v = adj->targets[p];
        if (d[v] == UINT32_MAX) {
          d[v] = d[u] + 1;
          q.push_back(v);
        }
      }
    }
    return OR_OK;
  });
}

int or_apply(int op, int monoid, int dtype, const void* a, const void* b, void* out) {
  if (op < 0 || op > 9) return OR_INVALID_ARGUMENT;
  switch (dtype) {
    case OR_F32: return apply_typed<float>(op, monoid, a, b, out);
    case OR_F64: return apply_typed<double>(op, monoid, a, b, out);
    case OR_I64: return apply_typed<int64_t>(op, monoid, a, b, out);
    case OR_BOOL: return apply_typed<uint8_t>(op, monoid, a, b, out);
    case OR_U32: return apply_typed<uint32_t>(op, monoid, a, b, out);
  }
  return OR_INVALID_ARGUMENT;
}

int or_from_weight(int dtype, double w, void* out) {
  Err e;
  switch (dtype) {
    case OR_F32: { float r = Elem<float>::from_weight(w, e); if (!e.code) *(float*)out = r; break; }
    case OR_F64: *(double*)out = w; break;
    case OR_I64: { int64_t r = Elem<int64_t>::from_weight(w, e); if (!e.code) *(int64_t*)out = r; break; }
    case OR_BOOL: *(uint8_t*)out = Elem<uint8_t>::from_weight(w, e); break;
    case OR_U32: { uint32_t r = Elem<uint32_t>::from_weight(w, e); if (!e.code) *(uint32_t*)out = r; break; }
    default: return OR_INVALID_ARGUMENT;
  }
  return e.code;
}

double or_uniform(uint64_t seed, uint64_t stream, uint64_t node) {
  const uint64_t GOLDEN = 0x9e3779b97f4a7c15ull;
  uint64_t bits = splitmix_mix(splitmix_mix(seed ^ splitmix_mix(stream * GOLDEN)) ^ (node * GOLDEN));
  return (double)(bits >> 11) / (double)(1ull << 52) - 1.0;
}

}  // extern "C"

// Adjoint of PageRank with respect to the edge weights (no reference
// counterpart; the forward is compute/pagerank/mod.rs:96-287). With s = the
// out-weight of each row, p = the personalisation and lt = lambda at live
// rows (0 at tombstoned ones), iterate from lambda = g:
//   lambda_i = d * (s_i > 0 ? (A lt)_i / s_i : p . lt) + g_i
// until the L1 change < live * tol (or max_iter), then for edge a -> b with
// s_a > 0: dL/dw_ab = (d x_a / s_a) (lt_b - (A lt)_a / s_a), else 0. All
// sums in ascending order, f64.
int or_pagerank_backward(const or_csr* fwd, const uint8_t* tomb, double damping, double tol,
                         int has_tol, uint32_t max_iter, const double* seed, const double* x,
                         const double* g, double* lambda, double* dldw, uint32_t* iterations,
                         int threads) {
  const uint64_t n = fwd->n;
  if (!fwd->weights) return OR_INVALID_ARGUMENT;
  auto dead = [&](uint64_t i) { return tomb && tomb[i]; };
  uint64_t live = 0;
  for (uint64_t i = 0; i < n; ++i) live += !dead(i);
  const double nodes = (double)live;
  Err e;
  const double start = sqlf(DIV_, 1.0, nodes, e);
  if (e.code) return e.code;
  std::vector<double> ones(n, 1.0), out(n, 0.0), p(n, 0.0), lt(n, 0.0), u(n, 0.0);
  ProductArgs pf{PLUS, FIRST, 1, nullptr, 0, 0, 0, -1, 1};
  int c = pull_width1<double>(fwd, pf, ones.data(), out.data(), threads);
  if (c) return c;
  if (seed) {
    double sum = 0.0;
    for (uint64_t i = 0; i < n; ++i)
      if (!dead(i)) sum += seed[i];
    const double total = sqlf(DIV_, 1.0, sum, e);
    if (e.code) return e.code;
    for (uint64_t i = 0; i < n; ++i)
      if (!dead(i)) p[i] = seed[i] * total;
  } else {
    for (uint64_t i = 0; i < n; ++i) p[i] = dead(i) ? 0.0 : start;
  }
  const double bound = has_tol ? nodes * tol : 0.0;
  double pi = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    lambda[i] = g[i];
    lt[i] = dead(i) ? 0.0 : g[i];
    pi += p[i] * lt[i];
  }
  ProductArgs pt{PLUS, TIMES, 1, nullptr, 0, 0, 0, -1, 1};
  uint32_t it = 0;
  while (it < max_iter) {
    ++it;
    c = pull_width1<double>(fwd, pt, lt.data(), u.data(), threads);
    if (c) return c;
    double change = 0.0, pin = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      const double z = sql_cmp_f(out[i], 0.0) > 0 ? u[i] / out[i] : pi;
      const double l = damping * z + g[i];
      change += std::fabs(l - lambda[i]);
      lambda[i] = l;
      lt[i] = dead(i) ? 0.0 : l;
      pin += p[i] * lt[i];
    }
    pi = pin;
    if (has_tol && sql_cmp_f(change, bound) < 0) break;
  }
  c = pull_width1<double>(fwd, pt, lt.data(), u.data(), threads);
  if (c) return c;
  for (uint64_t a = 0; a < n; ++a) {
    const bool flows = sql_cmp_f(out[a], 0.0) > 0;
    const double coef = flows ? damping * x[a] / out[a] : 0.0;
    const double pa = flows ? u[a] / out[a] : 0.0;
    for (uint64_t k = fwd->offsets[a]; k < fwd->offsets[a + 1]; ++k)
      dldw[k] = flows ? coef * (lt[fwd->targets[k]] - pa) : 0.0;
  }
  if (iterations) *iterations = it;
  return OR_OK;
}
