// Device-side semiring algebra over the five element types, as
// warren-core defines it (compute/semiring.rs:236-388, compute/cast.rs:87-108,
// compute/dtype.rs:142-150):
//   f32/f64  plain IEEE arithmetic; min/max in SQL order (NaN above every
//            number, the first operand kept on a tie); no error is raised.
//   i64      checked arithmetic (overflow -> BigintOutOfRange); the min/max
//            identities stand for +-infinity and absorb plus, and times by a
//            weight above zero.
//   bool     or/and algebra on 0/1 bytes.
//   u32      i64 restricted to [0, 2^32-2] with 2^32-1 standing for
//            i64::MAX: each operation runs the i64 rule, and a result outside
//            that image is BigintOutOfRange.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace wg {

enum Op { FIRST = 0, SECOND, PAIR, PLUS, TIMES, MIN, MAX, LOR, LAND, ANY };

// Status codes raised from kernels (the WG_* numbering).
constexpr uint32_t kErrOverflowF = 8, kErrUnderflowF = 9, kErrBigint = 10;

constexpr int64_t kI64Max = INT64_MAX;
constexpr int64_t kI64Min = INT64_MIN;
constexpr uint32_t kU32Inf = 0xffffffffu;

template <class E>
constexpr bool is_float_v = std::is_same<E, float>::value || std::is_same<E, double>::value;

// SQL order for floats: b < a and b > a with NaN highest, NaN == NaN.
template <class F>
__device__ __forceinline__ bool sql_less(F b, F a) {
  if (b != b) return false;
  if (a != a) return true;
  return b < a;
}
template <class F>
__device__ __forceinline__ bool sql_greater(F b, F a) {
  if (a != a) return false;
  if (b != b) return true;
  return b > a;
}

__device__ __forceinline__ int64_t u2i(uint32_t v) { return v == kU32Inf ? kI64Max : (int64_t)v; }
__device__ __forceinline__ uint32_t i2u(int64_t v, uint32_t& err) {
  if (v == kI64Max) return kU32Inf;
  if (v < 0 || v > (int64_t)0xfffffffeLL) {
    err = kErrBigint;
    return 0;
  }
  return (uint32_t)v;
}

template <int OP>
__device__ __forceinline__ int64_t apply_i64(int64_t a, int64_t b, uint32_t& err) {
  if constexpr (OP == FIRST || OP == ANY) return a;
  else if constexpr (OP == SECOND) return b;
  else if constexpr (OP == PAIR) return 1;
  else if constexpr (OP == PLUS) {
    int64_t r;
    if (__builtin_add_overflow(a, b, &r)) { err = kErrBigint; return 0; }
    return r;
  } else if constexpr (OP == TIMES) {
    int64_t r;
    if (__builtin_mul_overflow(a, b, &r)) { err = kErrBigint; return 0; }
    return r;
  } else if constexpr (OP == MIN) return a < b ? a : b;
  else if constexpr (OP == MAX) return a > b ? a : b;
  else if constexpr (OP == LOR) return (a != 0 || b != 0) ? 1 : 0;
  else return (a != 0 && b != 0) ? 1 : 0;  // LAND
}

// BinaryOp applied as SemiringElement::apply.
template <int OP, class E>
__device__ __forceinline__ E apply(E a, E b, uint32_t& err) {
  if constexpr (is_float_v<E>) {
    if constexpr (OP == FIRST || OP == ANY) return a;
    else if constexpr (OP == SECOND) return b;
    else if constexpr (OP == PAIR) return (E)1;
    else if constexpr (OP == PLUS) return a + b;
    else if constexpr (OP == TIMES) return a * b;
    else if constexpr (OP == MIN) return sql_less(b, a) ? b : a;
    else if constexpr (OP == MAX) return sql_greater(b, a) ? b : a;
    else if constexpr (OP == LOR) return (a != 0 || b != 0) ? (E)1 : (E)0;
    else return (a != 0 && b != 0) ? (E)1 : (E)0;
  } else if constexpr (std::is_same<E, int64_t>::value) {
    return apply_i64<OP>(a, b, err);
  } else if constexpr (std::is_same<E, uint8_t>::value) {
    if constexpr (OP == FIRST || OP == ANY) return a;
    else if constexpr (OP == SECOND) return b;
    else if constexpr (OP == PAIR) return 1;
    else if constexpr (OP == PLUS || OP == MAX || OP == LOR) return (a || b) ? 1 : 0;
    else return (a && b) ? 1 : 0;
  } else {  // uint32_t
    return i2u(apply_i64<OP>(u2i(a), u2i(b), err), err);
  }
}

// apply with the operator known only at run time (accumulate).
template <class E>
__device__ __forceinline__ E apply_rt(int op, E a, E b, uint32_t& err) {
  switch (op) {
    case FIRST: return apply<FIRST>(a, b, err);
    case SECOND: return apply<SECOND>(a, b, err);
    case PAIR: return apply<PAIR>(a, b, err);
    case PLUS: return apply<PLUS>(a, b, err);
    case TIMES: return apply<TIMES>(a, b, err);
    case MIN: return apply<MIN>(a, b, err);
    case MAX: return apply<MAX>(a, b, err);
    case LOR: return apply<LOR>(a, b, err);
    case LAND: return apply<LAND>(a, b, err);
    default: return apply<ANY>(a, b, err);
  }
}

// The identity of the monoid ADD: the value of "no entry".
template <class E, int ADD>
__host__ __device__ __forceinline__ E identity() {
  if constexpr (is_float_v<E>) {
    if constexpr (ADD == MIN) return (E)__builtin_huge_val();
    else if constexpr (ADD == MAX) return (E)-__builtin_huge_val();
    else return (E)0;
  } else if constexpr (std::is_same<E, int64_t>::value) {
    if constexpr (ADD == MIN) return kI64Max;
    else if constexpr (ADD == MAX) return kI64Min;
    else return 0;
  } else if constexpr (std::is_same<E, uint8_t>::value) {
    return ADD == MIN ? 1 : 0;
  } else {
    return ADD == MIN ? kU32Inf : 0;  // max over u32 is refused on the host
  }
}

// (x) (weight, value) in a semiring whose (+) is ADD: apply, except that the
// i64 (and u32) min/max identities absorb plus, and times by w > 0.
template <class E, int ADD, int MUL>
__device__ __forceinline__ E multiply(E w, E v, uint32_t& err) {
  if constexpr (std::is_same<E, int64_t>::value || std::is_same<E, uint32_t>::value) {
    int64_t vi, wi;
    if constexpr (std::is_same<E, uint32_t>::value) { vi = u2i(v); wi = u2i(w); }
    else { vi = v; wi = w; }
    bool infinite = (ADD == MIN && vi == kI64Max) || (ADD == MAX && vi == kI64Min);
    bool absorbs = MUL == PLUS || (MUL == TIMES && wi > 0);
    if (infinite && absorbs) return v;
    return apply<MUL>(w, v, err);
  } else {
    return apply<MUL>(w, v, err);
  }
}

// ---- folds: how a row's products combine, in ascending column order ------
// Fold<E, ADD>::A is the partial result of a contiguous run of a row's
// products; combine(a, b) joins the run a with the run that follows it.
// Every fold below gives exactly the result (and error) of the reference's
// left-to-right fold (compute/product/kernel.rs:100-109), except float plus,
// whose rounding depends on the grouping.
template <class E, int ADD>
struct Fold {
  using A = E;
  __device__ static A lift(E p) { return p; }
  __device__ static A combine(A a, A b, uint32_t& err) { return apply<ADD>(a, b, err); }
  __device__ static E finish(A a, uint32_t&) { return a; }
};

// i64 plus: the reference errors if any prefix sum overflows; a run is kept
// as (total, largest prefix, smallest prefix) in 128 bits, which composes
// exactly.
template <>
struct Fold<int64_t, PLUS> {
  struct A {
    __int128 t, hi, lo;
  };
  __device__ static A lift(int64_t p) { return A{p, p, p}; }
  __device__ static A combine(A a, A b, uint32_t&) {
    A r;
    r.t = a.t + b.t;
    __int128 h = a.t + b.hi, l = a.t + b.lo;
    r.hi = a.hi > h ? a.hi : h;
    r.lo = a.lo < l ? a.lo : l;
    return r;
  }
  __device__ static int64_t finish(A a, uint32_t& err) {
    if (a.hi > (__int128)kI64Max || a.lo < (__int128)kI64Min) { err = kErrBigint; return 0; }
    return (int64_t)a.t;
  }
};

// u32 min: the u32 image keeps the i64 order (u32::MAX, standing for
// i64::MAX, is the largest), so the smaller u32 is the i64 min's image and
// no range check can fail.
template <>
struct Fold<uint32_t, MIN> {
  using A = uint32_t;
  __device__ static A lift(uint32_t p) { return p; }
  __device__ static A combine(A a, A b, uint32_t&) { return b < a ? b : a; }
  __device__ static uint32_t finish(A a, uint32_t&) { return a; }
};

// u32 plus: terms are non-negative, so prefix sums only grow: the run keeps
// its finite sum and how many infinities it holds.
template <>
struct Fold<uint32_t, PLUS> {
  struct A {
    uint64_t f;
    uint32_t inf;
    uint32_t pad;
  };
  __device__ static A lift(uint32_t p) { return p == kU32Inf ? A{0, 1, 0} : A{p, 0, 0}; }
  __device__ static A combine(A a, A b, uint32_t&) { return A{a.f + b.f, a.inf + b.inf, 0}; }
  __device__ static uint32_t finish(A a, uint32_t& err) {
    if (a.inf == 0) {
      if (a.f > 0xfffffffeULL) { err = kErrBigint; return 0; }
      return (uint32_t)a.f;
    }
    if (a.inf == 1 && a.f == 0) return kU32Inf;
    err = kErrBigint;
    return 0;
  }
};

// ---- shuffles of any trivially copyable value, 32 bits at a time ---------
template <class T>
__device__ __forceinline__ T shfl_down(T v, unsigned delta, int width) {
  constexpr int W = (sizeof(T) + 3) / 4;
  union U {
    T t;
    int w[W];
    __device__ U() {}
  } in, out;
  in.t = v;
#pragma unroll
  for (int i = 0; i < W; ++i) out.w[i] = __shfl_down(in.w[i], delta, width);
  return out.t;
}

// ---- the weight of an edge as an element (compute/cast.rs:87-108) --------
// Returns 0 and sets err on a failing cast.
template <class E>
__device__ __forceinline__ E weight_cast(double w, uint32_t& err) {
  if constexpr (std::is_same<E, double>::value) {
    return w;
  } else if constexpr (std::is_same<E, float>::value) {
    float r = (float)w;
    if (__builtin_isinf(r) && !__builtin_isinf(w)) { err = kErrOverflowF; return 0; }
    if (r == 0.0f && w != 0.0) { err = kErrUnderflowF; return 0; }
    return r;
  } else if constexpr (std::is_same<E, uint8_t>::value) {
    return w != 0.0 ? 1 : 0;  // NaN reads as true
  } else {
    double r = __builtin_rint(w);  // round half to even
    if (!(r >= -9223372036854775808.0 && r < 9223372036854775808.0)) { err = kErrBigint; return 0; }
    int64_t i = (int64_t)r;
    if constexpr (std::is_same<E, int64_t>::value) return i;
    else return i2u(i, err);
  }
}

// Whether a value mask selects a row (compute/mask.rs:35-45): the mask's
// element is non-zero (NaN included), flipped by complement.
__device__ __forceinline__ bool mask_selects(const void* mask, int dtype, int complement,
                                             uint64_t row) {
  bool set;
  switch (dtype) {
    case 0: set = ((const float*)mask)[row] != 0.0f; break;
    case 1: set = ((const double*)mask)[row] != 0.0; break;
    case 2: set = ((const int64_t*)mask)[row] != 0; break;
    case 4: set = ((const uint32_t*)mask)[row] != 0; break;
    default: set = ((const uint8_t*)mask)[row] != 0; break;
  }
  return set != (complement != 0);
}

// How a row's result is written (Write::row, compute/product/kernel.rs:219-239).
struct WriteArgs {
  const void* mask;
  int mask_dtype;
  int complement;
  int replace;
  int accumulate;  // BinaryOp or -1
};

template <class E, int ADD>
__device__ __forceinline__ void write_elem(const WriteArgs& wa, uint64_t row, E* out,
                                           bool present, E t, uint32_t& err) {
  if (wa.mask && !mask_selects(wa.mask, wa.mask_dtype, wa.complement, row)) {
    if (wa.replace) *out = identity<E, ADD>();
    return;
  }
  if (wa.accumulate < 0) {
    *out = present ? t : identity<E, ADD>();
  } else if (present) {
    *out = apply_rt<E>(wa.accumulate, *out, t, err);
  }
}

// Raises (index, status) on the error word; the smallest index wins, so the
// reported error does not depend on scheduling.
__device__ __forceinline__ void raise_error(unsigned long long* word, uint64_t index,
                                            uint32_t status) {
  atomicMin(word, (unsigned long long)((index << 8) | status));
}

}  // namespace wg

namespace wg {

constexpr uint32_t kErrDivZero = 7;

// SQL float8 operators with their checks (compute/map/arithmetic.rs:49-79):
// division by zero (unless the dividend is NaN), overflow to infinity from
// finite operands, and underflow to zero from non-zero operands.
enum Arith { ADD_ = 0, SUB_, MUL_, DIV_ };

template <int OP>
__host__ __device__ __forceinline__ double sqlf(double a, double b, uint32_t& err) {
  if (OP == DIV_ && b == 0.0 && !(a != a)) {
    err = kErrDivZero;
    return 0.0;
  }
  double r = OP == ADD_ ? a + b : OP == SUB_ ? a - b : OP == MUL_ ? a * b : a / b;
  bool ainf = __builtin_isinf(a), binf = __builtin_isinf(b);
  bool finite_operands = OP == DIV_ ? !ainf : (!ainf && !binf);
  if (__builtin_isinf(r) && finite_operands) {
    err = kErrOverflowF;
    return 0.0;
  }
  bool under = OP == MUL_ ? (a != 0.0 && b != 0.0) : OP == DIV_ ? (a != 0.0 && !binf) : false;
  if (r == 0.0 && under) {
    err = kErrUnderflowF;
    return 0.0;
  }
  return r;
}

// a < b in SQL float order (NaN above every number).
__host__ __device__ __forceinline__ bool sql_lt(double a, double b) {
  if (a != a) return false;
  if (b != b) return true;
  return a < b;
}

}  // namespace wg
