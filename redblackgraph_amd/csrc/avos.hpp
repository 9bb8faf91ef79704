// AVOS semiring primitives shared by host and device code.
//
// Values live in a "widened" signed domain: int32_t for element widths 8/16/32
// and int64_t for width 64.  RED_ONE is -1 in every width; for unsigned element
// types the facade maps the all-ones sentinel (2^W - 1) to -1 before the data
// reaches the device, so one code path serves all ten integer element types.
//
// Semantics follow the C core of the reference:
//   avos_sum      redblackgraph/core/src/redblackgraph/rbg_math.c.in:30-36
//   MSB           redblackgraph/core/src/redblackgraph/rbg_math.c.in:38-46
//   avos_product  redblackgraph/core/src/redblackgraph/rbg_math.c.in:48-121
// The reference reports overflow through PyErr_Format and keeps computing; here
// the same conditions are OR-ed into an event word the caller turns into an
// exception.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rbg {

// Event bits reported by the arithmetic.
enum : uint32_t {
  EV_OVERFLOW_WIDTH = 1u,  // MSB(x) + MSB(y) >= W            (rbg_math.c.in:104-111)
  EV_OVERFLOW_RED = 2u,    // result collides with RED_ONE     (rbg_math.c.in:113-119)
  EV_INVALID = 4u,         // operand < -1 in the general case (the reference loops forever)
  EV_SIGN = 8u,            // signed result does not fit W-1 bits (the reference wraps negative)
};

template <typename T> struct unsigned_of;
template <> struct unsigned_of<int32_t> { using type = uint32_t; };
template <> struct unsigned_of<int64_t> { using type = uint64_t; };

__host__ __device__ inline int msb_u(uint32_t x) { return x ? 31 - __builtin_clz(x) : 0; }
__host__ __device__ inline int msb_u(uint64_t x) { return x ? 63 - __builtin_clzll(x) : 0; }

// floor(log2(x)) for x >= 1, 0 for x <= 1 (and for RED_ONE).  Matches the C
// loop for every value the reference terminates on.
template <typename T>
__host__ __device__ inline int msb(T x) {
  using U = typename unsigned_of<T>::type;
  return x > 1 ? msb_u((U)x) : 0;
}

// Non-zero minimum with RED_ONE absorbing (unsigned compare in the widened
// domain; identical to the signed rule for every value >= -1).
template <typename T>
__host__ __device__ inline T avos_sum(T a, T b) {
  using U = typename unsigned_of<T>::type;
  if (a == 0 || b == (T)-1) return b;
  if (b == 0 || a == (T)-1) return a;
  return ((U)a < (U)b) ? a : b;
}

// Narrow a W-bit result held in U back into the widened signed domain.
template <typename T>
__host__ __device__ inline T narrow(typename unsigned_of<T>::type r, int W, bool is_unsigned,
                                    uint32_t& ev) {
  using U = typename unsigned_of<T>::type;
  constexpr int BITS = (int)sizeof(T) * 8;
  if (W < BITS) {
    const U mask = (U(1) << W) - 1;
    r &= mask;
    if (r == mask) return (T)-1;  // the all-ones pattern is RED_ONE in every width
    if (!is_unsigned && (r >> (W - 1))) {
      ev |= EV_SIGN;
      return (T)(r | ~mask);  // sign-extend, as the narrow signed store would
    }
    return (T)r;
  }
  if (!is_unsigned && (T)r < -1) ev |= EV_SIGN;
  return (T)r;
}

// AVOS product.  W is the element bit width (8/16/32/64), is_unsigned selects the
// unsigned element types (their valid range reaches 2^W - 2).
template <typename T>
__host__ __device__ inline T avos_product(T lhs, T rhs, int W, bool is_unsigned, uint32_t& ev) {
  using U = typename unsigned_of<T>::type;
  if (lhs == 0 || rhs == 0) return 0;
  if (lhs == (T)-1 && rhs == (T)-1) return (T)-1;
  if (lhs == 1 && rhs == 1) return 1;
  if (lhs == (T)-1 && rhs == 1) return 0;
  if (lhs == 1 && rhs == (T)-1) return 0;
  const U x = (lhs == (T)-1) ? (U)1 : (U)lhs;
  if (rhs == (T)-1) return (lhs & 1) ? (T)0 : (T)x;
  if (rhs == 1) return (lhs & 1) ? (T)x : (T)0;
  if (lhs < (T)-1 || rhs < (T)-1) ev |= EV_INVALID;
  const U y = (U)rhs;
  const int m = msb_u(y);
  const int mx = msb_u(x);
  if (mx + m >= W) ev |= EV_OVERFLOW_WIDTH;
  U r = (y & ((U(1) << m) - 1)) | (x << m);
  {
    constexpr int BITS = (int)sizeof(T) * 8;
    const U all = (W < BITS) ? ((U(1) << W) - 1) : ~U(0);
    if ((r & all) == all) ev |= EV_OVERFLOW_RED;
  }
  return narrow<T>(r, W, is_unsigned, ev);
}

// Native-width shorthand (W = bits of T, signed element type).
template <typename T>
__host__ __device__ inline T avos_product(T lhs, T rhs, uint32_t& ev) {
  return avos_product<T>(lhs, rhs, (int)sizeof(T) * 8, false, ev);
}

}  // namespace rbg
