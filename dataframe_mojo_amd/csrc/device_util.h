// device_util.h — integer, fixed-point and float-ordering helpers shared by
// the kernels (and, where marked __host__, by the host finish code).
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#include <stdint.h>
#else
// hiprtc declares the fixed-width integers in its own namespace
using __hip_internal::int8_t;
using __hip_internal::int16_t;
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace dfr {

#define DFR_HD __host__ __device__ __forceinline__

DFR_HD uint64_t f64_bits(double x) {
  uint64_t u;
  __builtin_memcpy(&u, &x, 8);
  return u;
}
DFR_HD double bits_f64(uint64_t u) {
  double x;
  __builtin_memcpy(&x, &u, 8);
  return x;
}

// Ordered key of a non-NaN double: unsigned order of keys is numeric order,
// and -0.0 shares +0.0's key (the two compare equal).
DFR_HD uint64_t ord_key(uint64_t bits) {
  if ((bits << 1) == 0) bits = 0;  // -0.0 -> +0.0
  return (bits >> 63) ? ~bits : (bits | (1ull << 63));
}
DFR_HD uint64_t ord_key_inverse(uint64_t key) {
  return (key >> 63) ? (key & ~(1ull << 63)) : ~key;
}

// The canonical word of a Float64 group key: -0.0 and 0.0 are one key, as
// are all NaNs.
DFR_HD uint64_t canon_key_bits(uint64_t bits) {
  if ((bits << 1) == 0) return 0;
  if ((bits & 0x7FF0000000000000ull) == 0x7FF0000000000000ull && (bits << 12) != 0) return 0x7FF8000000000000ull;
  return bits;
}

// 128-bit shifts, shift amounts in [0, 127]; larger right shifts give 0.
DFR_HD void shl128(uint64_t &lo, uint64_t &hi, int s) {
  if (s <= 0) return;
  if (s >= 64) {
    hi = lo << (s - 64);
    lo = 0;
  } else {
    hi = (hi << s) | (lo >> (64 - s));
    lo <<= s;
  }
}
DFR_HD void shr128(uint64_t &lo, uint64_t &hi, int s) {
  if (s <= 0) return;
  if (s >= 128) {
    lo = hi = 0;
  } else if (s >= 64) {
    lo = hi >> (s - 64);
    hi = 0;
  } else {
    lo = (lo >> s) | (hi << (64 - s));
    hi >>= s;
  }
}

// Shifts of the 64-bit m into a 128-bit result without branches: left by
// sh in [0, 127] (bits above 127 are dropped) or right by -sh.
DFR_HD void shift_u64_to_128(uint64_t m, int sh, uint64_t &lo, uint64_t &hi) {
  const int l = sh < 0 ? 0 : sh;      // left amount
  const int r = sh < 0 ? -sh : 0;     // right amount
  const int l6 = l & 63, r6 = r & 63;
  const uint64_t ml = m << l6;
  const uint64_t spill = l6 ? (m >> (64 - l6)) : 0ull;
  const uint64_t mr = m >> r6;
  uint64_t L = (l < 64) ? ml : 0ull;
  uint64_t H = (l < 64) ? spill : (l < 128 ? ml : 0ull);
  if (sh < 0) {
    L = r < 64 ? mr : 0ull;
    H = 0;
  }
  lo = L;
  hi = H;
}

// Fixed-point image of a finite double: trunc(|x| * 2^S) with x's sign, as
// a two's-complement 128-bit integer. The caller's S keeps the magnitude
// below 2^127.
DFR_HD void fix_of(uint64_t bits, int S, uint64_t &lo, uint64_t &hi) {
  const int eb = (int)((bits >> 52) & 0x7ff);
  const uint64_t m = (bits & ((1ull << 52) - 1)) | (eb ? (1ull << 52) : 0ull);
  const int e = eb ? eb : 1;
  uint64_t l, h;
  shift_u64_to_128(m, e - 1075 + S, l, h);  // value = m * 2^(e - 1075)
  // two's complement when negative, without branches: with neg all ones,
  // (x ^ neg) - neg = ~x + 1; the low word's +1 carries into the high word
  // exactly when the low word is 0
  const uint64_t neg = (uint64_t)0 - (bits >> 63);
  lo = (l ^ neg) - neg;
  hi = (h ^ neg) + ((neg & 1ull) & (l == 0 ? 1ull : 0ull));
}

// Fixed-point image of x*x: trunc(x^2 * 2^S2), non-negative.
DFR_HD void fixsq_of(uint64_t bits, int S2, uint64_t &lo, uint64_t &hi) {
  int e = (int)((bits >> 52) & 0x7ff);
  uint64_t m = bits & ((1ull << 52) - 1);
  if (e == 0)
    e = 1;
  else
    m |= 1ull << 52;
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t ph = __umul64hi(m, m);
#else
  uint64_t ph = (uint64_t)(((unsigned __int128)m * m) >> 64);
#endif
  uint64_t pl = m * m;
  int sh = 2 * (e - 1075) + S2;
  if (sh >= 0)
    shl128(pl, ph, sh);
  else
    shr128(pl, ph, -sh);
  lo = pl;
  hi = ph;
}

// Correctly rounded (nearest-even) conversion of an unsigned 64-bit value.
DFR_HD double u64_to_f64(uint64_t u) {
  double h = (double)(uint32_t)(u >> 32);
  double l = (double)(uint32_t)u;
  return h * 4294967296.0 + l;  // both operands exact: one rounding
}

// Correctly rounded conversion of a 128-bit two's-complement integer,
// times 2^-scale.
DFR_HD double i128_to_f64(uint64_t lo, uint64_t hi, int scale) {
  bool neg = (int64_t)hi < 0;
  if (neg) {
    lo = ~lo + 1;
    hi = ~hi + (lo == 0 ? 1 : 0);
  }
  double d;
  if (hi == 0) {
    d = u64_to_f64(lo);
    d = ldexp(d, -scale);
  } else {
    int lz = __builtin_clzll(hi);
    int s = 64 - lz;  // right shift that leaves 64 significant bits
    uint64_t top = (lz == 0) ? hi : ((hi << lz) | (lo >> s));
    uint64_t dropped = (s == 64) ? lo : (lo << (64 - s));
    top |= (dropped != 0) ? 1 : 0;  // round to odd keeps the final rounding exact
    d = ldexp(u64_to_f64(top), s - scale);
  }
  return neg ? -d : d;
}

// ---- fixed point wider than 128 bits (2 to 4 words; used when a column's
// exponent range needs them, so kept simple rather than fast)

// 64 bits of the little-endian word array src[0..ns) starting at bit `o`
// (o may be negative or past the end: missing bits read as zero).
DFR_HD uint64_t bits_at(const uint64_t *src, int ns, int o) {
  if (o <= -64 || o >= 64 * ns) return 0;
  const int q = o >> 6, r = o & 63;  // floor division
  const uint64_t lo = (q >= 0 && q < ns) ? src[q] : 0ull;
  const uint64_t hi = (q + 1 >= 0 && q + 1 < ns) ? src[q + 1] : 0ull;
  return r ? ((lo >> r) | (hi << (64 - r))) : lo;
}

// Two's-complement negation of an N-word value.
template <int N>
DFR_HD void neg_words(uint64_t (&w)[N]) {
  uint64_t c = 1;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint64_t t = ~w[i] + c;
    c = (c && t == 0) ? 1ull : 0ull;
    w[i] = t;
  }
}

// Fixed-point image of a finite double in N words: trunc(|x| * 2^S) with
// x's sign, two's complement. The caller's S keeps the magnitude below
// 2^(64N - 1).
template <int N>
DFR_HD void fix_words(uint64_t bits, int S, uint64_t (&w)[N]) {
  const int eb = (int)((bits >> 52) & 0x7ff);
  uint64_t m[1] = {(bits & ((1ull << 52) - 1)) | (eb ? (1ull << 52) : 0ull)};
  const int sh = (eb ? eb : 1) - 1075 + S;  // value = m * 2^(e - 1075)
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = bits_at(m, 1, 64 * i - sh);
  if (bits >> 63) neg_words<N>(w);
}

// Fixed-point image of x*x in N words: trunc(x^2 * 2^S2), non-negative.
template <int N>
DFR_HD void fixsq_words(uint64_t bits, int S2, uint64_t (&w)[N]) {
  const int eb = (int)((bits >> 52) & 0x7ff);
  const uint64_t m = (bits & ((1ull << 52) - 1)) | (eb ? (1ull << 52) : 0ull);
  uint64_t p[2];
#if defined(__HIP_DEVICE_COMPILE__)
  p[1] = __umul64hi(m, m);
#else
  p[1] = (uint64_t)(((unsigned __int128)m * m) >> 64);
#endif
  p[0] = m * m;
  const int sh = 2 * ((eb ? eb : 1) - 1075) + S2;
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = bits_at(p, 2, 64 * i - sh);
}

// Word i of a 4-word value for a run-time i (0 outside 0..3). A select
// chain: an array indexed at run time would live in scratch (or LDS).
DFR_HD uint64_t pick4(const uint64_t (&w)[4], int i) {
  return i == 0 ? w[0] : i == 1 ? w[1] : i == 2 ? w[2] : i == 3 ? w[3] : 0ull;
}

// Correctly rounded value of the non-negative 4-word x times 2^-scale
// (round to odd into 64 bits, then one rounding to double).
DFR_HD double words4_to_f64(const uint64_t (&x)[4], int scale) {
  const int top = x[3] ? 3 : x[2] ? 2 : x[1] ? 1 : 0;
  if (top == 0) return ldexp(u64_to_f64(x[0]), -scale);
  const uint64_t hi = pick4(x, top), lo = pick4(x, top - 1);
  const int lz = __builtin_clzll(hi);
  uint64_t t = lz ? ((hi << lz) | (lo >> (64 - lz))) : hi;
  uint64_t sticky = lz ? (lo << lz) : lo;
#pragma unroll
  for (int i = 0; i < 2; ++i) sticky |= (i <= top - 2) ? x[i] : 0ull;
  t |= sticky ? 1 : 0;
  return ldexp(u64_to_f64(t), 64 * top - lz - scale);
}

// Correctly rounded value of the non-negative n-word x times 2^-scale
// (round to odd into 64 bits, then one rounding to double).
DFR_HD double words_to_f64(const uint64_t *x, int n, int scale) {
  int top = n - 1;
  while (top > 0 && x[top] == 0) --top;
  if (top == 0) return ldexp(u64_to_f64(x[0]), -scale);
  const int lz = __builtin_clzll(x[top]);
  const uint64_t hi = x[top], lo = x[top - 1];
  uint64_t t = lz ? ((hi << lz) | (lo >> (64 - lz))) : hi;
  uint64_t sticky = lz ? (lo << lz) : lo;
  for (int i = top - 2; i >= 0; --i) sticky |= x[i];
  t |= sticky ? 1 : 0;
  return ldexp(u64_to_f64(t), 64 * top - lz - scale);
}

// Correctly rounded value of the two's-complement n-word x (n <= 4) times
// 2^-scale.
DFR_HD double sword_to_f64(const uint64_t *x, int n, int scale) {
  uint64_t a[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = i < n ? x[i] : 0ull;
  const bool neg = (int64_t)pick4(a, n - 1) < 0;
  if (neg) {
    neg_words<4>(a);
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = i < n ? a[i] : 0ull;  // the n-word magnitude
  }
  const double d = words4_to_f64(a, scale);
  return neg ? -d : d;
}

// 256-bit unsigned arithmetic for the variance finish.
// q = n / d, r = n % d for d >= 1, given inv = 1 / d rounded to double.
// Below 2^53 the double estimate of n / d is within one of the quotient (n
// is exact, and the two roundings change n / d by at most 2 / d), so one
// correction step makes it exact; larger values use integer division.
DFR_HD void divmod_u64(uint64_t n, uint64_t d, double inv, uint64_t &q, uint64_t &r) {
  if (n < (1ull << 53) && d < (1ull << 53)) {
    uint64_t e = (uint64_t)((double)n * inv);
    int64_t rr = (int64_t)(n - e * d);
    if (rr < 0) {
      e -= 1;
      rr += (int64_t)d;
    } else if (rr >= (int64_t)d) {
      e += 1;
      rr -= (int64_t)d;
    }
    q = e;
    r = (uint64_t)rr;
  } else {
    q = n / d;
    r = n - q * d;
  }
}

struct U256 {
  uint64_t w[4];
};

DFR_HD void mul64(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  hi = __umul64hi(a, b);
  lo = a * b;
#else
  unsigned __int128 p = (unsigned __int128)a * b;
  lo = (uint64_t)p;
  hi = (uint64_t)(p >> 64);
#endif
}

// (a_lo, a_hi) * (b_lo, b_hi) as 256 bits.
DFR_HD U256 mul128(uint64_t alo, uint64_t ahi, uint64_t blo, uint64_t bhi) {
  U256 r = {{0, 0, 0, 0}};
  uint64_t a[2] = {alo, ahi}, b[2] = {blo, bhi};
  for (int i = 0; i < 2; ++i) {
    uint64_t carry = 0;
    for (int j = 0; j < 2; ++j) {
      uint64_t plo, phi;
      mul64(a[i], b[j], plo, phi);
      uint64_t t = r.w[i + j] + plo;
      uint64_t c1 = t < plo ? 1 : 0;
      uint64_t t2 = t + carry;
      uint64_t c2 = t2 < t ? 1 : 0;
      r.w[i + j] = t2;
      carry = phi + c1 + c2;
    }
    r.w[i + 2] += carry;
  }
  return r;
}

DFR_HD U256 shl256(U256 x, int s) {
  U256 r;
  const int q = s >> 6, b = s & 63;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint64_t hi = pick4(x.w, i - q), lo = pick4(x.w, i - q - 1);
    r.w[i] = b ? ((hi << b) | (lo >> (64 - b))) : hi;
  }
  return r;
}

// a - b; returns false when the difference is negative (then r = 0).
DFR_HD bool sub256(U256 a, U256 b, U256 &r) {
  uint64_t borrow = 0;
  for (int i = 0; i < 4; ++i) {
    uint64_t d = a.w[i] - b.w[i];
    uint64_t b1 = a.w[i] < b.w[i] ? 1 : 0;
    uint64_t d2 = d - borrow;
    uint64_t b2 = d < borrow ? 1 : 0;
    r.w[i] = d2;
    borrow = b1 | b2;
  }
  if (borrow) {
    r.w[0] = r.w[1] = r.w[2] = r.w[3] = 0;
    return false;
  }
  return true;
}

// Correctly rounded value of x * 2^-scale.
DFR_HD double u256_to_f64(U256 x, int scale) { return words4_to_f64(x.w, scale); }

// The variance finish of two-word totals (sum T at scale S, two's
// complement; sum of squares Q at scale S2 <= 2S): n * Q * 2^(2S - S2) - T^2
// exactly, in units of 2^-2S, rounded once, over n (n - ddof).
DFR_HD double var_2words(uint64_t tlo, uint64_t thi, int S, uint64_t qlo, uint64_t qhi, int S2, uint64_t n,
                         uint64_t ddof) {
  if ((int64_t)thi < 0) {
    tlo = ~tlo + 1;
    thi = ~thi + (tlo == 0 ? 1 : 0);
  }
  const U256 nq = mul128(qlo, qhi, n, 0);
  const U256 a = shl256(nq, 2 * S - S2);
  const U256 t2 = mul128(tlo, thi, tlo, thi);
  U256 r;
  sub256(a, t2, r);
  const double num = u256_to_f64(r, 2 * S);
  const double den = (double)n * (double)(n - ddof);
  return num / den;
}

// ---- 512-bit unsigned arithmetic for the variance finish of accumulators
// wider than 128 bits (little-endian words).
struct U512 {
  uint64_t w[8];
};

// a (na words) * b (nb words), na + nb <= 8.
DFR_HD U512 mul_words(const uint64_t *a, int na, const uint64_t *b, int nb) {
  U512 r;
  for (int i = 0; i < 8; ++i) r.w[i] = 0;
  for (int i = 0; i < na; ++i) {
    uint64_t carry = 0;
    for (int j = 0; j < nb; ++j) {
      uint64_t plo, phi;
      mul64(a[i], b[j], plo, phi);
      const uint64_t t = r.w[i + j] + plo;
      const uint64_t c1 = t < plo ? 1 : 0;
      const uint64_t t2 = t + carry;
      const uint64_t c2 = t2 < t ? 1 : 0;
      r.w[i + j] = t2;
      carry = phi + c1 + c2;
    }
    for (int k = i + nb; k < 8 && carry; ++k) {
      const uint64_t t = r.w[k] + carry;
      carry = t < carry ? 1 : 0;
      r.w[k] = t;
    }
  }
  return r;
}

// x << s for s >= 0 (bits past 512 are dropped).
DFR_HD U512 shl512(const U512 &x, int s) {
  U512 r;
  for (int i = 0; i < 8; ++i) r.w[i] = bits_at(x.w, 8, 64 * i - s);
  return r;
}

// a - b; returns false when the difference is negative (then r = 0).
DFR_HD bool sub512(const U512 &a, const U512 &b, U512 &r) {
  uint64_t borrow = 0;
  for (int i = 0; i < 8; ++i) {
    const uint64_t d = a.w[i] - b.w[i];
    const uint64_t b1 = a.w[i] < b.w[i] ? 1 : 0;
    const uint64_t d2 = d - borrow;
    const uint64_t b2 = d < borrow ? 1 : 0;
    r.w[i] = d2;
    borrow = b1 | b2;
  }
  if (borrow) {
    for (int i = 0; i < 8; ++i) r.w[i] = 0;
    return false;
  }
  return true;
}

// n * sum(x^2) and sum(x)^2 of a fixed-point column block (sum: nf words
// at scale S, two's complement; sum of squares: nq words at scale S2),
// brought to one scale, subtracted exactly, and rounded once:
// (n * Q - T^2) / (n * (n - ddof)) as a double. Both products stay below
// 2^508 for the widths and scales the host chooses (at most 4 words each).
DFR_HD double var_words(const uint64_t *sum, int nf, int S, const uint64_t *sq, int nq, int S2, uint64_t n,
                        uint64_t ddof) {
  uint64_t t[4] = {0, 0, 0, 0};
  for (int i = 0; i < nf; ++i) t[i] = sum[i];
  if ((int64_t)t[nf - 1] < 0) {
    // |T| (sign-extended to 4 words first)
    for (int i = nf; i < 4; ++i) t[i] = ~0ull;
    neg_words<4>(t);
  }
  const U512 nq_ = mul_words(sq, nq, &n, 1);
  const U512 t2 = mul_words(t, 4, t, 4);
  const int U = (2 * S > S2) ? 2 * S : S2;  // the common scale 2^-U
  const U512 a = shl512(nq_, U - S2);
  const U512 b = shl512(t2, U - 2 * S);
  U512 r;
  sub512(a, b, r);
  const double num = words_to_f64(r.w, 8, U);
  const double den = (double)n * (double)(n - ddof);
  return num / den;
}

// var_words out of line, for device code: its run-time indexed arrays then
// live in its own frame, and a kernel that never takes the wide path pays
// for them neither in registers nor in LDS.
static __host__ __device__ __attribute__((noinline)) double var_words_at(const uint64_t *sum, int nf, int S, const uint64_t *sq, int nq, int S2, uint64_t n,
                    uint64_t ddof) {
  return var_words(sum, nf, S, sq, nq, S2, n, ddof);
}

}  // namespace dfr
