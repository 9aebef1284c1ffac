// kernels.h — CDNA4 (gfx950) kernels of the filtered group-by path.
//
//   k_col_stats      one pass over a column at hand-off: min/max, exponent
//                    range, NaN/inf/signed-zero flags, null count
//   k_init_state     fills the HBM group table with each word's identity
//   k_aggregate      the fused scan: loads column values and validity words,
//                    runs the predicate program, packs the key, and folds
//                    each kept row into an LDS-private group table (direct
//                    or first-come hashed) or, when that misses, into the
//                    HBM table (direct-addressed or open-addressed with
//                    64-bit CAS); LDS entries are merged into HBM at the end
//   k_occupancy      slot occupancy -> bitmap (one ballot per 64 slots)
//   k_scan_*         exclusive prefix of bitmap popcounts
//   k_compact        slot order -> dense group order
//   k_mark_first,
//   k_rank_first     maintain_order: group order by first kept row, by
//                    ranking first rows in a row bitmap (no sort)
//   k_finalize       keys and aggregate results, validity as packed bits
#pragma once
// DFR_JIT: this header is being compiled at run time (hiprtc) for one
// query's pass A, which needs only the row helpers below.
#ifndef DFR_JIT
#define DFR_JIT 0
#endif
// DFR_COL_LEAVES 0: no predicate leaf compares two columns. A kernel
// compiled at run time for a plan with such leaves gets jit_leaf_rhs from
// its generated source: leaf l's right-hand column by constant index (a
// select over every column per unrolled leaf multiplies compile time).
#ifndef DFR_COL_LEAVES
#define DFR_COL_LEAVES 1
#endif
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

#include "device_util.h"
#include "plan.h"
#if !DFR_JIT
#include "runtime.h"
#endif

namespace dfr {

constexpr int kBlock = 256;

// The plan is read through the constant address space: it does not change
// during a launch, so uniform plan fields become scalar loads held in SGPRs.
#define DFR_CONSTANT __attribute__((address_space(4)))
typedef const DFR_CONSTANT DevPlan &PlanRef;
__device__ __forceinline__ PlanRef plan_ref(const DevPlan *p) { return *(const DFR_CONSTANT DevPlan *)p; }

// A register copy of one plan record (dword by dword, as scalar loads).
template <class T>
__device__ __forceinline__ T cget(const DFR_CONSTANT T &x) {
  static_assert(sizeof(T) % 4 == 0, "plan records are dword multiples");
  T out;
  const DFR_CONSTANT uint32_t *src = (const DFR_CONSTANT uint32_t *)&x;
  uint32_t *dst = (uint32_t *)&out;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
  return out;
}

// A plan record of a register copy of the plan (runtime-compiled pass A).
template <class T>
__device__ __forceinline__ T cget(const T &x) {
  return x;
}

// A register copy of the plan (the runtime-compiled kernels) is indexed
// only at compile-time positions, so it stays in registers: loops over its
// arrays are unrolled to their maximum length there.
template <class PV>
struct RegPlan {
  static constexpr bool value = false;
};
template <>
struct RegPlan<DevPlan> {
  static constexpr bool value = true;
};

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Column data, validity words and records live in device (global) memory:
// reading them through the global address space keeps them off the flat
// path, whose operations count against both the memory and the LDS
// counters.
#define DFR_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const DFR_GLOBAL T *gp(const T *p) {
  return (const DFR_GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ DFR_GLOBAL T *gpw(T *p) {
  return (DFR_GLOBAL T *)p;
}

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint64_t hash_index(uint64_t key, int log2) {
  return (key * 0x9E3779B97F4A7C15ull) >> (64 - log2);
}

__device__ __forceinline__ uint64_t atomic_load_u64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// `v2`, `valid2`: the right-hand column's value under DFR_LEAF_COLUMN.
__device__ __forceinline__ bool leaf_true(const DevLeaf &L, bool is_float, uint64_t v, bool valid,
                                          uint64_t v2, bool valid2) {
  if (L.op == 6) return !valid;  // DFR_IS_NULL
  if (L.op == 7) return valid;   // DFR_IS_NOT_NULL
  const bool by_column = (L.flags & 4) != 0;
  if (!valid || (by_column && !valid2)) return false;
  const uint64_t rhs = by_column ? v2 : (uint64_t)L.lit;
  bool t;
  if (is_float) {
    double x = bits_f64(v), y = bits_f64(rhs);
    switch (L.op) {
      case 0: t = x == y; break;
      case 1: t = x != y; break;
      case 2: t = x < y; break;
      case 3: t = x <= y; break;
      case 4: t = x > y; break;
      default: t = x >= y; break;
    }
  } else {
    int64_t x = (int64_t)v;
    if (L.flags & 1) {
      int64_t m = L.mod;
      if (m == 0) return false;  // modulo zero is null
      int64_t q = (m == -1) ? 0 : x % m;
      if (q != 0 && ((q < 0) != (m < 0))) q += m;
      x = q;
    }
    int64_t y = (int64_t)rhs;
    switch (L.op) {
      case 0: t = x == y; break;
      case 1: t = x != y; break;
      case 2: t = x < y; break;
      case 3: t = x <= y; break;
      case 4: t = x > y; break;
      default: t = x >= y; break;
    }
  }
  return (L.flags & 2) ? !t : t;
}

#if DFR_COL_LEAVES && DFR_JIT
template <int NC>
__device__ __forceinline__ void jit_leaf_rhs(int l, const uint64_t (&vals)[NC], uint32_t vbits, uint64_t &v, bool &ok);
#endif

// Column k's value and validity from a row's values (k uniform: a select
// chain, no scratch).
template <int NC>
__device__ __forceinline__ void pick_col(const uint64_t (&vals)[NC], uint32_t vbits, int k, uint64_t &v, bool &ok) {
#pragma unroll
  for (int j = 0; j < NC; ++j)
    if (j == k) {
      v = vals[j];
      ok = ((vbits >> j) & 1u) != 0;
    }
}

__device__ __forceinline__ uint32_t truth_bit(PlanRef P, uint32_t m) {
  return (P.truth[m >> 5] >> (m & 31)) & 1u;
}

// HBM slot of a packed key (inserting it in the hash table when new).
__device__ __noinline__ uint64_t global_slot_hash(uint64_t *keys_tab, uint64_t cap, int cap_log2, int max_probe,
                                                  uint32_t *fault, uint64_t key) {
  uint64_t mask = cap - 1;
  uint64_t i = hash_index(key, cap_log2);
  for (int p = 0; p < max_probe; ++p) {
    uint64_t k = atomic_load_u64(&keys_tab[i]);
    if (k == key) return i;
    if (k == kEmpty) {
      unsigned long long old =
          atomicCAS((unsigned long long *)&keys_tab[i], (unsigned long long)kEmpty, (unsigned long long)key);
      if (old == kEmpty || old == key) return i;
    }
    i = (i + 1) & mask;
  }
  atomicOr(&fault[0], 1u);
  return kEmpty;
}

// The HBM slot of a packed key: the key itself (direct table) or its slot
// in the open-addressing table; kEmpty when that table is full.
template <class PV>
__device__ __forceinline__ uint64_t global_slot(const PV &P, uint64_t key) {
  if (P.mode == MODE_DIRECT) return key;
  return global_slot_hash(P.keys_tab, P.cap, P.cap_log2, P.max_probe, P.fault, key);
}

__device__ __forceinline__ void add128(uint64_t *p, uint64_t lo, uint64_t hi) {
  unsigned long long old = atomicAdd((unsigned long long *)p, (unsigned long long)lo);
  uint64_t carry = ((uint64_t)old + lo) < (uint64_t)old ? 1 : 0;
  uint64_t h = hi + carry;
  if (h) atomicAdd((unsigned long long *)(p + 1), (unsigned long long)h);
}

// An n-word add (n <= 4) by atomics, word by word from the lowest: each
// word's returned old value gives its carry into the next. Whatever the
// interleaving with other adds, the final words are the exact sum modulo
// 2^(64n): word i + 1 receives one carry for every wrap of word i.
__device__ __forceinline__ void add_words(uint64_t *p, const uint64_t *w, int n) {
  uint64_t carry = 0;
  for (int i = 0; i < n; ++i) {
    const uint64_t x = w[i] + carry;
    const uint64_t c1 = x < carry ? 1ull : 0ull;  // w[i] all ones plus a carry
    if (i == n - 1) {
      if (x) atomicAdd((unsigned long long *)(p + i), (unsigned long long)x);
      break;
    }
    uint64_t c2 = 0;
    if (x) {
      const uint64_t old = atomicAdd((unsigned long long *)(p + i), (unsigned long long)x);
      c2 = (old + x) < old ? 1ull : 0ull;
    }
    carry = c1 | c2;
  }
}

// Folds one finite value's fixed-point images into the sum (nfw words at
// `fsum`, scale S) or the sum of squares (nqw words at `sq`, scale S2) of a
// column block, for those wider than 128 bits (0: not folded here).
__device__ __noinline__ void add_wide_at(uint64_t *fsum, int nfw, int S, uint64_t *sq, int nqw, int S2,
                                         uint64_t bits) {
  uint64_t w[4];
  if (nfw > 2) {
    fix_words<4>(bits, S, w);
    add_words(fsum, w, nfw);
  }
  if (nqw > 2) {
    fixsq_words<4>(bits, S2, w);
    add_words(sq, w, nqw);
  }
}

__device__ __forceinline__ void add_wide(uint64_t *b, const DevColAgg &A, uint64_t bits) {
  const uint32_t m = A.mask;
  add_wide_at(b + A.o_fsum, (m & M_FSUM) ? A.nfw : 0, A.fshift, b + A.o_sq, (m & M_FSUMSQ) ? A.nqw : 0,
              A.fshift2, bits);
}

// ------------------------------------------------------------ finalize

struct AggOut {
  uint64_t bits;  // int64 value, or double bits
  bool valid;
};

__device__ __forceinline__ double finish_fsum(const DevColAgg &A, const uint64_t *b, uint64_t flags) {
  if ((flags & FL_NAN) || ((flags & FL_PINF) && (flags & FL_NINF))) return __builtin_nan("");
  if (flags & FL_PINF) return __builtin_inf();
  if (flags & FL_NINF) return -__builtin_inf();
  if (A.nfw > 2) return sword_to_f64(b + A.o_fsum, A.nfw, A.fshift);
  return i128_to_f64(b[A.o_fsum], b[A.o_fsum + 1], A.fshift);
}

__device__ __forceinline__ double finish_var(const DevColAgg &A, const uint64_t *b, uint64_t n,
                                             int ddof, uint64_t flags) {
  if (flags & (FL_NAN | FL_PINF | FL_NINF)) return __builtin_nan("");
  if (A.nfw > 2 || A.nqw > 2)
    return var_words_at(b + A.o_fsum, A.nfw, A.fshift, b + A.o_sq, A.nqw, A.fshift2, n, (uint64_t)ddof);
  return var_2words(b[A.o_fsum], b[A.o_fsum + 1], A.fshift, b[A.o_sq], b[A.o_sq + 1], A.fshift2, n,
                    (uint64_t)ddof);
}

// Column `col`'s aggregate block and whether it is Float64. A register
// copy of the plan is indexed by a static loop only (an index read from the
// plan itself would keep the whole copy in scratch memory).
template <class PV>
__device__ __forceinline__ void col_agg_of(const PV &P, int col, DevColAgg &A, bool &is_float) {
  if constexpr (RegPlan<PV>::value) {
    // field by field: a select of whole records becomes a copy through a
    // selected address, which also keeps the plan copy in memory
    A = DevColAgg{};
    is_float = false;
#pragma unroll
    for (int c = 0; c < kMaxCols; ++c) {
      const bool hit = c == col;
      A.mask = hit ? P.cagg[c].mask : A.mask;
      A.gbase = hit ? P.cagg[c].gbase : A.gbase;
      A.lbase = hit ? P.cagg[c].lbase : A.lbase;
      A.o_count = hit ? P.cagg[c].o_count : A.o_count;
      A.o_isum = hit ? P.cagg[c].o_isum : A.o_isum;
      A.o_fsum = hit ? P.cagg[c].o_fsum : A.o_fsum;
      A.o_sq = hit ? P.cagg[c].o_sq : A.o_sq;
      A.o_min = hit ? P.cagg[c].o_min : A.o_min;
      A.o_max = hit ? P.cagg[c].o_max : A.o_max;
      A.o_flags = hit ? P.cagg[c].o_flags : A.o_flags;
      A.o_zfirst = hit ? P.cagg[c].o_zfirst : A.o_zfirst;
      A.fshift = hit ? P.cagg[c].fshift : A.fshift;
      A.fshift2 = hit ? P.cagg[c].fshift2 : A.fshift2;
      A.int_as_float = hit ? P.cagg[c].int_as_float : A.int_as_float;
      A.nfw = hit ? P.cagg[c].nfw : A.nfw;
      A.nqw = hit ? P.cagg[c].nqw : A.nqw;
      is_float = hit ? (P.cols[c].is_float != 0) : is_float;
    }
  } else {
    A = cget(P.cagg[col]);
    is_float = P.cols[col].is_float;
  }
}

template <class PV>
__device__ __forceinline__ AggOut finish_agg(const PV &P, const DevAgg &G, const uint64_t *slot) {
  AggOut o = {0, true};
  uint64_t rows = slot[0];
  if (G.op == 0) {  // LEN
    o.bits = rows;
    return o;
  }
  DevColAgg A;
  bool is_float;
  col_agg_of(P, G.col, A, is_float);
  const uint64_t *b = slot + A.gbase;
  uint64_t n = (A.mask & M_VCOUNT) ? b[A.o_count] : rows;
  uint64_t flags = (A.mask & M_FLAGS) ? b[A.o_flags] : 0;
  switch (G.op) {
    case 1:  // COUNT
      o.bits = n;
      return o;
    case 2: {  // SUM
      o.valid = (int64_t)n >= G.min_count;
      if (is_float) {
        o.bits = o.valid ? f64_bits(finish_fsum(A, b, flags)) : 0;
      } else if (A.mask & M_ISUM64) {
        int64_t t = (int64_t)b[A.o_isum];
        if (!o.valid) t = 0;
        if (G.out_dtype == T_INT32 && (t > 2147483647ll || t < -2147483648ll)) atomicOr(&P.fault[2], 1u);
        o.bits = (uint64_t)t;
      } else {
        uint64_t lo = b[A.o_isum], hi = b[A.o_isum + 1];
        bool fits = (hi == 0 && (int64_t)lo >= 0) || (hi == ~0ull && (int64_t)lo < 0);
        if (!o.valid) lo = 0;
        else if (!fits) atomicOr(&P.fault[G.out_dtype == T_INT32 ? 2 : 1], 1u);
        int64_t t = (int64_t)lo;
        if (o.valid && G.out_dtype == T_INT32 && (t > 2147483647ll || t < -2147483648ll))
          atomicOr(&P.fault[2], 1u);
        o.bits = (uint64_t)t;
      }
      return o;
    }
    case 3:
    case 4: {  // MIN / MAX
      const bool is_min = G.op == 3;
      if (is_float) {
        uint64_t k = b[is_min ? A.o_min : A.o_max];
        bool seen = is_min ? (k != kEmpty) : (k != 0);
        if (!seen) {
          o.valid = (flags & FL_NAN) != 0;  // only NaN values: the result is NaN
          o.bits = o.valid ? f64_bits(__builtin_nan("")) : 0;
          return o;
        }
        uint64_t v = ord_key_inverse(k);
        if ((v << 1) == 0 && (A.mask & M_ZFIRST)) {
          uint64_t z = b[A.o_zfirst];
          if (z != kEmpty) v = (z & 1) << 63;  // the first zero's sign
        }
        o.bits = v;
      } else {
        o.valid = n > 0;
        o.bits = o.valid ? b[is_min ? A.o_min : A.o_max] : 0;
      }
      return o;
    }
    case 5: {  // MEAN
      o.valid = n > 0;
      if (!o.valid) return o;
      double total;
      if (A.mask & M_ISUM64)
        total = (double)(int64_t)b[A.o_isum];
      else if (A.mask & M_ISUM128)
        total = i128_to_f64(b[A.o_isum], b[A.o_isum + 1], 0);
      else
        total = finish_fsum(A, b, flags);
      o.bits = f64_bits(total / (double)n);
      return o;
    }
    default: {  // VAR / STD
      o.valid = (int64_t)n - (int64_t)G.ddof > 0;
      if (!o.valid) return o;
      double v = finish_var(A, b, n, G.ddof, flags);
      if (G.op == 7) v = sqrt(v);
      o.bits = f64_bits(v);
      return o;
    }
  }
}

__device__ __forceinline__ void store_value(void *dst, int dtype, uint64_t i, uint64_t bits) {
  if (dtype == T_INT32 || dtype == T_DICT)
    ((uint32_t *)dst)[i] = (uint32_t)bits;
  else
    ((uint64_t *)dst)[i] = bits;
}

// Whether truth-table index m keeps its row. The truth words are uniform;
// up to five leaves need only the first.
template <class PV>
__device__ __forceinline__ bool truth_keeps(const PV &P, uint32_t m) {
  uint32_t w = P.truth[0];
  if (P.n_leaves > 5) {
    const uint32_t q = m >> 5;
#pragma unroll
    for (int i = 1; i < 8; ++i)
      if (q == (uint32_t)i) w = P.truth[i];
  }
  return ((w >> (m & 31)) & 1u) != 0;
}

// The raw loads of U row pairs per thread: values as loaded (32-bit
// columns not yet widened) and the validity words holding each pair's
// bits. Nothing here reads a loaded register, so a tile's loads stay in
// flight until fold_tile uses them.
template <int NC, int U>
struct ScanTile {
  uint64_t v[U][2][NC];
  uint64_t vw[U][NC];  // validity word of row r (all ones without a bitmap)
  uint32_t vx[U][NC];  // low half of the word after it (row r+1's bit when r's is bit 63)
  bool any[U], two[U];
};

// Tile rows base + u * 2 * blockDim + 2 * thread (+1) of the even-length
// range [b0, end), end - b0 >= 2. A pair past the range loads the range's
// last pair instead (its rows are not folded), so every load is issued
// unconditionally: a branch around a load, or between two load forms,
// makes the compiler wait for the load where the paths merge.
template <int NC, int U, class PV>
__device__ __forceinline__ void load_tile(const PV &P, ScanTile<NC, U> &t, int64_t base, int64_t end) {
  const int64_t stride = 2 * (int64_t)blockDim.x;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t r = base + u * stride + 2 * (int64_t)threadIdx.x;
    t.any[u] = t.two[u] = r < end;
    const int64_t rc = min(r, end - 2);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      t.v[u][0][c] = t.v[u][1][c] = 0;
      t.vw[u][c] = ~0ull;
      t.vx[u][c] = ~0u;
      if (c >= P.n_cols) continue;
      const DevCol C = cget(P.cols[c]);
      if (C.dtype == T_INT64 || C.dtype == T_FLOAT64) {
        const DFR_GLOBAL uint64_t *p = gp((const uint64_t *)C.values) + rc;
        if (P.vec_ok) {
          const u64x2 q = __builtin_nontemporal_load((const DFR_GLOBAL u64x2 *)p);
          t.v[u][0][c] = q.x;
          t.v[u][1][c] = q.y;
        } else {
          t.v[u][0][c] = __builtin_nontemporal_load(p);
          t.v[u][1][c] = __builtin_nontemporal_load(p + 1);
        }
      } else {
        const DFR_GLOBAL uint32_t *p = gp((const uint32_t *)C.values) + rc;
        if (P.vec_ok) {
          const u32x2 q = __builtin_nontemporal_load((const DFR_GLOBAL u32x2 *)p);
          t.v[u][0][c] = q.x;
          t.v[u][1][c] = q.y;
        } else {
          t.v[u][0][c] = __builtin_nontemporal_load(p);
          t.v[u][1][c] = __builtin_nontemporal_load(p + 1);
        }
      }
      if (C.valid) {
        const DFR_GLOBAL uint64_t *vp = gp(C.valid);
        const int64_t w = (C.vbit + rc) >> 6, wlast = (C.vbit + end - 1) >> 6;
        t.vw[u][c] = vp[w];
        t.vx[u][c] = ((const DFR_GLOBAL uint32_t *)vp)[2 * min(w + 1, wlast)];
      }
    }
  }
}

// The last row of a range of odd length: thread 0's row pair 0 holds it.
template <int NC, class PV>
__device__ __forceinline__ void load_last_row(const PV &P, ScanTile<NC, 1> &t, int64_t r) {
  t.any[0] = threadIdx.x == 0;
  t.two[0] = false;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    t.v[0][0][c] = t.v[0][1][c] = 0;
    t.vw[0][c] = ~0ull;
    t.vx[0][c] = ~0u;
    if (c >= P.n_cols || !t.any[0]) continue;
    const DevCol C = cget(P.cols[c]);
    if (C.dtype == T_INT64 || C.dtype == T_FLOAT64)
      t.v[0][0][c] = gp((const uint64_t *)C.values)[r];
    else
      t.v[0][0][c] = gp((const uint32_t *)C.values)[r];
    if (C.valid) t.vw[0][c] = gp(C.valid)[(C.vbit + r) >> 6];
  }
}

// A raw 32-bit value as the column's 64-bit value (Int32 sign-extends).
__device__ __forceinline__ uint64_t widen(const DevCol C, uint64_t x) {
  if (C.dtype == T_INT32) return (uint64_t)(int64_t)(int32_t)(uint32_t)x;
  if (C.dtype == T_INT64 || C.dtype == T_FLOAT64) return x;
  return (uint32_t)x;
}

// Folds a loaded tile's 2U rows: applies the predicate and hands each row to
// `sink(kept, row, vals, validity bits)`, called by every lane (kept or
// not). Rolled (ROLL), one loop over the rows keeps a single copy of the
// row code and picks the row's values from the tile registers with
// selects. The pipelined scan unrolls it: a loop that issues no loads but
// reads registers loaded before it makes the compiler wait for every load
// in flight at its entry, the next tile's included.
//
// TL > 0 is the number of loads the next tile has in flight behind this
// one: the fold first waits until only those remain, so this tile's
// registers are complete on every path through it (otherwise a register
// whose load a branch skipped waiting for is still pending where it is
// next written, and that write waits for every load in flight). The
// compiler still inserts any wait a use needs: the count only saves waits.
template <int NC, int U, bool ROLL, int TL = 0, class PV, class Sink>
__device__ __forceinline__ void fold_tile(const PV &P, const ScanTile<NC, U> &t, int64_t base, Sink sink) {
  const int64_t stride = 2 * (int64_t)blockDim.x;
  if constexpr (TL > 0 && TL < 64)  // gfx9 s_waitcnt: vmcnt in bits 3:0 and 15:14; expcnt and lgkmcnt not waited
    __builtin_amdgcn_s_waitcnt((TL & 15) | (7 << 4) | (15 << 8) | ((TL >> 4) << 14));
#pragma unroll(ROLL ? 1 : 2 * U)
  for (int i = 0; i < 2 * U; ++i) {
    uint64_t vals[NC];
    uint32_t vbits = 0;
    bool in = false;
    const int64_t row = base + (i >> 1) * stride + 2 * (int64_t)threadIdx.x + (i & 1);
#pragma unroll
    for (int c = 0; c < NC; ++c) vals[c] = 0;
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        if (i != 2 * u + j) continue;
        in = j ? t.two[u] : t.any[u];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (c >= P.n_cols) continue;
          const DevCol C = cget(P.cols[c]);
          vals[c] = widen(C, t.v[u][j][c]);
          uint32_t bitv;
          if (!C.valid) {
            bitv = 1u;
          } else {
            const int sh = (int)((C.vbit + row - j) & 63) + j;
            bitv = sh < 64 ? (uint32_t)(t.vw[u][c] >> sh) & 1u : t.vx[u][c] & 1u;
          }
          vbits |= bitv << c;
        }
      }
    bool kept = in;
    if (in && P.n_leaves) {
      uint32_t m = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c >= P.n_cols || !((P.leaf_cols >> c) & 1)) continue;
        const bool fl = P.cols[c].is_float;
        if constexpr (RegPlan<PV>::value) {
#pragma unroll
          for (int l = 0; l < kMaxLeaves; ++l) {
            if (l < P.leaf_begin[c] || l >= P.leaf_begin[c + 1]) continue;
            const DevLeaf L = P.leaves[l];
            uint64_t v2 = 0;
            bool ok2 = false;
#if DFR_COL_LEAVES && DFR_JIT
            jit_leaf_rhs<NC>(l, vals, vbits, v2, ok2);
#elif DFR_COL_LEAVES
            if (L.col2 >= 0) pick_col<NC>(vals, vbits, L.col2, v2, ok2);
#endif
            m |= (uint32_t)leaf_true(L, fl, vals[c], (vbits >> c) & 1u, v2, ok2) << L.bit;
          }
        } else {
          for (int l = P.leaf_begin[c]; l < P.leaf_begin[c + 1]; ++l) {
            const DevLeaf L = cget(P.leaves[l]);
            uint64_t v2 = 0;
            bool ok2 = false;
            if (L.col2 >= 0) pick_col<NC>(vals, vbits, L.col2, v2, ok2);
            m |= (uint32_t)leaf_true(L, fl, vals[c], (vbits >> c) & 1u, v2, ok2) << L.bit;
          }
        }
      }
      kept = truth_keeps(P, m);
    }
    sink(kept, row, vals, vbits);
  }
}

#if !DFR_JIT
// The rare-event words (FLAGS, ZFIRST) live in HBM only.
__device__ __noinline__ void rare_update(PlanRef P, uint64_t packed, int gword,
                                         uint64_t value, bool is_min) {
  uint64_t g = global_slot(P, packed);
  if (g == kEmpty) return;
  uint64_t *p = P.state + g * (uint64_t)P.words + gword;
  if (is_min)
    atomicMin((unsigned long long *)p, (unsigned long long)value);
  else
    atomicOr((unsigned long long *)p, (unsigned long long)value);
}

// Folds one kept row into a group's words. `LDS` selects the ds_* form.
template <bool LDS, int NC>
__device__ __forceinline__ void fold_row(PlanRef P, uint64_t *st, int64_t row,
                                         const uint64_t (&vals)[NC], uint32_t validbits,
                                         uint64_t packed) {
  atomicAdd((unsigned long long *)&st[0], 1ull);
  if (P.has_first) {
    if ((uint64_t)row < st[1]) atomicMin((unsigned long long *)&st[1], (unsigned long long)row);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c >= P.n_cols) continue;
    const DevColAgg A = cget(P.cagg[c]);
    uint32_t m = A.mask;
    if (!m || !((validbits >> c) & 1u)) continue;
    uint64_t *b = st + (LDS ? A.lbase : A.gbase);
    uint64_t v = vals[c];
    if (m & M_VCOUNT) atomicAdd((unsigned long long *)&b[A.o_count], 1ull);
    if (m & M_ISUM64) atomicAdd((unsigned long long *)&b[A.o_isum], (unsigned long long)v);
    if (m & M_ISUM128) add128(&b[A.o_isum], v, ((int64_t)v < 0) ? ~0ull : 0ull);
    if (m & M_IMIN) {
      if ((int64_t)v < (int64_t)b[A.o_min]) atomicMin((long long *)&b[A.o_min], (long long)v);
    }
    if (m & M_IMAX) {
      if ((int64_t)v > (int64_t)b[A.o_max]) atomicMax((long long *)&b[A.o_max], (long long)v);
    }
    if (m & (M_FSUM | M_FSUMSQ | M_FMIN | M_FMAX | M_FLAGS | M_ZFIRST)) {
      uint64_t bits = A.int_as_float ? f64_bits((double)(int64_t)v) : v;
      uint32_t ex = (uint32_t)(bits >> 52) & 0x7ffu;
      bool nan = false;
      if (ex == 0x7ffu) {
        nan = (bits << 12) != 0;
        uint64_t fl = nan ? FL_NAN : ((bits >> 63) ? FL_NINF : FL_PINF);
        if (m & M_FLAGS) rare_update(P, packed, A.gbase + A.o_flags, fl, false);
      } else {
        if ((m & M_FSUM) && A.nfw == 2) {
          uint64_t lo, hi;
          fix_of(bits, A.fshift, lo, hi);
          add128(&b[A.o_fsum], lo, hi);
        }
        if ((m & M_FSUMSQ) && A.nqw == 2) {
          uint64_t lo, hi;
          fixsq_of(bits, A.fshift2, lo, hi);
          add128(&b[A.o_sq], lo, hi);
        }
        if (A.nfw > 2 || A.nqw > 2) add_wide(b, A, bits);
        if ((m & M_ZFIRST) && (bits << 1) == 0)
          rare_update(P, packed, A.gbase + A.o_zfirst, ((uint64_t)row << 1) | (bits >> 63), true);
      }
      if (!nan) {
        uint64_t k = ord_key(bits);
        if (m & M_FMIN) {
          if (k < b[A.o_min]) atomicMin((unsigned long long *)&b[A.o_min], (unsigned long long)k);
        }
        if (m & M_FMAX) {
          if (k > b[A.o_max]) atomicMax((unsigned long long *)&b[A.o_max], (unsigned long long)k);
        }
      }
    }
  }
}

template <int LDSM>
__device__ __forceinline__ int lds_find(PlanRef P, uint64_t *lkeys, uint64_t key) {
  if (LDSM == LDS_DIRECT) return (int)key;
  const int E = P.lds_entries;
  int h = (int)hash_index(key, P.lds_log2);
  for (int p = 0; p < P.lds_probe; ++p) {
    uint64_t k = lkeys[h];
    if (k == key) return h;
    if (k == kEmpty) {
      unsigned long long old = atomicCAS((unsigned long long *)&lkeys[h],
                                         (unsigned long long)kEmpty, (unsigned long long)key);
      if (old == kEmpty || old == key) return h;
    }
    h = (h + 1) & (E - 1);
  }
  return -1;
}

template <int LDSM, int NC>
__device__ __forceinline__ void fold_kept(PlanRef P, uint64_t *lkeys, uint64_t *lst,
                                          int64_t row, const uint64_t (&vals)[NC],
                                          uint32_t validbits) {
  uint64_t packed = 0;
  if (P.keys[0].is_float) {
    packed = (validbits & 1u) ? canon_key_bits(vals[0]) : kNullFloatKey;
  } else {
#pragma unroll
    for (int k = 0; k < kMaxKeys && k < NC; ++k) {
      if (k >= P.n_keys) continue;
      const DevKey K = cget(P.keys[k]);
      uint64_t digit = ((validbits >> k) & 1u) ? (vals[k] - (uint64_t)K.minv) : (K.radix - 1);
      packed += digit * K.stride;
    }
  }
  if (LDSM != LDS_NONE) {
    int e = lds_find<LDSM>(P, lkeys, packed);
    if (e >= 0) {
      fold_row<true, NC>(P, lst + (int64_t)e * P.lds_words, row, vals, validbits, packed);
      return;
    }
  }
  uint64_t g = global_slot(P, packed);
  if (g == kEmpty) return;
  fold_row<false, NC>(P, P.state + g * (uint64_t)P.words, row, vals, validbits, packed);
}

// Merges one LDS entry into its HBM slot.
__device__ __forceinline__ void merge_entry(PlanRef P, const uint64_t *s, uint64_t key) {
  uint64_t g = global_slot(P, key);
  if (g == kEmpty) return;
  uint64_t *d = P.state + g * (uint64_t)P.words;
  atomicAdd((unsigned long long *)&d[0], (unsigned long long)s[0]);
  if (P.has_first) atomicMin((unsigned long long *)&d[1], (unsigned long long)s[1]);
  for (int c = 0; c < P.n_cols; ++c) {
    const DevColAgg A = cget(P.cagg[c]);
    uint32_t m = A.mask;
    if (!m) continue;
    const uint64_t *a = s + A.lbase;
    uint64_t *b = d + A.gbase;
    if ((m & M_VCOUNT) && a[A.o_count]) atomicAdd((unsigned long long *)&b[A.o_count], (unsigned long long)a[A.o_count]);
    if (m & M_ISUM64) atomicAdd((unsigned long long *)&b[A.o_isum], (unsigned long long)a[A.o_isum]);
    if (m & M_ISUM128) add128(&b[A.o_isum], a[A.o_isum], a[A.o_isum + 1]);
    if (m & M_FSUM) add_words(&b[A.o_fsum], &a[A.o_fsum], A.nfw);
    if (m & M_FSUMSQ) add_words(&b[A.o_sq], &a[A.o_sq], A.nqw);
    if (m & M_IMIN) atomicMin((long long *)&b[A.o_min], (long long)a[A.o_min]);
    if (m & M_IMAX) atomicMax((long long *)&b[A.o_max], (long long)a[A.o_max]);
    if ((m & M_FMIN) && a[A.o_min] != kEmpty) atomicMin((unsigned long long *)&b[A.o_min], (unsigned long long)a[A.o_min]);
    if ((m & M_FMAX) && a[A.o_max] != 0) atomicMax((unsigned long long *)&b[A.o_max], (unsigned long long)a[A.o_max]);
  }
}

// ------------------------------------------------------------- the scan

// One workgroup folds rows [b0, b1) into its LDS table, then merges the
// table into the HBM table. The caller syncs before reusing `smem`.
template <int LDSM, int NC>
__device__ __forceinline__ void aggregate_tile(PlanRef P, uint64_t *smem, int64_t b0, int64_t b1) {
  uint64_t *lkeys = smem;
  uint64_t *lst = smem + (LDSM == LDS_HASHED ? P.lds_entries : 0);
  if (LDSM != LDS_NONE) {
    const int E = P.lds_entries, LW = P.lds_words;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
      if (LDSM == LDS_HASHED) lkeys[i] = kEmpty;
      for (int w = 0; w < LW; ++w) lst[i * LW + w] = P.ident[P.lds_to_g[w]];
    }
    __syncthreads();
  }
  // whole row pairs with unconditional loads (every lane's loads in flight
  // together), then an odd range's last row
  auto fold = [&](bool kept, int64_t row, const uint64_t(&vals)[NC], uint32_t vbits) {
    if (kept) fold_kept<LDSM, NC>(P, lkeys, lst, row, vals, vbits);
  };
  const int64_t step = 2 * (int64_t)blockDim.x;
  const int64_t b1e = b0 + (max(b1 - b0, (int64_t)0) & ~(int64_t)1);
  for (int64_t base = b0; base < b1e; base += step) {
    ScanTile<NC, 1> t;
    load_tile<NC, 1>(P, t, base, b1e);
    fold_tile<NC, 1, true>(P, t, base, fold);
  }
  if (b1e < b1) {
    ScanTile<NC, 1> t;
    load_last_row<NC>(P, t, b1e);
    fold_tile<NC, 1, true>(P, t, b1e, fold);
  }
  if (LDSM != LDS_NONE) {
    __syncthreads();
    const int E = P.lds_entries, LW = P.lds_words;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      const uint64_t *s = lst + (int64_t)e * LW;
      if (s[0] == 0) continue;
      merge_entry(P, s, LDSM == LDS_DIRECT ? (uint64_t)e : lkeys[e]);
    }
  }
}

// Each workgroup folds rows [row_begin + blockIdx.x * chunk, +chunk). chunk
// is a multiple of 2 * blockDim.x.
template <int LDSM, int NC>
__global__ void __launch_bounds__(1024) k_aggregate(const DevPlan *__restrict__ plan,
                                                    int64_t row_begin, int64_t row_end,
                                                    int64_t chunk) {
  extern __shared__ uint64_t smem[];
  PlanRef P = plan_ref(plan);
  const int64_t b0 = row_begin + (int64_t)blockIdx.x * chunk;
  aggregate_tile<LDSM, NC>(P, smem, b0, min(b0 + chunk, row_end));
}

// ------------------------------------------------------ table set-up

__global__ void k_init_state(const DevPlan *__restrict__ plan) {
  PlanRef P = plan_ref(plan);
  const uint32_t W = (uint32_t)P.words;
  // each block covers kBlock slots, coalesced over their words
  for (uint64_t s0 = (uint64_t)blockIdx.x * kBlock; s0 < P.cap; s0 += (uint64_t)gridDim.x * kBlock) {
    uint64_t nslots = min((uint64_t)kBlock, P.cap - s0);
    uint32_t total = (uint32_t)nslots * W;
    uint64_t *base = P.state + s0 * W;
    for (uint32_t j = threadIdx.x; j < total; j += kBlock) base[j] = P.ident[j % W];
    if (P.mode == MODE_HASH && threadIdx.x < nslots) P.keys_tab[s0 + threadIdx.x] = kEmpty;
  }
}

// -------------------------------------------------- compaction / order

__global__ void k_occupancy(const DevPlan *__restrict__ plan, uint64_t *bits, uint64_t nwords) {
  PlanRef P = plan_ref(plan);
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool occ = s < P.cap && P.state[s * (uint64_t)P.words] != 0;
  uint64_t ball = __ballot(occ);
  if ((threadIdx.x & 63) == 0 && (s >> 6) < nwords) bits[s >> 6] = ball;
}

constexpr int kScanItems = 8;  // words per thread
constexpr int kScanChunk = kBlock * kScanItems;

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t x, uint64_t *total) {
  __shared__ uint64_t sh[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = x;
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t y = __shfl_up((unsigned long long)inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) sh[wave] = inc;
  __syncthreads();
  uint64_t off = 0, tot = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) off += sh[w];
    tot += sh[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - x;
}

__global__ void k_scan_blocks(const uint64_t *bits, uint64_t n, uint64_t *block_sums) {
  uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint64_t c = 0;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) c += __popcll(bits[base + i]);
  uint64_t tot;
  block_exclusive_scan(c, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// One workgroup: exclusive scan of block sums in place; total at [nb].
__global__ void k_scan_top(uint64_t *block_sums, uint64_t nb) {
  uint64_t carry = 0;
  for (uint64_t b0 = 0; b0 < nb; b0 += kBlock) {
    uint64_t i = b0 + threadIdx.x;
    uint64_t x = i < nb ? block_sums[i] : 0;
    uint64_t tot;
    uint64_t ex = block_exclusive_scan(x, &tot);
    if (i < nb) block_sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) block_sums[nb] = carry;
}

__global__ void k_scan_apply(const uint64_t *bits, uint64_t n, const uint64_t *block_sums,
                             uint64_t *prefix) {
  uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint64_t c = 0;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) c += __popcll(bits[base + i]);
  uint64_t tot;
  uint64_t run = block_sums[blockIdx.x] + block_exclusive_scan(c, &tot);
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) {
      prefix[base + i] = run;
      run += __popcll(bits[base + i]);
    }
}

__device__ __forceinline__ uint64_t rank_of(const uint64_t *bits, const uint64_t *prefix, uint64_t i) {
  uint64_t w = bits[i >> 6];
  uint64_t below = (i & 63) ? (w & ((1ull << (i & 63)) - 1)) : 0;
  return prefix[i >> 6] + __popcll(below);
}

__global__ void k_compact(const DevPlan *__restrict__ plan, const uint64_t *bits,
                          const uint64_t *prefix, uint64_t *order) {
  PlanRef P = plan_ref(plan);
  uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.cap) return;
  if ((bits[s >> 6] >> (s & 63)) & 1) order[rank_of(bits, prefix, s)] = s;
}

__global__ void k_mark_first(const DevPlan *__restrict__ plan, const uint64_t *order,
                             uint64_t groups, uint64_t *rowbits) {
  PlanRef P = plan_ref(plan);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= groups) return;
  uint64_t r = P.state[(order ? order[i] : i) * (uint64_t)P.words + 1];
  atomicOr((unsigned long long *)&rowbits[r >> 6], (unsigned long long)(1ull << (r & 63)));
}

__global__ void k_rank_first(const DevPlan *__restrict__ plan, const uint64_t *order_in,
                             uint64_t groups, const uint64_t *rowbits, const uint64_t *rowprefix,
                             uint64_t *order_out) {
  PlanRef P = plan_ref(plan);
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= groups) return;
  uint64_t s = order_in ? order_in[i] : i;
  uint64_t r = P.state[s * (uint64_t)P.words + 1];
  order_out[rank_of(rowbits, rowprefix, r)] = s;
}

// Output group i (of G) from table slot s. The 64 lanes of a wave hold
// groups 64k .. 64k + 63 of one result (inactive lanes past G), so lane 0
// writes each validity word whole.
__device__ __forceinline__ void finalize_group(PlanRef P, uint64_t i, uint64_t s, bool active, uint64_t G) {
  const uint64_t nwords = (G + 63) >> 6;
  const bool writer = (threadIdx.x & 63) == 0 && (i >> 6) < nwords;
  const uint64_t *slot = P.state + s * (uint64_t)P.words;
  uint64_t packed = 0;
  if (active) packed = (P.mode == MODE_DIRECT) ? s : P.keys_tab[s];
  if (active && P.out_first) P.out_first[i] = slot[1];
  uint64_t rest = packed;  // digits from the first key up
  for (int k = 0; k < P.n_keys; ++k) {
    const DevKey K = cget(P.keys[k]);
    bool valid;
    uint64_t v;
    if (K.is_float) {
      // the group's first row's own value (FIRST is tracked for such keys)
      valid = packed != kNullFloatKey;
      v = (active && valid) ? ((const uint64_t *)P.cols[k].values)[slot[1]] : 0;
    } else {
      uint64_t digit;
      divmod_u64(rest, K.radix, K.inv_radix, rest, digit);
      valid = !(K.nullable && digit == K.radix - 1);
      v = valid ? (uint64_t)K.minv + digit : 0;
    }
    if (active) store_value(P.out_keys[k], P.cols[k].dtype, i, v);
    uint64_t ball = __ballot(active && valid);
    if (writer) P.out_key_valid[k][i >> 6] = ball;
  }
  for (int a = 0; a < P.n_aggs; ++a) {
    const DevAgg Ga = cget(P.aggs[a]);
    AggOut o = {0, false};
    if (active) o = finish_agg(P, Ga, slot);
    if (active) store_value(Ga.out_values, Ga.out_dtype, i, o.bits);
    uint64_t ball = __ballot(active && o.valid);
    if (writer) Ga.out_valid[i >> 6] = ball;
  }
}

__global__ void k_finalize(const DevPlan *__restrict__ plan) {
  PlanRef P = plan_ref(plan);
  const uint64_t G = *P.groups;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < G;
  const uint64_t s = active ? (P.order ? P.order[i] : i) : 0;
  finalize_group(P, i, s, active, G);
}

// ------------------------------------------------------- column stats

// A HyperLogLog sketch of a column's distinct values (its group-key words:
// Float64 values canonical, as grouping sees them): kHllRegs registers, each
// the largest rank (leading zeros + 1) of the hashed words it received.
constexpr int kHllBits = 12;
constexpr int kHllRegs = 1 << kHllBits;

__device__ __forceinline__ void hll_add(uint32_t *regs, uint64_t word) {
  uint64_t h = word * 0x9E3779B97F4A7C15ull;  // SplitMix64's finalizer
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  const uint32_t j = (uint32_t)(h >> (64 - kHllBits));
  const uint32_t rank = (uint32_t)__builtin_clzll((h << kHllBits) | (1ull << (kHllBits - 1))) + 1u;
  if (rank > regs[j]) atomicMax(&regs[j], rank);
}

// One pass over a column: min/max, exponent range, flags, valid count, and
// (when `hll` is given: kHllRegs zeroed registers) the distinct sketch.
__global__ void k_col_stats(DevCol C, int64_t n, ColStats *out, uint32_t *hll) {
  __shared__ uint32_t regs[kHllRegs];
  if (hll) {
    for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x) regs[i] = 0;
    __syncthreads();
  }
  int64_t lo = INT64_MAX, hi = INT64_MIN;
  uint32_t emin = 0x7ff, emax = 0, flags = 0;
  uint64_t cnt = 0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    if (C.valid) {
      int64_t bit = C.vbit + r;
      if (!((C.valid[bit >> 6] >> (bit & 63)) & 1)) continue;
    }
    ++cnt;
    if (C.dtype == T_FLOAT64) {
      uint64_t b = ((const uint64_t *)C.values)[r];
      if (hll) hll_add(regs, canon_key_bits(b));
      uint32_t e = (uint32_t)(b >> 52) & 0x7ff;
      if (e == 0x7ff) {
        flags |= (b << 12) ? 1u : ((b >> 63) ? 4u : 2u);
      } else if ((b << 1) == 0) {
        flags |= (b >> 63) ? 8u : 16u;
      } else {
        flags |= 32u;
        uint32_t eb = e ? e : 1;
        emin = min(emin, eb);
        emax = max(emax, e ? e : 1u);
      }
    } else {
      int64_t v;
      if (C.dtype == T_INT64)
        v = ((const int64_t *)C.values)[r];
      else if (C.dtype == T_INT32)
        v = ((const int32_t *)C.values)[r];
      else
        v = ((const uint32_t *)C.values)[r];
      if (hll) hll_add(regs, (uint64_t)v);
      lo = min(lo, v);
      hi = max(hi, v);
    }
  }
  if (hll) {
    __syncthreads();
    for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x)
      if (regs[i]) atomicMax(&hll[i], regs[i]);
  }
  for (int d = 32; d >= 1; d >>= 1) {
    lo = min(lo, (int64_t)__shfl_xor((long long)lo, d, 64));
    hi = max(hi, (int64_t)__shfl_xor((long long)hi, d, 64));
    emin = min(emin, (uint32_t)__shfl_xor((unsigned int)emin, d, 64));
    emax = max(emax, (uint32_t)__shfl_xor((unsigned int)emax, d, 64));
    flags |= (uint32_t)__shfl_xor((unsigned int)flags, d, 64);
    cnt += (uint64_t)__shfl_xor((unsigned long long)cnt, d, 64);
  }
  // the block's waves combined in LDS: one set of device atomics per block
  // (same-address device atomics serialize)
  __shared__ ColStats part[kBlock / 64];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[wave].imin = lo;
    part[wave].imax = hi;
    part[wave].emin = emin;
    part[wave].emax = emax;
    part[wave].flags = flags;
    part[wave].valid = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
      lo = min(lo, part[w].imin);
      hi = max(hi, part[w].imax);
      emin = min(emin, part[w].emin);
      emax = max(emax, part[w].emax);
      flags |= part[w].flags;
      cnt += part[w].valid;
    }
    atomicMin((long long *)&out->imin, (long long)lo);
    atomicMax((long long *)&out->imax, (long long)hi);
    atomicMin(&out->emin, emin);
    atomicMax(&out->emax, emax);
    atomicOr(&out->flags, flags);
    atomicAdd((unsigned long long *)&out->valid, (unsigned long long)cnt);
  }
}

// Statistics of the values in the rows the query's predicate keeps, for the
// columns of `cols` (bit c: plan column c): integer min/max, and the
// exponent range of finite non-zero doubles (flags bit 32 when one was
// seen). Run only when a column's own statistics leave no fixed-point
// width that keeps its Float64 sums within the parity budget (a filtered
// out row may hold a value far larger than every kept one).
template <int NC>
__global__ void __launch_bounds__(kBlock) k_kept_stats(const DevPlan *__restrict__ plan, uint32_t cols,
                                                      ColStats *out) {
  PlanRef P = plan_ref(plan);
  int64_t lo[NC], hi[NC];
  uint32_t emin[NC], emax[NC], fl[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    lo[c] = INT64_MAX;
    hi[c] = INT64_MIN;
    emin[c] = 0x7ff;
    emax[c] = 0;
    fl[c] = 0;
  }
  auto sink = [&](bool kept, int64_t, const uint64_t(&vals)[NC], uint32_t vbits) {
    if (!kept) return;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (!((cols >> c) & 1u) || !((vbits >> c) & 1u)) continue;
      const uint64_t v = vals[c];
      if (P.cols[c].is_float) {
        const uint32_t e = (uint32_t)(v >> 52) & 0x7ffu;
        if (e == 0x7ffu || (v << 1) == 0) continue;
        const uint32_t eb = e ? e : 1u;
        emin[c] = min(emin[c], eb);
        emax[c] = max(emax[c], eb);
        fl[c] |= 32u;
      } else {
        lo[c] = min(lo[c], (int64_t)v);
        hi[c] = max(hi[c], (int64_t)v);
      }
    }
  };
  const int64_t rows = P.rows;
  const int64_t even = rows & ~(int64_t)1;
  const int64_t step = 2 * (int64_t)blockDim.x;
  if (even >= 2) {
    for (int64_t base = (int64_t)blockIdx.x * step; base < even; base += (int64_t)gridDim.x * step) {
      ScanTile<NC, 1> t;
      load_tile<NC, 1>(P, t, base, even);
      fold_tile<NC, 1, true>(P, t, base, sink);
    }
  }
  if ((rows & 1) && blockIdx.x == 0) {
    ScanTile<NC, 1> t;
    load_last_row<NC>(P, t, even);
    fold_tile<NC, 1, true>(P, t, even, sink);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (!((cols >> c) & 1u)) continue;
    for (int d = 32; d >= 1; d >>= 1) {
      lo[c] = min(lo[c], (int64_t)__shfl_xor((long long)lo[c], d, 64));
      hi[c] = max(hi[c], (int64_t)__shfl_xor((long long)hi[c], d, 64));
      emin[c] = min(emin[c], (uint32_t)__shfl_xor((unsigned int)emin[c], d, 64));
      emax[c] = max(emax[c], (uint32_t)__shfl_xor((unsigned int)emax[c], d, 64));
      fl[c] |= (uint32_t)__shfl_xor((unsigned int)fl[c], d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      atomicMin((long long *)&out[c].imin, (long long)lo[c]);
      atomicMax((long long *)&out[c].imax, (long long)hi[c]);
      atomicMin(&out[c].emin, emin[c]);
      atomicMax(&out[c].emax, emax[c]);
      atomicOr(&out[c].flags, fl[c]);
    }
  }
}

#endif  // !DFR_JIT

}  // namespace dfr
