// kernels_part.h — partitioned aggregation: no device-scope atomics in HBM
// on the per-row path.
//
//   pass A  k_scan_spill      fused scan; each workgroup folds kept rows into
//                             a first-come LDS table and spills the rows that
//                             miss it as compact records into its own
//                             segment; at the end the LDS table leaves as
//                             partial-state records. LDS histograms count the
//                             records of each key partition.
//   pass B  k_part_offsets,   per-(segment, partition) offsets (scan over
//           k_part_base,      segments, then over partitions), and a scatter
//           k_part_scatter    of every record into partition order
//   pass C  k_part_aggregate  one workgroup per partition folds its raw
//                             records and merges its partial states in an LDS
//                             table, then stores each group's final state
//                             into the HBM group table with plain stores
//
// Every state word merges associatively and commutatively (integer adds,
// min/max, or), so the result does not depend on which workgroup saw which
// row or in what order.
#pragma once
#include "kernels.h"

namespace dfr {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr int kMaxPartEntries = 16384;  // entries of a pass C LDS table, at most

// A group's output row (keys decoded, aggregates finished), defined below.
template <int NA, class PV>
__device__ __forceinline__ void emit_group(const PV &P, uint64_t i, const uint64_t *slot, uint64_t packed);

__device__ __forceinline__ uint32_t key32(uint64_t key) { return (uint32_t)key ^ (uint32_t)(key >> 32); }

template <class PV>
__device__ __forceinline__ uint32_t part_of(const PV &P, uint64_t key) {
  return P.part_log2 ? ((key32(key) * 0x9E3779B1u) >> (32 - P.part_log2)) : 0u;
}

// LDS index inside a partition's table: a second multiplicative hash
template <class PV>
__device__ __forceinline__ int part_slot(const PV &P, uint64_t key) {
  return (int)((key32(key) * 0x85EBCA77u) >> (32 - P.c_log2));
}

// A 128-bit add in two steps: the low word's atomic, whose returned old
// value add128_finish needs for the carry into the high word. Other updates
// can be issued between the two, so the wave waits for the value once.
__device__ __forceinline__ uint64_t add128_issue(uint64_t *p, uint64_t lo) {
  return atomicAdd((unsigned long long *)p, (unsigned long long)lo);
}
__device__ __forceinline__ void add128_finish(uint64_t *p, uint64_t old, uint64_t lo, uint64_t hi) {
  const uint64_t h = hi + ((old + lo < old) ? 1ull : 0ull);
  if (h) atomicAdd((unsigned long long *)(p + 1), (unsigned long long)h);
}

// Folds one kept row into a full-layout state (every word, FLAGS and ZFIRST
// included). `val(c)` and `ok(c)` give column c's value and validity. Per
// column, the 128-bit sums' low-word atomics are issued first and their
// carries last.
template <int NC = kMaxCols, class Val, class Ok, class PV>
__device__ __forceinline__ void fold_full(const PV &P, uint64_t *st, int64_t row, Val val, Ok ok) {
  atomicAdd((unsigned long long *)&st[0], 1ull);
  if (P.has_first) atomicMin((unsigned long long *)&st[1], (unsigned long long)row);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c >= P.n_cols) continue;
    const DevColAgg A = cget(P.cagg[c]);
    const uint32_t m = A.mask;
    if (!m || !ok(c)) continue;
    uint64_t *b = st + A.gbase;
    const uint64_t v = val(c);
    const uint64_t vhi = ((int64_t)v < 0) ? ~0ull : 0ull;
    uint64_t i_old = 0, f_old = 0, q_old = 0, flo = 0, fhi = 0, qlo = 0, qhi = 0;
    bool finite = false;
    if (m & M_ISUM128) i_old = add128_issue(&b[A.o_isum], v);
    if (m & (M_FSUM | M_FSUMSQ | M_FMIN | M_FMAX | M_FLAGS | M_ZFIRST)) {
      const uint64_t bits = A.int_as_float ? f64_bits((double)(int64_t)v) : v;
      const uint32_t ex = (uint32_t)(bits >> 52) & 0x7ffu;
      bool nan = false;
      if (ex == 0x7ffu) {
        nan = (bits << 12) != 0;
        if (m & M_FLAGS)
          atomicOr((unsigned long long *)&b[A.o_flags],
                   (unsigned long long)(nan ? FL_NAN : ((bits >> 63) ? FL_NINF : FL_PINF)));
      } else {
        finite = true;
        if ((m & M_FSUM) && A.nfw == 2) {
          fix_of(bits, A.fshift, flo, fhi);
          f_old = add128_issue(&b[A.o_fsum], flo);
        }
        if ((m & M_FSUMSQ) && A.nqw == 2) {
          fixsq_of(bits, A.fshift2, qlo, qhi);
          q_old = add128_issue(&b[A.o_sq], qlo);
        }
        if (A.nfw > 2 || A.nqw > 2) add_wide(b, A, bits);
        if ((m & M_ZFIRST) && (bits << 1) == 0)
          atomicMin((unsigned long long *)&b[A.o_zfirst],
                    (unsigned long long)(((uint64_t)row << 1) | (bits >> 63)));
      }
      if (!nan) {
        const uint64_t k = ord_key(bits);
        if (m & M_FMIN) atomicMin((unsigned long long *)&b[A.o_min], (unsigned long long)k);
        if (m & M_FMAX) atomicMax((unsigned long long *)&b[A.o_max], (unsigned long long)k);
      }
    }
    if (m & M_VCOUNT) atomicAdd((unsigned long long *)&b[A.o_count], 1ull);
    if (m & M_ISUM64) atomicAdd((unsigned long long *)&b[A.o_isum], (unsigned long long)v);
    if (m & M_IMIN) atomicMin((long long *)&b[A.o_min], (long long)v);
    if (m & M_IMAX) atomicMax((long long *)&b[A.o_max], (long long)v);
    if (m & M_ISUM128) add128_finish(&b[A.o_isum], i_old, v, vhi);
    if (finite && (m & M_FSUM) && A.nfw == 2) add128_finish(&b[A.o_fsum], f_old, flo, fhi);
    if (finite && (m & M_FSUMSQ) && A.nqw == 2) add128_finish(&b[A.o_sq], q_old, qlo, qhi);
  }
}

// An n-word (3 or 4) fixed-point total of a partial state added into `d`.
// Not inlined: where a plan's word layout is a run-time value, its inlined
// form (a carry chain of returning atomics under a run-time word count)
// crashed the AMDGPU backend (RenameIndependentSubregs) in some shapes.
template <class SrcPtr>
__device__ __noinline__ void merge_wide(uint64_t *d, SrcPtr s, int n) {
  uint64_t x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] = i < n ? (uint64_t)s[i] : 0ull;
  add_words(d, x, n);
}

// Merges a partial state `s` (W words) into `d`, word by word.
template <class SrcPtr, class PV>
__device__ __forceinline__ void merge_word(const PV &P, uint64_t *d, SrcPtr s, int w) {
  {
    const uint64_t v = s[w];
    switch (P.wop[w]) {
      case W_ADD:
        if (v) atomicAdd((unsigned long long *)&d[w], (unsigned long long)v);
        break;
      case W_ADD128:
        add128(&d[w], v, s[w + 1]);
        break;
      case W_ADD192:
      case W_ADD256:
        merge_wide(&d[w], &s[w], P.wop[w] == W_ADD192 ? 3 : 4);
        break;
      case W_MIN_S:
        if ((int64_t)v < (int64_t)d[w]) atomicMin((long long *)&d[w], (long long)v);
        break;
      case W_MAX_S:
        if ((int64_t)v > (int64_t)d[w]) atomicMax((long long *)&d[w], (long long)v);
        break;
      case W_MIN_U:
        if (v < d[w]) atomicMin((unsigned long long *)&d[w], (unsigned long long)v);
        break;
      case W_MAX_U:
        if (v > d[w]) atomicMax((unsigned long long *)&d[w], (unsigned long long)v);
        break;
      case W_OR:
        if (v) atomicOr((unsigned long long *)&d[w], (unsigned long long)v);
        break;
      default:  // W_HI: merged with its low word
        break;
    }
  }
}

template <class SrcPtr, class PV>
__device__ __forceinline__ void merge_full(const PV &P, uint64_t *d, SrcPtr s) {
#pragma unroll
  for (int w = 0; w < (RegPlan<PV>::value ? kMaxWords : 1); ++w) {
    if constexpr (RegPlan<PV>::value) {
      if (w < P.words) merge_word(P, d, s, w);
    } else {
      for (int x = 0; x < P.words; ++x) merge_word(P, d, s, x);
    }
  }
}

// Wave-aggregated append of one record slot per participating lane: returns
// the record index for lanes with `take`, from a shared LDS counter.
__device__ __forceinline__ uint32_t wave_append(bool take, uint32_t *counter) {
  const uint64_t m = __ballot(take);
  const int lane = __lane_id();
  const int rank = __popcll(m & ((1ull << lane) - 1ull));
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (take && lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  if (take) base = (uint32_t)__shfl((int)base, leader);
  return base + (uint32_t)rank;
}

// A row's packed group key (key columns first in `vals`).
template <int NC, class PV>
__device__ __forceinline__ uint64_t pack_key(const PV &P, const uint64_t (&vals)[NC], uint32_t vbits) {
  if (P.n_keys == 1 && P.keys[0].is_float) return (vbits & 1u) ? canon_key_bits(vals[0]) : kNullFloatKey;
  if (P.n_keys == 1) return (vbits & 1u) ? (vals[0] - (uint64_t)P.keys[0].minv) : (P.keys[0].radix - 1);
  uint64_t packed = 0;
#pragma unroll
  for (int k = 0; k < kMaxKeys && k < NC; ++k) {
    if (k >= P.n_keys) continue;
    const DevKey K = cget(P.keys[k]);
    const uint64_t digit = ((vbits >> k) & 1u) ? (vals[k] - (uint64_t)K.minv) : (K.radix - 1);
    packed += digit * K.stride;
  }
  return packed;
}

// The 4-slot bucket of a key in pass A's hashed LDS table of nb buckets.
__host__ __device__ __forceinline__ uint32_t bucket_of(uint64_t key, uint32_t nb) {
  const uint32_t k32 = (uint32_t)key ^ (uint32_t)(key >> 32);
  return (uint32_t)(((uint64_t)(k32 * 0x9E3779B1u) * nb) >> 32);
}

// Scan LDS table: buckets of 4 key slots, read together; a key goes to the
// first free slot of its bucket (first come), and a row whose bucket is
// full spills. One LDS round trip per lookup, hit or miss.
template <int LDSM, class PV>
__device__ __forceinline__ int bucket_find(const PV &P, uint64_t *lkeys, uint64_t key) {
  if (LDSM == LDS_DIRECT) return (int)key;
  const int b = (int)bucket_of(key, (uint32_t)P.lds_buckets);
  uint64_t *kb = lkeys + 4 * b;
  uint64_t k[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) k[j] = kb[j];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (k[j] == key) return 4 * b + j;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (k[j] != kEmpty) continue;
    const unsigned long long old =
        atomicCAS((unsigned long long *)&kb[j], (unsigned long long)kEmpty, (unsigned long long)key);
    if (old == kEmpty || old == key) return 4 * b + j;
  }
  return -1;
}

// One kept-or-not row of pass A: predicate already applied (`kept`); folds
// into the LDS table or spills a record. Every lane of the wave calls it.
template <int LDSM, int NC, class PV>
__device__ __forceinline__ void spill_row(const PV &P, uint64_t *lkeys, uint64_t *lst, uint64_t *seg,
                                          uint32_t &wcur, uint32_t *hraw, bool kept, int64_t row,
                                          const uint64_t (&vals)[NC], uint32_t vbits) {
  uint64_t packed = 0;
  int e = -1;
  if (kept) {
    packed = pack_key<NC>(P, vals, vbits);
    if (LDSM != LDS_NONE) e = bucket_find<LDSM>(P, lkeys, packed);
    if (e >= 0)
      fold_full<NC>(P, lst + (int64_t)e * P.lds_stride, row, [&](int c) { return vals[c]; },
                    [&](int c) { return ((vbits >> c) & 1u) != 0; });
  }
  // the wave appends to its own segment: cursor uniform across the wave
  const bool spill = kept && e < 0;
  const uint64_t sm = __ballot(spill);
  const uint32_t at = wcur + (uint32_t)__popcll(sm & ((1ull << __lane_id()) - 1ull));
  wcur += (uint32_t)__popcll(sm);
  if (spill) {
    DFR_GLOBAL uint64_t *rec = gpw(seg) + (uint64_t)at * P.rw;
    uint64_t vmask = 0, narrow = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c >= P.n_cols || P.val_of[c] < 0) continue;
      vmask |= (uint64_t)((vbits >> c) & 1u) << P.val_of[c];
      if (P.val_word[c] >= 0)
        rec[P.val_word[c]] = vals[c];
      else
        narrow |= ((vals[c] - (uint64_t)P.pk_base[c]) & ((1ull << P.pk_bits[c]) - 1)) << P.pk_shift[c];
    }
    rec[0] = P.vmask_in_key ? (packed | (vmask << P.key_bits) | narrow) : packed;
    if (P.rec_meta >= 0) rec[P.rec_meta] = vmask;
    if (P.rec_row >= 0) rec[P.rec_row] = (uint64_t)row;
    atomicAdd(&hraw[part_of(P, packed)], 1u);
  }
}

// Pass A. Workgroup s folds rows [s * chunk, (s + 1) * chunk); NC bounds
// the number of input columns, U the row pairs a thread loads at once, PF
// whether the next tile's loads overlap this tile's fold, TL the loads one
// tile issues (when known at compile time, else 0). PV is the plan's
// type: the plan in constant memory, or (runtime-compiled) a register copy
// whose fields of the query's shape are compile-time values.
template <int LDSM, int NC, int U, bool PF, int TL = 0, class PV>
__device__ __forceinline__ void scan_spill_body(const PV &P, int64_t chunk) {
  extern __shared__ uint64_t smem[];
  const int E = P.lds_entries, W = P.words, NP = P.n_parts;
  uint64_t *lkeys = smem;
  uint64_t *lst = smem + (LDSM == LDS_HASHED ? E : 0);
  const int S = P.lds_stride;
  // per-partition histograms: raw records during the scan, one per slice
  // of the workgroup's waves; then (once written out) the first one counts
  // the flushed partial states
  const int H = P.seg_slices;
  uint32_t *hraw = (uint32_t *)(lst + (LDSM == LDS_NONE ? 0 : (size_t)E * S));
  uint32_t *hpar = hraw;
  uint32_t *cnt = hraw + H * NP;  // [0] raw records, [1] partial records
  // (hashed table: the seeded hot keys, when there are, at their entries)
  const DFR_GLOBAL uint64_t *image = LDSM == LDS_HASHED && P.hot_image ? gp(P.hot_image) : nullptr;
  if (LDSM != LDS_NONE) {
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
      if (LDSM == LDS_HASHED) lkeys[i] = image ? image[i] : kEmpty;
      if constexpr (RegPlan<PV>::value) {
#pragma unroll
        for (int w = 0; w < kMaxWords; ++w)
          if (w < W) lst[i * S + w] = P.ident[w];
      } else {
        for (int w = 0; w < W; ++w) lst[i * S + w] = P.ident[w];
      }
    }
  }
  for (int i = threadIdx.x; i < H * NP + 2; i += blockDim.x) hraw[i] = 0;
  __syncthreads();
  const int s = blockIdx.x;
  // wave segment: the rows a wave sees, seg_rows records at most
  const int wave = threadIdx.x >> 6;
  const uint64_t wseg = (uint64_t)s * (blockDim.x >> 6) + wave;
  uint32_t *hw = hraw + (wave / ((blockDim.x >> 6) / H)) * NP;  // this wave's slice
  uint64_t *seg = P.spill + wseg * (uint64_t)P.seg_rows * P.rw;
  uint32_t wcur = 0;
  auto spill = [&](bool kept, int64_t row, const uint64_t(&vals)[NC], uint32_t vbits) {
    spill_row<LDSM, NC>(P, lkeys, lst, seg, wcur, hw, kept, row, vals, vbits);
  };
  const int64_t b0 = (int64_t)s * chunk;
  const int64_t b1 = min(b0 + chunk, P.rows);
  const int64_t step = 2 * (int64_t)blockDim.x * U;
  // whole row pairs through buffer loads; an odd range's last row after
  const int64_t b1e = b0 + (max(b1 - b0, (int64_t)0) & ~(int64_t)1);
  if constexpr (PF) {
    // software pipeline: the next tile's loads are issued before this
    // tile is folded, so they are in flight during the fold. Unrolled by
    // two so the tiles swap roles without register copies (a copy would
    // wait for the loads it copies).
    ScanTile<NC, U> ta, tb;
    if (b0 < b1e) load_tile<NC, U>(P, ta, b0, b1e);
    // (a tile past the range loads the range's last pair: in bounds)
    for (int64_t base = b0; base < b1e;) {
      load_tile<NC, U>(P, tb, base + step, b1e);
      fold_tile<NC, U, false, TL>(P, ta, base, spill);
      base += step;
      if (base >= b1e) break;
      load_tile<NC, U>(P, ta, base + step, b1e);
      fold_tile<NC, U, false, TL>(P, tb, base, spill);
      base += step;
    }
  } else {
    for (int64_t base = b0; base < b1e; base += step) {
      ScanTile<NC, U> t;
      load_tile<NC, U>(P, t, base, b1e);
      fold_tile<NC, U, true>(P, t, base, spill);
    }
  }
  if (b1e < b1) {
    ScanTile<NC, 1> t;
    load_last_row<NC>(P, t, b1e);
    fold_tile<NC, 1, true>(P, t, b1e, spill);
  }
  if ((threadIdx.x & 63) == 0) P.cnt_raw[wseg] = wcur;
  __syncthreads();
  for (int i = threadIdx.x; i < H * NP; i += blockDim.x) {
    P.hist_raw[(uint64_t)s * H * NP + i] = hraw[i];
    if (i < NP) hpar[i] = 0;
  }
  __syncthreads();
  if (LDSM != LDS_NONE) {
    uint64_t *pseg = P.partials + (uint64_t)s * (uint64_t)P.seg_partials * (1 + W);
    const int Eround = (E + 63) & ~63;
    for (int e = threadIdx.x; e < Eround; e += blockDim.x) {
      const uint64_t *st = lst + (int64_t)e * S;
      // a seeded entry's state goes to the dense hot states, whatever it holds
      const bool seeded = image && e < E && image[e] != kEmpty;
      if (seeded) {
        DFR_GLOBAL uint64_t *hs = gpw(P.hot_states) + ((uint64_t)s * E + e) * W;
        for (int w = 0; w < W; ++w) hs[w] = st[w];
      }
      const bool occ = e < E && !seeded && st[0] != 0;
      const uint64_t key = (LDSM == LDS_DIRECT) ? (uint64_t)e : (e < E ? lkeys[e] : 0);
      const uint32_t at = wave_append(occ, &cnt[1]);
      if (occ) {
        DFR_GLOBAL uint64_t *rec = gpw(pseg) + (uint64_t)at * (1 + W);
        rec[0] = key;
        for (int w = 0; w < W; ++w) rec[1 + w] = st[w];
        atomicAdd(&hpar[part_of(P, key)], 1u);
      }
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < NP; p += blockDim.x) P.hist_par[(uint64_t)s * NP + p] = hpar[p];
  if (threadIdx.x == 0) P.cnt_par[s] = cnt[1];
}

// The precompiled pass A runs unpipelined (it serves inputs below the
// runtime-compiled kernel's row threshold, and the pipeline doubles its code
// for every instantiation).
constexpr bool kScanPrefetch = false;

template <int LDSM, int NC, int U>
__global__ void __launch_bounds__(1024) k_scan_spill(const DevPlan *__restrict__ plan, int64_t chunk) {
  scan_spill_body<LDSM, NC, U, kScanPrefetch>(plan_ref(plan), chunk);
}

#if !DFR_JIT
// ---------------------------------------------------------- hot keys
//
// Heavy-hitter keys, found from a sample of the kept rows and seeded into
// every pass A workgroup's LDS table at the same entries. On skewed keys
// (Zipf) they carry most rows: seeded, those rows never spill, and their
// states leave pass A as dense per-workgroup arrays merged by one coalesced
// reduction instead of as partial-state records that pass C would fold into
// the same LDS entry from every workgroup. Any seeding gives the same
// result: the seeded keys only decide which rows take the LDS table.

constexpr int kHotSampleThreads = 256;
constexpr int kHotCountEntries = 2048;  // per-block LDS pre-count table

// Block b counts the packed keys (counts[key], key < domain) of the kept
// rows in rows [b * stride, b * stride + span), through a small LDS table so
// that the hottest keys take one device atomic per block.
template <int NC>
__global__ void __launch_bounds__(kHotSampleThreads) k_hot_count(const DevPlan *__restrict__ plan, uint32_t *counts,
                                                                 uint64_t domain, int64_t stride, int64_t span) {
  __shared__ uint64_t tkeys[kHotCountEntries];
  __shared__ uint32_t tcnt[kHotCountEntries];
  PlanRef P = plan_ref(plan);
  for (int i = threadIdx.x; i < kHotCountEntries; i += blockDim.x) {
    tkeys[i] = kEmpty;
    tcnt[i] = 0;
  }
  __syncthreads();
  auto sink = [&](bool kept, int64_t, const uint64_t(&vals)[NC], uint32_t vbits) {
    if (!kept) return;
    const uint64_t key = pack_key<NC>(P, vals, vbits);
    if (key >= domain) return;  // (every packed key of the column ranges is below it)
    int h = (int)((key * kGolden) >> (64 - 11));
    for (int p = 0; p < 4; ++p, h = (h + 1) & (kHotCountEntries - 1)) {
      const uint64_t k = tkeys[h];
      if (k == key) {
        atomicAdd(&tcnt[h], 1u);
        return;
      }
      if (k == kEmpty) {
        const unsigned long long old =
            atomicCAS((unsigned long long *)&tkeys[h], (unsigned long long)kEmpty, (unsigned long long)key);
        if (old == kEmpty || old == key) {
          atomicAdd(&tcnt[h], 1u);
          return;
        }
      }
    }
    atomicAdd(&counts[key], 1u);
  };
  const int64_t b0 = (int64_t)blockIdx.x * stride;
  const int64_t b1 = min(b0 + span, P.rows);
  const int64_t even = b0 + (max(b1 - b0, (int64_t)0) & ~(int64_t)1);
  for (int64_t base = b0; base + 1 < even; base += 2 * (int64_t)blockDim.x) {
    ScanTile<NC, 1> t;
    load_tile<NC, 1>(P, t, base, even);
    fold_tile<NC, 1, true>(P, t, base, sink);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHotCountEntries; i += blockDim.x)
    if (tkeys[i] != kEmpty && tcnt[i]) atomicAdd(&counts[tkeys[i]], tcnt[i]);
}

// The column hand-off's sample: block b counts the valid values v of rows
// [b * stride, b * stride + span) of one integer or dictionary column as
// counts[v - minv] (v - minv < domain), through the same LDS pre-count.
__global__ void __launch_bounds__(kHotSampleThreads) k_col_hot_count(DevCol C, int64_t n, int64_t minv,
                                                                     uint64_t domain, uint32_t *counts,
                                                                     int64_t stride, int64_t span) {
  __shared__ uint64_t tkeys[kHotCountEntries];
  __shared__ uint32_t tcnt[kHotCountEntries];
  for (int i = threadIdx.x; i < kHotCountEntries; i += blockDim.x) {
    tkeys[i] = kEmpty;
    tcnt[i] = 0;
  }
  __syncthreads();
  const int64_t b0 = (int64_t)blockIdx.x * stride;
  const int64_t b1 = min(b0 + span, n);
  for (int64_t r = b0 + threadIdx.x; r < b1; r += blockDim.x) {
    if (C.valid && !((C.valid[(C.vbit + r) >> 6] >> ((C.vbit + r) & 63)) & 1ull)) continue;
    const int64_t v = C.dtype == T_INT64 ? ((const int64_t *)C.values)[r]
                      : C.dtype == T_INT32 ? (int64_t)((const int32_t *)C.values)[r]
                                           : (int64_t)((const uint32_t *)C.values)[r];
    const uint64_t key = (uint64_t)v - (uint64_t)minv;
    if (key >= domain) continue;
    int h = (int)((key * kGolden) >> (64 - 11));
    bool done = false;
    for (int p = 0; p < 4 && !done; ++p, h = (h + 1) & (kHotCountEntries - 1)) {
      const uint64_t k = tkeys[h];
      if (k == key) {
        atomicAdd(&tcnt[h], 1u);
        done = true;
      } else if (k == kEmpty) {
        const unsigned long long old =
            atomicCAS((unsigned long long *)&tkeys[h], (unsigned long long)kEmpty, (unsigned long long)key);
        if (old == kEmpty || old == key) {
          atomicAdd(&tcnt[h], 1u);
          done = true;
        }
      }
    }
    if (!done) atomicAdd(&counts[key], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHotCountEntries; i += blockDim.x)
    if (tkeys[i] != kEmpty && tcnt[i]) atomicAdd(&counts[tkeys[i]], tcnt[i]);
}

constexpr int kHotBins = 4096;   // sample counts binned 0..4095 (the last bin: 4095 and more)
constexpr int kHotList = 4096;   // candidates, at most
constexpr int kHotSelectThreads = 1024;

// hist[min(c, kHotBins - 1)] += 1 for every sampled count c >= 2 of the key
// domain (a key sampled once is no candidate), and hist[kHotBins] (two
// words) += every c: the kept rows sampled. Small counts, which most keys
// share, are added once per wave and value (a ballot per distinct value),
// not once per lane on one LDS word.
__global__ void k_hot_hist(const uint32_t *counts, uint64_t domain, uint32_t *hist) {
  __shared__ uint32_t h[kHotBins];
  __shared__ unsigned long long total;
  for (int i = threadIdx.x; i < kHotBins; i += blockDim.x) h[i] = 0;
  if (threadIdx.x == 0) total = 0;
  __syncthreads();
  unsigned long long mine = 0;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t first = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  // every lane of a wave runs the same number of rounds (ballots below)
  const uint64_t rounds = (domain + step - 1) / step;
  for (uint64_t r = 0; r < rounds; ++r) {
    const uint64_t k = first + r * step;
    const uint32_t c = k < domain ? counts[k] : 0;
    mine += c;
    bool small = c >= 2 && c < 64;
    if (c >= 64) atomicAdd(&h[min(c, (uint32_t)kHotBins - 1)], 1u);
    uint64_t pend = __ballot(small);
    while (pend) {
      const int leader = __ffsll((unsigned long long)pend) - 1;
      const uint32_t v = (uint32_t)__shfl((int)c, leader);
      const uint64_t m = __ballot(small && c == v);
      if (__lane_id() == leader) atomicAdd(&h[v], (uint32_t)__popcll(m));
      pend &= ~m;
    }
  }
  atomicAdd(&total, mine);
  __syncthreads();
  for (int i = threadIdx.x; i < kHotBins; i += blockDim.x)
    if (h[i]) atomicAdd(&hist[i], h[i]);
  if (threadIdx.x == 0) atomicAdd((unsigned long long *)(hist + kHotBins), total);
}

// The smallest count t >= 2 whose keys (count >= t) number `want` at most,
// from hist; every thread of the (kHotSelectThreads) workgroup returns it.
__device__ uint32_t hot_threshold(const uint32_t *hist, uint32_t want, uint32_t *part) {
  const int per = kHotBins / kHotSelectThreads;
  const int t = threadIdx.x;
  uint32_t own = 0;
  for (int j = 0; j < per; ++j) own += hist[t * per + j];
  part[t] = own;
  __syncthreads();
  // suffix sums: part[t] = keys in bins >= t * per
  for (int d = 1; d < kHotSelectThreads; d <<= 1) {
    const uint32_t add = t + d < kHotSelectThreads ? part[t + d] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  __shared__ uint32_t best;
  if (t == 0) best = kHotBins;
  __syncthreads();
  uint32_t above = t + 1 < kHotSelectThreads ? part[t + 1] : 0;
  for (int j = per - 1; j >= 0; --j) {
    above += hist[t * per + j];
    if (above <= want) atomicMin(&best, (uint32_t)(t * per + j));
  }
  __syncthreads();
  return max(best, 2u);
}

// Keys whose sampled count reaches the threshold, appended to list as
// (count << 32 | key) (keys below 2^32); the 32-bit counter at
// list[kHotList] counts the appended (one atomic per wave).
__global__ void __launch_bounds__(kHotSelectThreads) k_hot_collect(const uint32_t *counts, uint64_t domain,
                                                                   const uint32_t *hist, uint32_t want,
                                                                   uint64_t *list) {
  __shared__ uint32_t part[kHotSelectThreads];
  const uint32_t thresh = hot_threshold(hist, want, part);
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < domain; k += step) {
    const uint32_t c = counts[k];
    const bool take = c >= thresh;
    const uint32_t at = wave_append(take, (uint32_t *)&list[kHotList]);
    if (take && at < (uint32_t)kHotList) list[at] = ((uint64_t)c << 32) | (uint32_t)k;
  }
}

// One workgroup: the candidates placed into an image of pass A's bucket
// table (`entries` keys, entries / 4 buckets of 4), in bands of descending
// sampled count (band b: counts in [2^b, 2^(b+1))) so that a full bucket
// keeps the hotter keys; a candidate takes the next free slot of its bucket,
// or none. A candidate's packed key is its listed key + delta, and only
// packed keys in [0, limit) are placed (a column's list, whose keys are
// offsets from the column's minimum, for a query whose key range is a
// window of it). The image stays empty when the placed keys carry less than
// min_share of the `total` sampled rows (no heavy hitters: seeding would not
// pay). hot_out[0] = keys placed.
__global__ void __launch_bounds__(kHotSelectThreads) k_hot_select(const uint64_t *list,
                                                                  const unsigned long long *total_rows,
                                                                  int64_t delta, uint64_t limit, uint64_t *image,
                                                                  int entries, float min_share, uint32_t *hot_out) {
  constexpr int PER = kHotList / kHotSelectThreads;
  __shared__ uint32_t fill[4096];  // per bucket (entries <= 16384)
  __shared__ unsigned long long covered;
  __shared__ uint32_t placed, top;
  const uint32_t n = min(*(const uint32_t *)&list[kHotList], (uint32_t)kHotList);
  const int nb = entries / 4;
  uint64_t mine[PER];
  int band[PER];
  for (int j = 0; j < PER; ++j) {
    const uint32_t i = threadIdx.x + j * kHotSelectThreads;
    mine[j] = i < n ? list[i] : 0;
    const uint64_t key = (uint64_t)((int64_t)(uint32_t)mine[j] + delta);
    band[j] = (i < n && key < limit) ? 31 - __clz((uint32_t)(mine[j] >> 32)) : -1;
  }
  for (int i = threadIdx.x; i < nb; i += blockDim.x) fill[i] = 0;
  for (int i = threadIdx.x; i < entries; i += blockDim.x) image[i] = kEmpty;
  if (threadIdx.x == 0) {
    covered = 0;
    placed = 0;
    top = 0;
  }
  __syncthreads();
  for (int j = 0; j < PER; ++j)
    if (band[j] >= 0) atomicMax(&top, (uint32_t)band[j]);
  __syncthreads();
  for (int b = (int)top; b >= 1; --b) {
    for (int j = 0; j < PER; ++j) {
      if (band[j] != b) continue;
      const uint64_t key = (uint64_t)((int64_t)(uint32_t)mine[j] + delta);
      const uint32_t bk = bucket_of(key, (uint32_t)nb);
      const uint32_t slot = atomicAdd(&fill[bk], 1u);
      if (slot < 4) {
        image[4 * bk + slot] = key;
        atomicAdd(&covered, (unsigned long long)(mine[j] >> 32));
        atomicAdd(&placed, 1u);
      }
    }
    __syncthreads();
  }
  const unsigned long long total = *total_rows;
  const bool keep = covered > 0 && (double)covered >= (double)min_share * (double)total;
  if (!keep)
    for (int i = threadIdx.x; i < entries; i += blockDim.x) image[i] = kEmpty;
  if (threadIdx.x == 0) hot_out[0] = keep ? placed : 0;
}

// One word of a merge, by its operation (an n-word add: all n words).
__device__ __forceinline__ void hot_merge_word(int8_t op, int n, uint64_t (&acc)[4], const uint64_t (&x)[4]) {
  switch (op) {
    case W_ADD: acc[0] += x[0]; break;
    case W_MIN_S: acc[0] = (int64_t)x[0] < (int64_t)acc[0] ? x[0] : acc[0]; break;
    case W_MAX_S: acc[0] = (int64_t)x[0] > (int64_t)acc[0] ? x[0] : acc[0]; break;
    case W_MIN_U: acc[0] = x[0] < acc[0] ? x[0] : acc[0]; break;
    case W_MAX_U: acc[0] = x[0] > acc[0] ? x[0] : acc[0]; break;
    case W_OR: acc[0] |= x[0]; break;
    default: {  // an n-word two's-complement add
      uint64_t carry = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i >= n) break;
        const uint64_t t = acc[i] + x[i];
        const uint64_t c1 = t < x[i] ? 1ull : 0ull;
        const uint64_t t2 = t + carry;
        const uint64_t c2 = t2 < t ? 1ull : 0ull;
        acc[i] = t2;
        carry = c1 | c2;
      }
    }
  }
}

// Words of the merge a state word leads: 2 to 4 for a multi-word add, else 1.
__device__ __forceinline__ int hot_words(int8_t op) {
  return op == W_ADD128 ? 2 : op == W_ADD192 ? 3 : op == W_ADD256 ? 4 : 1;
}

// The seeded entries' states merged over pass A's workgroups in two steps.
// Step 0, thread (group g, entry e, word w): workgroups [g * n / groups,
// (g + 1) * n / groups) of hot_states ([workgroup][entry][word]) into the
// first of them. Step 1, thread (e, w): the groups' results into the group
// table at the key's slot (plain stores: seeded keys never reach pass C),
// when the key has kept rows. A multi-word add is merged by its low word's
// thread; the others return. Loads are issued 8 at a time.
__global__ void k_hot_reduce(const DevPlan *__restrict__ plan, int groups, int step) {
  PlanRef P = plan_ref(plan);
  const int W = P.words, E = P.lds_entries, NS = P.n_segs;
  const uint64_t ew = (uint64_t)E * W;
  const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ew * (step == 0 ? (uint64_t)groups : 1)) return;
  const int g = (int)(idx / ew);
  const int e = (int)((idx % ew) / W), w = (int)(idx % W);
  const uint64_t key = P.hot_image[e];
  const int8_t op = P.wop[w];
  if (key == kEmpty || op == W_HI) return;
  const int n = hot_words(op);
  uint64_t acc[4];
  for (int i = 0; i < 4; ++i) acc[i] = i < n ? P.ident[w + i] : 0;
  const DFR_GLOBAL uint64_t *src = gp(P.hot_states) + (uint64_t)e * W + w;
  // the workgroups merged: a range (step 0) or the groups' first ones (step 1)
  int s0 = 0, cnt = groups;
  if (step == 0) {
    s0 = (int)((int64_t)g * NS / groups);
    cnt = (int)((int64_t)(g + 1) * NS / groups) - s0;
    if (cnt <= 0) return;
  }
  auto seg_of = [&](int i) -> int { return step == 0 ? s0 + i : (int)((int64_t)i * NS / groups); };
  auto live = [&](int i) -> bool {
    return step == 0 || (int)((int64_t)(i + 1) * NS / groups) > (int)((int64_t)i * NS / groups);
  };
  uint64_t rows = 0;  // (step 1) the key's kept rows: word 0 of each group
  constexpr int B = 8;
  for (int i0 = 0; i0 < cnt; i0 += B) {
    uint64_t x[B][4];
    uint64_t c0[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int i = i0 + j;
      const bool ok = i < cnt && live(i);
      const DFR_GLOBAL uint64_t *v = src + (uint64_t)(ok ? seg_of(i) : 0) * ew;
#pragma unroll
      for (int q = 0; q < 4; ++q) x[j][q] = (ok && q < n) ? v[q] : 0;
      c0[j] = (ok && step == 1) ? v[-w] : 0;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int i = i0 + j;
      if (!(i < cnt && live(i))) continue;
      hot_merge_word(op, n, acc, x[j]);
      rows += c0[j];
    }
  }
  if (step == 0) {
    DFR_GLOBAL uint64_t *d = gpw(P.hot_states) + (uint64_t)s0 * ew + (uint64_t)e * W + w;
    for (int i = 0; i < n; ++i) d[i] = acc[i];
    return;
  }
  if (P.fused) {
    // the merged state, whole, into the spare segment NS (k_hot_emit)
    DFR_GLOBAL uint64_t *d = gpw(P.hot_states) + (uint64_t)NS * ew + (uint64_t)e * W + w;
    for (int i = 0; i < n; ++i) d[i] = acc[i];
    return;
  }
  if (rows == 0) return;  // no kept row of the key (a group only exists with one)
  const uint64_t slot = global_slot(P, key);
  if (slot == kEmpty) return;
  uint64_t *d = P.state + slot * (uint64_t)W + w;
  for (int i = 0; i < n; ++i) d[i] = acc[i];
}

// A fused query's seeded keys with kept rows (their merged states in
// hot_states segment n_segs) at consecutive dense slots: one reservation
// per wave.
__global__ void k_hot_emit(const DevPlan *__restrict__ plan) {
  PlanRef P = plan_ref(plan);
  const int W = P.words, E = P.lds_entries;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t *st = P.hot_states + ((uint64_t)P.n_segs * E + (uint64_t)(e < E ? e : 0)) * W;
  const uint64_t key = e < E ? P.hot_image[e] : kEmpty;
  const bool live = key != kEmpty && st[0] != 0;  // word 0: the group's kept rows
  const uint64_t ball = __ballot(live);
  const int lane = threadIdx.x & 63;
  uint64_t base = 0;
  if (lane == 0 && ball) base = atomicAdd((unsigned long long *)P.gcount, (unsigned long long)__popcll(ball));
  base = __shfl(base, 0, 64);
  if (!live) return;
  const uint64_t slot = base + (uint64_t)__popcll(ball & ((1ull << lane) - 1ull));
  P.keys_tab[slot] = key;
  for (int w = 0; w < W; ++w) P.state[slot * (uint64_t)W + w] = st[w];
}

// Pass B, step 1: per partition p, exclusive scan over segments of the
// histograms (in place), partition totals into base_*[p].
__global__ void k_part_offsets(const DevPlan *__restrict__ plan) {
  PlanRef P = plan_ref(plan);
  const int p = blockIdx.x, NP = P.n_parts;
  for (int which = 0; which < 2; ++which) {
    // raw records: a histogram per workgroup slice; partial states: per workgroup
    const int S = which ? P.n_segs : P.n_segs * P.seg_slices;
    uint32_t *h = which ? P.hist_par : P.hist_raw;
    uint64_t carry = 0;
    for (int s0 = 0; s0 < S; s0 += kBlock) {
      const int s = s0 + threadIdx.x;
      const uint64_t x = s < S ? h[(uint64_t)s * NP + p] : 0;
      uint64_t tot;
      const uint64_t ex = block_exclusive_scan(x, &tot);
      if (s < S) h[(uint64_t)s * NP + p] = (uint32_t)(carry + ex);
      carry += tot;
    }
    if (threadIdx.x == 0) (which ? P.base_par : P.base_raw)[p] = carry;
  }
}

// Pass B, step 2 (one workgroup): partition starts from the totals.
__global__ void k_part_base(const DevPlan *__restrict__ plan) {
  PlanRef P = plan_ref(plan);
  const int NP = P.n_parts;
  for (int which = 0; which < 2; ++which) {
    uint64_t *b = which ? P.base_par : P.base_raw;
    uint64_t carry = 0;
    for (int p0 = 0; p0 < NP; p0 += kBlock) {
      const int p = p0 + threadIdx.x;
      const uint64_t x = p < NP ? b[p] : 0;
      uint64_t tot;
      const uint64_t ex = block_exclusive_scan(x, &tot);
      if (p < NP) b[p] = carry + ex;
      carry += tot;
    }
    if (threadIdx.x == 0) b[NP] = carry;
  }
}

// Pass B, step 3: segment s moves its records into partition order.
//
// Records are taken kScatterChunk at a time, counting-sorted by partition in
// LDS, and written out in that order: each partition then receives runs of
// consecutive records instead of isolated ones, so HBM sees whole lines
// rather than partial writes. Every write position is the partition's
// cursor for this segment plus the record's rank, so the result is the same
// as moving the records one by one.
constexpr int kScatterChunk = 16384;  // records sorted per round (chunk index fits 16 bits)
constexpr int kScatterThreads = 1024;

// Exclusive scan of a[0..n) in place by a 1024-thread block; a[n] = total.
__device__ __forceinline__ void block_scan_u32(uint32_t *a, int n, uint32_t *wsum) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int per = (n + kScatterThreads - 1) / kScatterThreads;
  const int t0 = t * per;
  uint32_t sum = 0;
  for (int j = 0; j < per; ++j)
    if (t0 + j < n) sum += a[t0 + j];
  uint32_t inc = sum;
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t off = 0, total = 0;
  for (int w = 0; w < kScatterThreads / 64; ++w) {
    if (w < wave) off += wsum[w];
    total += wsum[w];
  }
  uint32_t ex = off + inc - sum;
  for (int j = 0; j < per; ++j)
    if (t0 + j < n) {
      const uint32_t v = a[t0 + j];
      a[t0 + j] = ex;
      ex += v;
    }
  if (t == 0) a[n] = total;
  __syncthreads();
}

// N record words moved as 16-byte pieces (records are 8-byte aligned;
// global loads and stores of 16 bytes need no more).
typedef uint64_t u64x2a8 __attribute__((ext_vector_type(2), aligned(8)));

template <int N>
__device__ __forceinline__ void load_words(const DFR_GLOBAL uint64_t *p, uint64_t *w) {
#pragma unroll
  for (int k = 0; k + 1 < N; k += 2) {
    const u64x2a8 q = *(const DFR_GLOBAL u64x2a8 *)(p + k);
    w[k] = q.x;
    w[k + 1] = q.y;
  }
  if constexpr (N % 2) w[N - 1] = p[N - 1];
}

template <int N>
__device__ __forceinline__ void store_words(DFR_GLOBAL uint64_t *p, const uint64_t *w) {
#pragma unroll
  for (int k = 0; k + 1 < N; k += 2) {
    u64x2a8 q;
    q.x = w[k];
    q.y = w[k + 1];
    *(DFR_GLOBAL u64x2a8 *)(p + k) = q;
  }
  if constexpr (N % 2) p[N - 1] = w[N - 1];
}

// Moves n records of rw words from `src` to dst[cur[b] ...] in bucket order
// (bucket_of(record word 0) < nb), CH records a round, advancing the
// cursors. LDS: cnt[nb + 1], idx[CH], wsum[16]. EXACT: rw == RW, so each
// record moves in 16-byte pieces.
template <int RW, int CH, bool EXACT, class BucketOf>
__device__ __forceinline__ void scatter_sorted(const DFR_GLOBAL uint64_t *src, uint32_t n, int rw, int nb,
                                               BucketOf bucket_of, unsigned long long *cur, uint32_t *cnt,
                                               uint32_t *idx, uint32_t *wsum, DFR_GLOBAL uint64_t *dst) {
  static_assert(CH <= 65536 && CH % kScatterThreads == 0, "chunk indices fit 16 bits");
  constexpr int IT = CH / kScatterThreads;
  for (uint32_t c0 = 0; c0 < n; c0 += CH) {
    const uint32_t m = min((uint32_t)CH, n - c0);
    for (int b = threadIdx.x; b <= nb; b += kScatterThreads) cnt[b] = 0;
    __syncthreads();
    uint32_t br[IT];  // bucket << 16 | rank within the bucket
    uint64_t w0[IT];  // first record words, all loaded before any is used
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = min((uint32_t)(it * kScatterThreads + threadIdx.x), m - 1);
      w0[it] = src[(uint64_t)(c0 + i) * rw];
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = it * kScatterThreads + threadIdx.x;
      br[it] = 0;
      if (i < m) {
        const uint32_t b = bucket_of(w0[it]);
        br[it] = (b << 16) | atomicAdd(&cnt[b], 1u);
      }
    }
    __syncthreads();
    block_scan_u32(cnt, nb, wsum);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = it * kScatterThreads + threadIdx.x;
      if (i < m) {
        const uint32_t b = br[it] >> 16;
        idx[cnt[b] + (br[it] & 0xffffu)] = (b << 16) | i;
      }
    }
    __syncthreads();
    if constexpr (RW <= 16) {
      // G records per thread at once: every load before the stores
      constexpr int G = RW <= 4 ? 4 : (RW <= 8 ? 2 : 1);
      for (uint32_t s0 = 0; s0 < m; s0 += G * kScatterThreads) {
        uint64_t w[G][RW];
        uint64_t pos[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const uint32_t slot = min(s0 + g * kScatterThreads + threadIdx.x, m - 1);
          const uint32_t v = idx[slot];
          const uint32_t b = v >> 16, i = v & 0xffffu;
          pos[g] = cur[b] + (slot - cnt[b]);
          const DFR_GLOBAL uint64_t *r = src + (uint64_t)(c0 + i) * rw;
          if constexpr (EXACT) {
            load_words<RW>(r, w[g]);
          } else {
#pragma unroll
            for (int k = 0; k < RW; ++k) w[g][k] = k < rw ? r[k] : 0;
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (s0 + g * kScatterThreads + threadIdx.x >= m) continue;
          DFR_GLOBAL uint64_t *d = dst + pos[g] * rw;
          if constexpr (EXACT) {
            store_words<RW>(d, w[g]);
          } else {
#pragma unroll
            for (int k = 0; k < RW; ++k)
              if (k < rw) d[k] = w[g][k];
          }
        }
      }
    } else {
      for (uint32_t slot = threadIdx.x; slot < m; slot += kScatterThreads) {
        const uint32_t v = idx[slot];
        const uint32_t b = v >> 16, i = v & 0xffffu;
        const uint64_t pos = cur[b] + (slot - cnt[b]);
        const DFR_GLOBAL uint64_t *r = src + (uint64_t)(c0 + i) * rw;
        DFR_GLOBAL uint64_t *d = dst + pos * rw;
        for (int k = 0; k < rw; ++k) d[k] = r[k];
      }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kScatterThreads) cur[b] += cnt[b + 1] - cnt[b];
    __syncthreads();
  }
}

// Moves n records of exactly RW words from `src` to dst[cur[b] ...] in
// bucket order, IT * 1024 records a round. Each thread loads its records
// whole (consecutive records across the lanes, so every line is read once,
// coalesced); the workgroup counts and ranks them by bucket, places them in
// LDS at their sorted positions, then streams them out in that order, so
// nothing is read twice. LDS: cnt[nb + 1], stage[IT * 1024 * RW], wsum[16].
template <int RW, int IT, class BucketOf>
__device__ __forceinline__ void scatter_staged(const DFR_GLOBAL uint64_t *src, uint32_t n, int nb, BucketOf bucket_of,
                                               unsigned long long *cur, uint32_t *cnt, uint64_t *stage,
                                               uint32_t *wsum, DFR_GLOBAL uint64_t *dst) {
  constexpr int CH = IT * kScatterThreads;
  static_assert(CH <= 65536, "ranks fit 16 bits");
  for (uint32_t c0 = 0; c0 < n; c0 += CH) {
    const uint32_t m = min((uint32_t)CH, n - c0);
    for (int b = threadIdx.x; b <= nb; b += kScatterThreads) cnt[b] = 0;
    uint64_t w[IT][RW];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = min((uint32_t)(it * kScatterThreads + threadIdx.x), m - 1);
      load_words<RW>(src + (uint64_t)(c0 + i) * RW, w[it]);
    }
    __syncthreads();
    uint32_t br[IT];  // bucket << 16 | rank within the bucket
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = it * kScatterThreads + threadIdx.x;
      br[it] = 0;
      if (i < m) {
        const uint32_t b = bucket_of(w[it][0]);
        br[it] = (b << 16) | atomicAdd(&cnt[b], 1u);
      }
    }
    __syncthreads();
    block_scan_u32(cnt, nb, wsum);
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const uint32_t i = it * kScatterThreads + threadIdx.x;
      if (i < m) {
        uint64_t *d = stage + (uint64_t)(cnt[br[it] >> 16] + (br[it] & 0xffffu)) * RW;
#pragma unroll
        for (int k = 0; k < RW; ++k) d[k] = w[it][k];
      }
    }
    __syncthreads();
    for (uint32_t slot = threadIdx.x; slot < m; slot += kScatterThreads) {
      uint64_t x[RW];
#pragma unroll
      for (int k = 0; k < RW; ++k) x[k] = stage[(uint64_t)slot * RW + k];
      const uint32_t b = bucket_of(x[0]);
      store_words<RW>(dst + (cur[b] + (slot - cnt[b])) * RW, x);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += kScatterThreads) cur[b] += cnt[b + 1] - cnt[b];
    __syncthreads();
  }
}

// Records per thread and round of the staged scatter, by record words (0:
// not staged).
__host__ __device__ constexpr int staged_it(int rw) { return rw == 2 ? 6 : rw == 3 ? 4 : rw == 4 ? 3 : 0; }

// LDS bytes of k_part_scatter_staged: cursors, counters, the staging area.
__host__ __device__ constexpr size_t part_scatter_staged_lds(int np, int rw) {
  return (size_t)2 * np * 8 + (((size_t)(np + 1) * 4 + 7) & ~(size_t)7) +
         (size_t)staged_it(rw) * kScatterThreads * rw * 8;
}

// LDS bytes of k_part_scatter for np partitions and chunks of ch records.
__host__ __device__ constexpr size_t part_scatter_lds(int np, int ch) {
  return (size_t)2 * np * 8 + (size_t)(np + 1) * 4 + (size_t)ch * 4;
}

// Pass A workgroup s's flushed partial states into partition order, from
// the cursors curp[n_parts].
__device__ __forceinline__ void scatter_partials(PlanRef P, int s, unsigned long long *curp) {
  const int W = P.words;
  const DFR_GLOBAL uint64_t *pseg = gp(P.partials) + (uint64_t)s * (uint64_t)P.seg_partials * (1 + W);
  const uint32_t n = P.cnt_par[s];
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
    const DFR_GLOBAL uint64_t *rec = pseg + (uint64_t)i * (1 + W);
    const uint32_t p = part_of(P, rec[0]);
    const uint64_t pos = atomicAdd(&curp[p], 1ull);
    DFR_GLOBAL uint64_t *d = gpw(P.part_par) + pos * (1 + W);
    for (int w = 0; w < 1 + W; ++w) d[w] = rec[w];
  }
}

// Workgroup (s, h) = blockIdx.x (s * seg_slices + h) moves the records of
// slice h of pass A workgroup s's waves (and, for h = 0, its partial
// states), from that slice's cursors.
template <int CH>
__global__ void __launch_bounds__(1024) k_part_scatter(const DevPlan *__restrict__ plan) {
  extern __shared__ unsigned long long cur[];  // [2 * n_parts], then cnt and idx
  __shared__ uint32_t wsum[kScatterThreads / 64];
  PlanRef P = plan_ref(plan);
  const int H = P.seg_slices, s = blockIdx.x / H, h = blockIdx.x % H, NP = P.n_parts;
  uint32_t *cnt = (uint32_t *)(cur + 2 * NP);
  uint32_t *idx = cnt + NP + 1;
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    cur[p] = P.base_raw[p] + P.hist_raw[((uint64_t)s * H + h) * NP + p];
    if (h == 0) cur[NP + p] = P.base_par[p] + P.hist_par[(uint64_t)s * NP + p];
  }
  __syncthreads();
  const uint64_t keymask = P.vmask_in_key ? ((1ull << P.key_bits) - 1) : ~0ull;
  auto bucket = [&](uint64_t w0) { return part_of(P, w0 & keymask); };
  const int per = P.seg_waves / H;
  for (int w = h * per; w < (h + 1) * per; ++w) {
    const uint64_t wseg = (uint64_t)s * P.seg_waves + w;
    const DFR_GLOBAL uint64_t *rseg = gp(P.spill) + wseg * (uint64_t)P.seg_rows * P.rw;
    const uint32_t n = P.cnt_raw[wseg];
    DFR_GLOBAL uint64_t *dst = gpw(P.part_raw);
    if (P.rw == 2)
      scatter_sorted<2, CH, true>(rseg, n, 2, NP, bucket, cur, cnt, idx, wsum, dst);
    else if (P.rw == 3)
      scatter_sorted<3, CH, true>(rseg, n, 3, NP, bucket, cur, cnt, idx, wsum, dst);
    else if (P.rw == 4)
      scatter_sorted<4, CH, true>(rseg, n, 4, NP, bucket, cur, cnt, idx, wsum, dst);
    else if (P.rw <= 8)
      scatter_sorted<8, CH, false>(rseg, n, P.rw, NP, bucket, cur, cnt, idx, wsum, dst);
    else
      scatter_sorted<kMaxCols + 3, CH, false>(rseg, n, P.rw, NP, bucket, cur, cnt, idx, wsum, dst);
  }
  if (h != 0) return;
  scatter_partials(P, s, cur + NP);
}

// k_part_scatter with staged chunks (scatter_staged) for records of 2 to 4
// words.
__global__ void __launch_bounds__(1024) k_part_scatter_staged(const DevPlan *__restrict__ plan) {
  extern __shared__ unsigned long long cur[];  // [2 * n_parts], then cnt and the staging area
  __shared__ uint32_t wsum[kScatterThreads / 64];
  PlanRef P = plan_ref(plan);
  const int H = P.seg_slices, s = blockIdx.x / H, h = blockIdx.x % H, NP = P.n_parts;
  uint32_t *cnt = (uint32_t *)(cur + 2 * NP);
  uint64_t *stage = (uint64_t *)((char *)cnt + (((size_t)(NP + 1) * 4 + 7) & ~(size_t)7));
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    cur[p] = P.base_raw[p] + P.hist_raw[((uint64_t)s * H + h) * NP + p];
    if (h == 0) cur[NP + p] = P.base_par[p] + P.hist_par[(uint64_t)s * NP + p];
  }
  __syncthreads();
  const uint64_t keymask = P.vmask_in_key ? ((1ull << P.key_bits) - 1) : ~0ull;
  auto bucket = [&](uint64_t w0) { return part_of(P, w0 & keymask); };
  const int per = P.seg_waves / H;
  DFR_GLOBAL uint64_t *dst = gpw(P.part_raw);
  for (int w = h * per; w < (h + 1) * per; ++w) {
    const uint64_t wseg = (uint64_t)s * P.seg_waves + w;
    const DFR_GLOBAL uint64_t *rseg = gp(P.spill) + wseg * (uint64_t)P.seg_rows * P.rw;
    const uint32_t n = P.cnt_raw[wseg];
    if (P.rw == 2)
      scatter_staged<2, staged_it(2)>(rseg, n, NP, bucket, cur, cnt, stage, wsum, dst);
    else if (P.rw == 3)
      scatter_staged<3, staged_it(3)>(rseg, n, NP, bucket, cur, cnt, stage, wsum, dst);
    else
      scatter_staged<4, staged_it(4)>(rseg, n, NP, bucket, cur, cnt, stage, wsum, dst);
  }
  if (h != 0) return;
  scatter_partials(P, s, cur + NP);
}
#endif  // !DFR_JIT

// LDS probe of a partition table (first-come, linear probing over the whole
// table); -1 only when every entry holds another key.
// `placed` (when given) counts the keys this call's wave places in empty
// entries: one LDS atomic per wave.
template <class PV>
__device__ __forceinline__ int part_find(const PV &P, uint64_t *lkeys, uint64_t key, uint32_t *placed = nullptr) {
  const int E = P.c_entries;
  int h = part_slot(P, key);
  int e = -1;
  bool won = false;
  for (int p = 0; p < E; ++p) {
    const uint64_t k = lkeys[h];
    if (k == key) {
      e = h;
      break;
    }
    if (k == kEmpty) {
      const unsigned long long old = atomicCAS((unsigned long long *)&lkeys[h], (unsigned long long)kEmpty,
                                               (unsigned long long)key);
      if (old == kEmpty || old == key) {
        won = old == kEmpty;
        e = h;
        break;
      }
    }
    h = (h + 1) & (E - 1);
  }
  if (placed) {
    const uint64_t wm = __ballot(won);  // the lanes that reached this call
    if (wm && __lane_id() == __ffsll((unsigned long long)wm) - 1) atomicAdd(placed, (uint32_t)__popcll(wm));
  }
  return e;
}

// U records per thread of a pass C tile: their words (those past rw
// unused) and whether each lies in the range.
template <int RW, int U>
struct RecTile {
  uint64_t w[U][RW];
  bool ok[U];
};

// Loads records i0 + u * blockDim.x + threadIdx.x (u < U) of a range that
// ends at b1 > i0. Past the range, the range's last record is loaded (and
// not folded): every load is unconditional, so all are in flight together.
template <int RW, int U, class PV>
__device__ __forceinline__ void load_recs(const PV &P, RecTile<RW, U> &t, const DFR_GLOBAL uint64_t *recs,
                                          uint64_t i0, uint64_t b1) {
  const int rw = P.rw;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint64_t i = i0 + (uint64_t)u * blockDim.x + threadIdx.x;
    t.ok[u] = i < b1;
    const uint64_t ic = min(i, b1 - 1);
#pragma unroll
    for (int k = 0; k < RW; ++k) t.w[u][k] = k < rw ? recs[ic * rw + k] : 0;
  }
}

// Folds a loaded tile into the LDS table. A key the LDS table cannot place
// goes to the HBM table, or, with DENSE (no HBM table), raises the
// sub-partition capacity fault.
template <int RW, int U, bool DENSE, class PV>
__device__ __forceinline__ void fold_recs(const PV &P, uint64_t *lkeys, uint64_t *lst, uint64_t keymask,
                                          const RecTile<RW, U> &t, uint32_t *placed = nullptr) {
  const int W = P.words;
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (!t.ok[u]) continue;
    const uint64_t *rec = t.w[u];
    // record words picked with selects (uniform indices, no scratch)
    auto pick = [&](int k) {
      uint64_t v = 0;
#pragma unroll
      for (int j = 0; j < RW; ++j)
        if (j == k) v = rec[j];
      return v;
    };
    const uint64_t key = rec[0] & keymask;
    const uint64_t vmask = P.vmask_in_key ? (rec[0] >> P.key_bits) : (P.rec_meta >= 0 ? pick(P.rec_meta) : ~0ull);
    const int64_t row = P.rec_row >= 0 ? (int64_t)pick(P.rec_row) : 0;
    auto val = [&](int c) {
      return P.val_word[c] >= 0 ? pick(P.val_word[c])
                                : ((rec[0] >> P.pk_shift[c]) & ((1ull << P.pk_bits[c]) - 1)) + (uint64_t)P.pk_base[c];
    };
    auto ok = [&](int c) { return ((vmask >> P.val_of[c]) & 1ull) != 0; };
    const int e = part_find(P, lkeys, key, placed);
    if (e >= 0) {
      fold_full(P, lst + (int64_t)e * P.c_stride, row, val, ok);
    } else if (DENSE || P.fused) {
      // no HBM table to take the key: the host runs the query again
      atomicOr(&P.fault[3], 1u);
    } else {
      const uint64_t g = global_slot(P, key);
      if (g != kEmpty) fold_full(P, P.state + g * (uint64_t)W, row, val, ok);
    }
  }
}

// Folds the raw records [b0, b1) of one partition, U per thread at once
// (loads first).
template <int RW, int U, bool DENSE, class PV>
__device__ __forceinline__ void part_fold_records(const PV &P, uint64_t *lkeys, uint64_t *lst, uint64_t keymask,
                                                  const DFR_GLOBAL uint64_t *recs, uint64_t b0, uint64_t b1,
                                                  uint32_t *placed = nullptr) {
  for (uint64_t i0 = b0; i0 < b1; i0 += (uint64_t)U * blockDim.x) {
    RecTile<RW, U> t;
    load_recs<RW, U>(P, t, recs, i0, b1);
    fold_recs<RW, U, DENSE>(P, lkeys, lst, keymask, t, placed);
  }
}

// Merges one partial state (its key, and its W words at `s`) into the LDS
// table; a key the table cannot place goes to the HBM table, or, with DENSE
// (no HBM table), raises the sub-partition capacity fault.
template <bool DENSE, class SrcPtr, class PV>
__device__ __forceinline__ void merge_one(const PV &P, uint64_t *lkeys, uint64_t *lst, uint64_t key, SrcPtr s,
                                          uint32_t *placed) {
  const int e = part_find(P, lkeys, key, placed);
  if (e >= 0) {
    merge_full(P, lst + (int64_t)e * P.c_stride, s);
  } else if (DENSE || P.fused) {
    atomicOr(&P.fault[3], 1u);
  } else {
    const uint64_t g = global_slot(P, key);
    if (g != kEmpty) merge_full(P, P.state + g * (uint64_t)P.words, s);
  }
}

// Merges the partial-state records [q0, q1) of `par` (key, then W words).
// With a register copy of the plan (W known), a record's words are loaded
// all at once first: one memory latency per record, not one per word.
template <bool DENSE, class PV>
__device__ __forceinline__ void merge_partials(const PV &P, uint64_t *lkeys, uint64_t *lst,
                                               const DFR_GLOBAL uint64_t *par, uint64_t q0, uint64_t q1,
                                               uint32_t *placed = nullptr) {
  const int W = P.words;
  for (uint64_t i = q0 + threadIdx.x; i < q1; i += blockDim.x) {
    const DFR_GLOBAL uint64_t *rec = par + i * (1 + W);
    if constexpr (RegPlan<PV>::value) {
      uint64_t r[kMaxWords + 1];
#pragma unroll
      for (int w = 0; w <= kMaxWords; ++w)
        if (w <= W) r[w] = rec[w];
      merge_one<DENSE>(P, lkeys, lst, r[0], (const uint64_t *)(r + 1), placed);
    } else {
      merge_one<DENSE>(P, lkeys, lst, rec[0], rec + 1, placed);
    }
  }
}

// Pass C: partition p.
// A fused pass C's groups (the occupied entries of its LDS table) placed at
// consecutive slots of the dense key and state tables: each 64-entry
// chunk's occupied count, a prefix over the chunks (wave 0), one
// reservation of the partition's total, then each group's key and words
// (plain stores; k_finalize finishes the tables in slot order). Finishing
// the groups here instead would put the finish code into this kernel,
// whose register copy of the plan then no longer stays in registers.
template <class PV>
__device__ __forceinline__ void place_table(const PV &P, const uint64_t *lkeys, const uint64_t *lst, int E, int S,
                                            uint32_t *wsum, uint64_t &gbase) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int nchunks = (E + 63) >> 6;
  const int W = P.words;
  for (int c = wave; c < nchunks; c += nwaves) {
    const int e = c * 64 + lane;
    const uint64_t ball = __ballot(e < E && lkeys[e] != kEmpty);
    if (lane == 0) wsum[c] = (uint32_t)__popcll(ball);
  }
  __syncthreads();
  if (wave == 0) {
    const int per = (nchunks + 63) >> 6;
    uint32_t own = 0;
    for (int j = 0; j < per; ++j) {
      const int c = lane * per + j;
      if (c < nchunks) own += wsum[c];
    }
    uint32_t inc = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(inc, d, 64);
      if (lane >= d) inc += y;
    }
    uint32_t run = inc - own;
    for (int j = 0; j < per; ++j) {
      const int c = lane * per + j;
      if (c < nchunks) {
        const uint32_t x = wsum[c];
        wsum[c] = run;
        run += x;
      }
    }
    const uint32_t total = __shfl(inc, 63, 64);
    if (lane == 0) gbase = total ? atomicAdd((unsigned long long *)P.gcount, (unsigned long long)total) : 0;
  }
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int c = wave; c < nchunks; c += nwaves) {
    const int e = c * 64 + lane;
    const uint64_t key = e < E ? lkeys[e] : kEmpty;
    const uint64_t ball = __ballot(key != kEmpty);
    if (key == kEmpty) continue;
    const uint64_t slot = gbase + wsum[c] + (uint64_t)__popcll(ball & lt);
    P.keys_tab[slot] = key;
    uint64_t *d = P.state + slot * (uint64_t)W;
    const uint64_t *src = lst + (int64_t)e * S;
    for (int w = 0; w < W; ++w) d[w] = src[w];
  }
}

template <class PV>
__device__ __forceinline__ void part_aggregate_body(const PV &P) {
  extern __shared__ uint64_t smem[];
  __shared__ uint32_t wsum[kMaxPartEntries / 64];
  __shared__ uint64_t gbase;
  const int p = blockIdx.x, E = P.c_entries, W = P.words, S = P.c_stride;
  uint64_t *lkeys = smem;
  uint64_t *lst = smem + E;
  for (int i = threadIdx.x; i < E; i += blockDim.x) {
    lkeys[i] = kEmpty;
    if constexpr (RegPlan<PV>::value) {
#pragma unroll
      for (int w = 0; w < kMaxWords; ++w)
        if (w < W) lst[i * S + w] = P.ident[w];
    } else {
      for (int w = 0; w < W; ++w) lst[i * S + w] = P.ident[w];
    }
  }
  __syncthreads();
  const uint64_t keymask = P.vmask_in_key ? ((1ull << P.key_bits) - 1) : ~0ull;
  const DFR_GLOBAL uint64_t *recs = gp(P.part_raw);
  const uint64_t b0 = P.base_raw[p], b1 = P.base_raw[p + 1];
  if (P.rw <= 4)
    part_fold_records<4, 4, false>(P, lkeys, lst, keymask, recs, b0, b1);
  else if (P.rw <= 8)
    part_fold_records<8, 2, false>(P, lkeys, lst, keymask, recs, b0, b1);
  else
    part_fold_records<kMaxCols + 3, 1, false>(P, lkeys, lst, keymask, recs, b0, b1);
  merge_partials<false>(P, lkeys, lst, gp(P.part_par), P.base_par[p], P.base_par[p + 1]);
  __syncthreads();
  if (P.fused) {
    // unordered: the partition's groups go to consecutive dense slots
    place_table(P, lkeys, lst, E, S, wsum, gbase);
    return;
  }
  // every key of this partition lives only here: plain stores
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    const uint64_t key = lkeys[e];
    if (key == kEmpty) continue;
    const uint64_t g = global_slot(P, key);
    if (g == kEmpty) continue;
    uint64_t *d = P.state + g * (uint64_t)W;
    const uint64_t *s = lst + (int64_t)e * S;
    for (int w = 0; w < W; ++w) d[w] = s[w];
  }
}

#if !DFR_JIT
__global__ void __launch_bounds__(1024) k_part_aggregate(const DevPlan *__restrict__ plan) {
  part_aggregate_body(plan_ref(plan));
}
#endif

#if !DFR_JIT
// ------------------------------------------------ second partition level

// Sub-partition of a key: the hash bits below the partition's, scaled to
// [0, n_sub) by a multiply-shift.
__device__ __forceinline__ uint32_t sub_of(PlanRef P, uint64_t key) {
  const uint32_t h = (key32(key) * 0x9E3779B1u) << P.part_log2;
  return (uint32_t)(((uint64_t)h * (uint32_t)P.n_sub) >> 32);
}

// Chunk j (of nch) of partition p's records [base[p], base[p + 1]).
__device__ __forceinline__ void chunk_range(const uint64_t *base, int p, int j, int nch, uint64_t &lo,
                                            uint64_t &hi) {
  const uint64_t b0 = base[p], n = base[p + 1] - b0;
  lo = b0 + n * (uint64_t)j / (uint64_t)nch;
  hi = b0 + n * (uint64_t)(j + 1) / (uint64_t)nch;
}

// Grid (n_parts, 2): partition p's records (y = 0 raw, y = 1 partial
// states), counted per (sub-partition, chunk) in LDS and written as one
// contiguous block of the [p][sub][chunk] histogram (n_sub * sub_chunks
// counters, at most 16384).
__global__ void __launch_bounds__(1024) k_sub_hist(const DevPlan *__restrict__ plan) {
  extern __shared__ uint32_t h2[];  // [n_sub][sub_chunks]
  PlanRef P = plan_ref(plan);
  const int p = blockIdx.x, which = blockIdx.y;
  const int NS = P.n_sub, CH = P.sub_chunks;
  for (int i = threadIdx.x; i < NS * CH; i += blockDim.x) h2[i] = 0;
  __syncthreads();
  const DFR_GLOBAL uint64_t *recs = gp(which ? P.part_par : P.part_raw);
  const uint64_t rw = which ? 1 + P.words : P.rw;
  const uint64_t keymask = (!which && P.vmask_in_key) ? ((1ull << P.key_bits) - 1) : ~0ull;
  constexpr int K = 8;  // records per thread at once, loads first
  for (int j = 0; j < CH; ++j) {
    uint64_t lo, hi;
    chunk_range(which ? P.base_par : P.base_raw, p, j, CH, lo, hi);
    for (uint64_t i0 = lo; i0 < hi; i0 += (uint64_t)K * blockDim.x) {
      uint64_t w0[K];
#pragma unroll
      for (int k = 0; k < K; ++k) w0[k] = recs[min(i0 + (uint64_t)k * blockDim.x + threadIdx.x, hi - 1) * rw];
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (i0 + (uint64_t)k * blockDim.x + threadIdx.x < hi)
          atomicAdd(&h2[sub_of(P, w0[k] & keymask) * CH + j], 1u);
    }
  }
  __syncthreads();
  uint32_t *out = (which ? P.hist2_par : P.hist2_raw) + (uint64_t)p * NS * CH;
  for (int i = threadIdx.x; i < NS * CH; i += blockDim.x) out[i] = h2[i];
}

// Exclusive scan of n 32-bit counts into 64-bit offsets, out[n] = total:
// per-block sums, k_scan_top over them, then the offsets.
__global__ void k_count_blocks(const uint32_t *c, uint64_t n, uint64_t *block_sums) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint64_t t = 0;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) t += c[base + i];
  uint64_t tot;
  block_exclusive_scan(t, &tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

__global__ void k_count_apply(const uint32_t *c, uint64_t n, const uint64_t *block_sums, uint64_t *out) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanChunk + (uint64_t)threadIdx.x * kScanItems;
  uint64_t t = 0;
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) t += c[base + i];
  uint64_t tot;
  uint64_t run = block_sums[blockIdx.x] + block_exclusive_scan(t, &tot);
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) {
      out[base + i] = run;
      run += c[base + i];
    }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = block_sums[gridDim.x];
}

// Same grid as k_sub_hist: chunk x of partition y moves its records into
// sub-partition order (counting-sorted in LDS, written as runs), from the
// cursors the scan of the counts gives.
constexpr int kSubScatterChunk = 8192;

__host__ __device__ constexpr size_t sub_scatter_lds(int ns) {
  return (size_t)ns * 8 + (size_t)(ns + 1) * 4 + (size_t)kSubScatterChunk * 4;
}

// LDS bytes of k_sub_scatter staging raw records of rw words (staged_it(rw)
// records per thread and round) for ns sub-partitions.
__host__ __device__ constexpr size_t sub_scatter_staged_lds(int ns, int rw) {
  return (size_t)ns * 8 + (((size_t)(ns + 1) * 4 + 7) & ~(size_t)7) + (size_t)staged_it(rw) * kScatterThreads * rw * 8;
}

// stage: raw records of 2 to 4 words move through scatter_staged (the LDS
// then holds sub_scatter_staged_lds bytes); partial states always gather.
__global__ void __launch_bounds__(1024) k_sub_scatter(const DevPlan *__restrict__ plan, int stage) {
  extern __shared__ unsigned long long cur2[];  // [n_sub], then cnt and idx (or the staging area)
  __shared__ uint32_t wsum[kScatterThreads / 64];
  PlanRef P = plan_ref(plan);
  const int j = blockIdx.x, p = blockIdx.y, which = blockIdx.z;
  const int NS = P.n_sub;
  uint32_t *cnt = (uint32_t *)(cur2 + NS);
  uint32_t *idx = cnt + NS + 1;
  const uint64_t *b2 = which ? P.base2_par : P.base2_raw;
  for (int i = threadIdx.x; i < NS; i += blockDim.x) cur2[i] = b2[((uint64_t)p * NS + i) * P.sub_chunks + j];
  __syncthreads();
  const DFR_GLOBAL uint64_t *src = gp(which ? P.part_par : P.part_raw);
  DFR_GLOBAL uint64_t *dst = gpw(which ? P.part2_par : P.part2_raw);
  const int rw = which ? 1 + P.words : P.rw;
  const uint64_t keymask = (!which && P.vmask_in_key) ? ((1ull << P.key_bits) - 1) : ~0ull;
  uint64_t lo, hi;
  chunk_range(which ? P.base_par : P.base_raw, p, j, P.sub_chunks, lo, hi);
  auto bucket = [&](uint64_t w0) { return sub_of(P, w0 & keymask); };
  const DFR_GLOBAL uint64_t *s0 = src + lo * rw;
  const uint32_t n = (uint32_t)(hi - lo);
  if (stage && !which) {
    uint64_t *st = (uint64_t *)((char *)cnt + (((size_t)(NS + 1) * 4 + 7) & ~(size_t)7));
    if (rw == 2)
      scatter_staged<2, staged_it(2)>(s0, n, NS, bucket, cur2, cnt, st, wsum, dst);
    else if (rw == 3)
      scatter_staged<3, staged_it(3)>(s0, n, NS, bucket, cur2, cnt, st, wsum, dst);
    else
      scatter_staged<4, staged_it(4)>(s0, n, NS, bucket, cur2, cnt, st, wsum, dst);
    return;
  }
  if (rw == 2)
    scatter_sorted<2, kSubScatterChunk, true>(s0, n, 2, NS, bucket, cur2, cnt, idx, wsum, dst);
  else if (rw == 3)
    scatter_sorted<3, kSubScatterChunk, true>(s0, n, 3, NS, bucket, cur2, cnt, idx, wsum, dst);
  else if (rw == 4)
    scatter_sorted<4, kSubScatterChunk, true>(s0, n, 4, NS, bucket, cur2, cnt, idx, wsum, dst);
  else if (rw == 5)
    scatter_sorted<5, kSubScatterChunk, true>(s0, n, 5, NS, bucket, cur2, cnt, idx, wsum, dst);
  else if (rw <= 8)
    scatter_sorted<8, kSubScatterChunk, false>(s0, n, rw, NS, bucket, cur2, cnt, idx, wsum, dst);
  else
    scatter_sorted<kMaxWords + 1, kSubScatterChunk, false>(s0, n, rw, NS, bucket, cur2, cnt, idx, wsum, dst);
}
#endif  // !DFR_JIT

// Output row i of a group: its key columns from the packed key, its
// aggregates from its state words `slot` (in LDS). Validity words start all
// ones: a null key or aggregate clears its bit (workgroups' output runs
// share words, so with an atomic; such values are rare).
//
// NA > 0 (a register copy of the plan) is the exact number of aggregates,
// so their loop unrolls completely.
// f(P.aggs[a]) for a = A .. NA - 1, each with a constant index: a loop over
// them that the compiler declines to unroll (its body, the finish of any
// aggregate, is large) indexes a register copy of the plan at run time,
// which keeps the whole copy in scratch memory.
template <int A, int NA, class PV, class F>
__device__ __forceinline__ void each_agg(const PV &P, F &f) {
  if constexpr (A < NA) {
    f(P.aggs[A]);
    each_agg<A + 1, NA>(P, f);
  }
}

template <int NA, class PV>
__device__ __forceinline__ void emit_group(const PV &P, uint64_t i, const uint64_t *slot, uint64_t packed) {
  const uint64_t clear = ~(1ull << (i & 63));
  uint64_t rest = packed;  // digits from the first key up
#pragma unroll
  for (int k = 0; k < kMaxKeys; ++k) {
    if (k >= P.n_keys) continue;
    const DevKey K = cget(P.keys[k]);
    bool valid;
    uint64_t v;
    if (K.is_float) {
      // the group's first row's own value (FIRST is tracked for such keys)
      valid = packed != kNullFloatKey;
      v = valid ? gp((const uint64_t *)P.cols[k].values)[slot[1]] : 0;
    } else {
      uint64_t digit;
      divmod_u64(rest, K.radix, K.inv_radix, rest, digit);
      valid = !(K.nullable && digit == K.radix - 1);
      v = valid ? (uint64_t)K.minv + digit : 0;
    }
    store_value(P.out_keys[k], P.cols[k].dtype, i, v);
    if (!valid) atomicAnd((unsigned long long *)&P.out_key_valid[k][i >> 6], (unsigned long long)clear);
  }
  auto one = [&](const DevAgg &G) {
    const AggOut o = finish_agg(P, G, slot);
    store_value(G.out_values, G.out_dtype, i, o.bits);
    if (!o.valid) atomicAnd((unsigned long long *)&G.out_valid[i >> 6], (unsigned long long)clear);
  };
  if constexpr (RegPlan<PV>::value) {
    each_agg<0, NA>(P, one);
  } else {
    for (int a = 0; a < P.n_aggs; ++a) one(cget(P.aggs[a]));
  }
}

// Pass C over sub-partitions x = p * n_sub + sub, from blockIdx.x on in
// steps of gridDim.x: folds the sub-partition's raw records and merges its
// partial states in an LDS table, then gives each group the next free slot
// of the dense HBM tables (one device atomic per sub-partition) and stores
// its key and state there, or, fused, its output row. The first tile of the
// next sub-partition is loaded while this one's groups are placed.
template <int NA, int RW, int U, class PV>
__device__ __forceinline__ void sub_aggregate_loop(const PV &P, uint32_t *wsum, uint64_t &gbase, uint32_t &placed) {
  extern __shared__ uint64_t smem[];
  const int E = P.c_entries, W = P.words, S = P.c_stride, CH = P.sub_chunks;
  uint64_t *lkeys = smem;
  uint64_t *lst = smem + E;
  const uint64_t keymask = P.vmask_in_key ? ((1ull << P.key_bits) - 1) : ~0ull;
  const DFR_GLOBAL uint64_t *recs = gp(P.part2_raw);
  const uint64_t nsp = (uint64_t)P.n_parts * (uint64_t)P.n_sub;
  const uint64_t tile = (uint64_t)U * blockDim.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int nchunks = E >> 6;
  RecTile<RW, U> t;
  // the sub-partition's raw and partial ranges, read (and its first tile
  // loaded) one sub-partition ahead
  uint64_t b0 = 0, b1 = 0, q0 = 0, q1 = 0;
  uint64_t sp = blockIdx.x;
  if (sp < nsp) {
    b0 = P.base2_raw[sp * CH];
    b1 = P.base2_raw[(sp + 1) * CH];
    q0 = P.base2_par[sp * CH];
    q1 = P.base2_par[(sp + 1) * CH];
    if (b0 < b1) load_recs<RW, U>(P, t, recs, b0, b1);
  }
  // (every iteration starts with its tile in registers: no fold waits for
  // the loads issued at the loop's top)
  __builtin_amdgcn_s_waitcnt(0xF70);
  for (; sp < nsp; sp += gridDim.x) {
    // the next sub-partition's ranges, in flight during this one's fold
    const uint64_t nx = sp + gridDim.x;
    uint64_t nb0 = 0, nb1 = 0, nq0 = 0, nq1 = 0;
    if (nx < nsp) {
      nb0 = P.base2_raw[nx * CH];
      nb1 = P.base2_raw[(nx + 1) * CH];
      nq0 = P.base2_par[nx * CH];
      nq1 = P.base2_par[(nx + 1) * CH];
    }
    if (threadIdx.x == 0) placed = 0;
    for (int i = threadIdx.x; i < E; i += blockDim.x) {
      lkeys[i] = kEmpty;
      if constexpr (RegPlan<PV>::value) {
#pragma unroll
        for (int w = 0; w < kMaxWords; ++w)
          if (w < W) lst[i * S + w] = P.ident[w];
      } else {
        for (int w = 0; w < W; ++w) lst[i * S + w] = P.ident[w];
      }
    }
    __syncthreads();
    if (b0 < b1) {
      fold_recs<RW, U, true>(P, lkeys, lst, keymask, t, &placed);
      if (b0 + tile < b1) part_fold_records<RW, U, true>(P, lkeys, lst, keymask, recs, b0 + tile, b1, &placed);
    }
    merge_partials<true>(P, lkeys, lst, gp(P.part2_par), q0, q1, &placed);
    // and its first tile, in flight while this one's groups are placed
    if (nb0 < nb1) load_recs<RW, U>(P, t, recs, nb0, nb1);
    b0 = nb0;
    b1 = nb1;
    q0 = nq0;
    q1 = nq1;
    __syncthreads();
    // dense slots: one reservation for the sub-partition's groups (counted
    // as their keys were placed), issued first so that it is in flight
    // while each 64-entry chunk's occupied count and a prefix over the
    // chunks (wave 0) give each group its place in the reservation
    uint64_t reserved = 0;
    if (threadIdx.x == 0 && placed) reserved = atomicAdd((unsigned long long *)P.gcount, (unsigned long long)placed);
    for (int c = wave; c < nchunks; c += nwaves) {
      const uint64_t ball = __ballot(lkeys[c * 64 + lane] != kEmpty);
      if (lane == 0) wsum[c] = (uint32_t)__popcll(ball);
    }
    __syncthreads();
    if (wave == 0) {
      const int per = (nchunks + 63) >> 6;
      uint32_t own = 0;
      for (int j = 0; j < per; ++j) {
        const int c = lane * per + j;
        if (c < nchunks) own += wsum[c];
      }
      uint32_t inc = own;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
      }
      uint32_t run = inc - own;
      for (int j = 0; j < per; ++j) {
        const int c = lane * per + j;
        if (c < nchunks) {
          const uint32_t x = wsum[c];
          wsum[c] = run;
          run += x;
        }
      }
      if (lane == 0) gbase = reserved;
    }
    __syncthreads();
    // the next tile has arrived by now: waiting for it here, before this
    // sub-partition's stores, keeps the next fold from waiting for them
    // (loads and stores share one in-order counter: gfx9 s_waitcnt
    // vmcnt(0), expcnt and lgkmcnt not waited)
    __builtin_amdgcn_s_waitcnt(0xF70);
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int c = wave; c < nchunks; c += nwaves) {
      const int e = c * 64 + lane;
      const uint64_t key = lkeys[e];
      const uint64_t ball = __ballot(key != kEmpty);
      if (key == kEmpty) continue;
      const uint64_t slot = gbase + wsum[c] + (uint64_t)__popcll(ball & lt);
      if (P.fused) {
        emit_group<NA>(P, slot, lst + (int64_t)e * S, key);
        continue;
      }
      P.keys_tab[slot] = key;
      uint64_t *d = P.state + slot * (uint64_t)W;
      const uint64_t *s = lst + (int64_t)e * S;
      for (int w = 0; w < W; ++w) d[w] = s[w];
    }
    __syncthreads();  // (the table is set up again for the next sub-partition)
  }
}

// The second-level pass C's registers, held for three 512-thread
// workgroups per CU (what their default 48 KiB LDS tables allow): its loop
// keeps the next sub-partition's first records in registers.
#define DFR_SUB_WAVES __attribute__((amdgpu_waves_per_eu(6, 8)))

template <int NA = 0, class PV>
__device__ __forceinline__ void sub_aggregate_body(const PV &P) {
  __shared__ uint32_t wsum[kMaxPartEntries / 64];
  __shared__ uint64_t gbase;
  __shared__ uint32_t placed;
  if (P.rw <= 4)
    sub_aggregate_loop<NA, 4, 2>(P, wsum, gbase, placed);
  else if (P.rw <= 8)
    sub_aggregate_loop<NA, 8, 1>(P, wsum, gbase, placed);
  else
    sub_aggregate_loop<NA, kMaxCols + 3, 1>(P, wsum, gbase, placed);
}

#if !DFR_JIT
__global__ void __launch_bounds__(512) DFR_SUB_WAVES k_sub_aggregate(const DevPlan *__restrict__ plan) {
  sub_aggregate_body(plan_ref(plan));
}

// After a fused pass C: the validity bits of rows G.. (set by the all-ones
// fill) cleared in each output column's last word.
__global__ void k_clear_tail(const DevPlan *__restrict__ plan, uint64_t G) {
  PlanRef P = plan_ref(plan);
  if (threadIdx.x != 0 || (G & 63) == 0) return;
  const uint64_t keep = (1ull << (G & 63)) - 1;
  for (int k = 0; k < P.n_keys; ++k) P.out_key_valid[k][G >> 6] &= keep;
  for (int a = 0; a < P.n_aggs; ++a) P.aggs[a].out_valid[G >> 6] &= keep;
}
#endif

}  // namespace dfr
