// plan.h — the device-side form of one filtered group-by query.
//
// The host compiles a dfr_query into a DevPlan (one per query, kept in device
// memory). Everything the kernels need is here: the input columns (key
// columns first), the predicate program, the packed-key layout, and the
// per-group state layout in HBM and in LDS.
//
// Per-group state is a row of 64-bit words ("slot"), stored array-of-structs
// so that one group's words share cache lines:
//   word 0            ROWS   rows of the group that passed the filter
//   word 1 (optional) FIRST  smallest row index of the group (maintain_order)
//   then, for each input column with aggregates, a block:
//     VCOUNT  non-null values                        (1 word)
//     SUM     int64 total (narrow) or 128-bit total  (1 or 2 words);
//             Float64: fixed point of 2, 3 or 4 words
//     SUMSQ   total of squares (fixed point)         (2, 3 or 4 words)
//     MIN     int64 or ordered float key             (1 word)
//     MAX                                            (1 word)
//     FLAGS   NaN / +inf / -inf seen                 (1 word, HBM only)
//     ZFIRST  first zero's (row << 1 | sign)         (1 word, HBM only)
// An LDS entry holds the same words except FLAGS and ZFIRST, which are
// rare-event words updated in HBM directly. Float sums are fixed-point
// integers of 128 to 256 bits (width and scale 2^shift chosen per column
// from its exponent range and the row count, so that the sum is exact or,
// when no width up to 256 bits is, its error is bounded below the parity
// budget), so every state is order independent and results are bitwise
// reproducible whatever the launch geometry or batching.
#pragma once
#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

namespace dfr {

constexpr int kMaxCols = 12;
constexpr int kMaxKeys = 4;
constexpr int kMaxLeaves = 8;
constexpr int kMaxWords = 64;     // HBM words per group slot
constexpr int kMaxAggs = 32;
constexpr uint64_t kEmpty = ~0ull;  // empty hash-table key

// column block masks
enum : uint32_t {
  M_VCOUNT = 1u << 0,
  M_ISUM64 = 1u << 1,
  M_ISUM128 = 1u << 2,
  M_FSUM = 1u << 3,
  M_FSUMSQ = 1u << 4,
  M_IMIN = 1u << 5,
  M_IMAX = 1u << 6,
  M_FMIN = 1u << 7,
  M_FMAX = 1u << 8,
  M_FLAGS = 1u << 9,
  M_ZFIRST = 1u << 10,
};

// flag bits of the FLAGS word
enum : uint64_t { FL_NAN = 1, FL_PINF = 2, FL_NINF = 4 };

enum : int32_t { MODE_DIRECT = 0, MODE_HASH = 1 };

// merge operation of each state word (partial states are folded word by word)
enum : int8_t {
  W_ADD = 0,     // 64-bit add
  W_ADD128 = 1,  // low word of a 128-bit add (the next word is its high word)
  W_HI = 2,      // high word of a 128-bit add (merged with its low word)
  W_MIN_S = 3,   // signed min
  W_MAX_S = 4,   // signed max
  W_MIN_U = 5,   // unsigned min
  W_MAX_U = 6,   // unsigned max
  W_OR = 7,
  W_ADD192 = 8,  // low word of a 192-bit add (the next two words are W_HI)
  W_ADD256 = 9,  // low word of a 256-bit add (the next three words are W_HI)
};

// How kept rows reach the group table.
enum : int32_t {
  AGG_ATOMIC = 0,     // LDS table, misses fold into HBM with device atomics
  AGG_PARTITION = 1,  // LDS table, misses spill to records that are
                      // radix-partitioned by key and folded per partition
};
enum : int32_t { LDS_NONE = 0, LDS_DIRECT = 1, LDS_HASHED = 2 };

// dtypes (same values as dfrocm.h)
enum : int32_t { T_INT32 = 1, T_INT64 = 2, T_FLOAT64 = 3, T_DICT = 4 };

struct DevCol {
  const void *values;     // first element of the window
  const uint64_t *valid;  // LSB-first validity words, or nullptr
  int64_t vbit;           // bit of row 0 in `valid`
  int32_t dtype;
  int32_t is_float;       // 1 for FLOAT64
};

// A key column's digit in the mixed-radix packed key:
// digit = null ? radix - 1 : value - minv; packed = sum(digit * stride).
struct DevKey {
  int64_t minv;
  uint64_t radix;
  uint64_t stride;   // product of the radixes of the keys before this one
  double inv_radix;  // 1 / radix, for decoding digits without integer division
  int32_t nullable;
  int32_t is_float;  // a Float64 key (the only key): packed as its canonical word
};

// Packed key of a null Float64 key: a NaN pattern canonical words never take
// (every NaN canonicalises to 0x7FF8000000000000).
constexpr uint64_t kNullFloatKey = 0x7FF0000000000001ull;

struct DevLeaf {
  int32_t op;
  int32_t flags;
  int32_t bit;   // position of this leaf's is-true bit in the truth index
  int32_t col2;  // the right-hand plan column under DFR_LEAF_COLUMN (flag 4), else -1
  int64_t lit;
  int64_t mod;
};

// Aggregate blocks of one column.
struct DevColAgg {
  uint32_t mask;
  int16_t gbase;   // first HBM word of the block
  int16_t lbase;   // first LDS word of the block (LDS entry without its key)
  // word offsets inside the block (-1 when absent)
  int8_t o_count, o_isum, o_fsum, o_sq, o_min, o_max, o_flags, o_zfirst;
  int32_t fshift;   // scale of the fixed-point sum
  int32_t fshift2;  // scale of the fixed-point sum of squares
  int32_t int_as_float;  // an integer column whose VAR goes through doubles
  int8_t nfw;       // words of the fixed-point sum (2 to 4)
  int8_t nqw;       // words of the fixed-point sum of squares (2 to 4)
  int16_t pad2;
};

// Output aggregate (finalize).
struct DevAgg {
  int32_t op;         // DFR_AGG_*
  int32_t col;        // index into cols
  int32_t ddof;
  int32_t out_dtype;  // T_INT32 / T_INT64 / T_FLOAT64
  int64_t min_count;
  void *out_values;
  uint64_t *out_valid;
};

struct DevPlan {
  int64_t rows;
  int32_t n_cols, n_keys, n_leaves, n_aggs;
  int32_t words;       // HBM words per slot
  int32_t lds_words;   // LDS words per entry (not counting the key word)
  int32_t mode;        // MODE_DIRECT / MODE_HASH
  int32_t lds_mode;    // LDS_NONE / LDS_DIRECT / LDS_HASHED
  int32_t lds_entries; // power of two (LDS_HASHED) or >= domain (LDS_DIRECT)
  int32_t lds_log2;
  int32_t lds_probe;   // probes before a row goes to HBM
  int32_t has_first;   // FIRST word present (maintain_order)
  int32_t vec_ok;      // every column window is 16-byte aligned
  int32_t leaf_cols;   // bit c set: column c carries predicate leaves
  int32_t max_probe;   // HBM hash table probes before a capacity fault
  int32_t cap_log2;
  uint32_t truth[8];
  uint64_t cap;         // slots
  uint64_t domain;      // product of key radices
  uint64_t *keys_tab;   // MODE_HASH: packed key per slot
  uint64_t *state;      // cap * words
  uint32_t *fault;      // [0] capacity, [1] Int64 sum overflow, [2] int32 sum overflow,
                        // [3] a sub-partition outgrew its LDS table
  uint64_t ident[kMaxWords];  // identity value of each HBM word
  int8_t lds_to_g[kMaxWords];  // HBM word of each LDS word
  DevCol cols[kMaxCols];
  DevColAgg cagg[kMaxCols];
  DevKey keys[kMaxKeys];
  int8_t leaf_begin[kMaxCols + 1];  // leaves sorted by column
  DevLeaf leaves[kMaxLeaves];
  DevAgg aggs[kMaxAggs];
  // ---- partitioned aggregation (AGG_PARTITION)
  int32_t agg_mode;
  int32_t rw;            // words per spilled row record
  int32_t key_bits;      // bits of the packed key (vmask sits above them when it fits)
  int32_t vmask_in_key;  // record word 0 = key | vmask << key_bits
  int32_t n_vals;        // value words per record (columns with aggregates)
  int32_t rec_row;       // word index of the row (has_first) or -1
  int32_t rec_meta;      // word index of the validity mask word or -1
  int32_t rec_val0;      // first value word
  int8_t val_of[kMaxCols];  // record value index of each column or -1
  // A value's record word, or -1 when it rides in word 0 above the key and
  // the validity mask: (value - pk_base) in pk_bits bits from bit pk_shift
  // (the column's range fits; a null's bits are masked and never read).
  int8_t val_word[kMaxCols];
  int8_t pk_shift[kMaxCols];
  int8_t pk_bits[kMaxCols];
  int64_t pk_base[kMaxCols];
  int8_t wop[kMaxWords];    // merge op of each state word
  int32_t part_log2;     // partitions = 1 << part_log2 (hash mode) ...
  int32_t n_parts;       // number of partitions
  int32_t pk_log2;       // direct mode: keys per partition = 1 << pk_log2
  int32_t c_entries;     // LDS entries of a partition table (hash mode)
  int32_t c_log2;
  int32_t n_segs;        // workgroups of pass A
  int32_t seg_waves;     // waves per workgroup: one spill segment each
  int32_t seg_slices;    // raw-record histograms per workgroup, one per run of
                         // seg_waves / seg_slices waves (scattered separately)
  int64_t seg_rows;      // record capacity of a wave's spill segment
  int64_t seg_partials;  // partial-state capacity of a segment
  uint64_t *spill;       // n_segs * seg_rows * rw
  uint64_t *partials;    // n_segs * seg_partials * (1 + words)
  uint32_t *hist_raw;    // [n_segs][n_parts]
  uint32_t *hist_par;
  uint32_t *cnt_raw;     // [n_segs * seg_waves]
  uint32_t *cnt_par;
  uint64_t *base_raw;    // [n_parts + 1] partition starts
  uint64_t *base_par;
  uint64_t *part_raw;    // records in partition order
  uint64_t *part_par;
  // ---- second partition level (high cardinality; two_level != 0)
  // Partition p splits into n_sub sub-partitions by the hash bits below the
  // partition's. Its records are cut into sub_chunks equal chunks; hist2_*
  // holds [p][sub][chunk] counts, base2_* their exclusive scan (+ total),
  // and part2_* the records in sub-partition order. Pass C then runs per
  // sub-partition and emits groups densely: slots 0..G-1 of the state and
  // key tables, G counted in *gcount (no global hash table).
  int32_t two_level;
  int32_t n_sub;
  int32_t sub_chunks;
  // words between consecutive entries of the partitioned path's LDS tables:
  // words rounded up to odd, so 64-bit accesses of entries spread over
  // every LDS bank pair
  int32_t lds_stride;
  // the same for the pass C tables (partition and sub-partition)
  int32_t c_stride;
  // pass A's hashed LDS table: 4-entry buckets, any count (a key's bucket
  // is the high half of its 32-bit hash times the count)
  int32_t lds_buckets;
  // second level, unordered output: pass C writes each group's output row
  // (keys and aggregates) at its dense slot, with no state or key tables
  int32_t fused;
  uint32_t *hist2_raw;
  uint32_t *hist2_par;
  uint64_t *base2_raw;
  uint64_t *base2_par;
  uint64_t *part2_raw;
  uint64_t *part2_par;
  uint64_t *gcount;
  // seeded hot keys (one partition level, hashed pass A table): every pass A
  // workgroup starts its LDS table from hot_image (lds_entries keys, empty
  // where nothing is seeded), so the seeded keys sit at the same entries in
  // every workgroup and never spill; at the end a workgroup writes its
  // seeded entries' states densely to hot_states[workgroup][entry][words],
  // and k_hot_reduce merges them across workgroups into the group table.
  // Null: no seeding (first-come entries only).
  uint64_t *hot_image;
  uint64_t *hot_states;
  // outputs of finalize
  void *out_keys[kMaxKeys];
  uint64_t *out_key_valid[kMaxKeys];
  uint64_t *order;      // output index -> slot (null: the identity, dense tables)
  uint64_t *out_first;  // first kept row per output group (when has_first)
  uint64_t *groups;     // device scalar: number of groups
};

}  // namespace dfr
