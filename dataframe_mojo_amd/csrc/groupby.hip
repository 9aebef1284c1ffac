// groupby.hip — query compiler (dfr_query -> DevPlan) and the launch
// sequence of one filtered group-by, plus the column statistics pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/dfrocm.h"
#include "groupby.h"
#include "kernels.h"
#include "kernels_part.h"
#include "runtime.h"

using namespace dfr;

#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return dfr_fail(error, DFR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

namespace {

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t value) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = value;
}

int bitlen_u64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

int log2_exact(uint64_t p) { return 63 - __builtin_clzll(p); }

int env_int(const char *name, int dflt) {
  const char *v = getenv(name);
  return v && *v ? atoi(v) : dflt;
}

// Bounds of a column's values for its fixed-point sums: every value has
// |x| < 2^B and is a multiple of 2^lb. Integers (whose variance goes
// through doubles) are multiples of 1; a double of an integer below 2^b can
// round up to 2^b, hence one more bit.
void fix_bounds(const ColStats &st, int dtype, int &B, int &lb) {
  if (dtype == DFR_FLOAT64) {
    if (!(st.flags & 32u)) {  // zeros and non-finite values only
      B = 0;
      lb = 0;
      return;
    }
    B = (int)st.emax - 1022;
    lb = (int)st.emin - 1075;
    return;
  }
  lb = 0;
  if (st.imin > st.imax) {  // no valid value
    B = 0;
    return;
  }
  const int64_t lo = st.imin, hi = st.imax;
  const uint64_t a = std::max(lo < 0 ? (uint64_t)(-(lo + 1)) + 1 : (uint64_t)lo,
                              hi < 0 ? (uint64_t)(-(hi + 1)) + 1 : (uint64_t)hi);
  B = bitlen_u64(a) + 1;
}

// Widths (2 to 4 words) and scales of a column block's fixed-point sum and
// sum of squares for `rows` values with bounds (B, lb). With T = 64 * words,
// the scale S = T - 2 - B - L (L = bit length of the row count) keeps every
// partial sum below 2^(T - 2), so nothing overflows.
//   - A sum is exact when S + lb >= 0 (every value is a whole number of
//     units 2^-S). The narrowest exact width is taken; when none up to 4
//     words is exact, 4 words are taken if the truncation bound
//     rows * 2^-S <= 2^-46 holds.
//   - A variance (n * Q - T^2) / (n (n - ddof)), computed exactly from the
//     two fixed-point totals, is off by at most
//     (ddof + 1) * (2^-S2 + 2^(B + 1 - S) + 2^-2S) (each value truncated
//     once per total; 2^(bitlen(ddof) + 2 + max(-S2, B + 1 - S)) bounds it).
//     The narrowest widths for which it is exact or that bound is at most
//     2^-84 are taken: its square root (std) is then within 2^-42.
// Both bounds are far inside the parity budget |a - e| <= 1e-12 * max(1, |e|).
// Returns false when no width fits.
bool choose_fixed(int B, int lb, int64_t rows, bool need_sum, bool need_sq, int ddof_max, int &nf, int &nq,
                  int &S, int &S2) {
  const int L = std::max(1, bitlen_u64((uint64_t)rows));
  const int dbits = bitlen_u64((uint64_t)std::max(ddof_max, 0));
  int exact_nf = 0;
  for (int w = 2; w <= 4 && !exact_nf; ++w)
    if (64 * w - 2 - B - L + lb >= 0) exact_nf = w;
  int best = 99;
  for (int f = 2; f <= 4; ++f) {
    const int s = 64 * f - 2 - B - L;
    const bool sum_exact = s + lb >= 0;
    if (need_sum) {
      const bool ok = exact_nf ? f >= exact_nf : (f == 4 && L - s <= -46);
      if (!ok) continue;
    }
    if (!need_sq) {
      if (f < best) {
        best = f;
        nf = f;
        nq = 2;
        S = s;
        S2 = 0;
      }
      continue;
    }
    for (int q = 2; q <= 4; ++q) {
      const int s2 = 64 * q - 2 - 2 * B - L;
      const bool exact = sum_exact && s2 + 2 * lb >= 0;
      const int log2e = dbits + 2 + std::max(-s2, B + 1 - s);
      if (!(exact || log2e <= -84)) continue;
      if (f + q < best) {
        best = f + q;
        nf = f;
        nq = q;
        S = s;
        S2 = s2;
      }
    }
  }
  return best != 99;
}

}  // namespace

Tuning default_tuning(dfr_context *ctx) {
  Tuning t;
  t.block = env_int("DFR_BLOCK", 512);
  t.wg_per_cu = env_int("DFR_WG_PER_CU", 2);
  t.lds_bytes = std::min(env_int("DFR_LDS_BYTES", 65536), ctx->lds_per_block);
  t.lds_probe = env_int("DFR_LDS_PROBE", 4);
  t.lds_mode = env_int("DFR_LDS_MODE", -1);
  t.agg_mode = env_int("DFR_AGG_MODE", AGG_PARTITION);
  t.part_lds_bytes = std::min(env_int("DFR_PART_LDS_BYTES", 131072), ctx->lds_per_block);
  t.scan_unroll = env_int("DFR_SCAN_UNROLL", 1);
  t.two_level = env_int("DFR_TWO_LEVEL", -1);
  t.sub_lds_bytes = std::min(env_int("DFR_SUB_LDS_BYTES", 49152), ctx->lds_per_block - 2048);
  t.sub_block = std::max(64, std::min(512, env_int("DFR_SUB_BLOCK", 512) / 64 * 64));
  t.sub_wg_per_cu = std::max(-1, env_int("DFR_SUB_WG_PER_CU", -1));
  t.sub_fill = std::max(10, std::min(90, env_int("DFR_SUB_FILL", 50)));
  t.jit = env_int("DFR_JIT", 1);
  t.scan_prefetch = env_int("DFR_SCAN_PREFETCH", 1) ? 1 : 0;
  t.lds_odd_stride = env_int("DFR_LDS_ODD_STRIDE", 0) ? 1 : 0;
  t.part_odd_stride = env_int("DFR_PART_ODD_STRIDE", 1) ? 1 : 0;
  t.part_block = env_int("DFR_PART_BLOCK", 1024);
  t.two_level_parts = env_int("DFR_TWO_LEVEL_PARTS", 512);
  t.scatter_slices = env_int("DFR_SCATTER_SLICES", 1);
  t.scatter_chunk = env_int("DFR_SCATTER_CHUNK", 16384) == 8192 ? 8192 : 16384;
  t.scatter_stage = env_int("DFR_SCATTER_STAGE", 1) ? 1 : 0;
  t.hot_keys = std::max(-1, std::min(1, env_int("DFR_HOT_KEYS", -1)));
  t.rec_pack = env_int("DFR_REC_PACK", 1) ? 1 : 0;
  t.fuse_one_level = env_int("DFR_FUSE_ONE_LEVEL", 1) ? 1 : 0;
  t.batch_part_rows = env_int("DFR_BATCH_PART_ROWS", 1 << 22);
  t.scan_block = env_int("DFR_SCAN_BLOCK", 1024);
  t.scan_wg_per_cu = env_int("DFR_SCAN_WG_PER_CU", 1);
  t.scan_lds_bytes = std::min(env_int("DFR_SCAN_LDS_BYTES", 159744), ctx->lds_per_block);
  return t;
}

namespace {

ColStats stats_identity() {
  ColStats init;
  init.imin = INT64_MAX;
  init.imax = INT64_MIN;
  init.emin = 0x7ff;
  init.emax = 0;
  init.flags = 0;
  init.pad = 0;
  init.valid = 0;
  init.distinct = 0;
  return init;
}

// An upper estimate of the distinct values a HyperLogLog sketch of kHllRegs
// registers saw (Flajolet et al.'s estimator with its small-range
// correction; standard error 1.04 / sqrt(m), so three of them are added),
// at most `valid`.
uint64_t hll_upper(const uint32_t *regs, uint64_t valid) {
  const double m = kHllRegs;
  double inv = 0.0;
  int zeros = 0;
  for (int j = 0; j < kHllRegs; ++j) {
    inv += ldexp(1.0, -(int)regs[j]);
    zeros += regs[j] == 0;
  }
  double e = (0.7213 / (1.0 + 1.079 / m)) * m * m / inv;
  if (e <= 2.5 * m && zeros > 0) e = m * log(m / zeros);
  const double hi = e * (1.0 + 3.0 * 1.04 / sqrt(m)) + 16.0;
  return std::min<uint64_t>(valid, (uint64_t)std::ceil(hi));
}

// Statistics of rows [0, col->length) of the column's window, computed on
// the context's device into *out (the column is left as it is: a slice
// keeps its parent's statistics, which other threads may be reading).
int window_stats(dfr_context *ctx, const dfr_column *col, ColStats *out, char **error) {
  ColStats *d = nullptr;
  int st = ctx->pool.alloc(sizeof(ColStats), (void **)&d, error);
  if (st) return st;
  ColStats *h = (ColStats *)ctx->staging;
  *h = stats_identity();
  hipError_t e = hipMemcpyAsync(d, h, sizeof(ColStats), hipMemcpyHostToDevice, ctx->stream);
  const int64_t n = col->length;
  const int blocks = (int)std::min<int64_t>((n + 16 * kBlock - 1) / (16 * kBlock), (int64_t)ctx->cus * 2);
  if (e == hipSuccess && blocks > 0) {
    hipLaunchKernelGGL(k_col_stats, dim3(blocks), dim3(kBlock), 0, ctx->stream, col->dev(), n, d, (uint32_t *)nullptr);
    ctx->obs.launches++;
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(ColStats), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->obs.waits++;
  ctx->pool.release(d);
  if (e != hipSuccess) return dfr_fail(error, DFR_DEVICE, "window statistics: %s", hipGetErrorString(e));
  *out = *h;
  return DFR_OK;
}

// Statistics of the values in the rows the plan's predicate keeps, for the
// plan columns of `cols`, into st[c] (the other entries are left alone).
int kept_stats(dfr_context *ctx, const DevPlan &P, uint32_t cols, ColStats *st, char **error) {
  const size_t plan_bytes = (sizeof(DevPlan) + 255) / 256 * 256;
  const size_t stat_bytes = sizeof(ColStats) * kMaxCols;
  void *d = nullptr;
  int rc = ctx->pool.alloc(plan_bytes + stat_bytes, &d, error);
  if (rc) return rc;
  char *h = ctx->staging;
  memcpy(h, &P, sizeof(DevPlan));
  ColStats *hs = (ColStats *)(h + plan_bytes);
  for (int c = 0; c < kMaxCols; ++c) hs[c] = stats_identity();
  const DevPlan *dplan = (const DevPlan *)d;
  ColStats *ds = (ColStats *)((char *)d + plan_bytes);
  hipError_t e = hipMemcpyAsync(d, h, plan_bytes + stat_bytes, hipMemcpyHostToDevice, ctx->stream);
  const int64_t tiles = (P.rows + 2 * kBlock - 1) / (2 * kBlock);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)ctx->cus * 4));
  if (e == hipSuccess && P.rows > 0) {
    if (P.n_cols <= 4)
      hipLaunchKernelGGL(k_kept_stats<4>, dim3(blocks), dim3(kBlock), 0, ctx->stream, dplan, cols, ds);
    else if (P.n_cols <= 8)
      hipLaunchKernelGGL(k_kept_stats<8>, dim3(blocks), dim3(kBlock), 0, ctx->stream, dplan, cols, ds);
    else
      hipLaunchKernelGGL(k_kept_stats<kMaxCols>, dim3(blocks), dim3(kBlock), 0, ctx->stream, dplan, cols, ds);
    ctx->obs.launches++;
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hs, ds, stat_bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->obs.waits++;
  ctx->pool.release(d);
  if (e != hipSuccess) return dfr_fail(error, DFR_DEVICE, "kept-row statistics: %s", hipGetErrorString(e));
  for (int c = 0; c < P.n_cols; ++c)
    if ((cols >> c) & 1u) st[c] = hs[c];
  return DFR_OK;
}

}  // namespace

size_t column_hot_words() { return (size_t)kHotList + 1 + (kHotBins + 2) / 2; }

namespace {

int fill(dfr_context *ctx, uint64_t *p, uint64_t n, uint64_t value);

// The heavy hitters of a sample of an integer or dictionary column's valid
// values at hand-off (columns of hot_rows rows or more whose value range is
// at most hot_domain): counted like a query's sample, without a filter, and
// kept as the column's candidate list for queries grouped by it alone.
int column_hot(dfr_context *ctx, dfr_column *col, char **error) {
  const Tuning &T = ctx->tune;
  const ColStats &st = col->stats;
  if (T.hot_keys == 0 || col->length < T.hot_rows || st.valid == 0) return DFR_OK;
  if (col->dtype != DFR_INT32 && col->dtype != DFR_INT64 && col->dtype != DFR_DICT) return DFR_OK;
  const uint64_t range = (uint64_t)st.imax - (uint64_t)st.imin + 1;
  if (range == 0 || range > (uint64_t)T.hot_domain) return DFR_OK;
  void *hot = nullptr, *counts = nullptr;
  int rc = ctx->pool.alloc(column_hot_words() * 8, &hot, error);
  if (rc) return rc;
  if ((rc = ctx->pool.alloc((range + 2) * 4, &counts, error))) {
    ctx->pool.release(hot);
    return rc;
  }
  uint64_t *list = (uint64_t *)hot;
  uint32_t *hist = (uint32_t *)(list + kHotList + 1);
  fill(ctx, (uint64_t *)counts, (range + 1) / 2, 0);
  fill(ctx, (uint64_t *)hist, (kHotBins + 2) / 2, 0);
  fill(ctx, list + kHotList, 1, 0);
  const int64_t span = 1024, n = col->length;
  const int64_t nblk = std::max<int64_t>(1, std::min<int64_t>((n + span - 1) / span, (int64_t)T.hot_sample / span));
  const int64_t stride = (nblk == 1 || nblk * span >= n) ? span : (n - span) / (nblk - 1);
  hipLaunchKernelGGL(k_col_hot_count, dim3((unsigned)nblk), dim3(kHotSampleThreads), 0, ctx->stream, col->dev(), n,
                     st.imin, range, (uint32_t *)counts, stride, span);
  const unsigned hgrid =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((range + 255) / 256, (uint64_t)ctx->cus * 4));
  hipLaunchKernelGGL(k_hot_hist, dim3(hgrid), dim3(256), 0, ctx->stream, (const uint32_t *)counts, range, hist);
  const unsigned cgrid = (unsigned)std::max<uint64_t>(
      1, std::min<uint64_t>((range + kHotSelectThreads - 1) / kHotSelectThreads, (uint64_t)ctx->cus));
  hipLaunchKernelGGL(k_hot_collect, dim3(cgrid), dim3(kHotSelectThreads), 0, ctx->stream, (const uint32_t *)counts,
                     range, (const uint32_t *)hist, (uint32_t)kHotList, list);
  ctx->obs.launches += 6;
  // complete before the column is handed out: other contexts' streams read it
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx->obs.waits++;
  ctx->obs.completed = ctx->obs.launches;
  ctx->pool.release(counts);
  if (e != hipSuccess) {
    ctx->pool.release(hot);
    return dfr_fail(error, DFR_DEVICE, "column heavy hitters: %s", hipGetErrorString(e));
  }
  col->hot = list;
  col->hot_min = st.imin;
  return DFR_OK;
}

}  // namespace

int column_stats(dfr_context *ctx, dfr_column *col, char **error) {
  // the statistics, then the distinct sketch's registers
  const size_t stat_bytes = (sizeof(ColStats) + 255) / 256 * 256;
  void *d = nullptr;
  int st = ctx->pool.alloc(stat_bytes + kHllRegs * 4, &d, error);
  if (st) return st;
  ColStats *ds = (ColStats *)d;
  uint32_t *dh = (uint32_t *)((char *)d + stat_bytes);
  ColStats *h = (ColStats *)ctx->staging;
  *h = stats_identity();
  const uint32_t *hregs = (const uint32_t *)(ctx->staging + stat_bytes);
  HIP_TRY(hipMemcpyAsync(ds, h, sizeof(ColStats), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_fill_u64, dim3(kHllRegs / 2 / kBlock), dim3(kBlock), 0, ctx->stream, (uint64_t *)dh,
                     (uint64_t)(kHllRegs / 2), 0ull);
  DevCol C = col->dev();
  int64_t n = col->length;
  // at least 16 rows a thread, at most two blocks a CU
  int blocks = (int)std::min<int64_t>((n + 16 * kBlock - 1) / (16 * kBlock), (int64_t)ctx->cus * 2);
  if (blocks > 0) {
    hipLaunchKernelGGL(k_col_stats, dim3(blocks), dim3(kBlock), 0, ctx->stream, C, n, ds, dh);
    ctx->obs.launches++;
  }
  HIP_TRY(hipMemcpyAsync(h, ds, sizeof(ColStats), hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipMemcpyAsync((void *)hregs, dh, kHllRegs * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->obs.waits++;
  ctx->obs.completed = ctx->obs.launches;
  h->distinct = hll_upper(hregs, h->valid);
  col->stats = *h;
  col->nulls = n - (int64_t)h->valid;
  ctx->pool.release(d);
  return column_hot(ctx, col, error);
}

// ------------------------------------------------------------ compiler

int compile_query(dfr_context *ctx, const dfr_query *q, Compiled &C, char **error, bool may_allocate) {
  if (!q) return dfr_fail(error, DFR_INVALID, "query is null");
  if (q->n_keys < 1 || q->n_keys > kMaxKeys || !q->keys)
    return dfr_fail(error, DFR_UNSUPPORTED, "the device path groups by 1 to %d keys", kMaxKeys);
  if (q->n_columns < 0 || (q->n_columns > 0 && !q->columns))
    return dfr_fail(error, DFR_INVALID, "bad column list");
  if (q->n_aggs < 0 || q->n_aggs > kMaxAggs || (q->n_aggs > 0 && !q->aggs))
    return dfr_fail(error, DFR_UNSUPPORTED, "at most %d aggregates", kMaxAggs);
  if (q->predicate.n_leaves < 0 || q->predicate.n_leaves > kMaxLeaves)
    return dfr_fail(error, DFR_UNSUPPORTED, "at most %d predicate leaves", kMaxLeaves);

  DevPlan &P = C.plan;
  memset(&P, 0, sizeof P);
  C.cols.clear();
  for (int k = 0; k < q->n_keys; ++k) {
    if (!q->keys[k]) return dfr_fail(error, DFR_INVALID, "key %d is null", k);
    C.cols.push_back(q->keys[k]);
  }
  std::vector<int> colmap(q->n_columns);
  for (int i = 0; i < q->n_columns; ++i) {
    dfr_column *c = q->columns[i];
    if (!c) return dfr_fail(error, DFR_INVALID, "column %d is null", i);
    int at = -1;
    for (size_t j = q->n_keys; j < C.cols.size(); ++j)
      if (C.cols[j] == c) at = (int)j;
    if (at < 0) {
      // a key column used as an input is loaded once, as the key
      for (int j = 0; j < q->n_keys; ++j)
        if (C.cols[j] == c) at = j;
    }
    if (at < 0) {
      if ((int)C.cols.size() >= kMaxCols)
        return dfr_fail(error, DFR_UNSUPPORTED, "at most %d distinct input columns", kMaxCols);
      at = (int)C.cols.size();
      C.cols.push_back(c);
    }
    colmap[i] = at;
  }
  const int64_t rows = C.cols[0]->length;
  for (dfr_column *c : C.cols) {
    if (c->length != rows) return dfr_fail(error, DFR_INVALID, "columns differ in length");
    if (c->ctx->device != ctx->device) return dfr_fail(error, DFR_INVALID, "a column lives on another device");
  }
  C.rows = rows;
  P.rows = rows;
  P.n_cols = (int32_t)C.cols.size();
  P.n_keys = q->n_keys;
  P.vec_ok = 1;
  for (int c = 0; c < P.n_cols; ++c) {
    P.cols[c] = C.cols[c]->dev();
    uintptr_t a = (uintptr_t)P.cols[c].values;
    int need = (P.cols[c].dtype == DFR_INT64 || P.cols[c].dtype == DFR_FLOAT64) ? 16 : 8;
    if (a % need) P.vec_ok = 0;
  }

  // keys: mixed-radix packing from the column statistics
  unsigned __int128 domain = 1;
  C.key_dtypes.clear();
  bool float_key = false;
  for (int k = 0; k < q->n_keys; ++k) {
    dfr_column *c = C.cols[k];
    DevKey &K = P.keys[k];
    if (c->dtype == DFR_FLOAT64) {
      // a Float64 key groups by its canonical 64-bit word: the whole word
      // range, so it is the query's only key
      if (q->n_keys != 1)
        return dfr_fail(error, DFR_UNSUPPORTED, "a Float64 group key is the only key on the device path");
      K.is_float = 1;
      K.nullable = c->nulls > 0 ? 1 : 0;
      K.minv = 0;
      K.radix = 0;
      K.stride = 1;
      float_key = true;
      C.key_dtypes.push_back(c->dtype);
      continue;
    }
    bool has_values = c->stats.valid > 0;
    K.minv = has_values ? c->stats.imin : 0;
    unsigned __int128 span =
        has_values ? (unsigned __int128)((uint64_t)c->stats.imax - (uint64_t)c->stats.imin) + 1 : 0;
    K.nullable = c->nulls > 0 ? 1 : 0;
    unsigned __int128 radix = span + (unsigned __int128)K.nullable;
    if (radix == 0) radix = 1;
    if (radix > ((unsigned __int128)1 << 62))
      return dfr_fail(error, DFR_UNSUPPORTED, "key %d spans more than 2^62 values", k);
    K.radix = (uint64_t)radix;
    K.inv_radix = 1.0 / (double)K.radix;
    K.stride = (uint64_t)domain;
    domain *= radix;
    if (domain > ((unsigned __int128)1 << 62))
      return dfr_fail(error, DFR_UNSUPPORTED, "the combined key domain exceeds 2^62");
    C.key_dtypes.push_back(c->dtype);
  }
  P.domain = float_key ? ~0ull : (uint64_t)domain;

  // predicate leaves, grouped by column; leaf i keeps truth bit i
  const dfr_predicate &pr = q->predicate;
  P.n_leaves = pr.n_leaves;
  for (int i = 0; i < 8; ++i) P.truth[i] = pr.truth[i];
  {
    std::vector<std::pair<int, int>> order;
    for (int i = 0; i < pr.n_leaves; ++i) {
      const dfr_leaf &L = pr.leaves[i];
      if (L.column < 0 || L.column >= q->n_columns)
        return dfr_fail(error, DFR_INVALID, "leaf %d names column %d", i, L.column);
      if (L.op < DFR_EQ || L.op > DFR_IS_NOT_NULL)
        return dfr_fail(error, DFR_UNSUPPORTED, "leaf %d: unknown comparison %d", i, L.op);
      int pc = colmap[L.column];
      if ((L.flags & DFR_LEAF_MOD) && P.cols[pc].dtype == DFR_FLOAT64)
        return dfr_fail(error, DFR_UNSUPPORTED, "leaf %d: Float64 modulo is not on the device path", i);
      if (L.flags & DFR_LEAF_COLUMN) {
        if (L.literal < 0 || L.literal >= q->n_columns)
          return dfr_fail(error, DFR_INVALID, "leaf %d compares with column %lld", i, (long long)L.literal);
        const int32_t a = P.cols[pc].dtype, b = P.cols[colmap[L.literal]].dtype;
        const bool ints = (a == DFR_INT32 || a == DFR_INT64) && (b == DFR_INT32 || b == DFR_INT64);
        if ((L.flags & DFR_LEAF_MOD) || L.op > DFR_GE || !(ints || (a == DFR_FLOAT64 && b == DFR_FLOAT64)))
          return dfr_fail(error, DFR_UNSUPPORTED,
                          "leaf %d: column comparisons take two integer or two Float64 columns", i);
      }
      order.push_back({pc, i});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    int at = 0;
    for (int c = 0; c <= kMaxCols; ++c) {
      P.leaf_begin[c] = (int8_t)at;
      while (at < (int)order.size() && order[at].first == c) {
        const dfr_leaf &L = pr.leaves[order[at].second];
        DevLeaf &D = P.leaves[at];
        D.op = L.op;
        D.flags = L.flags;
        D.bit = order[at].second;
        D.lit = L.literal;
        D.mod = L.modulus;
        D.col2 = (L.flags & DFR_LEAF_COLUMN) ? colmap[L.literal] : -1;
        P.leaf_cols |= 1 << c;
        ++at;
      }
    }
  }

  // aggregates -> per-column blocks
  uint32_t mask[kMaxCols] = {0};
  int int_as_float[kMaxCols] = {0};
  bool want_fsum[kMaxCols] = {false};  // a Float64 SUM or MEAN reads the column's fixed-point sum
  int ddof_max[kMaxCols] = {0};
  P.n_aggs = q->n_aggs;
  C.agg_dtypes.clear();
  for (int a = 0; a < q->n_aggs; ++a) {
    const dfr_agg &G = q->aggs[a];
    DevAgg &D = P.aggs[a];
    D.op = G.op;
    D.ddof = G.ddof;
    D.min_count = G.min_count;
    if (G.op == DFR_AGG_LEN) {
      D.col = 0;
      D.out_dtype = T_INT64;
      C.agg_dtypes.push_back(DFR_INT64);
      continue;
    }
    if (G.op < DFR_AGG_LEN || G.op > DFR_AGG_STD)
      return dfr_fail(error, DFR_UNSUPPORTED, "aggregate %d: unknown op %d", a, G.op);
    if (G.column < 0 || G.column >= q->n_columns)
      return dfr_fail(error, DFR_INVALID, "aggregate %d names column %d", a, G.column);
    int c = colmap[G.column];
    D.col = c;
    dfr_column *col = C.cols[c];
    const bool nulls = col->nulls > 0;
    const bool fl = col->dtype == DFR_FLOAT64;
    const uint32_t fstats = col->stats.flags;
    const bool nonfinite = (fstats & 7u) != 0;
    if (col->dtype == DFR_DICT && G.op != DFR_AGG_COUNT)
      return dfr_fail(error, DFR_UNSUPPORTED, "aggregate %d: only count() of a dictionary column", a);
    if (nulls && G.op != DFR_AGG_LEN) mask[c] |= M_VCOUNT;
    int out = T_FLOAT64;
    // an integer total fits 64 bits when |value| * rows < 2^63
    unsigned __int128 maxabs = 0;
    if (!fl && col->stats.valid > 0) {
      int64_t lo = col->stats.imin, hi = col->stats.imax;
      unsigned __int128 alo = lo < 0 ? (unsigned __int128)(-(__int128)lo) : (unsigned __int128)lo;
      unsigned __int128 ahi = hi < 0 ? (unsigned __int128)(-(__int128)hi) : (unsigned __int128)hi;
      maxabs = std::max(alo, ahi);
    }
    const bool narrow = maxabs * (unsigned __int128)(rows + 1) < ((unsigned __int128)1 << 63);
    switch (G.op) {
      case DFR_AGG_COUNT:
        out = T_INT64;
        break;
      case DFR_AGG_SUM:
        if (fl) {
          want_fsum[c] = true;
          mask[c] |= M_FSUM | (nonfinite ? M_FLAGS : 0);
          out = T_FLOAT64;
        } else {
          mask[c] |= narrow ? M_ISUM64 : M_ISUM128;
          out = col->dtype == DFR_INT32 ? T_INT32 : T_INT64;
        }
        break;
      case DFR_AGG_MIN:
      case DFR_AGG_MAX: {
        const bool mn = G.op == DFR_AGG_MIN;
        if (fl) {
          mask[c] |= (mn ? M_FMIN : M_FMAX) | ((fstats & 1u) ? M_FLAGS : 0) |
                     ((fstats & 8u) ? M_ZFIRST : 0);
          out = T_FLOAT64;
        } else {
          mask[c] |= mn ? M_IMIN : M_IMAX;
          out = col->dtype == DFR_INT32 ? T_INT32 : T_INT64;
        }
        break;
      }
      case DFR_AGG_MEAN:
        if (fl) {
          want_fsum[c] = true;
          mask[c] |= M_FSUM | (nonfinite ? M_FLAGS : 0);
        } else
          mask[c] |= narrow ? M_ISUM64 : M_ISUM128;
        out = T_FLOAT64;
        break;
      default:  // VAR, STD
        if (G.ddof < 0) return dfr_fail(error, DFR_INVALID, "aggregate %d: negative ddof", a);
        ddof_max[c] = std::max(ddof_max[c], (int)G.ddof);
        mask[c] |= M_FSUM | M_FSUMSQ | (nonfinite ? M_FLAGS : 0);
        if (!fl) int_as_float[c] = 1;
        out = T_FLOAT64;
        break;
    }
    D.out_dtype = out;
    C.agg_dtypes.push_back(out);
  }

  // fixed-point widths and scales of the columns' Float64 sums, from the
  // column statistics; for a slice (whose statistics are its parent's),
  // then from its own rows; then from the rows the predicate keeps; and
  // when still no width fits, the query is not run on the device
  int fnf[kMaxCols] = {0}, fnq[kMaxCols] = {0}, fS[kMaxCols] = {0}, fS2[kMaxCols] = {0};
  {
    ColStats st[kMaxCols];
    for (int c = 0; c < P.n_cols; ++c) st[c] = C.cols[c]->stats;
    auto settle = [&](uint32_t &bad) {
      bad = 0;
      for (int c = 0; c < P.n_cols; ++c) {
        if (!(mask[c] & (M_FSUM | M_FSUMSQ))) continue;
        int B, lb;
        fix_bounds(st[c], C.cols[c]->dtype, B, lb);
        if (!choose_fixed(B, lb, rows, want_fsum[c], (mask[c] & M_FSUMSQ) != 0, ddof_max[c], fnf[c], fnq[c], fS[c],
                          fS2[c]))
          bad |= 1u << c;
      }
    };
    uint32_t bad;
    settle(bad);
    if (bad && may_allocate) {
      for (int c = 0; c < P.n_cols; ++c) {
        if (!((bad >> c) & 1u) || C.cols[c]->stats_window) continue;
        int st2 = window_stats(ctx, C.cols[c], &st[c], error);
        if (st2) return st2;
      }
      settle(bad);
    }
    if (bad && may_allocate) {
      int st2 = kept_stats(ctx, P, bad, st, error);
      if (st2) return st2;
      settle(bad);
    }
    if (bad) {
      int c = 0;
      while (!((bad >> c) & 1u)) ++c;
      return dfr_fail(error, DFR_UNSUPPORTED,
                      "the Float64 values of input column %d span too wide a range for an exact device %s over %lld "
                      "rows",
                      c, (mask[c] & M_FSUMSQ) ? "variance" : "sum", (long long)rows);
    }
  }

  // slot layout
  int w = 1;  // ROWS
  // FIRST is tracked for ordered output, and for a Float64 key, whose
  // output is its group's first row's value
  P.has_first = (q->maintain_order || float_key) ? 1 : 0;
  if (P.has_first) w = 2;
  int lw = w;
  int l2g[kMaxWords];
  for (int i = 0; i < w; ++i) l2g[i] = i;
  uint64_t ident[kMaxWords];
  int8_t wop[kMaxWords];
  ident[0] = 0;
  wop[0] = W_ADD;
  if (P.has_first) {
    ident[1] = kEmpty;
    wop[1] = W_MIN_U;
  }
  for (int c = 0; c < P.n_cols; ++c) {
    DevColAgg &A = P.cagg[c];
    memset(&A, 0, sizeof A);
    A.o_count = A.o_isum = A.o_fsum = A.o_sq = A.o_min = A.o_max = A.o_flags = A.o_zfirst = -1;
    A.mask = mask[c];
    if (!mask[c]) continue;
    if (w + 16 > kMaxWords) return dfr_fail(error, DFR_UNSUPPORTED, "too many aggregate words");
    A.gbase = (int16_t)w;
    A.lbase = (int16_t)lw;
    int o = 0;
    auto put = [&](int8_t &field, int n, uint64_t id, bool lds, int8_t op) {
      field = (int8_t)o;
      const int8_t low = n == 2 ? W_ADD128 : (n == 3 ? W_ADD192 : W_ADD256);
      for (int i = 0; i < n; ++i) {
        ident[w] = id;
        wop[w] = (n >= 2) ? (i == 0 ? low : W_HI) : op;
        if (lds) l2g[lw++] = w;
        ++w;
        ++o;
      }
    };
    if (mask[c] & M_VCOUNT) put(A.o_count, 1, 0, true, W_ADD);
    if (mask[c] & M_ISUM64) put(A.o_isum, 1, 0, true, W_ADD);
    if (mask[c] & M_ISUM128) put(A.o_isum, 2, 0, true, W_ADD128);
    if (mask[c] & M_FSUM) put(A.o_fsum, fnf[c], 0, true, W_ADD128);
    if (mask[c] & M_FSUMSQ) put(A.o_sq, fnq[c], 0, true, W_ADD128);
    if (mask[c] & M_IMIN) put(A.o_min, 1, (uint64_t)INT64_MAX, true, W_MIN_S);
    if (mask[c] & M_IMAX) put(A.o_max, 1, (uint64_t)INT64_MIN, true, W_MAX_S);
    if (mask[c] & M_FMIN) put(A.o_min, 1, kEmpty, true, W_MIN_U);
    if (mask[c] & M_FMAX) put(A.o_max, 1, 0, true, W_MAX_U);
    if (mask[c] & M_FLAGS) put(A.o_flags, 1, 0, false, W_OR);
    if (mask[c] & M_ZFIRST) put(A.o_zfirst, 1, kEmpty, false, W_MIN_U);
    if (mask[c] & (M_FSUM | M_FSUMSQ)) {
      A.fshift = fS[c];
      A.fshift2 = fS2[c];
      A.nfw = (int8_t)fnf[c];
      A.nqw = (int8_t)((mask[c] & M_FSUMSQ) ? fnq[c] : 0);
      A.int_as_float = int_as_float[c];
    }
  }
  P.words = w;
  P.lds_words = lw;
  for (int i = 0; i < w; ++i) {
    P.ident[i] = ident[i];
    P.wop[i] = wop[i];
  }
  for (int i = 0; i < lw; ++i) P.lds_to_g[i] = (int8_t)l2g[i];
  C.maintain_order = q->maintain_order != 0;
  // spilled row records: [key | vmask << key_bits] [meta] [row] [values]
  P.n_vals = 0;
  for (int c = 0; c < kMaxCols; ++c) P.val_of[c] = -1;
  for (int c = 0; c < P.n_cols; ++c)
    if (P.cagg[c].mask) P.val_of[c] = (int8_t)P.n_vals++;
  P.key_bits = bitlen_u64(P.domain > 0 ? P.domain - 1 : 0);
  P.vmask_in_key = (P.key_bits < 64 && P.key_bits + P.n_vals <= 64) ? 1 : 0;
  // the row index travels with a record when the state orders by row:
  // FIRST (maintain_order) or a first-zero word (signed-zero ties)
  bool need_row = P.has_first != 0;
  for (int c = 0; c < P.n_cols; ++c)
    if (P.cagg[c].mask & M_ZFIRST) need_row = true;
  int rw = 1;
  P.rec_meta = P.vmask_in_key ? -1 : rw++;
  P.rec_row = need_row ? rw++ : -1;
  P.rec_val0 = rw;
  // narrow integer values ride in word 0's free bits (narrowest first)
  for (int c = 0; c < kMaxCols; ++c) {
    P.val_word[c] = -1;
    P.pk_shift[c] = P.pk_bits[c] = 0;
    P.pk_base[c] = 0;
  }
  int shift = P.key_bits + P.n_vals;
  bool packed[kMaxCols] = {};
  while (P.vmask_in_key && ctx->tune.rec_pack && shift < 64) {
    int best = -1, best_bits = 65;
    for (int c = 0; c < P.n_cols; ++c) {
      const dfr_column *col = C.cols[c];
      if (P.val_of[c] < 0 || packed[c] || col->dtype == DFR_FLOAT64) continue;
      const int bits = col->stats.valid > 0 ? std::max(1, bitlen_u64((uint64_t)col->stats.imax - (uint64_t)col->stats.imin)) : 1;
      if (bits < best_bits) {
        best = c;
        best_bits = bits;
      }
    }
    if (best < 0 || shift + best_bits > 64) break;
    packed[best] = true;
    P.pk_shift[best] = (int8_t)shift;
    P.pk_bits[best] = (int8_t)best_bits;
    P.pk_base[best] = C.cols[best]->stats.valid > 0 ? C.cols[best]->stats.imin : 0;
    shift += best_bits;
  }
  for (int c = 0; c < P.n_cols; ++c)
    if (P.val_of[c] >= 0 && !packed[c]) P.val_word[c] = (int8_t)rw++;
  P.rw = rw;
  return DFR_OK;
}

// An upper bound of the query's groups from its key columns' distinct
// estimates (made at hand-off; a null key is one more group): their
// product, or UINT64_MAX when a key has none.
static uint64_t distinct_bound(const Compiled &C) {
  unsigned __int128 b = 1;
  for (int k = 0; k < C.plan.n_keys; ++k) {
    const dfr_column *c = C.cols[k];
    if (c->stats.distinct == 0 && c->stats.valid > 0) return UINT64_MAX;
    b *= (unsigned __int128)c->stats.distinct + (c->nulls > 0 ? 1 : 0);
    if (b > ((unsigned __int128)1 << 62)) return UINT64_MAX;
  }
  return (uint64_t)std::max<unsigned __int128>(b, 1);
}

// Table mode and launch shape from the tuning and the plan. The groups a
// hash table and the partition fan-out are sized for are at most the key
// domain, the rows and the distinct bound; a table that fills up anyway
// raises the capacity fault, and the query runs again with a larger one.
void choose_layout(dfr_context *ctx, Compiled &C, int64_t hash_slots, int sub_boost, const Tuning *tune) {
  DevPlan &P = C.plan;
  const Tuning &T = tune ? *tune : ctx->tune;
  const int64_t rows = C.rows;
  const uint64_t D = P.domain;
  const uint64_t G_hi = std::min<uint64_t>(std::min<uint64_t>(D, (uint64_t)std::max<int64_t>(rows, 1)), distinct_bound(C));
  const size_t slot_bytes = (size_t)P.words * 8;
  P.lds_stride = T.lds_odd_stride ? (P.words | 1) : P.words;
  P.c_stride = T.part_odd_stride ? (P.words | 1) : P.words;
  const bool direct_fits = D <= (uint64_t)T.direct_max_slots &&
                           (int64_t)(D * slot_bytes) <= std::max<int64_t>(ctx->pool.budget / 4, 1 << 20);
  if (direct_fits) {
    P.mode = MODE_DIRECT;
    P.cap = std::max<uint64_t>(D, 1);
    P.cap_log2 = 1;
  } else {
    P.mode = MODE_HASH;
    uint64_t want = hash_slots > 0 ? (uint64_t)hash_slots : next_pow2(std::max<uint64_t>(64, 2 * G_hi));
    P.cap = next_pow2(want);
    P.cap_log2 = log2_exact(P.cap);
  }
  P.max_probe = (int32_t)std::min<uint64_t>((uint64_t)T.hash_probe, P.cap);
  // LDS table
  const int lds_bytes = std::min(T.lds_bytes, ctx->lds_per_block);
  const int LW = P.lds_words;
  int mode = T.lds_mode;
  const uint64_t direct_entries = (uint64_t)lds_bytes / (8ull * LW);
  if (mode < 0) mode = (D <= direct_entries) ? LDS_DIRECT : LDS_HASHED;
  if (mode == LDS_DIRECT && D > direct_entries) mode = LDS_HASHED;
  if (mode == LDS_DIRECT) {
    P.lds_mode = LDS_DIRECT;
    P.lds_entries = (int32_t)std::max<uint64_t>(D, 1);
    P.lds_log2 = 1;
  } else if (mode == LDS_HASHED) {
    uint64_t e = 1;
    while (e * 2 * 8ull * (LW + 1) <= (uint64_t)lds_bytes && e < 8192) e *= 2;
    if (e < 64) {
      P.lds_mode = LDS_NONE;
      P.lds_entries = 0;
    } else {
      P.lds_mode = LDS_HASHED;
      P.lds_entries = (int32_t)e;
      P.lds_log2 = log2_exact(e);
    }
  } else {
    P.lds_mode = LDS_NONE;
    P.lds_entries = 0;
  }
  P.lds_probe = T.lds_probe;
  P.agg_mode = T.agg_mode;
  if (P.agg_mode == AGG_PARTITION) {
    // partition tables: E entries of (key + W words, at an odd stride)
    // within part_lds_bytes
    const int W = P.words;
    const int S = P.c_stride;
    uint64_t ce = 64;
    // (2 KiB of LDS stay free for the second level's static arrays)
    const uint64_t c_budget = (uint64_t)std::min(T.part_lds_bytes, ctx->lds_per_block - 2048);
    while (ce * 2 * 8ull * (S + 1) <= c_budget && ce < (uint64_t)kMaxPartEntries) ce *= 2;
    P.c_entries = (int32_t)ce;
    P.c_log2 = log2_exact(ce);
    const uint64_t groups_est = std::max<uint64_t>(1, G_hi);
    const uint64_t max_parts = (uint64_t)T.max_parts;
    uint64_t np = next_pow2((2 * groups_est + ce - 1) / ce);
    // pass C runs one workgroup per partition: however few the groups, a
    // query of many rows (whose misses may all spill) gets partitions
    // enough to fill the CUs (10M rows over 1e4 keys: 32 -> 256 partitions,
    // 1.9 -> 0.4 ms)
    const uint64_t np_rows = std::min<uint64_t>(next_pow2(std::max<uint64_t>(1, (uint64_t)rows / 32768)),
                                                next_pow2((uint64_t)std::max(1, ctx->cus)));
    np = std::max(np, np_rows);
    np = std::min<uint64_t>(std::max<uint64_t>(np, 1), max_parts);
    P.n_parts = (int32_t)np;
    P.part_log2 = log2_exact(np);
    // a second level when one level's tables may not hold every group; its
    // pass C emits groups into dense tables instead of a global one. The
    // sub-partition count is set once the first level's record totals are
    // known (second_level_layout).
    P.two_level = 0;
    P.n_sub = 0;
    P.sub_chunks = 0;
    const bool two = T.two_level > 0 || (T.two_level < 0 && 2 * groups_est > max_parts * ce);
    if (two) {
      P.two_level = 1;
      const uint64_t np2 = std::min<uint64_t>(max_parts, (uint64_t)T.two_level_parts);
      P.n_parts = (int32_t)np2;
      P.part_log2 = log2_exact(np2);
      // smaller tables than one level's: more workgroups per CU
      uint64_t e2 = 64;
      while (e2 * 2 * 8ull * (S + 1) <= (uint64_t)T.sub_lds_bytes && e2 < (uint64_t)kMaxPartEntries) e2 *= 2;
      P.c_entries = (int32_t)e2;
      P.c_log2 = log2_exact(e2);
      P.mode = MODE_HASH;
      P.cap = std::max<uint64_t>(1, std::min<uint64_t>(D, (uint64_t)std::max<int64_t>(rows, 1)));
    }
    // the scan's LDS table (full state words) next to its histograms
    const int waves = std::max(64, std::min(1024, T.scan_block)) / 64;
    int H = 1;  // a power of two that divides the waves
    while (H * 2 <= std::min(T.scatter_slices, waves) && waves % (H * 2) == 0) H *= 2;
    P.seg_slices = H;
    const int hist_bytes = (H * P.n_parts + 2) * 4;
    const int avail = std::max(0, std::min(T.scan_lds_bytes, ctx->lds_per_block) - hist_bytes);
    const uint64_t direct_e = (uint64_t)avail / (8ull * S);
    int m = T.lds_mode;
    if (m < 0) m = (D <= direct_e) ? LDS_DIRECT : LDS_HASHED;
    if (m == LDS_DIRECT && D > direct_e) m = LDS_HASHED;
    if (m == LDS_DIRECT) {
      P.lds_mode = LDS_DIRECT;
      P.lds_entries = (int32_t)std::max<uint64_t>(D, 1);
      P.lds_log2 = 1;
    } else if (m == LDS_HASHED) {
      // as many 4-entry buckets as the LDS holds (the table caches the
      // first keys to arrive; a full bucket's rows spill)
      const uint64_t nb = std::min<uint64_t>((uint64_t)avail / (4 * 8ull * (S + 1)), 16384 / 4);
      if (nb < 16) {
        P.lds_mode = LDS_NONE;
        P.lds_entries = 0;
      } else {
        P.lds_mode = LDS_HASHED;
        P.lds_buckets = (int32_t)nb;
        P.lds_entries = (int32_t)(4 * nb);
        P.lds_log2 = 0;
      }
    } else {
      P.lds_mode = LDS_NONE;
      P.lds_entries = 0;
    }
  }
  // launch shape: persistent workgroups over contiguous row chunks (the
  // partitioned pass A and the one-pass kernel have their own)
  const bool part = P.agg_mode == AGG_PARTITION;
  C.block = std::max(64, std::min(1024, part ? T.scan_block : T.block));
  int64_t step = 8 * (int64_t)C.block;  // 2 rows x up to 4 pairs per thread
  int64_t blocks = (int64_t)ctx->cus * std::max(1, part ? T.scan_wg_per_cu : T.wg_per_cu);
  int64_t chunk = (rows + blocks - 1) / std::max<int64_t>(blocks, 1);
  chunk = std::max<int64_t>(step, (chunk + step - 1) / step * step);
  // pass A addresses a workgroup's rows with 32-bit buffer offsets
  const int64_t max_chunk = ((int64_t)1 << 28) / step * step;
  chunk = std::min(chunk, max_chunk);
  C.chunk = chunk;
  C.grid = (int)std::max<int64_t>(1, (rows + chunk - 1) / chunk);
  C.lds_bytes = P.lds_mode == LDS_NONE
                    ? 0
                    : (size_t)P.lds_entries * 8 * (LW + (P.lds_mode == LDS_HASHED ? 1 : 0));
  if (P.agg_mode == AGG_PARTITION) {
    P.n_segs = C.grid;
    P.seg_waves = C.block / 64;
    P.seg_rows = chunk / P.seg_waves;  // a wave sees chunk / waves rows
    P.seg_partials = std::max(P.lds_entries, 1);
    const int S = P.lds_stride;
    C.lds_bytes = (P.lds_mode == LDS_NONE ? 0
                                          : (size_t)P.lds_entries * 8 * (S + (P.lds_mode == LDS_HASHED ? 1 : 0))) +
                  (size_t)(P.seg_slices * P.n_parts + 2) * 4;
    C.c_lds_bytes = (size_t)P.c_entries * 8 * (P.c_stride + 1);
  }
}

// ------------------------------------------------------------ launch

namespace {

struct Scratch {
  dfr_context *ctx;
  std::vector<void *> held;
  explicit Scratch(dfr_context *c) : ctx(c) {}
  ~Scratch() {
    for (void *p : held) ctx->pool.release(p);
  }
  template <class T>
  int get(size_t count, T **out, char **error) {
    void *p = nullptr;
    int st = ctx->pool.alloc(std::max<size_t>(count * sizeof(T), 8), &p, error);
    if (st) return st;
    held.push_back(p);
    *out = (T *)p;
    return DFR_OK;
  }
};

int fill(dfr_context *ctx, uint64_t *p, uint64_t n, uint64_t value) {
  if (n == 0) return DFR_OK;
  int blocks = (int)std::min<uint64_t>((n + kBlock - 1) / kBlock, (uint64_t)ctx->cus * 16);
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks), dim3(kBlock), 0, ctx->stream, p, n, value);
  ctx->obs.launches++;
  return DFR_OK;
}

// Exclusive prefix of popcounts of `n` bitmap words; the total lands at
// sums[nb] where nb = number of scan blocks.
int scan_bits(dfr_context *ctx, Scratch &S, const uint64_t *bits, uint64_t n, uint64_t *prefix,
              uint64_t **total, char **error) {
  uint64_t nb = std::max<uint64_t>(1, (n + kScanChunk - 1) / kScanChunk);
  uint64_t *sums;
  int st = S.get(nb + 1, &sums, error);
  if (st) return st;
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, bits, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, ctx->stream, sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, bits, n, sums,
                     prefix);
  ctx->obs.launches += 3;
  *total = sums + nb;
  return DFR_OK;
}

template <class K>
int set_lds(K kernel, size_t bytes, char **error) {
  if (bytes > 65536)
    HIP_TRY(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return DFR_OK;
}

// The template shape pass A runs with: LDS table mode, column bound, row
// pairs per thread.
void scan_shape(const dfr_context *ctx, const Compiled &C, int *ldsm, int *nc, int *u) {
  const int n = C.plan.n_cols, t = ctx->tune.scan_unroll;
  *ldsm = C.plan.lds_mode;
  *nc = n <= 4 ? 4 : (n <= 8 ? 8 : kMaxCols);
  *u = n <= 4 ? (t == 4 ? 4 : (t == 2 ? 2 : 1)) : (n <= 8 ? (t >= 2 ? 2 : 1) : 1);
}

// Launches the kernel `which` compiled at run time for the query's shape;
// false when the tuning leaves it off or it cannot run here (the caller then
// launches the precompiled kernel).
// Workgroups of a loop over `items` work items: one per item with tune
// per_cu 0, else per_cu per CU (-1: as many as fit on a CU at once, from the
// kernel's occupancy `occ`), at most one per item.
unsigned persistent_grid(const dfr_context *ctx, int per_cu, int occ, uint64_t items) {
  if (per_cu == 0) return (unsigned)items;
  const uint64_t k = per_cu > 0 ? (uint64_t)per_cu : (uint64_t)std::max(1, occ);
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(items, k * (uint64_t)ctx->cus));
}

// grid 0: a persistent loop over `items` work items (persistent_grid with
// tune sub_wg_per_cu).
bool jit_launch(dfr_context *ctx, const Compiled &C, int which, unsigned grid, unsigned block, size_t lds,
                void **args, uint64_t items = 0) {
  const Tuning &T = ctx->tune;
  if (!(T.jit == 2 || (T.jit == 1 && (C.rows >= T.jit_rows || C.jit_any_size)))) return false;
  int ldsm, nc, u;
  scan_shape(ctx, C, &ldsm, &nc, &u);
  hipFunction_t fn = jit_function(ctx, C.plan, ldsm, nc, u, which);
  if (!fn) return false;
  if (lds > 65536 &&
      hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  if (grid == 0) {
    int occ = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, (int)block, lds) != hipSuccess) {
      (void)hipGetLastError();
      occ = 1;
    }
    grid = persistent_grid(ctx, T.sub_wg_per_cu, occ, items);
  }
  if (hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, (unsigned)lds, ctx->stream, args, nullptr) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  ctx->obs.jit_launches++;
  return true;
}

template <int LDSM, int NC, int U>
int launch_scan(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  int st;
  {
    int64_t chunk = C.chunk;
    void *args[] = {(void *)&dplan, (void *)&chunk};
    if (jit_launch(ctx, C, JIT_SCAN, C.grid, C.block, C.lds_bytes, args)) return DFR_OK;
  }
  if ((st = set_lds(k_scan_spill<LDSM, NC, U>, C.lds_bytes, error))) return st;
  hipLaunchKernelGGL((k_scan_spill<LDSM, NC, U>), dim3(C.grid), dim3(C.block), C.lds_bytes, ctx->stream, dplan,
                     C.chunk);
  return DFR_OK;
}

template <int LDSM>
int launch_scan_cols(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  int ldsm, nc, u;
  scan_shape(ctx, C, &ldsm, &nc, &u);
  if (nc == 4) {
    if (u == 4) return launch_scan<LDSM, 4, 4>(ctx, C, dplan, error);
    if (u == 2) return launch_scan<LDSM, 4, 2>(ctx, C, dplan, error);
    return launch_scan<LDSM, 4, 1>(ctx, C, dplan, error);
  }
  if (nc == 8) {
    if (u == 2) return launch_scan<LDSM, 8, 2>(ctx, C, dplan, error);
    return launch_scan<LDSM, 8, 1>(ctx, C, dplan, error);
  }
  return launch_scan<LDSM, kMaxCols, 1>(ctx, C, dplan, error);
}

// Whether a query's pass A seeds its table with sampled heavy-hitter keys:
// a one-level partitioned query with a hashed LDS table over a packed key
// domain small enough to count every key of a sample.
bool hot_wanted(const dfr_context *ctx, const Compiled &C, bool dense, bool col_list) {
  const DevPlan &P = C.plan;
  const Tuning &T = ctx->tune;
  if (T.hot_keys == 0 || dense || P.agg_mode != AGG_PARTITION || P.lds_mode != LDS_HASHED) return false;
  if (P.lds_entries < 4 || C.rows <= 0) return false;
  if (!col_list && (P.domain == 0 || P.domain > (uint64_t)T.hot_domain)) return false;
  return T.hot_keys == 1 || C.rows >= T.hot_rows;
}

template <int NC>
void launch_hot_count(dfr_context *ctx, const DevPlan *dplan, uint32_t *counts, uint64_t domain, unsigned blocks,
                      int64_t stride, int64_t span) {
  hipLaunchKernelGGL(k_hot_count<NC>, dim3(blocks), dim3(kHotSampleThreads), 0, ctx->stream, dplan, counts, domain,
                     stride, span);
}

// The seeded keys: sampled counts, their histogram, the candidates above
// the count threshold, and the image of pass A's table they are placed in.
int launch_hot_select(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  const DevPlan &P = C.plan;
  const Tuning &T = ctx->tune;
  const uint64_t D = P.domain;
  const unsigned long long *total = (const unsigned long long *)(C.hot_hist + kHotBins);
  if (!C.hot_sampled) {
    // the key column's list
    hipLaunchKernelGGL(k_hot_select, dim3(1), dim3(kHotSelectThreads), 0, ctx->stream, C.hot_list, total,
                       C.hot_delta, C.hot_limit, P.hot_image, P.lds_entries, (float)T.hot_share / 100.0f,
                       C.hot_out);
    ctx->obs.launches += 1;
    HIP_TRY(hipGetLastError());
    return DFR_OK;
  }
  fill(ctx, (uint64_t *)C.hot_counts, (D + 1) / 2, 0);
  fill(ctx, (uint64_t *)C.hot_hist, (kHotBins + 2) / 2, 0);
  fill(ctx, C.hot_list + kHotList, 1, 0);
  // blocks of `span` rows spread evenly over the rows (even row offsets, as
  // the paired loads need)
  const int64_t span = 1024, rows = C.rows;
  const int64_t nblk =
      std::max<int64_t>(1, std::min<int64_t>((rows + span - 1) / span, (int64_t)T.hot_sample / span));
  const int64_t stride = (nblk == 1 || nblk * span >= rows) ? span : ((rows - span) / (nblk - 1)) & ~(int64_t)1;
  int ldsm, nc, u;
  scan_shape(ctx, C, &ldsm, &nc, &u);
  if (nc == 4)
    launch_hot_count<4>(ctx, dplan, C.hot_counts, D, (unsigned)nblk, stride, span);
  else if (nc == 8)
    launch_hot_count<8>(ctx, dplan, C.hot_counts, D, (unsigned)nblk, stride, span);
  else
    launch_hot_count<kMaxCols>(ctx, dplan, C.hot_counts, D, (unsigned)nblk, stride, span);
  const unsigned hgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((D + 255) / 256, (uint64_t)ctx->cus * 4));
  hipLaunchKernelGGL(k_hot_hist, dim3(hgrid), dim3(256), 0, ctx->stream, C.hot_counts, D, C.hot_hist);
  const unsigned cgrid =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((D + kHotSelectThreads - 1) / kHotSelectThreads, (uint64_t)ctx->cus));
  // about twice as many candidates as the table has entries
  const uint32_t want = (uint32_t)std::min<int64_t>(kHotList, 2 * (int64_t)P.lds_entries);
  hipLaunchKernelGGL(k_hot_collect, dim3(cgrid), dim3(kHotSelectThreads), 0, ctx->stream, C.hot_counts, D,
                     C.hot_hist, want, C.hot_list);
  hipLaunchKernelGGL(k_hot_select, dim3(1), dim3(kHotSelectThreads), 0, ctx->stream, C.hot_list, total, (int64_t)0,
                     D, P.hot_image, P.lds_entries, (float)T.hot_share / 100.0f, C.hot_out);
  ctx->obs.launches += 7;
  HIP_TRY(hipGetLastError());
  return DFR_OK;
}

// The seeded entries' states over pass A's workgroups, into the group table.
int launch_hot_reduce(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  const DevPlan &P = C.plan;
  const int groups = std::max(1, std::min(16, P.n_segs));
  const uint64_t ew = (uint64_t)P.lds_entries * (uint64_t)P.words;
  hipLaunchKernelGGL(k_hot_reduce, dim3((unsigned)((ew * groups + 255) / 256)), dim3(256), 0, ctx->stream, dplan,
                     groups, 0);
  hipLaunchKernelGGL(k_hot_reduce, dim3((unsigned)((ew + 255) / 256)), dim3(256), 0, ctx->stream, dplan, groups, 1);
  ctx->obs.launches += 2;
  if (P.fused) {
    hipLaunchKernelGGL(k_hot_emit, dim3((unsigned)((P.lds_entries + 255) / 256)), dim3(256), 0, ctx->stream, dplan);
    ctx->obs.launches += 1;
  }
  HIP_TRY(hipGetLastError());
  return DFR_OK;
}

// Passes A, B and C of the partitioned aggregation. `events`: record the
// phase boundary (off under graph capture).
int launch_partitioned(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error, bool events = true) {
  const DevPlan &P = C.plan;
  int st;
  if (P.hot_image && (st = launch_hot_select(ctx, C, dplan, error))) return st;
  switch (P.lds_mode) {
    case LDS_DIRECT: st = launch_scan_cols<LDS_DIRECT>(ctx, C, dplan, error); break;
    case LDS_HASHED: st = launch_scan_cols<LDS_HASHED>(ctx, C, dplan, error); break;
    default: st = launch_scan_cols<LDS_NONE>(ctx, C, dplan, error);
  }
  if (st) return st;
  HIP_TRY(hipGetLastError());
  if (events) HIP_TRY(hipEventRecord(ctx->evp[1], ctx->stream));
  hipLaunchKernelGGL(k_part_offsets, dim3(P.n_parts), dim3(kBlock), 0, ctx->stream, dplan);
  hipLaunchKernelGGL(k_part_base, dim3(1), dim3(kBlock), 0, ctx->stream, dplan);
  // k_part_scatter: 64-bit write cursors per partition and kind, then its
  // chunk sort's counters and record indices
  const dim3 sgrid((unsigned)(P.n_segs * P.seg_slices));
  const size_t staged_bytes = part_scatter_staged_lds(P.n_parts, P.rw);
  if (ctx->tune.scatter_stage && staged_it(P.rw) > 0 && staged_bytes + 1024 <= (size_t)ctx->lds_per_block) {
    if ((st = set_lds(k_part_scatter_staged, staged_bytes, error))) return st;
    hipLaunchKernelGGL(k_part_scatter_staged, sgrid, dim3(kScatterThreads), staged_bytes, ctx->stream, dplan);
  } else if (ctx->tune.scatter_chunk == 8192) {
    const size_t cur_bytes = part_scatter_lds(P.n_parts, 8192);
    if ((st = set_lds(k_part_scatter<8192>, cur_bytes, error))) return st;
    hipLaunchKernelGGL(k_part_scatter<8192>, sgrid, dim3(kScatterThreads), cur_bytes, ctx->stream, dplan);
  } else {
    const size_t cur_bytes = part_scatter_lds(P.n_parts, kScatterChunk);
    if ((st = set_lds(k_part_scatter<kScatterChunk>, cur_bytes, error))) return st;
    hipLaunchKernelGGL(k_part_scatter<kScatterChunk>, sgrid, dim3(kScatterThreads), cur_bytes, ctx->stream, dplan);
  }
  ctx->obs.launches += 3;
  if (!P.two_level) {
    void *args[] = {(void *)&dplan};
    const int pb = ctx->tune.part_block;
    if (!jit_launch(ctx, C, JIT_PART, P.n_parts, pb, C.c_lds_bytes, args)) {
      if ((st = set_lds(k_part_aggregate, C.c_lds_bytes, error))) return st;
      hipLaunchKernelGGL(k_part_aggregate, dim3(P.n_parts), dim3(pb), C.c_lds_bytes, ctx->stream, dplan);
    }
    ctx->obs.launches += 1;
    if (P.hot_image && (st = launch_hot_reduce(ctx, C, dplan, error))) return st;
  }
  HIP_TRY(hipGetLastError());
  return DFR_OK;
}

// Sub-partitions for `records` first-level records (raw + partial): on
// average sub_fill percent of a table's entries each, twice as many
// sub-partitions per retry (boost).
void second_level_layout(dfr_context *ctx, Compiled &C, uint64_t records, int boost) {
  DevPlan &P = C.plan;
  const uint64_t NP = (uint64_t)P.n_parts;
  const uint64_t per_sub = std::max<uint64_t>(1, (uint64_t)P.c_entries * (uint64_t)ctx->tune.sub_fill / 100);
  uint64_t ns = (std::max<uint64_t>(records, 1) + NP * per_sub - 1) / (NP * per_sub);
  // (the second level's LDS cursors, 8 bytes per sub-partition, stay
  // within 64 KiB)
  ns = std::min<uint64_t>(std::max<uint64_t>(ns << boost, 2), 8192);
  P.n_sub = (int32_t)ns;
  // chunks of about 8192 records, at most 64 per partition, and a
  // partition's (sub-partition, chunk) counters within 64 KiB of LDS
  uint64_t ch = next_pow2(std::max<uint64_t>(1, records / (NP * 8192)));
  ch = std::min<uint64_t>(ch, 64);
  while (ch > 1 && ns * ch > 16384) ch >>= 1;
  P.sub_chunks = (int32_t)ch;
}

// The second level after the first scatter: per-chunk sub-partition counts,
// their scans, the scatter into sub-partition order, then pass C per
// sub-partition.
// Whether the second-level pass C compiled at run time for this plan (its
// fused form: each group's output row written by pass C) keeps more than a
// few spill slots in scratch memory (kSubSpillBytes per lane). Such a kernel
// can run many times slower than placing the groups in dense tables for the
// finalize: when a register copy of the plan ended in scratch, 200M Zipf
// rows with 8 aggregates took 63 ms in this pass. False when the
// precompiled kernel would run.
constexpr int kSubSpillBytes = 256;
bool sub_fused_spills(dfr_context *ctx, const Compiled &C) {
  const Tuning &T = ctx->tune;
  if (!(T.jit == 2 || (T.jit == 1 && (C.rows >= T.jit_rows || C.jit_any_size)))) return false;
  int ldsm, nc, u;
  scan_shape(ctx, C, &ldsm, &nc, &u);
  hipFunction_t fn = jit_function(ctx, C.plan, ldsm, nc, u, JIT_SUB);
  if (!fn) return false;
  int local = 0;
  if (hipFuncGetAttribute(&local, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return local > kSubSpillBytes;
}

int launch_second_level(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  const DevPlan &P = C.plan;
  int st;
  const uint64_t NS = (uint64_t)P.n_sub;
  const uint64_t N2 = (uint64_t)P.n_parts * NS * (uint64_t)P.sub_chunks;
  const dim3 grid2(P.sub_chunks, P.n_parts, 2);
  const size_t h2_bytes = NS * (uint64_t)P.sub_chunks * 4;
  if ((st = set_lds(k_sub_hist, h2_bytes, error))) return st;
  hipLaunchKernelGGL(k_sub_hist, dim3(P.n_parts, 2), dim3(1024), h2_bytes, ctx->stream, dplan);
  const uint64_t nb = (N2 + kScanChunk - 1) / kScanChunk;
  for (int which = 0; which < 2; ++which) {
    uint64_t *bs = which ? C.scan2_par : C.scan2_raw;
    const uint32_t *h = which ? P.hist2_par : P.hist2_raw;
    uint64_t *b2 = which ? P.base2_par : P.base2_raw;
    hipLaunchKernelGGL(k_count_blocks, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, h, N2, bs);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, ctx->stream, bs, nb);
    hipLaunchKernelGGL(k_count_apply, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, h, N2, bs, b2);
  }
  // raw records staged through LDS when they fit (partial states gather)
  const size_t staged = sub_scatter_staged_lds((int)NS, P.rw);
  const int stage = ctx->tune.scatter_stage && staged_it(P.rw) > 0 && staged + 1024 <= (size_t)ctx->lds_per_block;
  const size_t sc_bytes = std::max(sub_scatter_lds((int)NS), stage ? staged : (size_t)0);
  if ((st = set_lds(k_sub_scatter, sc_bytes, error))) return st;
  hipLaunchKernelGGL(k_sub_scatter, grid2, dim3(kScatterThreads), sc_bytes, ctx->stream, dplan, stage);
  void *args[] = {(void *)&dplan};
  const uint64_t n_sub_parts = (uint64_t)P.n_parts * NS;
  if (!jit_launch(ctx, C, JIT_SUB, 0, ctx->tune.sub_block, C.c_lds_bytes, args, n_sub_parts)) {
    if ((st = set_lds(k_sub_aggregate, C.c_lds_bytes, error))) return st;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_sub_aggregate, ctx->tune.sub_block, C.c_lds_bytes) !=
        hipSuccess) {
      (void)hipGetLastError();
      occ = 1;
    }
    const unsigned grid = persistent_grid(ctx, ctx->tune.sub_wg_per_cu, occ, n_sub_parts);
    hipLaunchKernelGGL(k_sub_aggregate, dim3(grid), dim3(ctx->tune.sub_block), C.c_lds_bytes, ctx->stream, dplan);
  }
  ctx->obs.launches += 9;
  HIP_TRY(hipGetLastError());
  return DFR_OK;
}

template <int LDSM, int NC>
int launch_aggregate_nc(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  int st;
  const size_t lds = LDSM == LDS_NONE ? 0 : C.lds_bytes;
  if ((st = set_lds(k_aggregate<LDSM, NC>, lds, error))) return st;
  hipLaunchKernelGGL((k_aggregate<LDSM, NC>), dim3(C.grid), dim3(C.block), lds, ctx->stream, dplan, (int64_t)0,
                     C.rows, C.chunk);
  return DFR_OK;
}

template <int LDSM>
int launch_aggregate_mode(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  if (C.plan.n_cols <= 4) return launch_aggregate_nc<LDSM, 4>(ctx, C, dplan, error);
  return launch_aggregate_nc<LDSM, kMaxCols>(ctx, C, dplan, error);
}

int launch_aggregate(dfr_context *ctx, const Compiled &C, const DevPlan *dplan, char **error) {
  int st;
  switch (C.plan.lds_mode) {
    case LDS_DIRECT: st = launch_aggregate_mode<LDS_DIRECT>(ctx, C, dplan, error); break;
    case LDS_HASHED: st = launch_aggregate_mode<LDS_HASHED>(ctx, C, dplan, error); break;
    default: st = launch_aggregate_mode<LDS_NONE>(ctx, C, dplan, error);
  }
  if (st) return st;
  ctx->obs.launches++;
  HIP_TRY(hipGetLastError());
  return DFR_OK;
}

}  // namespace

// Pinned staging bytes: a plan on its way to the device (hipMemcpyAsync reads
// it when the stream reaches the copy, which can be after the host has
// enqueued the whole query) in [0, kReadback); every read-back (group
// count, fault flags, record totals) at kReadback or later. The host writes
// the plan bytes only after a stream synchronisation.
constexpr size_t kReadback = (sizeof(DevPlan) + 255) / 256 * 256;
static_assert(kReadback + 64 + 32 <= kStagingBytes, "staging holds a plan and its read-backs");
static uint64_t *readback(dfr_context *ctx) { return (uint64_t *)(ctx->staging + kReadback); }

int run_query(dfr_context *ctx, Compiled &C, dfr_result **out, char **error) {
  DevPlan &P = C.plan;
  int64_t hash_slots = ctx->tune.hash_slots;
  int sub_boost = 0;
  bool no_fuse = false;  // a fused one-level pass C's table overflowed: run unfused
  for (int attempt = 0;; ++attempt) {
    choose_layout(ctx, C, hash_slots, sub_boost);
    const bool dense = P.two_level != 0;
    // unordered second level: pass C writes the output rows itself, so
    // there are no state or key tables and no finalize pass
    bool fused = dense && !C.maintain_order;
    // one level: pass C and the seeded keys' merge put each group at the
    // next slot of dense key and state tables sized by the most groups the
    // query can have (its packed key domain, at most its rows): no table
    // set-up, occupancy scan or compaction; the finalize reads the tables
    // in slot order, or ordered, in the first-row order of those slots
    const uint64_t one_cap = std::min<uint64_t>(P.domain, (uint64_t)std::max<int64_t>(C.rows, 0));
    const bool fused1 = !dense && P.agg_mode == AGG_PARTITION && C.rows > 0 &&
                        ctx->tune.fuse_one_level && !no_fuse && one_cap > 0 && one_cap <= ((uint64_t)1 << 28) &&
                        one_cap * (uint64_t)(P.words + 1) * 8 <=
                            std::min<uint64_t>((uint64_t)1 << 31, (uint64_t)ctx->pool.budget / 8);
    if (fused1) {
      P.mode = MODE_HASH;  // (keys from keys_tab: the finalize decodes them there)
      P.cap = one_cap;
    }
    P.fused = 0;
    Scratch S(ctx);
    int st;
    DevPlan *dplan;
    uint32_t *fault;
    dfr_result *r = nullptr;  // released on every early return
    struct Release {
      dfr_result *&r;
      ~Release() {
        if (r) dfr_result_release(r);
      }
    } release{r};
    // the result's columns (and first rows when ordered) for n groups at
    // most, in one pool block; their addresses go into the plan
    auto alloc_result = [&](uint64_t n) -> int {
      r = new dfr_result();
      r->ctx = ctx;
      r->groups = (int64_t)n;
      r->key_dtypes = C.key_dtypes;
      r->agg_dtypes = C.agg_dtypes;
      const uint64_t vwords = (n + 63) / 64 + 1;
      size_t bytes = 0;
      auto col_bytes = [&](int dtype) {
        size_t b = (size_t)n * ((dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8);
        return (b + 255) / 256 * 256 + vwords * 8;
      };
      for (int32_t d : C.key_dtypes) bytes += col_bytes(d);
      for (int32_t d : C.agg_dtypes) bytes += col_bytes(d);
      if (C.maintain_order) bytes += (size_t)n * 8 + 256;
      int st2 = ctx->pool.alloc(bytes, &r->buffer, error);
      if (st2) {
        delete r;
        r = nullptr;
        return st2;
      }
      r->bytes = bytes;
      char *cur = (char *)r->buffer;
      auto take = [&](int dtype, void **vals, uint64_t **valid) {
        size_t b = (size_t)n * ((dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8);
        *vals = cur;
        cur += (b + 255) / 256 * 256;
        *valid = (uint64_t *)cur;
        cur += vwords * 8;
      };
      for (size_t k = 0; k < C.key_dtypes.size(); ++k) {
        void *v;
        uint64_t *m;
        take(C.key_dtypes[k], &v, &m);
        r->key_values.push_back(v);
        r->key_valid.push_back(m);
        P.out_keys[k] = v;
        P.out_key_valid[k] = m;
      }
      for (size_t a = 0; a < C.agg_dtypes.size(); ++a) {
        void *v;
        uint64_t *m;
        take(C.agg_dtypes[a], &v, &m);
        r->agg_values.push_back(v);
        r->agg_valid.push_back(m);
        P.aggs[a].out_values = v;
        P.aggs[a].out_valid = m;
      }
      if (C.maintain_order) {
        r->first_rows = (uint64_t *)cur;
        P.out_first = r->first_rows;
      } else {
        P.out_first = nullptr;
      }
      return DFR_OK;
    };
    if ((st = S.get(1, &dplan, error)) || (st = S.get(4, &fault, error))) return st;
    P.state = nullptr;
    P.keys_tab = nullptr;
    if (!fused) {
      if ((st = S.get(P.cap * (uint64_t)P.words, &P.state, error))) return st;
      if (P.mode == MODE_HASH && (st = S.get(P.cap, &P.keys_tab, error))) return st;
    }
    if (P.agg_mode == AGG_PARTITION) {
      const uint64_t S_ = (uint64_t)P.n_segs, NP = (uint64_t)P.n_parts, W = (uint64_t)P.words;
      const uint64_t raw_words = S_ * (uint64_t)P.seg_waves * (uint64_t)P.seg_rows * (uint64_t)P.rw;
      const uint64_t par_words = S_ * (uint64_t)P.seg_partials * (1 + W);
      if ((st = S.get(raw_words, &P.spill, error)) || (st = S.get(raw_words, &P.part_raw, error)) ||
          (st = S.get(par_words, &P.partials, error)) || (st = S.get(par_words, &P.part_par, error)) ||
          (st = S.get(S_ * (uint64_t)P.seg_slices * NP, &P.hist_raw, error)) ||
          (st = S.get(S_ * NP, &P.hist_par, error)) ||
          (st = S.get(S_ * (uint64_t)P.seg_waves, &P.cnt_raw, error)) || (st = S.get(S_, &P.cnt_par, error)) ||
          (st = S.get(NP + 1, &P.base_raw, error)) || (st = S.get(NP + 1, &P.base_par, error)))
        return st;
      P.hot_image = P.hot_states = nullptr;
      // a sole integer or dictionary key with a hand-off list, else a sample
      const dfr_column *kc = P.n_keys == 1 && !P.keys[0].is_float ? C.cols[0] : nullptr;
      const bool col_list = kc && kc->hot;
      if (hot_wanted(ctx, C, dense, col_list)) {
        const uint64_t E = (uint64_t)P.lds_entries;
        C.hot_sampled = !col_list;
        if (col_list) {
          C.hot_list = kc->hot;
          C.hot_hist = (uint32_t *)(kc->hot + kHotList + 1);
          C.hot_delta = kc->hot_min - P.keys[0].minv;
          C.hot_limit = P.keys[0].radix - (uint64_t)P.keys[0].nullable;
        } else {
          C.hot_delta = 0;
          C.hot_limit = P.domain;
          if ((st = S.get(P.domain + 2, &C.hot_counts, error)) || (st = S.get(kHotBins + 2, &C.hot_hist, error)) ||
              (st = S.get(kHotList + 1, &C.hot_list, error)))
            return st;
        }
        // (a fused query merges the seeded states into one more segment)
        if ((st = S.get(1, &C.hot_out, error)) || (st = S.get(E, &P.hot_image, error)) ||
            (st = S.get((S_ + (fused1 ? 1 : 0)) * E * W, &P.hot_states, error)))
          return st;
      }
      if (fused1 && (st = S.get(1, &P.gcount, error))) return st;
      if (dense) {
        // second level: the spill segments are dead after the first
        // scatter and take the sub-partition order
        if ((st = S.get(1, &P.gcount, error))) return st;
        P.part2_raw = P.spill;
        P.part2_par = P.partials;
      }
    }
    P.fault = fault;
    const uint64_t occ_words = (P.cap + 63) / 64;
    uint64_t *occ = nullptr, *occ_prefix = nullptr, *order = nullptr;
    const uint64_t max_groups = std::min<uint64_t>(P.cap, (uint64_t)std::max<int64_t>(C.rows, 1));
    if (!dense && !fused1 &&
        ((st = S.get(occ_words, &occ, error)) || (st = S.get(occ_words, &occ_prefix, error)) ||
         (st = S.get(max_groups + 1, &order, error))))
      return st;
    if (fused1) P.fused = 1;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    memcpy(ctx->staging, &P, sizeof P);
    HIP_TRY(hipMemcpyAsync(dplan, ctx->staging, sizeof P, hipMemcpyHostToDevice, ctx->stream));
    fill(ctx, (uint64_t *)fault, 2, 0);
    if (fused1 || dense) {
      fill(ctx, P.gcount, 1, 0);  // dense tables need no set-up: pass C writes every word
    } else {
      int blocks = (int)std::min<uint64_t>((P.cap + kBlock - 1) / kBlock, (uint64_t)ctx->cus * 8);
      hipLaunchKernelGGL(k_init_state, dim3(std::max(blocks, 1)), dim3(kBlock), 0, ctx->stream, dplan);
      ctx->obs.launches++;
    }
    HIP_TRY(hipEventRecord(ctx->evp[0], ctx->stream));
    if (P.agg_mode == AGG_PARTITION) {
      if (C.rows > 0 && (st = launch_partitioned(ctx, C, dplan, error))) return st;
      if (C.rows == 0) HIP_TRY(hipEventRecord(ctx->evp[1], ctx->stream));
      if (dense && C.rows > 0) {
        // the first level's record totals size the second level
        uint64_t *tot = readback(ctx);
        HIP_TRY(hipMemcpyAsync(tot, P.base_raw + P.n_parts, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(tot + 1, P.base_par + P.n_parts, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        ctx->obs.waits++;
        second_level_layout(ctx, C, tot[0] + tot[1], sub_boost);
        const uint64_t N2 = (uint64_t)P.n_parts * (uint64_t)P.n_sub * (uint64_t)P.sub_chunks;
        const uint64_t nb = (N2 + kScanChunk - 1) / kScanChunk;
        if ((st = S.get(N2, &P.hist2_raw, error)) || (st = S.get(N2, &P.hist2_par, error)) ||
            (st = S.get(N2 + 1, &P.base2_raw, error)) || (st = S.get(N2 + 1, &P.base2_par, error)) ||
            (st = S.get(nb + 1, &C.scan2_raw, error)) || (st = S.get(nb + 1, &C.scan2_par, error)))
          return st;
        if (fused) {
          P.fused = 1;
          if (sub_fused_spills(ctx, C)) {
            // place the groups in dense tables and finish them afterwards
            P.fused = 0;
            fused = false;
            if ((st = S.get(P.cap * (uint64_t)P.words, &P.state, error)) ||
                (st = S.get(P.cap, &P.keys_tab, error)))
              return st;
          }
        }
        if (fused) {
          // every group has a record: the record total bounds the groups
          const uint64_t records = tot[0] + tot[1];
          if ((st = alloc_result(std::min<uint64_t>(records, P.cap)))) return st;
          for (int k = 0; k < P.n_keys; ++k) fill(ctx, P.out_key_valid[k], ((uint64_t)r->groups + 63) / 64 + 1, ~0ull);
          for (int a = 0; a < P.n_aggs; ++a) fill(ctx, P.aggs[a].out_valid, ((uint64_t)r->groups + 63) / 64 + 1, ~0ull);
        }
        memcpy(ctx->staging, &P, sizeof P);
        HIP_TRY(hipMemcpyAsync(dplan, ctx->staging, sizeof P, hipMemcpyHostToDevice, ctx->stream));
        if ((st = launch_second_level(ctx, C, dplan, error))) return st;
      }
    } else {
      if (C.rows > 0 && (st = launch_aggregate(ctx, C, dplan, error))) return st;
      HIP_TRY(hipEventRecord(ctx->evp[1], ctx->stream));
    }
    uint64_t *total;
    if (dense || fused1) {
      total = P.gcount;  // slots 0..G-1, in slot order
    } else {
      uint64_t threads = occ_words * 64;
      hipLaunchKernelGGL(k_occupancy, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         ctx->stream, dplan, occ, occ_words);
      ctx->obs.launches++;
      if ((st = scan_bits(ctx, S, occ, occ_words, occ_prefix, &total, error))) return st;
      hipLaunchKernelGGL(k_compact, dim3((unsigned)((P.cap + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         ctx->stream, dplan, occ, occ_prefix, order);
      ctx->obs.launches++;
    }
    uint64_t *hs = readback(ctx);
    HIP_TRY(hipMemcpyAsync(hs, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(hs + 1, fault, 16, hipMemcpyDeviceToHost, ctx->stream));
    hs[3] = 0;
    if (P.hot_image && C.rows > 0) HIP_TRY(hipMemcpyAsync(hs + 3, C.hot_out, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->obs.waits++;
    ctx->obs.completed = ctx->obs.launches;
    ctx->obs.hot_keys += (int64_t)(uint32_t)hs[3];
    const uint64_t G = hs[0];
    const uint32_t *hf = (const uint32_t *)(hs + 1);
    if (hf[3] && fused1) {
      // a partition's keys outgrew its LDS table with no HBM table behind it
      no_fuse = true;
      continue;  // (the guard releases this attempt's result)
    }
    if (hf[3]) {
      // a sub-partition outgrew its LDS table: twice the sub-partitions
      if (sub_boost >= 6 || P.n_sub >= 8192)
        return dfr_fail(error, DFR_CAPACITY, "group table capacity exceeded");
      ++sub_boost;
      continue;  // (the guard releases this attempt's result)
    }
    if (hf[0]) {
      // a full hash table: retry once at load <= 1/2 of every possible group
      uint64_t need = next_pow2(std::max<uint64_t>(64, 2 * std::min<uint64_t>(P.domain, (uint64_t)C.rows)));
      if (P.mode != MODE_HASH || attempt >= 4 || P.cap >= need)
        return dfr_fail(error, DFR_CAPACITY, "group table capacity exceeded");
      hash_slots = (int64_t)std::max<uint64_t>(P.cap * 4, need);
      continue;
    }
    P.groups = total;
    P.order = (dense || fused1) ? nullptr : order;
    if (C.maintain_order && G > 0) {
      const uint64_t rwords = ((uint64_t)C.rows + 63) / 64;
      uint64_t *rowbits, *rowprefix, *order2;
      if ((st = S.get(rwords, &rowbits, error)) || (st = S.get(rwords, &rowprefix, error)) ||
          (st = S.get(G + 1, &order2, error)))
        return st;
      fill(ctx, rowbits, rwords, 0);
      unsigned gb = (unsigned)((G + kBlock - 1) / kBlock);
      hipLaunchKernelGGL(k_mark_first, dim3(gb), dim3(kBlock), 0, ctx->stream, dplan, P.order, G, rowbits);
      uint64_t *rtotal;
      if ((st = scan_bits(ctx, S, rowbits, rwords, rowprefix, &rtotal, error))) return st;
      hipLaunchKernelGGL(k_rank_first, dim3(gb), dim3(kBlock), 0, ctx->stream, dplan, P.order, G, rowbits,
                         rowprefix, order2);
      ctx->obs.launches += 2;
      P.order = order2;
    }
    HIP_TRY(hipEventRecord(ctx->evp[2], ctx->stream));
    if (!fused) {
      if ((st = alloc_result(G))) return st;
      memcpy(ctx->staging, &P, sizeof P);
      if (G > 0) {
        HIP_TRY(hipMemcpyAsync(dplan, ctx->staging, sizeof P, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_finalize, dim3((unsigned)((G + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           ctx->stream, dplan);
        ctx->obs.launches++;
      }
    } else {
      if (!r && (st = alloc_result(0))) return st;  // no rows: no pass C ran
      r->groups = (int64_t)G;
      if (G > 0) {
        hipLaunchKernelGGL(k_clear_tail, dim3(1), dim3(64), 0, ctx->stream, dplan, G);
        ctx->obs.launches++;
      }
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t *hf2 = (uint32_t *)(readback(ctx) + 8);
    HIP_TRY(hipMemcpyAsync(hf2, fault, 16, hipMemcpyDeviceToHost, ctx->stream));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    ctx->obs.waits++;
    if (e != hipSuccess) return dfr_fail(error, DFR_DEVICE, "group-by kernels: %s", hipGetErrorString(e));
    ctx->obs.completed = ctx->obs.launches;
    ctx->obs.queries++;
    float ms = 0;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    r->device_ms = ms;
    hipEvent_t marks[5] = {ctx->ev0, ctx->evp[0], ctx->evp[1], ctx->evp[2], ctx->ev1};
    for (int ph = 0; ph < 4; ++ph) {
      float t = 0;
      hipEventElapsedTime(&t, marks[ph], marks[ph + 1]);
      r->phase_ms[ph] = t;
    }
    if (hf2[1]) return dfr_fail(error, DFR_DATA_FAULT, "Int64 expression sum overflow");
    if (hf2[2]) return dfr_fail(error, DFR_DATA_FAULT, "int32 expression sum overflow");
    *out = r;
    r = nullptr;
    return DFR_OK;
  }
}

extern "C" int dfr_groupby(dfr_context *ctx, const dfr_query *query, dfr_result **out, char **error) {
  if (!ctx || !out) return dfr_fail(error, DFR_INVALID, "null argument");
  HIP_TRY(hipSetDevice(ctx->device));
  Compiled C;
  int st = compile_query(ctx, query, C, error);
  if (st) return st;
  return run_query(ctx, C, out, error);
}

#include "batch.h"
#include "noalloc.h"
#include "values.h"
