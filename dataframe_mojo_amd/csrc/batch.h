// batch.h — many filtered group-by queries in one persistent launch
// (dfr_groupby_batch). Included at the end of groupby.hip: it shares that
// file's query compiler, scratch pool and scan helpers.
//
// Launch sequence of a batch (no host synchronisation until the end, so the
// whole sequence can be captured as one HIP graph):
//   k_batch_init       every query's group table set to its identity words
//   k_batch_aggregate  persistent workgroups drain a device work queue of
//                      (query, row tile) items: each item folds its rows
//                      into an LDS table laid out for that query, then
//                      merges the table into the query's HBM table
//   k_batch_occupancy  occupied slots of every table as one bit array (each
//                      query's slots start at a multiple of 64)
//   scan               one exclusive popcount scan over all of it
//   k_batch_counts     groups per query, and padded output offsets
//   k_batch_compact    slot of each group, per query
//   k_batch_finalize   every group of every query finished into its result
// Output buffers are sized by each query's table capacity, so they exist
// before the group counts do.
#pragma once

namespace {

struct WorkItem {
  int32_t q;
  int32_t pad;
  int64_t b0, b1;
};

__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// The per-table kernels (init, occupancy, compaction) give each query
// blocks of kSlotSpan table slots: query q owns blocks [boff[q], boff[q + 1])
// (at least one), so a table of 1e5 slots is not left to one workgroup.
constexpr uint64_t kSlotSpan = 8 * kBlock;  // a multiple of 64 and of kBlock

// The query of this block (the last q with boff[q] <= blockIdx.x) and the
// first slot of its span.
__device__ __forceinline__ int span_query(const uint64_t *__restrict__ boff, int n, uint64_t *s0) {
  const uint64_t b = blockIdx.x;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (boff[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  *s0 = (b - boff[lo]) * kSlotSpan;
  return __builtin_amdgcn_readfirstlane(lo);
}

// A span of a query's table set to its identity words (and empty keys).
__global__ void k_batch_init(const DevPlan *__restrict__ plans, const uint64_t *__restrict__ boff, int n) {
  uint64_t s0;
  const int q = span_query(boff, n, &s0);
  PlanRef P = plan_ref(plans + q);
  if (s0 >= P.cap) return;
  const uint64_t s1 = min(P.cap, s0 + kSlotSpan);
  const uint32_t W = (uint32_t)P.words;
  const uint64_t total = (s1 - s0) * W;
  uint64_t *base = P.state + s0 * W;  // slot-aligned: word j of the span is word j % W of its slot
  uint32_t w = threadIdx.x % W;
  const uint32_t stride_w = blockDim.x % W;
  for (uint64_t j = threadIdx.x; j < total; j += blockDim.x) {
    base[j] = P.ident[w];
    w += stride_w;
    if (w >= W) w -= W;
  }
  if (P.mode == MODE_HASH)
    for (uint64_t j = s0 + threadIdx.x; j < s1; j += blockDim.x) P.keys_tab[j] = kEmpty;
}

// Persistent workgroups: each takes the next work item until the queue is
// drained. Every wave reaches the exit test after each item.
template <int NC>
__global__ void __launch_bounds__(512) k_batch_aggregate(const DevPlan *__restrict__ plans,
                                                          const WorkItem *__restrict__ items, int64_t n_items,
                                                          unsigned long long *next) {
  extern __shared__ uint64_t smem[];
  __shared__ int64_t s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = (int64_t)atomicAdd(next, 1ull);
    __syncthreads();
    const int64_t it = uniform_i64(s_item);
    if (it >= n_items) break;
    const int q = __builtin_amdgcn_readfirstlane(items[it].q);
    const int64_t b0 = uniform_i64(items[it].b0), b1 = uniform_i64(items[it].b1);
    PlanRef P = plan_ref(plans + q);
    switch (P.lds_mode) {
      case LDS_DIRECT: aggregate_tile<LDS_DIRECT, NC>(P, smem, b0, b1); break;
      case LDS_HASHED: aggregate_tile<LDS_HASHED, NC>(P, smem, b0, b1); break;
      default: aggregate_tile<LDS_NONE, NC>(P, smem, b0, b1);
    }
    __syncthreads();  // the LDS table and s_item are reused by the next item
  }
}

// A span's occupancy bits, at word woff[q] of query q's bit array onward.
__global__ void k_batch_occupancy(const DevPlan *__restrict__ plans, const uint64_t *__restrict__ woff,
                                  const uint64_t *__restrict__ boff, int n, uint64_t *bits) {
  uint64_t s0;
  const int q = span_query(boff, n, &s0);
  PlanRef P = plan_ref(plans + q);
  const uint64_t nw = (P.cap + 63) / 64;
  const uint64_t w_end = min(nw, (s0 + kSlotSpan) >> 6);  // this span's words end here
  uint64_t *out = bits + woff[q];
  for (uint64_t sb = s0; (sb >> 6) < w_end; sb += blockDim.x) {
    const uint64_t s = sb + threadIdx.x;
    const bool occ = s < P.cap && P.state[s * (uint64_t)P.words] != 0;
    const uint64_t ball = __ballot(occ);
    if ((threadIdx.x & 63) == 0 && (s >> 6) < w_end) out[s >> 6] = ball;
  }
}

// Thread q: groups of query q (from the prefix at its first word and the
// next query's), its first global rank, and its output span padded to 64.
__global__ void k_batch_counts(const uint64_t *__restrict__ woff, const uint64_t *__restrict__ prefix, int n,
                               uint64_t *groups, uint64_t *gbase, uint64_t *pad_off) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const uint64_t r0 = prefix[woff[q]], r1 = prefix[woff[q + 1]];
  groups[q] = r1 - r0;
  gbase[q] = r0;
  pad_off[q] = (r1 - r0 + 63) / 64 * 64;
}

// order[global rank] = local slot of each occupied slot of a span; query
// q's groups take ranks gbase[q] .. gbase[q] + groups[q] - 1.
__global__ void k_batch_compact(const DevPlan *__restrict__ plans, const uint64_t *__restrict__ woff,
                                const uint64_t *__restrict__ boff, int n, const uint64_t *__restrict__ bits,
                                const uint64_t *__restrict__ prefix, uint64_t *order) {
  uint64_t s0;
  const int q = span_query(boff, n, &s0);
  PlanRef P = plan_ref(plans + q);
  const uint64_t w0 = woff[q];
  const uint64_t s1 = min(P.cap, s0 + kSlotSpan);
  for (uint64_t s = s0 + threadIdx.x; s < s1; s += blockDim.x) {
    const uint64_t w = bits[w0 + (s >> 6)];
    if (!((w >> (s & 63)) & 1)) continue;
    const uint64_t below = (s & 63) ? (w & ((1ull << (s & 63)) - 1)) : 0;
    order[prefix[w0 + (s >> 6)] + __popcll(below)] = s;  // global rank
  }
}

// Thread t of the padded output space: query q owns [pad_off[q],
// pad_off[q + 1]), a multiple of 64 wide, so a wave serves one query.
__global__ void k_batch_finalize(const DevPlan *__restrict__ plans, int n, const uint64_t *__restrict__ pad_off,
                                 const uint64_t *__restrict__ groups, const uint64_t *__restrict__ gbase,
                                 const uint64_t *__restrict__ order) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t t0 = t & ~63ull;  // the wave's first index: the same query for the whole wave
  if (t0 >= pad_off[n]) return;
  int lo = 0, hi = n - 1;  // last q with pad_off[q] <= t0
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pad_off[mid] <= t0) lo = mid;
    else hi = mid - 1;
  }
  const int q = __builtin_amdgcn_readfirstlane(lo);
  PlanRef P = plan_ref(plans + q);
  const uint64_t G = groups[q];
  const uint64_t i = t - pad_off[q];
  const bool active = i < G;
  const uint64_t s = active ? order[gbase[q] + i] : 0;
  finalize_group(P, i, s, active, G);
}

}  // namespace

// ------------------------------------------------------------------ host

namespace {

struct BatchBlock {
  dfr_context *ctx;
  void *p;
  ~BatchBlock() {
    if (p) ctx->pool.release(p);
  }
};

int ensure_batch_staging(dfr_context *ctx, size_t bytes, char **error) {
  if (ctx->bstaging_bytes >= bytes) return DFR_OK;
  if (ctx->bstaging) hipHostFree(ctx->bstaging);
  ctx->bstaging = nullptr;
  ctx->bstaging_bytes = 0;
  size_t want = std::max<size_t>(bytes, (size_t)1 << 24);
  HIP_TRY(hipHostMalloc((void **)&ctx->bstaging, want, hipHostMallocDefault));
  ctx->bstaging_bytes = want;
  return DFR_OK;
}

// The context's first n side streams and the fork and join events.
int ensure_side_streams(dfr_context *ctx, int n, char **error) {
  for (int i = 0; i <= n; ++i)
    if (!ctx->side_ev[i]) HIP_TRY(hipEventCreateWithFlags(&ctx->side_ev[i], hipEventDisableTiming));
  for (int i = 0; i < n; ++i)
    if (!ctx->side[i]) HIP_TRY(hipStreamCreateWithFlags(&ctx->side[i], hipStreamNonBlocking));
  return DFR_OK;
}

size_t out_col_bytes(int dtype, uint64_t cap) {
  const size_t b = (size_t)cap * ((dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8);
  return (b + 255) / 256 * 256 + ((cap + 63) / 64 + 1) * 8;
}

}  // namespace

extern "C" int dfr_groupby_batch(dfr_context *ctx, const dfr_query *queries, int32_t n, dfr_result **results,
                                 int32_t *statuses, char **errors, char **error) {
  if (!ctx || (n > 0 && (!queries || !results || !statuses))) return dfr_fail(error, DFR_INVALID, "null argument");
  if (n < 0) return dfr_fail(error, DFR_INVALID, "negative query count");
  HIP_TRY(hipSetDevice(ctx->device));
  for (int i = 0; i < n; ++i) {
    results[i] = nullptr;
    statuses[i] = DFR_OK;
    if (errors) errors[i] = nullptr;
  }
  if (n == 0) return DFR_OK;
  // compile every query for the atomic table path (one pass, no spill)
  std::vector<Compiled> C(n);
  std::vector<int> live;
  Tuning atomic_tune = ctx->tune;
  atomic_tune.agg_mode = AGG_ATOMIC;
  for (int i = 0; i < n; ++i) {
    char *msg = nullptr;
    int st = compile_query(ctx, &queries[i], C[i], &msg);
    if (st == DFR_OK && queries[i].maintain_order)
      st = dfr_fail(&msg, DFR_UNSUPPORTED, "maintain_order runs through dfr_groupby, not in a batch");
    if (st) {
      statuses[i] = st;
      if (errors) errors[i] = msg;
      else free(msg);
      continue;
    }
    choose_layout(ctx, C[i], ctx->tune.hash_slots, 0, &atomic_tune);
    live.push_back(i);
  }
  const int m = (int)live.size();
  if (m == 0) return DFR_OK;
  // Large queries whose groups outgrow the batch kernel's LDS table (their
  // rows would fold into HBM with device atomics) take the partitioned
  // passes instead, launched after the batch kernel into the same tables.
  // Under graph capture their runtime-compiled kernels are built first,
  // before the capture begins (a compile cannot be captured). On
  // batch_4096 (rows >= 4M, 4096 queries) this takes 172 ms to 140 ms.
  const bool graph = ctx->tune.batch_graph != 0;
  std::vector<char> parted(m, 0);
  uint64_t part_bytes = 0;
  auto part_sizes = [](const DevPlan &P, uint64_t *w) {
    const uint64_t S_ = (uint64_t)P.n_segs, NP = (uint64_t)P.n_parts, W = (uint64_t)P.words;
    w[0] = S_ * (uint64_t)P.seg_waves * (uint64_t)P.seg_rows * (uint64_t)P.rw;  // spill, part_raw
    w[1] = S_ * (uint64_t)P.seg_partials * (1 + W);                             // partials, part_par
    w[2] = S_ * (uint64_t)P.seg_slices * NP;                                   // hist_raw (u32)
    w[3] = S_ * NP;                                                             // hist_par (u32)
    w[4] = S_ * (uint64_t)P.seg_waves;                                          // cnt_raw (u32)
    w[5] = S_;                                                                  // cnt_par (u32)
    w[6] = NP + 1;                                                              // base_raw, base_par
  };
  auto bytes_of = [](const uint64_t *w) {
    auto r = [](uint64_t b) { return (b + 255) / 256 * 256; };
    return 2 * r(w[0] * 8) + 2 * r(w[1] * 8) + r(w[2] * 4) + r(w[3] * 4) + r(w[4] * 4) + r(w[5] * 4) +
           2 * r(w[6] * 8);
  };
  if (ctx->tune.batch_part_rows > 0) {
    for (int j = 0; j < m; ++j) {
      Compiled &c = C[live[j]];
      if (c.plan.lds_mode == LDS_DIRECT || c.rows < ctx->tune.batch_part_rows) continue;
      Compiled t = c;
      choose_layout(ctx, t, ctx->tune.hash_slots, 0);  // the context's (partitioned) tuning
      if (t.plan.agg_mode != AGG_PARTITION || t.plan.two_level) continue;
      if (!ctx->tune.jit_wait || graph) {
        // routed only once the shape's kernels exist, since the precompiled
        // ones lose to the batch kernel: a lane that never waits for a
        // compile (a server's) takes them only when compiled already; a
        // graph capture compiles them now, outside the capture
        int ldsm, nc, u;
        scan_shape(ctx, t, &ldsm, &nc, &u);
        if (graph && ctx->tune.jit_wait && ctx->tune.jit) {
          jit_function(ctx, t.plan, ldsm, nc, u, JIT_SCAN);
          jit_function(ctx, t.plan, ldsm, nc, u, JIT_PART);
        }
        if (!jit_compiled(ctx, t.plan, ldsm, nc, u, JIT_SCAN) || !jit_compiled(ctx, t.plan, ldsm, nc, u, JIT_PART))
          continue;
      }
      // a batch's queries share a few shapes: their runtime-compiled kernels
      // serve every size (the precompiled ones run these sizes slower)
      t.jit_any_size = true;
      uint64_t w[7];
      part_sizes(t.plan, w);
      part_bytes += bytes_of(w);
      c = t;
      parted[j] = 1;
    }
  }

  // device layout: tables, occupancy word offsets, outputs, work items
  Scratch S(ctx);
  int st;
  uint64_t state_words = 0, key_slots = 0, occ_words = 0, out_bytes = 0, pad_total = 0;
  size_t lds_max = 0;
  std::vector<uint64_t> woff(m + 1);
  for (int j = 0; j < m; ++j) {
    const Compiled &c = C[live[j]];
    const DevPlan &P = c.plan;
    state_words += P.cap * (uint64_t)P.words;
    if (P.mode == MODE_HASH) key_slots += P.cap;
    woff[j] = occ_words;
    occ_words += (P.cap + 63) / 64;
    pad_total += (P.cap + 63) / 64 * 64;
    for (int32_t d : c.key_dtypes) out_bytes += out_col_bytes(d, P.cap);
    for (int32_t d : c.agg_dtypes) out_bytes += out_col_bytes(d, P.cap);
    if (!parted[j]) lds_max = std::max(lds_max, c.lds_bytes);
  }
  woff[m] = occ_words;
  // blocks of the per-table kernels, kSlotSpan slots each (woff and boff
  // travel as one array)
  woff.resize(2 * ((size_t)m + 1));
  uint64_t *boff = woff.data() + m + 1;
  uint64_t span_blocks = 0;
  for (int j = 0; j < m; ++j) {
    boff[j] = span_blocks;
    span_blocks += std::max<uint64_t>(1, (C[live[j]].plan.cap + kSlotSpan - 1) / kSlotSpan);
  }
  boff[m] = span_blocks;
  const int block = std::min(512, ctx->tune.block);  // k_batch_aggregate's launch bound
  const int grid = ctx->cus * std::max(1, ctx->tune.wg_per_cu);
  // Work items are at most batch_tile rows. A batch of few rows is cut
  // finer, so that its items still give every workgroup of the grid
  // batch_fill of them (at batch_tile, a 1e6-row query alone would keep 8
  // workgroups busy). A query's item never gets shorter than its LDS
  // table: each item clears the table and merges its occupied entries into
  // HBM once. More items per workgroup lose: every item's merge lands on
  // the same few HBM entries of a low-cardinality query (a 1.8e6-row query
  // of 21 groups: 0.21 ms at one item per workgroup, 0.35 ms at two).
  const int64_t unit = 2 * (int64_t)block;
  const int64_t tile_max = std::max<int64_t>(unit, ctx->tune.batch_tile / unit * unit);
  int64_t rows_all = 0;
  for (int j = 0; j < m; ++j)
    if (!parted[j]) rows_all += C[live[j]].rows;
  const int64_t want = (int64_t)ctx->tune.batch_fill * grid;
  const int64_t tile_fill = want > 0 ? std::min(tile_max, std::max(unit, rows_all / want / unit * unit)) : tile_max;
  std::vector<WorkItem> items;
  for (int j = 0; j < m; ++j) {
    if (parted[j]) continue;
    const DevPlan &P = C[live[j]].plan;
    const int64_t rows = C[live[j]].rows;
    const int64_t table = P.lds_mode == LDS_NONE ? 0 : (P.lds_entries + unit - 1) / unit * unit;
    const int64_t tile = std::min(tile_max, std::max(tile_fill, table));
    for (int64_t b = 0; b < rows; b += tile) items.push_back({j, 0, b, std::min(rows, b + tile)});
  }
  // larger items first: the queue's tail is made of the smallest
  std::stable_sort(items.begin(), items.end(),
                   [](const WorkItem &a, const WorkItem &b) { return a.b1 - a.b0 > b.b1 - b.b0; });

  DevPlan *dplans;
  WorkItem *ditems;
  uint64_t *state, *keys = nullptr, *dwoff, *occ, *prefix, *counts, *gbase, *pad_off, *order, *next;
  uint32_t *faults;
  uint64_t *sums;
  void *outblock = nullptr;
  const uint64_t scan_n = occ_words + 1;
  const uint64_t scan_nb = std::max<uint64_t>(1, (scan_n + kScanChunk - 1) / kScanChunk);
  const uint64_t n_items = items.size();
  if ((st = S.get(m, &dplans, error)) || (st = S.get(std::max<uint64_t>(n_items, 1), &ditems, error)) ||
      (st = S.get(std::max<uint64_t>(state_words, 1), &state, error)) ||
      (key_slots && (st = S.get(key_slots, &keys, error))) || (st = S.get(2 * ((uint64_t)m + 1), &dwoff, error)) ||
      (st = S.get(occ_words + 1, &occ, error)) || (st = S.get(occ_words + 1, &prefix, error)) ||
      (st = S.get(m, &counts, error)) || (st = S.get(m, &gbase, error)) || (st = S.get(m + 1, &pad_off, error)) ||
      (st = S.get(std::max<uint64_t>(pad_total, 1), &order, error)) || (st = S.get(1, &next, error)) ||
      (st = S.get(4 * (uint64_t)m, &faults, error)) || (st = S.get(scan_nb + 1, &sums, error)))
    return st;
  const uint64_t *dboff = dwoff + m + 1;
  char *partblock = nullptr;
  if (part_bytes && (st = S.get(part_bytes, &partblock, error))) return st;
  if ((st = ctx->pool.alloc(std::max<uint64_t>(out_bytes, 256), &outblock, error))) return st;
  // the partitioned queries' spill, partition and histogram buffers
  {
    char *pc = partblock;
    auto carve = [&](uint64_t bytes) {
      char *p = pc;
      pc += (bytes + 255) / 256 * 256;
      return p;
    };
    for (int j = 0; j < m; ++j) {
      if (!parted[j]) continue;
      DevPlan &P = C[live[j]].plan;
      uint64_t w[7];
      part_sizes(P, w);
      P.spill = (uint64_t *)carve(w[0] * 8);
      P.part_raw = (uint64_t *)carve(w[0] * 8);
      P.partials = (uint64_t *)carve(w[1] * 8);
      P.part_par = (uint64_t *)carve(w[1] * 8);
      P.hist_raw = (uint32_t *)carve(w[2] * 4);
      P.hist_par = (uint32_t *)carve(w[3] * 4);
      P.cnt_raw = (uint32_t *)carve(w[4] * 4);
      P.cnt_par = (uint32_t *)carve(w[5] * 4);
      P.base_raw = (uint64_t *)carve(w[6] * 8);
      P.base_par = (uint64_t *)carve(w[6] * 8);
    }
  }
  auto outs = std::make_shared<BatchBlock>();
  outs->ctx = ctx;
  outs->p = outblock;

  // per-query plans pointing into the shared blocks; results alongside
  std::vector<dfr_result *> res(m, nullptr);
  {
    uint64_t so = 0, ko = 0;
    char *cur = (char *)outblock;
    for (int j = 0; j < m; ++j) {
      Compiled &c = C[live[j]];
      DevPlan &P = c.plan;
      P.state = state + so;
      so += P.cap * (uint64_t)P.words;
      P.keys_tab = P.mode == MODE_HASH ? keys + ko : nullptr;
      if (P.mode == MODE_HASH) ko += P.cap;
      P.fault = faults + 4 * (uint64_t)j;
      P.groups = counts + j;
      P.order = nullptr;
      P.out_first = nullptr;
      dfr_result *r = new dfr_result();
      r->ctx = ctx;
      r->key_dtypes = c.key_dtypes;
      r->agg_dtypes = c.agg_dtypes;
      r->shared = outs;
      auto take = [&](int dtype, void **vals, uint64_t **valid) {
        const size_t b = (size_t)P.cap * ((dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8);
        *vals = cur;
        cur += (b + 255) / 256 * 256;
        *valid = (uint64_t *)cur;
        cur += ((P.cap + 63) / 64 + 1) * 8;
        r->bytes += out_col_bytes(dtype, P.cap);
      };
      for (size_t k = 0; k < c.key_dtypes.size(); ++k) {
        void *v;
        uint64_t *vm;
        take(c.key_dtypes[k], &v, &vm);
        r->key_values.push_back(v);
        r->key_valid.push_back(vm);
        P.out_keys[k] = v;
        P.out_key_valid[k] = vm;
      }
      for (size_t a = 0; a < c.agg_dtypes.size(); ++a) {
        void *v;
        uint64_t *vm;
        take(c.agg_dtypes[a], &v, &vm);
        r->agg_values.push_back(v);
        r->agg_valid.push_back(vm);
        P.aggs[a].out_values = v;
        P.aggs[a].out_valid = vm;
      }
      res[j] = r;
    }
  }
  auto drop_all = [&]() {
    for (dfr_result *r : res) delete r;
  };

  // one pinned upload: plans, items, word offsets
  const size_t plan_bytes = sizeof(DevPlan) * (size_t)m, item_bytes = sizeof(WorkItem) * n_items,
               woff_bytes = 16 * (size_t)(m + 1);
  const size_t up_bytes = plan_bytes + item_bytes + woff_bytes;
  const size_t down_bytes = 8 * (size_t)m + 16 * (size_t)m;
  if ((st = ensure_batch_staging(ctx, std::max(up_bytes, down_bytes), error))) {
    drop_all();
    return st;
  }
  {
    char *hb = ctx->bstaging;
    for (int j = 0; j < m; ++j) memcpy(hb + sizeof(DevPlan) * j, &C[live[j]].plan, sizeof(DevPlan));
    memcpy(hb + plan_bytes, items.data(), item_bytes);
    memcpy(hb + plan_bytes + item_bytes, woff.data(), woff_bytes);
  }
  // the widest plan of the batch picks the column count the kernel unrolls
  int nc_max = 0;
  for (int j = 0; j < m; ++j) nc_max = std::max(nc_max, (int)C[live[j]].plan.n_cols);
  auto kagg = nc_max <= 4 ? k_batch_aggregate<4> : k_batch_aggregate<kMaxCols>;
  if ((st = set_lds(kagg, lds_max, error))) {
    drop_all();
    return st;
  }
  // the partitioned queries run on side streams, beside the batch kernel and
  // each other (they write only their own tables and buffers)
  int n_parted = 0;
  for (int j = 0; j < m; ++j) n_parted += parted[j] && C[live[j]].rows > 0;
  const int ns = std::min(std::min(ctx->tune.batch_streams, dfr_context::kSideStreams), n_parted);
  if (ns >= 2 && (st = ensure_side_streams(ctx, ns, error))) {
    drop_all();
    return st;
  }
  hipGraph_t g = nullptr;
  hipGraphExec_t ge = nullptr;
  HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
  if (graph) HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  {
    char *hb = ctx->bstaging;
    hipMemcpyAsync(dplans, hb, plan_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (item_bytes) hipMemcpyAsync(ditems, hb + plan_bytes, item_bytes, hipMemcpyHostToDevice, ctx->stream);
    hipMemcpyAsync(dwoff, hb + plan_bytes + item_bytes, woff_bytes, hipMemcpyHostToDevice, ctx->stream);
    fill(ctx, next, 1, 0);
    fill(ctx, (uint64_t *)faults, 2 * (uint64_t)m, 0);
    fill(ctx, occ + occ_words, 1, 0);
    hipLaunchKernelGGL(k_batch_init, dim3((unsigned)span_blocks), dim3(kBlock), 0, ctx->stream, dplans, dboff, m);
    if (!graph) hipEventRecord(ctx->evp[0], ctx->stream);
    hipStream_t main = ctx->stream;
    if (ns >= 2) {
      hipEventRecord(ctx->side_ev[0], main);  // after the tables' set-up
      for (int k = 0; k < ns; ++k) hipStreamWaitEvent(ctx->side[k], ctx->side_ev[0], 0);
    }
    for (int j = 0, k = 0; j < m; ++j) {
      if (!parted[j] || C[live[j]].rows == 0) continue;
      if (ns >= 2) ctx->stream = ctx->side[k++ % ns];
      st = launch_partitioned(ctx, C[live[j]], dplans + j, error, !graph && ns < 2);
      ctx->stream = main;
      if (st) {
        hipStreamSynchronize(main);
        for (int q = 0; q < ns; ++q) hipStreamSynchronize(ctx->side[q]);
        drop_all();
        return st;
      }
    }
    if (n_items) {
      hipLaunchKernelGGL(kagg, dim3(grid), dim3(block), lds_max, ctx->stream, dplans,
                         ditems, (int64_t)n_items, (unsigned long long *)next);
    }
    if (ns >= 2) {
      for (int k = 0; k < ns; ++k) {
        hipEventRecord(ctx->side_ev[1 + k], ctx->side[k]);
        hipStreamWaitEvent(main, ctx->side_ev[1 + k], 0);
      }
    }
    if (!graph) hipEventRecord(ctx->evp[1], ctx->stream);
    hipLaunchKernelGGL(k_batch_occupancy, dim3((unsigned)span_blocks), dim3(kBlock), 0, ctx->stream, dplans, dwoff,
                       dboff, m, occ);
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)scan_nb), dim3(kBlock), 0, ctx->stream, occ, scan_n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, ctx->stream, sums, scan_nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)scan_nb), dim3(kBlock), 0, ctx->stream, occ, scan_n, sums,
                       prefix);
    hipLaunchKernelGGL(k_batch_counts, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, ctx->stream, dwoff,
                       prefix, m, counts, gbase, pad_off);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, ctx->stream, pad_off, (uint64_t)m);
    hipLaunchKernelGGL(k_batch_compact, dim3((unsigned)span_blocks), dim3(kBlock), 0, ctx->stream, dplans, dwoff,
                       dboff, m, occ, prefix, order);
    if (!graph) hipEventRecord(ctx->evp[2], ctx->stream);
    if (pad_total)
      hipLaunchKernelGGL(k_batch_finalize, dim3((unsigned)((pad_total + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         ctx->stream, dplans, m, pad_off, counts, gbase, order);
    hipMemcpyAsync(ctx->bstaging, counts, 8 * (size_t)m, hipMemcpyDeviceToHost, ctx->stream);
    hipMemcpyAsync(ctx->bstaging + 8 * (size_t)m, faults, 16 * (size_t)m, hipMemcpyDeviceToHost, ctx->stream);
  }
  if (graph) {
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(ge, ctx->stream);
    if (e != hipSuccess) {
      if (ge) hipGraphExecDestroy(ge);
      if (g) hipGraphDestroy(g);
      drop_all();
      return dfr_fail(error, DFR_DEVICE, "batch graph: %s", hipGetErrorString(e));
    }
  }
  HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
  ctx->obs.launches += 14;
  hipError_t e = hipStreamSynchronize(ctx->stream);
  ctx->obs.waits++;
  if (ge) hipGraphExecDestroy(ge);
  if (g) hipGraphDestroy(g);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) {
    drop_all();
    return dfr_fail(error, DFR_DEVICE, "batch kernels: %s", hipGetErrorString(e));
  }
  ctx->obs.completed = ctx->obs.launches;
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  float phase[4] = {0, 0, 0, 0};
  if (!graph) {
    hipEvent_t marks[5] = {ctx->ev0, ctx->evp[0], ctx->evp[1], ctx->evp[2], ctx->ev1};
    for (int ph = 0; ph < 4; ++ph) hipEventElapsedTime(&phase[ph], marks[ph], marks[ph + 1]);
  }
  const uint64_t *hg = (const uint64_t *)ctx->bstaging;
  const uint32_t *hf = (const uint32_t *)(ctx->bstaging + 8 * (size_t)m);
  for (int j = 0; j < m; ++j) {
    const int i = live[j];
    dfr_result *r = res[j];
    res[j] = nullptr;
    r->groups = (int64_t)hg[j];
    r->device_ms = ms;  // the whole batch's
    for (int ph = 0; ph < 4; ++ph) r->phase_ms[ph] = phase[ph];
    const uint32_t *f = hf + 4 * j;
    if (f[0]) {
      // a full hash table: this query alone, through the retrying path
      delete r;
      char *msg = nullptr;
      dfr_result *one = nullptr;
      statuses[i] = dfr_groupby(ctx, &queries[i], &one, &msg);
      results[i] = one;
      if (errors) errors[i] = msg;
      else free(msg);
      continue;
    }
    if (f[1] || f[2]) {
      delete r;
      statuses[i] = DFR_DATA_FAULT;
      const char *text = f[1] ? "Int64 expression sum overflow" : "int32 expression sum overflow";
      if (errors) dfr_fail(&errors[i], DFR_DATA_FAULT, "%s", text);
      continue;
    }
    results[i] = r;
    ctx->obs.queries++;
  }
  return DFR_OK;
}
