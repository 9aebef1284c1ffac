// noalloc.h — dfr_groupby_into: a filtered group-by that allocates no
// device memory, for callers that cannot (code under HIP graph capture,
// serving loops that forbid allocation). Included at the end of
// groupby.hip: it shares that file's compiler, kernels and launch helpers.
//
// The caller passes the inputs (device columns) and the outputs (device
// buffers of a stated group capacity). Everything else the query needs —
// its device plan, fault flags, the group table, the compaction arrays and,
// under maintain_order, the row bitmap and its prefix — lives in `scratch`,
// a fixed device array of kNoAllocScratchBytes on each GPU, held until the
// process ends. Every call on one GPU uses that GPU's copy: calls that
// overlap on one GPU race. What must start at zero (fault flags, the row
// bitmap) is cleared by fill kernels at each call; the table is set to its
// identity words by k_init_state. The query runs through the one-pass
// aggregation (LDS table, HBM table with device atomics), whose scratch is
// bounded by the table, not by the rows.
#pragma once

#include <mutex>

namespace {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;

// The device array. Host code never passes its name to a kernel or a copy
// (on the host it is a stand-in, not a device address): it uses the
// per-device address noalloc_init looked up.
static __device__ __align__(256) uint64_t scratch[kNoAllocScratchBytes / 8];

std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};

}  // namespace

// Looks up this GPU's address of `scratch` (once per device, with the
// device current); called when a context is created on the device.
int noalloc_init(int device, char **error) {
  if (device < 0 || device >= kNoAllocMaxDevices)
    return dfr_fail(error, DFR_INVALID, "device %d beyond the no-alloc scratch table", device);
  std::lock_guard<std::mutex> g(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return DFR_OK;
  HIP_TRY(hipSetDevice(device));
  void *ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_noalloc_ptr[device] = ptr;
  return DFR_OK;
}

namespace {

// Bump allocation inside the scratch array (256-byte aligned pieces).
struct ScratchCarve {
  char *base;
  size_t used = 0;
  bool over = false;
  template <class T>
  T *take(size_t count) {
    const size_t bytes = (std::max<size_t>(count * sizeof(T), 8) + 255) / 256 * 256;
    if (used + bytes > kNoAllocScratchBytes) {
      over = true;
      return nullptr;
    }
    T *p = (T *)(base + used);
    used += bytes;
    return p;
  }
};

// Exclusive popcount prefix of n bitmap words into `prefix` (block sums in
// `sums`, nb + 1 words); the total lands at sums[nb].
void scan_bits_into(dfr_context *ctx, const uint64_t *bits, uint64_t n, uint64_t *prefix, uint64_t *sums,
                    uint64_t nb) {
  hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, bits, n, sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, ctx->stream, sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kBlock), 0, ctx->stream, bits, n, sums, prefix);
  ctx->obs.launches += 3;
}

uint64_t scan_blocks_of(uint64_t n) { return std::max<uint64_t>(1, (n + kScanChunk - 1) / kScanChunk); }

}  // namespace

extern "C" int dfr_groupby_into(dfr_context *ctx, const dfr_query *query, const dfr_output *out, int64_t *groups,
                                char **error) {
  if (!ctx || !query || !out || !groups) return dfr_fail(error, DFR_INVALID, "null argument");
  if (out->capacity < 0 || (out->capacity > 0 && ((query->n_keys > 0 && !out->keys) ||
                                                  (query->n_aggs > 0 && !out->aggs))))
    return dfr_fail(error, DFR_INVALID, "output buffers missing");
  HIP_TRY(hipSetDevice(ctx->device));
  int st = noalloc_init(ctx->device, error);
  if (st) return st;
  char *base = (char *)g_noalloc_ptr[ctx->device];
  Compiled C;
  if ((st = compile_query(ctx, query, C, error, false))) return st;
  if (C.maintain_order && !out->first_rows && out->capacity > 0)
    return dfr_fail(error, DFR_INVALID, "maintain_order needs first_rows output");
  DevPlan &P = C.plan;
  Tuning tune = ctx->tune;
  tune.agg_mode = AGG_ATOMIC;
  const uint64_t rows = (uint64_t)std::max<int64_t>(C.rows, 0);
  int64_t hash_slots = 0;
  for (int attempt = 0;; ++attempt) {
    choose_layout(ctx, C, hash_slots, 0, &tune);
    // the scratch layout of this attempt
    ScratchCarve S{base};
    DevPlan *dplan = S.take<DevPlan>(1);
    uint32_t *fault = S.take<uint32_t>(4);
    P.state = S.take<uint64_t>(P.cap * (uint64_t)P.words);
    P.keys_tab = P.mode == MODE_HASH ? S.take<uint64_t>(P.cap) : nullptr;
    const uint64_t occ_words = (P.cap + 63) / 64;
    const uint64_t max_groups = std::min<uint64_t>(P.cap, std::max<uint64_t>(rows, 1));
    uint64_t *occ = S.take<uint64_t>(occ_words);
    uint64_t *occ_prefix = S.take<uint64_t>(occ_words);
    uint64_t *order = S.take<uint64_t>(max_groups + 1);
    const uint64_t occ_nb = scan_blocks_of(occ_words);
    uint64_t *occ_sums = S.take<uint64_t>(occ_nb + 1);
    const uint64_t rwords = (rows + 63) / 64;
    uint64_t *rowbits = nullptr, *rowprefix = nullptr, *order2 = nullptr, *row_sums = nullptr;
    const uint64_t row_nb = scan_blocks_of(rwords);
    if (C.maintain_order) {
      rowbits = S.take<uint64_t>(rwords);
      rowprefix = S.take<uint64_t>(rwords);
      order2 = S.take<uint64_t>(max_groups + 1);
      row_sums = S.take<uint64_t>(row_nb + 1);
    }
    if (S.over)
      return dfr_fail(error, DFR_CAPACITY,
                      "the query needs more than the %zu bytes of no-alloc scratch (its group table has %llu slots of "
                      "%d words)",
                      kNoAllocScratchBytes, (unsigned long long)P.cap, P.words);
    P.fault = fault;
    P.order = nullptr;
    P.out_first = nullptr;
    HIP_TRY(hipEventRecord(ctx->ev0, ctx->stream));
    memcpy(ctx->staging, &P, sizeof P);
    HIP_TRY(hipMemcpyAsync(dplan, ctx->staging, sizeof P, hipMemcpyHostToDevice, ctx->stream));
    fill(ctx, (uint64_t *)fault, 2, 0);
    {
      const int blocks = (int)std::min<uint64_t>((P.cap + kBlock - 1) / kBlock, (uint64_t)ctx->cus * 8);
      hipLaunchKernelGGL(k_init_state, dim3(std::max(blocks, 1)), dim3(kBlock), 0, ctx->stream, dplan);
      ctx->obs.launches++;
    }
    HIP_TRY(hipEventRecord(ctx->evp[0], ctx->stream));
    if (rows > 0 && (st = launch_aggregate(ctx, C, dplan, error))) return st;
    HIP_TRY(hipEventRecord(ctx->evp[1], ctx->stream));
    {
      const uint64_t threads = occ_words * 64;
      hipLaunchKernelGGL(k_occupancy, dim3((unsigned)((threads + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         ctx->stream, dplan, occ, occ_words);
      ctx->obs.launches++;
    }
    scan_bits_into(ctx, occ, occ_words, occ_prefix, occ_sums, occ_nb);
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((P.cap + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                       dplan, occ, occ_prefix, order);
    ctx->obs.launches++;
    uint64_t *hs = readback(ctx);
    HIP_TRY(hipMemcpyAsync(hs, occ_sums + occ_nb, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(hs + 1, fault, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->obs.waits++;
    const uint64_t G = hs[0];
    const uint32_t *hf = (const uint32_t *)(hs + 1);
    if (hf[0]) {
      // a full hash table: once more with room for every possible group, if
      // the scratch holds it
      const uint64_t need = next_pow2(std::max<uint64_t>(64, 2 * std::min<uint64_t>(P.domain, rows)));
      if (P.mode != MODE_HASH || attempt >= 4 || P.cap >= need)
        return dfr_fail(error, DFR_CAPACITY, "group table capacity exceeded");
      hash_slots = (int64_t)std::max<uint64_t>(P.cap * 4, need);
      continue;
    }
    if (G > (uint64_t)out->capacity)
      return dfr_fail(error, DFR_CAPACITY, "%llu groups exceed the output capacity of %lld", (unsigned long long)G,
                      (long long)out->capacity);
    // outputs: the caller's buffers
    for (int k = 0; k < P.n_keys; ++k) {
      P.out_keys[k] = out->keys[k].values;
      P.out_key_valid[k] = out->keys[k].validity;
    }
    for (int a = 0; a < P.n_aggs; ++a) {
      P.aggs[a].out_values = out->aggs[a].values;
      P.aggs[a].out_valid = out->aggs[a].validity;
    }
    P.groups = occ_sums + occ_nb;
    P.order = order;
    if (C.maintain_order && G > 0) {
      fill(ctx, rowbits, rwords, 0);
      const unsigned gb = (unsigned)((G + kBlock - 1) / kBlock);
      hipLaunchKernelGGL(k_mark_first, dim3(gb), dim3(kBlock), 0, ctx->stream, dplan, P.order, G, rowbits);
      scan_bits_into(ctx, rowbits, rwords, rowprefix, row_sums, row_nb);
      hipLaunchKernelGGL(k_rank_first, dim3(gb), dim3(kBlock), 0, ctx->stream, dplan, P.order, G, rowbits, rowprefix,
                         order2);
      ctx->obs.launches += 2;
      P.order = order2;
      P.out_first = (uint64_t *)out->first_rows;
    }
    HIP_TRY(hipEventRecord(ctx->evp[2], ctx->stream));
    if (G > 0) {
      memcpy(ctx->staging, &P, sizeof P);
      HIP_TRY(hipMemcpyAsync(dplan, ctx->staging, sizeof P, hipMemcpyHostToDevice, ctx->stream));
      hipLaunchKernelGGL(k_finalize, dim3((unsigned)((G + kBlock - 1) / kBlock)), dim3(kBlock), 0, ctx->stream,
                         dplan);
      ctx->obs.launches++;
    }
    HIP_TRY(hipEventRecord(ctx->ev1, ctx->stream));
    uint32_t *hf2 = (uint32_t *)(readback(ctx) + 8);
    HIP_TRY(hipMemcpyAsync(hf2, fault, 16, hipMemcpyDeviceToHost, ctx->stream));
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    ctx->obs.waits++;
    if (e != hipSuccess) return dfr_fail(error, DFR_DEVICE, "group-by kernels: %s", hipGetErrorString(e));
    ctx->obs.completed = ctx->obs.launches;
    ctx->obs.queries++;
    if (hf2[1]) return dfr_fail(error, DFR_DATA_FAULT, "Int64 expression sum overflow");
    if (hf2[2]) return dfr_fail(error, DFR_DATA_FAULT, "int32 expression sum overflow");
    *groups = (int64_t)G;
    return DFR_OK;
  }
}
