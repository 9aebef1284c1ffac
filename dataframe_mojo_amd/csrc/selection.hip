// Columns that reach the device in pieces: the chunks of a chunked column,
// or the batches a CPU pipeline hands over as selections (a source window, a
// start and an optional row list per batch, as the reference's
// SelectionBatch, vector/selection_batches.mojo:13-75).
//
// The host packs, for every batch, the span of source rows it selects
// (values, and the validity bytes covering them), its row list rebased to
// that span, and a descriptor into one image, copied to the device in one
// transfer. k_select_gather then writes each batch's rows, in batch order,
// into one device column: one thread per output row, values stored
// coalesced, validity bits assembled per wave (each wave's 64 output rows
// straddle at most two 64-bit words, so a wave sets its bits with two
// atomic ORs; neighbouring batches share words at their boundaries).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/dfrocm.h"
#include "runtime.h"

namespace {

constexpr int kSelThreads = 256;
constexpr int kSelRowsPerBlock = kSelThreads * 4;

struct SelBatch {
  uint64_t values_off;  // image byte offset of the span's values
  int64_t valid_off;    // image byte offset of the span's validity bytes; -1: all valid
  int64_t rows_off;     // image byte offset of the rebased u32 row list; -1: contiguous
  uint64_t count;       // rows the batch contributes
  uint64_t dst;         // its first output row
  int32_t bit_shift;    // bit of span row 0 in the first validity byte
  int32_t pad;
};

// Zero-fills n 64-bit words.
__global__ void k_sel_zero(uint64_t *p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Block k gathers rows [r0, r0 + kSelRowsPerBlock) of the batch whose block
// range holds k (block_start: each batch's first block, nb + 1 entries).
template <int W>
__global__ void __launch_bounds__(kSelThreads)
    k_select_gather(const char *__restrict__ image, const SelBatch *__restrict__ batches,
                    const uint64_t *__restrict__ block_start, int nb, char *__restrict__ dst_values,
                    uint64_t *__restrict__ dst_valid) {
  const uint64_t k = blockIdx.x;
  int lo = 0, hi = nb - 1;  // the last batch whose first block is <= k
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block_start[mid] <= k)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SelBatch b = batches[lo];
  const uint64_t r0 = (k - block_start[lo]) * kSelRowsPerBlock;
  const uint32_t *rows = b.rows_off >= 0 ? (const uint32_t *)(image + b.rows_off) : nullptr;
  const unsigned char *vbytes = b.valid_off >= 0 ? (const unsigned char *)(image + b.valid_off) : nullptr;
  const char *src = image + b.values_off;
  const int lane = threadIdx.x & 63;
  // every lane runs every iteration (ballots need the whole wave)
  for (int it = 0; it < kSelRowsPerBlock / kSelThreads; ++it) {
    const uint64_t j = r0 + (uint64_t)it * kSelThreads + threadIdx.x;
    const bool live = j < b.count;
    bool valid = false;
    if (live) {
      const uint64_t s = rows ? (uint64_t)rows[j] : j;
      if (W == 4)
        *(uint32_t *)(dst_values + (b.dst + j) * 4) = *(const uint32_t *)(src + s * 4);
      else
        *(uint64_t *)(dst_values + (b.dst + j) * 8) = *(const uint64_t *)(src + s * 8);
      const uint64_t bit = (uint64_t)b.bit_shift + s;
      valid = vbytes ? ((vbytes[bit >> 3] >> (bit & 7)) & 1u) != 0 : true;
    }
    if (dst_valid) {
      // lane i of the wave writes output row out0 + i (lanes past the batch
      // end contribute no bit)
      const uint64_t ball = __ballot(valid);
      const uint64_t jw = r0 + (uint64_t)it * kSelThreads + (threadIdx.x & ~63u);
      if (lane == 0 && ball && jw < b.count) {
        const uint64_t out0 = b.dst + jw;
        const unsigned sh = (unsigned)(out0 & 63);
        atomicOr((unsigned long long *)&dst_valid[out0 >> 6], (unsigned long long)(ball << sh));
        if (sh && (ball >> (64 - sh)))
          atomicOr((unsigned long long *)&dst_valid[(out0 >> 6) + 1], (unsigned long long)(ball >> (64 - sh)));
      }
    }
  }
}

}  // namespace

#define SEL_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      st = dfr_fail(error, DFR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
      goto done;                                                                            \
    }                                                                                       \
  } while (0)

extern "C" int dfr_column_upload_selection(dfr_context *ctx, int32_t dtype, const dfr_selection *batches,
                                           int32_t n_batches, int64_t dict_size, dfr_column **out,
                                           char **error) {
  if (!ctx || !out || (n_batches > 0 && !batches) || n_batches < 0)
    return dfr_fail(error, DFR_INVALID, "null argument");
  if (dtype != DFR_INT32 && dtype != DFR_INT64 && dtype != DFR_FLOAT64 && dtype != DFR_DICT)
    return dfr_fail(error, DFR_UNSUPPORTED, "unsupported column type %d", dtype);
  const size_t W = (dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8;
  // the image: per batch its span's values, validity bytes and row list,
  // then the descriptors and the block starts
  std::vector<SelBatch> desc;
  std::vector<char> image;
  auto put = [&](const void *p, size_t bytes) {
    size_t at = (image.size() + 15) / 16 * 16;
    image.resize(at + bytes);
    if (bytes) memcpy(image.data() + at, p, bytes);
    return at;
  };
  uint64_t total = 0;
  bool any_valid = false;
  for (int32_t i = 0; i < n_batches; ++i) {
    const dfr_selection &b = batches[i];
    const dfr_source &s = b.source;
    if (b.count < 0 || b.start < 0 || s.offset < 0 || s.length < 0 || s.bit_offset < 0)
      return dfr_fail(error, DFR_INVALID, "selection batch %d has a negative size", i);
    if (b.count == 0) continue;
    if (!s.values) return dfr_fail(error, DFR_INVALID, "selection batch %d has no values", i);
    uint64_t lo = 0, hi = (uint64_t)b.count - 1;
    if (b.rows) {
      lo = UINT64_MAX;
      hi = 0;
      for (int64_t j = 0; j < b.count; ++j) {
        lo = std::min<uint64_t>(lo, b.rows[j]);
        hi = std::max<uint64_t>(hi, b.rows[j]);
      }
    }
    if ((uint64_t)b.start + hi >= (uint64_t)s.length)
      return dfr_fail(error, DFR_INVALID, "selection batch %d: row %llu is outside its source column of %lld rows",
                      i, (unsigned long long)((uint64_t)b.start + hi), (long long)s.length);
    const uint64_t span_lo = (uint64_t)b.start + lo, span_n = hi - lo + 1;
    SelBatch d{};
    d.values_off = put((const char *)s.values + (s.offset + span_lo) * W, span_n * W);
    d.valid_off = -1;
    if (s.validity) {
      const uint64_t b0 = (uint64_t)s.bit_offset + span_lo;
      d.valid_off = (int64_t)put(s.validity + (b0 >> 3), ((b0 & 7) + span_n + 7) / 8);
      d.bit_shift = (int32_t)(b0 & 7);
      any_valid = true;
    }
    d.rows_off = -1;
    if (b.rows) {
      std::vector<uint32_t> rebased(b.rows, b.rows + b.count);
      for (auto &r : rebased) r -= (uint32_t)lo;
      d.rows_off = (int64_t)put(rebased.data(), rebased.size() * 4);
    }
    d.count = (uint64_t)b.count;
    d.dst = total;
    total += (uint64_t)b.count;
    desc.push_back(d);
  }
  const int nb = (int)desc.size();
  std::vector<uint64_t> block_start(nb + 1, 0);
  for (int i = 0; i < nb; ++i)
    block_start[i + 1] = block_start[i] + (desc[i].count + kSelRowsPerBlock - 1) / kSelRowsPerBlock;
  const size_t desc_off = put(desc.data(), desc.size() * sizeof(SelBatch));
  const size_t bs_off = put(block_start.data(), block_start.size() * 8);
  int st = DFR_OK;
  void *dimage = nullptr;
  dfr_column *col = nullptr;
  if (hipSetDevice(ctx->device) != hipSuccess) return dfr_fail(error, DFR_DEVICE, "hipSetDevice failed");
  col = new dfr_column();
  col->ctx = ctx;
  col->dtype = dtype;
  col->length = (int64_t)total;
  col->dict_size = dict_size;
  if ((st = ctx->pool.alloc(std::max<size_t>(16, total * W), &col->owned_values, error))) goto done;
  col->values = col->owned_values;
  if (any_valid) {
    const uint64_t nwords = (total + 63) / 64 + 1;
    if ((st = ctx->pool.alloc(nwords * 8, &col->owned_valid, error))) goto done;
    col->valid = (const uint64_t *)col->owned_valid;
    col->vbit = 0;
    const unsigned g = (unsigned)std::min<uint64_t>((nwords + kSelThreads - 1) / kSelThreads, 4096);
    hipLaunchKernelGGL(k_sel_zero, dim3(std::max(1u, g)), dim3(kSelThreads), 0, ctx->stream,
                       (uint64_t *)col->owned_valid, nwords);
    ctx->obs.launches++;
  }
  if (nb > 0) {
    if ((st = ctx->pool.alloc(image.size(), &dimage, error))) goto done;
    SEL_TRY(hipMemcpyAsync(dimage, image.data(), image.size(), hipMemcpyHostToDevice, ctx->stream));
    ctx->obs.upload_bytes += (int64_t)image.size();
    const char *dimg = (const char *)dimage;
    const dim3 grid((unsigned)block_start[nb]);
    if (W == 4)
      hipLaunchKernelGGL(k_select_gather<4>, grid, dim3(kSelThreads), 0, ctx->stream, dimg,
                         (const SelBatch *)(dimg + desc_off), (const uint64_t *)(dimg + bs_off), nb,
                         (char *)col->owned_values, (uint64_t *)col->owned_valid);
    else
      hipLaunchKernelGGL(k_select_gather<8>, grid, dim3(kSelThreads), 0, ctx->stream, dimg,
                         (const SelBatch *)(dimg + desc_off), (const uint64_t *)(dimg + bs_off), nb,
                         (char *)col->owned_values, (uint64_t *)col->owned_valid);
    ctx->obs.launches++;
    SEL_TRY(hipGetLastError());
    // (the image is host memory that hipMemcpyAsync may still read: wait)
    SEL_TRY(hipStreamSynchronize(ctx->stream));
    ctx->obs.waits++;
  }
  if ((st = column_stats(ctx, col, error))) goto done;
done:
  if (dimage) ctx->pool.release(dimage);
  if (st) {
    if (col) dfr_column_release(col);
    return st;
  }
  *out = col;
  return DFR_OK;
}
