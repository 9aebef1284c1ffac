// dfrocm.hip — libdfrocm: the C ABI of include/dfrocm.h over the HIP
// kernels in kernels.h. One translation unit, built for gfx950 only.
//
// Host-side pieces: device discovery from the KFD topology (no HIP context),
// contexts (device, stream, caching HBM pool with a byte budget, pinned
// staging, observation counters), device columns with statistics, the query
// compiler (dfr_query -> DevPlan), and the launch sequence of one query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dfrocm.h"
#include "runtime.h"

using namespace dfr;

// ------------------------------------------------------------ errors

int dfr_fail(char **error, int status, const char *fmt, ...) {
  if (error) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    *error = strdup(buf);
  }
  return status;
}

#define HIP_TRY(expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return dfr_fail(error, DFR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ devices

namespace {

struct KfdNode {
  int node;
  uint64_t gfx_version;
  int cus;
  int64_t mem;
};

bool read_file(const std::string &path, std::string &out) {
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return false;
  char buf[4096];
  size_t n;
  out.clear();
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  fclose(f);
  return true;
}

uint64_t prop_value(const std::string &props, const char *key) {
  std::string k = std::string(key) + " ";
  size_t at = 0;
  while ((at = props.find(k, at)) != std::string::npos) {
    if (at == 0 || props[at - 1] == '\n') return strtoull(props.c_str() + at + k.size(), nullptr, 10);
    at += k.size();
  }
  return 0;
}

std::vector<KfdNode> kfd_gpus() {
  std::vector<KfdNode> out;
  const std::string root = "/sys/class/kfd/kfd/topology/nodes/";
  for (int n = 0; n < 256; ++n) {
    std::string props;
    if (!read_file(root + std::to_string(n) + "/properties", props)) {
      if (n > 0) break;
      continue;
    }
    uint64_t gfx = prop_value(props, "gfx_target_version");
    if (!gfx) continue;  // a CPU node
    KfdNode g;
    g.node = n;
    g.gfx_version = gfx;
    uint64_t simd = prop_value(props, "simd_count");
    uint64_t per = prop_value(props, "simd_per_cu");
    g.cus = per ? (int)(simd / per) : 0;
    g.mem = 0;
    for (int b = 0; b < 16; ++b) {
      std::string mp;
      if (!read_file(root + std::to_string(n) + "/mem_banks/" + std::to_string(b) + "/properties", mp)) break;
      g.mem += (int64_t)prop_value(mp, "size_in_bytes");
    }
    out.push_back(g);
  }
  return out;
}

}  // namespace

extern "C" int dfr_abi_version(void) { return DFR_ABI_VERSION; }

extern "C" void dfr_free(void *p) { free(p); }

extern "C" int dfr_device_count(int32_t *count, char **error) {
  if (!count) return dfr_fail(error, DFR_INVALID, "count is null");
  *count = (int32_t)kfd_gpus().size();
  return DFR_OK;
}

extern "C" int dfr_device_info_get(int32_t ordinal, dfr_device_info *info, char **error) {
  if (!info) return dfr_fail(error, DFR_INVALID, "info is null");
  std::vector<KfdNode> g = kfd_gpus();
  if (ordinal < 0 || ordinal >= (int32_t)g.size())
    return dfr_fail(error, DFR_INVALID, "device %d out of range (%zu devices)", ordinal, g.size());
  memset(info, 0, sizeof *info);
  info->ordinal = ordinal;
  info->compute_units = g[ordinal].cus;
  info->memory_bytes = g[ordinal].mem;
  uint64_t v = g[ordinal].gfx_version;  // e.g. 90500 -> gfx950
  snprintf(info->arch, sizeof info->arch, "gfx%d%x%x", (int)(v / 10000), (int)((v / 100) % 100),
           (int)(v % 100));
  const char *name = strcmp(info->arch, "gfx950") == 0 ? "AMD Instinct MI355X" : "AMD GPU";
  snprintf(info->name, sizeof info->name, "%s", name);
  return DFR_OK;
}

// ------------------------------------------------------------ context

int Pool::alloc(size_t bytes, void **out, char **error) {
  std::lock_guard<std::mutex> g(mu);
  size_t size = round_size(bytes);
  auto it = free_.find(size);
  if (it != free_.end() && !it->second.empty()) {
    *out = it->second.back();
    it->second.pop_back();
    cached -= size;
    live[*out] = size;
    return DFR_OK;
  }
  if (budget > 0 && reserved + size > (size_t)budget) trim();
  if (budget > 0 && reserved + size > (size_t)budget)
    return dfr_fail(error, DFR_CAPACITY,
                    "query needs %zu more device bytes; the budget is %lld and %zu are held", size,
                    (long long)budget, reserved);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, size);
  if (e != hipSuccess) {
    trim();
    e = hipMalloc(&p, size);
  }
  if (e != hipSuccess)
    return dfr_fail(error, DFR_CAPACITY, "hipMalloc(%zu) failed: %s", size, hipGetErrorString(e));
  reserved += size;
  peak = std::max(peak, reserved);
  live[p] = size;
  *out = p;
  return DFR_OK;
}

void Pool::release(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  auto it = live.find(p);
  if (it == live.end()) return;
  size_t size = it->second;
  live.erase(it);
  free_[size].push_back(p);
  cached += size;
}

void Pool::trim() {
  for (auto &kv : free_)
    for (void *p : kv.second) {
      hipFree(p);
      reserved -= kv.first;
      cached -= kv.first;
    }
  free_.clear();
}

size_t Pool::round_size(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes <= (1u << 20)) {
    size_t s = 256;
    while (s < bytes) s <<= 1;
    return s;
  }
  // above 1 MiB: multiples of 1/8 of the next power of two
  size_t p = 1;
  while (p < bytes) p <<= 1;
  size_t q = p / 8;
  return (bytes + q - 1) / q * q;
}

extern "C" int dfr_context_create(int32_t device, int64_t budget_bytes, dfr_context **out,
                                  char **error) {
  if (!out) return dfr_fail(error, DFR_INVALID, "out is null");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return dfr_fail(error, DFR_INVALID, "device %d out of range (%d HIP devices)", device, n);
  HIP_TRY(hipSetDevice(device));
  dfr_context *ctx = new dfr_context();
  ctx->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete ctx;
    return dfr_fail(error, DFR_DEVICE, "hipGetDeviceProperties: %s", hipGetErrorString(e));
  }
  ctx->cus = prop.multiProcessorCount;
  int lds = 0;
  hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  ctx->lds_per_block = lds > 0 ? lds : 65536;
  size_t free_b = 0, total_b = 0;
  hipMemGetInfo(&free_b, &total_b);
  ctx->pool.budget = budget_bytes > 0 ? budget_bytes : (int64_t)free_b;
  if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
      (e = hipHostMalloc((void **)&ctx->staging, kStagingBytes, hipHostMallocDefault)) != hipSuccess ||
      (e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
      (e = hipEventCreate(&ctx->evp[0])) != hipSuccess || (e = hipEventCreate(&ctx->evp[1])) != hipSuccess ||
      (e = hipEventCreate(&ctx->evp[2])) != hipSuccess || (e = hipEventCreate(&ctx->evp[3])) != hipSuccess) {
    delete ctx;
    return dfr_fail(error, DFR_DEVICE, "context set-up: %s", hipGetErrorString(e));
  }
  ctx->tune = default_tuning(ctx);
  {
    dfr_device_info info;
    char *ierr = nullptr;
    if (dfr_device_info_get(device, &info, &ierr) != DFR_OK) {
      free(ierr);
      snprintf(info.name, sizeof info.name, "%s", prop.name);
    }
    apply_calibration(ctx, prop, info.name);
  }
  int st = noalloc_init(device, error);
  if (st) {
    dfr_context_destroy(ctx);
    return st;
  }
  *out = ctx;
  return DFR_OK;
}

void context_hold(dfr_result *r) {
  if (!r || r->holds_ctx) return;
  std::lock_guard<std::mutex> g(r->ctx->hold_mu);
  r->ctx->holds++;
  r->holds_ctx = true;
}

void context_retire(dfr_context *ctx) {
  bool now;
  {
    std::lock_guard<std::mutex> g(ctx->hold_mu);
    ctx->retired = true;
    now = ctx->holds == 0;
  }
  if (now) dfr_context_destroy(ctx);
}

extern "C" void dfr_context_destroy(dfr_context *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (hipStream_t s : ctx->side)
    if (s) hipStreamSynchronize(s);
  for (auto &kv : ctx->pool.live) hipFree(kv.first);
  ctx->pool.live.clear();
  ctx->pool.trim();
  if (ctx->staging) hipHostFree(ctx->staging);
  if (ctx->bstaging) hipHostFree(ctx->bstaging);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  for (hipEvent_t ev : ctx->evp) hipEventDestroy(ev);
  for (hipStream_t s : ctx->side)
    if (s) hipStreamDestroy(s);
  for (hipEvent_t ev : ctx->side_ev)
    if (ev) hipEventDestroy(ev);
  hipStreamDestroy(ctx->stream);
  delete ctx;
}

extern "C" void dfr_context_trim(dfr_context *ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  std::lock_guard<std::mutex> g(ctx->pool.mu);
  ctx->pool.trim();
}

extern "C" int dfr_context_observe(const dfr_context *ctx, dfr_observation *out) {
  if (!ctx || !out) return DFR_INVALID;
  *out = ctx->obs;
  out->reserved_bytes = (int64_t)ctx->pool.reserved;
  out->peak_reserved_bytes = (int64_t)ctx->pool.peak;
  return DFR_OK;
}

extern "C" int dfr_context_tune(dfr_context *ctx, const char *key, int64_t value, char **error) {
  if (!ctx || !key) return dfr_fail(error, DFR_INVALID, "null argument");
  dfr::Tuning &t = ctx->tune;
  std::string k(key);
  if (k == "block" && value >= 64 && value <= 1024 && value % 64 == 0)
    t.block = (int)value;
  else if (k == "wg_per_cu" && value >= 1 && value <= 32)
    t.wg_per_cu = (int)value;
  else if (k == "lds_bytes" && value >= 0 && value <= ctx->lds_per_block)
    t.lds_bytes = (int)value;
  else if (k == "scan_block" && value >= 64 && value <= 1024 && value % 64 == 0)
    t.scan_block = (int)value;
  else if (k == "scan_wg_per_cu" && value >= 1 && value <= 32)
    t.scan_wg_per_cu = (int)value;
  else if (k == "scan_lds_bytes" && value >= 0 && value <= ctx->lds_per_block)
    t.scan_lds_bytes = (int)value;
  else if (k == "lds_probe" && value >= 1 && value <= 64)
    t.lds_probe = (int)value;
  else if (k == "lds_mode" && value >= -1 && value <= 2)
    t.lds_mode = (int)value;
  else if (k == "direct_max_slots" && value >= 0)
    t.direct_max_slots = value;
  else if (k == "hash_probe" && value >= 1)
    t.hash_probe = (int)std::min<int64_t>(value, 1 << 30);
  else if (k == "hash_slots" && value >= 0)
    t.hash_slots = value;
  else if (k == "agg_mode" && (value == 0 || value == 1))
    t.agg_mode = (int)value;
  else if (k == "part_lds_bytes" && value >= 4096 && value <= ctx->lds_per_block)
    t.part_lds_bytes = (int)value;
  else if (k == "scan_unroll" && (value == 1 || value == 2 || value == 4))
    t.scan_unroll = (int)value;
  else if (k == "two_level" && value >= -1 && value <= 1)
    t.two_level = (int)value;
  else if (k == "max_parts" && value >= 1 && value <= 4096 && (value & (value - 1)) == 0)
    t.max_parts = (int)value;
  else if (k == "sub_lds_bytes" && value >= 4096 && value <= ctx->lds_per_block - 2048)
    t.sub_lds_bytes = (int)value;
  else if (k == "sub_block" && value >= 64 && value <= 512 && value % 64 == 0)
    t.sub_block = (int)value;
  else if (k == "sub_fill" && value >= 10 && value <= 90)
    t.sub_fill = (int)value;
  else if (k == "sub_wg_per_cu" && value >= -1 && value <= 64)
    t.sub_wg_per_cu = (int)value;
  else if (k == "batch_tile" && value >= 2048 && value <= (1 << 30))
    t.batch_tile = (int)value;
  else if (k == "batch_streams" && value >= 0 && value <= dfr_context::kSideStreams)
    t.batch_streams = (int)value;
  else if (k == "batch_fill" && value >= 0 && value <= 64)
    t.batch_fill = (int)value;
  else if (k == "batch_graph" && (value == 0 || value == 1))
    t.batch_graph = (int)value;
  else if (k == "batch_part_rows" && value >= 0)
    t.batch_part_rows = value;
  else if (k == "scatter_slices" && value >= 1 && value <= 16 && (value & (value - 1)) == 0)
    t.scatter_slices = (int)value;
  else if (k == "scatter_stage" && (value == 0 || value == 1))
    t.scatter_stage = (int)value;
  else if (k == "hot_keys" && value >= -1 && value <= 1)
    t.hot_keys = (int)value;
  else if (k == "hot_rows" && value >= 0)
    t.hot_rows = value;
  else if (k == "hot_domain" && value >= 0 && value <= ((int64_t)1 << 28))
    t.hot_domain = value;
  else if (k == "hot_sample" && value >= 4096 && value <= (1 << 28))
    t.hot_sample = (int)value;
  else if (k == "hot_share" && value >= 0 && value <= 100)
    t.hot_share = (int)value;
  else if (k == "rec_pack" && (value == 0 || value == 1))
    t.rec_pack = (int)value;
  else if (k == "fuse_one_level" && (value == 0 || value == 1))
    t.fuse_one_level = (int)value;
  else if (k == "scatter_chunk" && (value == 8192 || value == 16384))
    t.scatter_chunk = (int)value;
  else if (k == "two_level_parts" && value >= 1 && value <= 4096 && (value & (value - 1)) == 0)
    t.two_level_parts = (int)value;
  else if (k == "part_block" && value >= 64 && value <= 1024 && value % 64 == 0)
    t.part_block = (int)value;
  else if (k == "lds_odd_stride" && (value == 0 || value == 1))
    t.lds_odd_stride = (int)value;
  else if (k == "part_odd_stride" && (value == 0 || value == 1))
    t.part_odd_stride = (int)value;
  else if (k == "scan_prefetch" && (value == 0 || value == 1))
    t.scan_prefetch = (int)value;
  else if (k == "jit" && value >= 0 && value <= 2)
    t.jit = (int)value;
  else if (k == "jit_rows" && value >= 0)
    t.jit_rows = value;
  else if (k == "jit_wait" && (value == 0 || value == 1))
    t.jit_wait = (int)value;
  else
    return dfr_fail(error, DFR_INVALID, "bad tuning %s=%lld", key, (long long)value);
  return DFR_OK;
}

extern "C" int dfr_context_tune_get(const dfr_context *ctx, const char *key, int64_t *value, char **error) {
  if (!ctx || !key || !value) return dfr_fail(error, DFR_INVALID, "null argument");
  const dfr::Tuning &t = ctx->tune;
  const std::map<std::string, int64_t> v = {
      {"block", t.block}, {"wg_per_cu", t.wg_per_cu}, {"lds_bytes", t.lds_bytes},
      {"scan_block", t.scan_block}, {"scan_wg_per_cu", t.scan_wg_per_cu}, {"scan_lds_bytes", t.scan_lds_bytes},
      {"lds_probe", t.lds_probe}, {"lds_mode", t.lds_mode}, {"direct_max_slots", t.direct_max_slots},
      {"hash_probe", t.hash_probe}, {"hash_slots", t.hash_slots}, {"agg_mode", t.agg_mode},
      {"part_lds_bytes", t.part_lds_bytes}, {"scan_unroll", t.scan_unroll}, {"two_level", t.two_level},
      {"max_parts", t.max_parts}, {"sub_lds_bytes", t.sub_lds_bytes}, {"sub_block", t.sub_block},
      {"sub_fill", t.sub_fill}, {"sub_wg_per_cu", t.sub_wg_per_cu}, {"batch_tile", t.batch_tile}, {"batch_fill", t.batch_fill}, {"batch_streams", t.batch_streams},
      {"batch_graph", t.batch_graph}, {"batch_part_rows", t.batch_part_rows},
      {"scatter_slices", t.scatter_slices}, {"scatter_stage", t.scatter_stage}, {"hot_keys", t.hot_keys},
      {"hot_rows", t.hot_rows}, {"hot_domain", t.hot_domain}, {"hot_sample", t.hot_sample},
      {"hot_share", t.hot_share}, {"rec_pack", t.rec_pack}, {"fuse_one_level", t.fuse_one_level}, {"scatter_chunk", t.scatter_chunk},
      {"two_level_parts", t.two_level_parts}, {"part_block", t.part_block}, {"lds_odd_stride", t.lds_odd_stride},
      {"part_odd_stride", t.part_odd_stride}, {"scan_prefetch", t.scan_prefetch}, {"jit", t.jit},
      {"jit_rows", t.jit_rows}, {"jit_wait", t.jit_wait}};
  auto it = v.find(key);
  if (it == v.end()) return dfr_fail(error, DFR_INVALID, "unknown tuning key %s", key);
  *value = it->second;
  return DFR_OK;
}

// ------------------------------------------------------------ columns

static int dtype_width(int32_t dtype) { return (dtype == DFR_INT32 || dtype == DFR_DICT) ? 4 : 8; }

static bool dtype_ok(int32_t dtype) {
  return dtype == DFR_INT32 || dtype == DFR_INT64 || dtype == DFR_FLOAT64 || dtype == DFR_DICT;
}

extern "C" int dfr_column_upload(dfr_context *ctx, int32_t dtype, const void *values, int64_t offset,
                                 int64_t length, const uint8_t *validity, int64_t bit_offset,
                                 int64_t dict_size, dfr_column **out, char **error) {
  if (!ctx || !out || (!values && length > 0)) return dfr_fail(error, DFR_INVALID, "null argument");
  if (!dtype_ok(dtype)) return dfr_fail(error, DFR_UNSUPPORTED, "unsupported column type %d", dtype);
  if (length < 0 || offset < 0 || bit_offset < 0) return dfr_fail(error, DFR_INVALID, "negative size");
  HIP_TRY(hipSetDevice(ctx->device));
  dfr_column *col = new dfr_column();
  col->ctx = ctx;
  col->dtype = dtype;
  col->length = length;
  col->dict_size = dict_size;
  size_t w = dtype_width(dtype);
  size_t vbytes = std::max<size_t>(16, (size_t)length * w);
  int st = ctx->pool.alloc(vbytes, &col->owned_values, error);
  if (st) {
    delete col;
    return st;
  }
  col->values = col->owned_values;
  if (length > 0) {
    hipError_t e = hipMemcpy(col->owned_values, (const char *)values + offset * w, (size_t)length * w,
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      dfr_column_release(col);
      return dfr_fail(error, DFR_DEVICE, "column upload: %s", hipGetErrorString(e));
    }
    ctx->obs.upload_bytes += (int64_t)(length * w);
  }
  if (validity) {
    // re-base the bitmap so that row 0 is bit 0 of word 0
    size_t nwords = ((size_t)length + 63) / 64 + 1;
    std::vector<uint64_t> words(nwords, 0);
    for (int64_t r = 0; r < length; ++r) {
      int64_t b = bit_offset + r;
      if ((validity[b >> 3] >> (b & 7)) & 1) words[r >> 6] |= 1ull << (r & 63);
    }
    st = ctx->pool.alloc(nwords * 8, &col->owned_valid, error);
    if (st) {
      dfr_column_release(col);
      return st;
    }
    hipError_t e = hipMemcpy(col->owned_valid, words.data(), nwords * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      dfr_column_release(col);
      return dfr_fail(error, DFR_DEVICE, "validity upload: %s", hipGetErrorString(e));
    }
    ctx->obs.upload_bytes += (int64_t)(nwords * 8);
    col->valid = (const uint64_t *)col->owned_valid;
    col->vbit = 0;
  }
  st = column_stats(ctx, col, error);
  if (st) {
    dfr_column_release(col);
    return st;
  }
  *out = col;
  return DFR_OK;
}

extern "C" int dfr_column_wrap(dfr_context *ctx, int32_t dtype, const void *device_values,
                               int64_t offset, int64_t length, const uint64_t *device_validity,
                               int64_t bit_offset, int64_t dict_size, dfr_column **out,
                               char **error) {
  if (!ctx || !out || (!device_values && length > 0))
    return dfr_fail(error, DFR_INVALID, "null argument");
  if (!dtype_ok(dtype)) return dfr_fail(error, DFR_UNSUPPORTED, "unsupported column type %d", dtype);
  if (length < 0 || offset < 0 || bit_offset < 0) return dfr_fail(error, DFR_INVALID, "negative size");
  HIP_TRY(hipSetDevice(ctx->device));
  dfr_column *col = new dfr_column();
  col->ctx = ctx;
  col->dtype = dtype;
  col->length = length;
  col->dict_size = dict_size;
  col->values = (const char *)device_values + offset * dtype_width(dtype);
  col->valid = device_validity;
  col->vbit = bit_offset;
  int st = column_stats(ctx, col, error);
  if (st) {
    dfr_column_release(col);  // (owns nothing of the caller's)
    return st;
  }
  *out = col;
  return DFR_OK;
}

extern "C" int dfr_column_slice(const dfr_column *parent, int64_t offset, int64_t length, dfr_column **out,
                                char **error) {
  if (!parent || !out) return dfr_fail(error, DFR_INVALID, "null argument");
  if (offset < 0 || length < 0 || offset + length > parent->length)
    return dfr_fail(error, DFR_INVALID, "slice [%lld, %lld) outside a column of %lld rows", (long long)offset,
                    (long long)(offset + length), (long long)parent->length);
  dfr_column *col = new dfr_column();
  col->ctx = parent->ctx;
  col->dtype = parent->dtype;
  col->length = length;
  col->dict_size = parent->dict_size;
  col->values = (const char *)parent->values + offset * dtype_width(parent->dtype);
  col->valid = parent->valid;
  col->vbit = parent->vbit + offset;
  // the parent's statistics bound the slice's (range, exponents); its null
  // count is an upper bound, so a slice of a nullable column stays nullable
  col->stats = parent->stats;
  col->stats_window = 0;
  col->nulls = parent->nulls;
  *out = col;
  return DFR_OK;
}

extern "C" void dfr_column_release(dfr_column *col) {
  if (!col) return;
  if (col->owned_values) col->ctx->pool.release(col->owned_values);
  if (col->owned_valid) col->ctx->pool.release(col->owned_valid);
  if (col->hot) col->ctx->pool.release(col->hot);
  delete col;
}

extern "C" int64_t dfr_column_length(const dfr_column *col) { return col ? col->length : -1; }
extern "C" int64_t dfr_column_null_count(const dfr_column *col) { return col ? col->nulls : -1; }
extern "C" int64_t dfr_column_distinct(const dfr_column *col) { return col ? (int64_t)col->stats.distinct : -1; }

// ------------------------------------------------------------ results

extern "C" int64_t dfr_result_groups(const dfr_result *r) { return r ? r->groups : -1; }

extern "C" int32_t dfr_result_key_dtype(const dfr_result *r, int32_t k) {
  if (!r || k < 0 || k >= (int32_t)r->key_dtypes.size()) return -1;
  return r->key_dtypes[k];
}

extern "C" int32_t dfr_result_agg_dtype(const dfr_result *r, int32_t a) {
  if (!r || a < 0 || a >= (int32_t)r->agg_dtypes.size()) return -1;
  return r->agg_dtypes[a];
}

static int copy_column(const dfr_result *r, const void *dvals, const uint64_t *dvalid, int dtype,
                       void *values, uint8_t *validity, char **error) {
  dfr_context *ctx = r->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  size_t n = (size_t)r->groups;
  if (n == 0) return DFR_OK;
  if (values) {
    HIP_TRY(hipMemcpy(values, dvals, n * dtype_width(dtype), hipMemcpyDeviceToHost));
    ctx->obs.download_bytes += (int64_t)(n * dtype_width(dtype));
  }
  if (validity) {
    std::vector<uint64_t> words((n + 63) / 64);
    HIP_TRY(hipMemcpy(words.data(), dvalid, words.size() * 8, hipMemcpyDeviceToHost));
    ctx->obs.download_bytes += (int64_t)(words.size() * 8);
    memcpy(validity, words.data(), (n + 7) / 8);
  }
  return DFR_OK;
}

extern "C" int dfr_result_copy_key(const dfr_result *r, int32_t k, void *values, uint8_t *validity,
                                   char **error) {
  if (!r || k < 0 || k >= (int32_t)r->key_dtypes.size())
    return dfr_fail(error, DFR_INVALID, "key index out of range");
  return copy_column(r, r->key_values[k], r->key_valid[k], r->key_dtypes[k], values, validity, error);
}

extern "C" int dfr_result_copy_agg(const dfr_result *r, int32_t a, void *values, uint8_t *validity,
                                   char **error) {
  if (!r || a < 0 || a >= (int32_t)r->agg_dtypes.size())
    return dfr_fail(error, DFR_INVALID, "aggregate index out of range");
  return copy_column(r, r->agg_values[a], r->agg_valid[a], r->agg_dtypes[a], values, validity, error);
}

extern "C" int dfr_result_copy_first_rows(const dfr_result *r, int64_t *rows, char **error) {
  if (!r || !rows) return dfr_fail(error, DFR_INVALID, "null argument");
  if (!r->first_rows)
    return dfr_fail(error, DFR_UNSUPPORTED, "first rows are kept only under maintain_order");
  if (r->groups == 0) return DFR_OK;
  HIP_TRY(hipSetDevice(r->ctx->device));
  HIP_TRY(hipMemcpy(rows, r->first_rows, (size_t)r->groups * 8, hipMemcpyDeviceToHost));
  return DFR_OK;
}

extern "C" int64_t dfr_result_device_bytes(const dfr_result *r) { return r ? (int64_t)r->bytes : 0; }

extern "C" double dfr_result_device_ms(const dfr_result *r) { return r ? r->device_ms : 0.0; }

extern "C" double dfr_result_phase_ms(const dfr_result *r, int32_t phase) {
  if (!r || phase < 0 || phase > 3) return -1.0;
  return r->phase_ms[phase];
}

extern "C" int dfr_result_buffer(dfr_result *r, int32_t kind, int32_t index, int32_t part, dfr_buffer **out,
                                 int64_t *bytes, char **error) {
  if (!r || !out) return dfr_fail(error, DFR_INVALID, "null argument");
  const void *dev = nullptr;
  int64_t n = 0;
  const int64_t g = r->groups;
  if (kind == DFR_BUFFER_KEY || kind == DFR_BUFFER_AGG) {
    const std::vector<int32_t> &dt = kind == DFR_BUFFER_KEY ? r->key_dtypes : r->agg_dtypes;
    if (index < 0 || index >= (int32_t)dt.size()) return dfr_fail(error, DFR_INVALID, "buffer index out of range");
    if (part == DFR_BUFFER_VALUES) {
      dev = kind == DFR_BUFFER_KEY ? r->key_values[index] : r->agg_values[index];
      n = g * dtype_width(dt[index]);
    } else if (part == DFR_BUFFER_VALIDITY) {
      dev = kind == DFR_BUFFER_KEY ? (const void *)r->key_valid[index] : (const void *)r->agg_valid[index];
      n = (g + 7) / 8;
    } else {
      return dfr_fail(error, DFR_INVALID, "unknown buffer part %d", part);
    }
  } else if (kind == DFR_BUFFER_FIRST_ROWS) {
    if (!r->first_rows) return dfr_fail(error, DFR_UNSUPPORTED, "first rows are kept only under maintain_order");
    dev = r->first_rows;
    n = g * 8;
  } else {
    return dfr_fail(error, DFR_INVALID, "unknown buffer kind %d", kind);
  }
  dfr_buffer *b = new dfr_buffer();
  b->result = r;
  b->device = dev;
  b->bytes = n;
  r->refs.fetch_add(1);
  *out = b;
  if (bytes) *bytes = n;
  return DFR_OK;
}

extern "C" int64_t dfr_buffer_copy_to_host(void *buffer, void *host_dst, int64_t bytes) {
  dfr_buffer *b = (dfr_buffer *)buffer;
  if (!b || bytes < 0 || bytes > b->bytes || (bytes > 0 && !host_dst)) return DFR_INVALID;
  if (bytes == 0) return DFR_OK;
  dfr_context *ctx = b->result->ctx;
  if (hipSetDevice(ctx->device) != hipSuccess) return DFR_DEVICE;
  if (hipMemcpy(host_dst, b->device, (size_t)bytes, hipMemcpyDeviceToHost) != hipSuccess) return DFR_DEVICE;
  return DFR_OK;
}

extern "C" void dfr_buffer_release(void *buffer) {
  dfr_buffer *b = (dfr_buffer *)buffer;
  if (!b) return;
  dfr_result_release(b->result);
  delete b;
}

extern "C" void dfr_result_release(dfr_result *r) {
  if (!r) return;
  if (r->refs.fetch_sub(1) != 1) return;  // a buffer of it is still held
  dfr_context *ctx = r->ctx;
  const bool held = r->holds_ctx;
  if (r->buffer) ctx->pool.release(r->buffer);
  delete r;  // (a batch's shared block goes back to the pool with its last result)
  if (!held) return;
  bool last;
  {
    std::lock_guard<std::mutex> g(ctx->hold_mu);
    last = --ctx->holds == 0 && ctx->retired;
  }
  if (last) dfr_context_destroy(ctx);
}
