// runtime.h — host-side objects behind the opaque handles of dfrocm.h.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <string>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/dfrocm.h"
#include "plan.h"

namespace dfr {

// Column statistics gathered on the device when a column is handed off.
struct ColStats {
  int64_t imin, imax;   // integer columns: min/max of valid values
  uint32_t emin, emax;  // FLOAT64: biased exponent range of finite non-zero values
  uint32_t flags;       // FLOAT64: 1 NaN, 2 +inf, 4 -inf, 8 -0.0, 16 +0.0, 32 finite non-zero
  uint32_t pad;
  uint64_t valid;       // non-null values
  uint64_t distinct;    // an upper estimate of distinct non-null values (0: not estimated)
};

// Launch and table parameters. Defaults come from the device; the autotuner
// overrides them per context (dfr_context_tune).
struct Tuning {
  int block = 512;             // threads per workgroup of the one-pass (atomic) scan and of a batch
  int wg_per_cu = 2;           // their persistent workgroups per CU
  int lds_bytes = 65536;       // their LDS table bytes per workgroup
  int scan_block = 1024;       // the partitioned pass A: threads per workgroup
  int scan_wg_per_cu = 1;      // its persistent workgroups per CU
  int scan_lds_bytes = 159744; // its LDS bytes per workgroup (table and histogram)
  int lds_probe = 4;           // LDS probes before a row goes to HBM
  int lds_mode = -1;           // -1 automatic, else LDS_NONE/LDS_DIRECT/LDS_HASHED
  int64_t direct_max_slots = int64_t(1) << 26;  // largest direct-addressed HBM table
  int hash_probe = 1 << 20;    // HBM probes before a capacity fault
  int64_t hash_slots = 0;      // 0: sized from the key domain and the rows
  int agg_mode = 1;            // AGG_PARTITION (1) or AGG_ATOMIC (0)
  int part_lds_bytes = 131072; // LDS bytes of a partition table (pass C)
  int scan_unroll = 1;         // row pairs a scan thread loads at once (1, 2, 4)
  int max_parts = 4096;        // first-level key partitions, at most (a power of two)
  int sub_lds_bytes = 49152;   // LDS bytes of a second-level pass C table
  int sub_block = 512;         // threads of a second-level pass C workgroup
  int sub_fill = 50;           // records per sub-partition, in percent of its table's entries
  int sub_wg_per_cu = -1;      // its workgroups per CU (-1: as many as fit at once; 0: one per sub-partition)
  int two_level = -1;          // second partition level: -1 when groups may outgrow one level, 0 never, 1 always
  int batch_tile = 131072;     // rows of a batch work item
  int batch_streams = 4;       // streams a batch's partitioned queries are spread over (0, 1: the context's own)
  int batch_fill = 1;          // a batch of few rows is cut into this many items per workgroup (0: never)
  int batch_graph = 0;         // 1: a batch's launch sequence runs as one captured HIP graph
  int64_t batch_part_rows = 1 << 22;  // a batch query of this many rows whose groups outgrow the batch
                                      // kernel's LDS table takes the partitioned passes (0: never)
  int jit = 1;                 // pass A compiled for the query's shape: 0 never, 1 from jit_rows rows, 2 always
  int64_t jit_rows = 1 << 24;  // rows from which jit = 1 compiles
  int jit_wait = 1;            // 0: a runtime-compiled kernel is used only once compiled (no compile waits)
  int scan_prefetch = 1;       // runtime-compiled pass A: 1 overlaps the next tile's loads with this tile's fold
  int part_odd_stride = 1;     // pass C tables: entries at an odd number of words
  int lds_odd_stride = 0;      // partitioned path: LDS table entries at an odd number of words
  int part_block = 1024;       // threads of a first-level pass C workgroup
  int two_level_parts = 512;   // first-level partitions of a two-level query (a power of two)
  int scatter_slices = 1;      // pass A workgroup's waves split into this many scatter workgroups
  int scatter_chunk = 16384;   // records a scatter workgroup sorts at once (8192 or 16384)
  int scatter_stage = 1;       // 1: the scatters stage whole records of 2 to 4 words in LDS
  int hot_keys = -1;           // pass A tables seeded with sampled heavy-hitter keys: -1 from hot_rows rows, 0 never, 1 always
  int64_t hot_rows = 1 << 24;  // rows from which hot_keys = -1 seeds
  int64_t hot_domain = 1 << 24;  // largest packed key domain that is sampled (a count per key)
  int hot_sample = 1 << 19;    // rows sampled
  int hot_share = 30;          // percent of the sampled kept rows the seeded keys must carry
  int rec_pack = 1;            // 1: narrow integer values of spilled records ride in the key word
  int fuse_one_level = 1;      // 1: an unordered one-level partitioned query's pass C fills dense slots
};

}  // namespace dfr

struct Pool {
  std::mutex mu;  // results may be released from another thread than the context's
  int64_t budget = 0;
  size_t reserved = 0;  // bytes held from hipMalloc (live + cached)
  size_t cached = 0;
  size_t peak = 0;
  std::map<void *, size_t> live;
  std::map<size_t, std::vector<void *>> free_;
  int alloc(size_t bytes, void **out, char **error);
  void release(void *p);
  void trim();
  static size_t round_size(size_t bytes);
};

constexpr size_t kStagingBytes = 1 << 20;

struct dfr_context {
  int device = 0;
  int cus = 0;
  int lds_per_block = 65536;
  hipStream_t stream = nullptr;
  Pool pool;
  char *staging = nullptr;  // pinned host bytes for plans and small read-backs
  char *bstaging = nullptr;  // pinned host bytes of batch uploads (grown on demand)
  size_t bstaging_bytes = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evp[4] = {nullptr, nullptr, nullptr, nullptr};  // phase boundaries
  // side streams of a batch's partitioned queries and their fork and join
  // events (created on first use)
  static constexpr int kSideStreams = 8;
  hipStream_t side[kSideStreams] = {};
  hipEvent_t side_ev[kSideStreams + 1] = {};  // [0] fork, [1 + s] side stream s's join
  dfr_observation obs = {};
  dfr::Tuning tune;
  // A server's lane context outlives its server while results it produced
  // are still held: each such result counts here, and the context is
  // destroyed by whichever comes last, the server or the last release.
  std::mutex hold_mu;
  int64_t holds = 0;
  bool retired = false;
  // what became of the device's calibration profile at creation
  std::string calibration;
};

// Counts a result of a lane context (see dfr_context::holds).
void context_hold(dfr_result *r);
// Marks a lane context for destruction once no result holds it.
void context_retire(dfr_context *ctx);

struct dfr_column {
  dfr_context *ctx = nullptr;
  int32_t dtype = 0;
  int64_t length = 0;
  int64_t dict_size = 0;
  int64_t nulls = 0;
  const void *values = nullptr;
  const uint64_t *valid = nullptr;
  int64_t vbit = 0;
  void *owned_values = nullptr;
  void *owned_valid = nullptr;
  dfr::ColStats stats = {};
  int32_t stats_window = 1;  // 0: the statistics are a parent's (a slice), bounds of this window's
  // heavy hitters of a sample of the values (integer and dictionary columns
  // of many rows): candidate words (count << 32 | value - hot_min), their
  // counter and the sample's histogram with its row total; device memory of
  // the column's context, or null
  uint64_t *hot = nullptr;
  int64_t hot_min = 0;
  dfr::DevCol dev() const {
    dfr::DevCol c;
    c.values = values;
    c.valid = valid;
    c.vbit = vbit;
    c.dtype = dtype;
    c.is_float = dtype == DFR_FLOAT64 ? 1 : 0;
    return c;
  }
};

struct dfr_result {
  dfr_context *ctx = nullptr;
  int64_t groups = 0;
  std::vector<int32_t> key_dtypes, agg_dtypes;
  std::vector<void *> key_values, agg_values;
  std::vector<uint64_t *> key_valid, agg_valid;
  uint64_t *first_rows = nullptr;
  void *buffer = nullptr;  // one pool block holding every output
  std::shared_ptr<void> shared;  // or a block shared by a batch's results
  size_t bytes = 0;
  bool holds_ctx = false;  // counted in ctx->holds (a server lane's result)
  std::atomic<int32_t> refs{1};  // the caller's handle plus each dfr_buffer of it
  double device_ms = 0.0;
  double phase_ms[4] = {0, 0, 0, 0};
};

// One output buffer of a result (a key's or an aggregate's values or
// validity, or the first rows), as a device residence: it keeps the result
// alive until released.
struct dfr_buffer {
  dfr_result *result = nullptr;
  const void *device = nullptr;
  int64_t bytes = 0;
};

int dfr_fail(char **error, int status, const char *fmt, ...);
// Looks up a device's address of the no-alloc scratch array (once per device).
int noalloc_init(int device, char **error);
dfr::Tuning default_tuning(dfr_context *ctx);
// Applies the device's identity-matched calibration profile (if any) to a
// new context's tuning.
void apply_calibration(dfr_context *ctx, const hipDeviceProp_t &prop, const char *device_name);
int column_stats(dfr_context *ctx, dfr_column *col, char **error);
