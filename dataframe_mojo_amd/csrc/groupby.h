// groupby.h — the compiled form of a query on the host side.
#pragma once
#include <vector>

#include "plan.h"
#include "runtime.h"

struct Compiled {
  dfr::DevPlan plan;
  std::vector<dfr_column *> cols;  // key columns first
  std::vector<int32_t> key_dtypes, agg_dtypes;
  int64_t rows = 0;
  bool maintain_order = false;
  bool jit_any_size = false;  // runtime-compiled kernels whatever the row count (tune jit 1)
  // launch shape
  int block = 512;
  int grid = 1;
  int64_t chunk = 0;
  size_t lds_bytes = 0;
  size_t c_lds_bytes = 0;  // partition tables (pass C)
  uint64_t *scan2_raw = nullptr;  // block sums of the second level's scans
  uint64_t *scan2_par = nullptr;
  // hot-key seeding of pass A (plan.hot_image set): sampled counts per
  // packed key, their histogram, the sorted candidates, keys placed
  uint32_t *hot_counts = nullptr;
  uint32_t *hot_hist = nullptr;
  uint64_t *hot_list = nullptr;
  uint32_t *hot_out = nullptr;
  bool hot_sampled = false;  // the query samples its own rows (else: its key column's list)
  int64_t hot_delta = 0;     // listed key + delta = packed key, placed when below hot_limit
  uint64_t hot_limit = 0;
};

// Words of a column's heavy-hitter buffer: candidates and their counter,
// then the sample's histogram (32-bit bins) and row total.
size_t column_hot_words();

// may_allocate false: the query's columns' own statistics must settle its
// fixed-point widths (no window or kept-row statistics pass, which needs
// device scratch); otherwise DFR_UNSUPPORTED.
int compile_query(dfr_context *ctx, const dfr_query *q, Compiled &C, char **error, bool may_allocate = true);
// `tune` overrides the context's tuning (null: the context's).
void choose_layout(dfr_context *ctx, Compiled &C, int64_t hash_slots, int sub_boost,
                   const dfr::Tuning *tune = nullptr);
int run_query(dfr_context *ctx, Compiled &C, dfr_result **out, char **error);

// Kernels compiled at run time for the plan's shape (jit.hip): pass A, and
// pass C at one or two partition levels; null when they cannot be built.
enum { JIT_SCAN = 0, JIT_PART = 1, JIT_SUB = 2 };
hipFunction_t jit_function(dfr_context *ctx, const dfr::DevPlan &P, int LDSM, int NC, int U, int which);
bool jit_compiled(dfr_context *ctx, const dfr::DevPlan &P, int LDSM, int NC, int U, int which);
