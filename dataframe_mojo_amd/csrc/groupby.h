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
};

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
