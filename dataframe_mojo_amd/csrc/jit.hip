// jit.hip — pass A compiled at run time (hiprtc) for one query's shape.
//
// The precompiled pass A reads every plan field from constant memory: the
// column types, which aggregates each column feeds, the state-word offsets
// and the predicate's structure are all run-time values there, and the
// kernel carries the branches and registers for every case. This file
// generates a kernel that builds a register copy of the plan in which those
// shape fields are literals, so the compiler folds them away; the fields
// that vary with the data or the query's constants (column pointers, key
// ranges, literals, fixed-point scales, table sizes) stay run-time values,
// so one compiled kernel serves every query of the same shape. Kernels are
// cached per device and shape for the life of the process. Each compile
// runs in a child process (dfr_jitc, csrc/jitc.cpp), so a compiler crash
// costs only the specialised kernel.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <dlfcn.h>
#include <spawn.h>
#include <sys/wait.h>
#include <unistd.h>

#include <cerrno>

#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dfrocm.h"
#include "groupby.h"
#include "runtime.h"

using namespace dfr;

extern char **environ;

namespace {

struct JitEntry {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
  bool failed = false;
  bool compiling = false;  // another thread is compiling it (outside the lock)
};

const char *const kJitNames[3] = {"dfr_scan_jit", "dfr_part_jit", "dfr_sub_jit"};  // JIT_SCAN, JIT_PART, JIT_SUB

std::mutex g_mu;
std::condition_variable g_cv;  // a compile finished
std::unordered_map<std::string, JitEntry> g_cache;

// The directory of the kernel headers: csrc/ beside this library.
std::string header_dir() {
  Dl_info info;
  if (!dladdr((const void *)&header_dir, &info) || !info.dli_fname) return "";
  std::string path = info.dli_fname;
  const size_t slash = path.rfind('/');
  return (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/csrc";
}

// The source of one kernel (JIT_SCAN, JIT_PART or JIT_SUB) for this plan's
// shape.
std::string shape_source(const DevPlan &P, int LDSM, int NC, int U, int PF, int which) {
  std::ostringstream o;
  bool col_leaves = false;
  for (int l = 0; l < P.n_leaves; ++l) col_leaves |= P.leaves[l].col2 >= 0;
  o << "#define DFR_JIT 1\n#define DFR_COL_LEAVES " << (col_leaves ? 1 : 0) << "\n#include \"kernels_part.h\"\nusing namespace dfr;\n";
  if (col_leaves) {
    // the right-hand column of each column-comparison leaf, by constant index
    o << "template <int NC>\n__device__ __forceinline__ void dfr::jit_leaf_rhs(int l, const uint64_t (&vals)[NC], "
         "uint32_t vbits, uint64_t &v, bool &ok) {\n  switch (l) {\n";
    for (int l = 0; l < P.n_leaves; ++l)
      if (P.leaves[l].col2 >= 0)
        o << "    case " << l << ":\n      if constexpr (" << P.leaves[l].col2 << " < NC) {\n        v = vals["
          << P.leaves[l].col2 << "];\n        ok = ((vbits >> " << P.leaves[l].col2 << ") & 1u) != 0;\n      }\n"
          << "      break;\n";
    o << "    default: break;\n  }\n}\n";
  }
  o
    << "__device__ __forceinline__ DevPlan make_plan(const DevPlan *plan) {\n"
    << "  PlanRef R = plan_ref(plan);\n"
    << "  DevPlan L{};\n";
  // values of the data and of the query's constants: from the plan
  o << "  L.rows = R.rows; L.lds_entries = R.lds_entries; L.lds_log2 = R.lds_log2; L.lds_buckets = R.lds_buckets;\n"
    << "  L.n_parts = R.n_parts; L.part_log2 = R.part_log2; L.key_bits = R.key_bits;\n"
    << "  L.spill = R.spill; L.partials = R.partials; L.seg_rows = R.seg_rows; L.seg_partials = R.seg_partials;\n"
    << "  L.hist_raw = R.hist_raw; L.hist_par = R.hist_par; L.cnt_raw = R.cnt_raw; L.cnt_par = R.cnt_par;\n"
    << "  L.c_entries = R.c_entries; L.c_log2 = R.c_log2; L.base_raw = R.base_raw; L.base_par = R.base_par;\n"
    << "  L.part_raw = R.part_raw; L.part_par = R.part_par; L.state = R.state; L.keys_tab = R.keys_tab;\n"
    << "  L.mode = R.mode; L.cap = R.cap; L.cap_log2 = R.cap_log2; L.max_probe = R.max_probe; L.fault = R.fault;\n"
    << "  L.n_sub = R.n_sub; L.sub_chunks = R.sub_chunks; L.base2_raw = R.base2_raw; L.base2_par = R.base2_par;\n"
    << "  L.part2_raw = R.part2_raw; L.part2_par = R.part2_par; L.gcount = R.gcount;\n"
    << "  L.hot_image = R.hot_image; L.hot_states = R.hot_states;\n";
  // a column without a validity bitmap has a null one in the shape, so its
  // validity code folds away
  for (int c = 0; c < P.n_cols; ++c)
    o << "  L.cols[" << c << "].values = R.cols[" << c << "].values; L.cols[" << c << "].valid = "
      << (P.cols[c].valid ? "R.cols[" + std::to_string(c) + "].valid" : std::string("nullptr")) << "; L.cols[" << c
      << "].vbit = R.cols[" << c << "].vbit;\n"
      << (P.cols[c].valid ? "  __builtin_assume(L.cols[" + std::to_string(c) + "].valid != nullptr);\n" : "")
      << "  L.cagg[" << c << "].fshift = R.cagg[" << c << "].fshift; L.cagg[" << c << "].fshift2 = R.cagg[" << c
      << "].fshift2;\n";
  for (int k = 0; k < P.n_keys; ++k)
    o << "  L.keys[" << k << "].minv = R.keys[" << k << "].minv; L.keys[" << k << "].radix = R.keys[" << k
      << "].radix; L.keys[" << k << "].stride = R.keys[" << k << "].stride; L.keys[" << k
      << "].inv_radix = R.keys[" << k << "].inv_radix;\n";
  for (int l = 0; l < P.n_leaves; ++l)
    o << "  L.leaves[" << l << "].lit = R.leaves[" << l << "].lit; L.leaves[" << l << "].mod = R.leaves[" << l
      << "].mod;\n";
  // outputs (a fused pass C writes them)
  for (int k = 0; k < P.n_keys; ++k)
    o << "  L.out_keys[" << k << "] = R.out_keys[" << k << "]; L.out_key_valid[" << k << "] = R.out_key_valid[" << k
      << "];\n";
  for (int a = 0; a < P.n_aggs; ++a)
    o << "  L.aggs[" << a << "].ddof = R.aggs[" << a << "].ddof; L.aggs[" << a << "].min_count = R.aggs[" << a
      << "].min_count; L.aggs[" << a << "].out_values = R.aggs[" << a << "].out_values; L.aggs[" << a
      << "].out_valid = R.aggs[" << a << "].out_valid;\n"
      << "  L.aggs[" << a << "].op = " << P.aggs[a].op << "; L.aggs[" << a << "].col = " << P.aggs[a].col
      << "; L.aggs[" << a << "].out_dtype = " << P.aggs[a].out_dtype << ";\n";
  // the query's shape: literals
  o << "  L.n_cols = " << P.n_cols << "; L.n_keys = " << P.n_keys << "; L.n_leaves = " << P.n_leaves
    << "; L.words = " << P.words << "; L.lds_words = " << P.lds_words << "; L.lds_mode = " << P.lds_mode
    << "; L.has_first = " << P.has_first << "; L.vec_ok = " << P.vec_ok << "; L.leaf_cols = " << P.leaf_cols
    << ";\n";
  o << "  L.rw = " << P.rw << "; L.n_vals = " << P.n_vals << "; L.rec_row = " << P.rec_row
    << "; L.rec_meta = " << P.rec_meta << "; L.rec_val0 = " << P.rec_val0 << "; L.vmask_in_key = " << P.vmask_in_key
    << "; L.lds_stride = " << P.lds_stride << "; L.c_stride = " << P.c_stride << "; L.n_aggs = " << P.n_aggs << "; L.fused = " << P.fused
    << "; L.seg_slices = " << P.seg_slices << ";\n";
  for (int i = 0; i < 8; ++i) o << "  L.truth[" << i << "] = " << P.truth[i] << "u;\n";
  for (int c = 0; c < P.n_cols; ++c) {
    const DevColAgg &A = P.cagg[c];
    o << "  L.cols[" << c << "].dtype = " << P.cols[c].dtype << "; L.cols[" << c << "].is_float = " << P.cols[c].is_float
      << "; L.val_of[" << c << "] = " << (int)P.val_of[c] << "; L.val_word[" << c << "] = " << (int)P.val_word[c]
      << "; L.pk_shift[" << c << "] = " << (int)P.pk_shift[c] << "; L.pk_bits[" << c << "] = " << (int)P.pk_bits[c]
      << "; L.pk_base[" << c << "] = R.pk_base[" << c << "];\n"
      << "  L.cagg[" << c << "].mask = " << A.mask << "u; L.cagg[" << c << "].gbase = " << A.gbase << "; L.cagg[" << c
      << "].lbase = " << A.lbase << "; L.cagg[" << c << "].int_as_float = " << A.int_as_float << "; L.cagg[" << c
      << "].nfw = " << (int)A.nfw << "; L.cagg[" << c << "].nqw = " << (int)A.nqw << ";\n"
      << "  L.cagg[" << c << "].o_count = " << (int)A.o_count << "; L.cagg[" << c << "].o_isum = " << (int)A.o_isum
      << "; L.cagg[" << c << "].o_fsum = " << (int)A.o_fsum << "; L.cagg[" << c << "].o_sq = " << (int)A.o_sq
      << "; L.cagg[" << c << "].o_min = " << (int)A.o_min << "; L.cagg[" << c << "].o_max = " << (int)A.o_max
      << "; L.cagg[" << c << "].o_flags = " << (int)A.o_flags << "; L.cagg[" << c << "].o_zfirst = " << (int)A.o_zfirst
      << ";\n";
  }
  for (int c = 0; c <= P.n_cols; ++c) o << "  L.leaf_begin[" << c << "] = " << (int)P.leaf_begin[c] << ";\n";
  for (int k = 0; k < P.n_keys; ++k)
    o << "  L.keys[" << k << "].nullable = " << P.keys[k].nullable << "; L.keys[" << k << "].is_float = "
      << P.keys[k].is_float << ";\n";
  for (int l = 0; l < P.n_leaves; ++l)
    o << "  L.leaves[" << l << "].op = " << P.leaves[l].op << "; L.leaves[" << l << "].flags = " << P.leaves[l].flags
      << "; L.leaves[" << l << "].bit = " << P.leaves[l].bit << "; L.leaves[" << l << "].col2 = " << P.leaves[l].col2
      << ";\n";
  for (int w = 0; w < P.words; ++w)
    o << "  L.ident[" << w << "] = " << P.ident[w] << "ull; L.wop[" << w << "] = " << (int)P.wop[w] << ";\n";
  o << "  return L;\n}\n";
  // loads of one pass A tile: per row pair and column, the values (one
  // vector load or two) and, with a bitmap, two validity words
  int tile_loads = 0;
  for (int c = 0; c < P.n_cols; ++c) tile_loads += (P.vec_ok ? 1 : 2) + (P.cols[c].valid ? 2 : 0);
  tile_loads *= U;
  // pass A, pass C (one level) or pass C (second level)
  if (which == JIT_SCAN)
    o << "extern \"C\" __global__ void __launch_bounds__(1024) dfr_scan_jit(const DevPlan *__restrict__ plan, "
         "int64_t chunk) {\n"
      << "  const DevPlan L = make_plan(plan);\n"
      << "  scan_spill_body<" << LDSM << ", " << NC << ", " << U << ", " << (PF ? "true" : "false") << ", "
      << tile_loads << ">(L, chunk);\n}\n";
  else if (which == JIT_PART)
    o << "extern \"C\" __global__ void __launch_bounds__(1024) dfr_part_jit(const DevPlan *__restrict__ plan) {\n"
      << "  const DevPlan L = make_plan(plan);\n  part_aggregate_body(L);\n}\n";
  else
    o << "extern \"C\" __global__ void __launch_bounds__(512) DFR_SUB_WAVES dfr_sub_jit(const DevPlan *__restrict__ "
         "plan) {\n"
      << "  const DevPlan L = make_plan(plan);\n  sub_aggregate_body<" << P.n_aggs << ">(L);\n}\n";
  return o.str();
}

// The compiler options of a generated kernel for the device's architecture.
std::vector<std::string> compile_options(int device) {
  hipDeviceProp_t prop;
  std::string arch = "--offload-arch=gfx950";
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    std::string name = prop.gcnArchName;
    const size_t colon = name.find(':');
    arch = "--offload-arch=" + name.substr(0, colon);
  }
  // (no atomic optimizer: it would turn the one-lane reservation atomic of
  // the second-level pass C into a wave-wide sequence that waits for the
  // returned value at once, where the kernel means to wait for it later;
  // the kernels' other uniform-address atomics are aggregated by hand)
  return {arch, "-O3", "-std=c++17", "-I" + header_dir(), "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-mllvm", "-amdgpu-atomic-optimizer-strategy=None"};
}

// hiprtc in this process (when the compiler program cannot be started).
bool compile_here(const std::string &src, const std::vector<std::string> &opts, std::vector<char> &code) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "dfr_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) return false;
  std::vector<const char *> o;
  for (const std::string &x : opts) o.push_back(x.c_str());
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)o.size(), o.data());
  size_t n = 0;
  bool ok = r == HIPRTC_SUCCESS && hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS && n > 0;
  if (ok) {
    code.resize(n);
    ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

// The compile in a child process (dfr_jitc beside this library), so that
// a compiler crash cannot end the caller's process. Returns 1 with the
// code object, 0 when the compile failed or crashed, -1 when the compiler
// program could not be started (then the caller compiles here).
int compile_child(const std::string &src, const std::vector<std::string> &opts, std::vector<char> &code) {
  if (const char *v = getenv("DFR_JIT_INPROCESS"))
    if (*v && *v != '0') return -1;
  const std::string dir = header_dir();
  if (dir.size() < 5) return -1;
  const std::string exe = dir.substr(0, dir.size() - 5) + "/dfr_jitc";  // (header_dir is <lib dir>/csrc)
  if (access(exe.c_str(), X_OK) != 0) return -1;
  const char *tmp = getenv("TMPDIR");
  std::string base = std::string(tmp && *tmp ? tmp : "/tmp") + "/dfr_jit_XXXXXX";
  std::vector<char> path(base.begin(), base.end());
  path.push_back('\0');
  const int fd = mkstemp(path.data());
  if (fd < 0) return -1;
  const std::string src_path = path.data(), out_path = src_path + ".co";
  bool wrote = write(fd, src.data(), src.size()) == (ssize_t)src.size();
  close(fd);
  int result = -1;
  if (wrote) {
    std::vector<std::string> args = {exe, src_path, out_path};
    args.insert(args.end(), opts.begin(), opts.end());
    std::vector<char *> argv;
    for (std::string &a : args) argv.push_back(&a[0]);
    argv.push_back(nullptr);
    pid_t pid;
    if (posix_spawn(&pid, exe.c_str(), nullptr, nullptr, argv.data(), environ) == 0) {
      int status = 0;
      while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
      }
      result = 0;
      if (WIFEXITED(status) && WEXITSTATUS(status) == 0) {
        if (FILE *f = fopen(out_path.c_str(), "rb")) {
          fseek(f, 0, SEEK_END);
          const long n = ftell(f);
          fseek(f, 0, SEEK_SET);
          if (n > 0) {
            code.resize((size_t)n);
            if (fread(code.data(), 1, (size_t)n, f) == (size_t)n) result = 1;
          }
          fclose(f);
        }
      }
    }
  }
  unlink(src_path.c_str());
  unlink(out_path.c_str());
  return result;
}

JitEntry compile(const std::string &src, int device, int which) {
  JitEntry e;
  e.failed = true;
  const std::vector<std::string> opts = compile_options(device);
  std::vector<char> code;
  int got = compile_child(src, opts, code);
  if (got < 0) got = compile_here(src, opts, code) ? 1 : 0;
  if (got == 1 && hipModuleLoadData(&e.module, code.data()) == hipSuccess &&
      hipModuleGetFunction(&e.fn, e.module, kJitNames[which]) == hipSuccess)
    e.failed = false;
  return e;
}

}  // namespace

// A runtime-compiled kernel (JIT_SCAN, JIT_PART, JIT_SUB) for this plan's
// shape on the context's device, or null when it cannot be built here (then
// the precompiled kernel runs). With tune jit_wait 0, only a kernel already
// compiled is returned: a new shape is not compiled (that takes seconds).
hipFunction_t jit_function(dfr_context *ctx, const DevPlan &P, int LDSM, int NC, int U, int which) {
  const std::string src = shape_source(P, LDSM, NC, U, ctx->tune.scan_prefetch, which);
  const std::string key = std::to_string(ctx->device) + "\n" + src;
  {
    // the lock guards the cache only: a compile (seconds) runs outside it,
    // so a caller that does not wait for compiles is never held up by one
    std::unique_lock<std::mutex> g(g_mu);
    for (;;) {
      auto it = g_cache.find(key);
      if (it == g_cache.end()) break;
      if (!it->second.compiling) return it->second.failed ? nullptr : it->second.fn;
      if (!ctx->tune.jit_wait) return nullptr;
      g_cv.wait(g);  // the same shape is being compiled by another thread
    }
    if (!ctx->tune.jit_wait) return nullptr;
    g_cache[key].compiling = true;
  }
  hipSetDevice(ctx->device);
  if (const char *dir = getenv("DFR_JIT_DUMP")) {
    // the generated source, for inspection
    const std::string path =
        std::string(dir) + "/" + kJitNames[which] + "_" + std::to_string(std::hash<std::string>()(src)) + ".hip";
    if (FILE *f = fopen(path.c_str(), "w")) {
      fputs(src.c_str(), f);
      fclose(f);
    }
  }
  const JitEntry e = compile(src, ctx->device, which);
  ctx->obs.jit_compiles++;
  {
    std::lock_guard<std::mutex> g(g_mu);
    g_cache[key] = e;
  }
  g_cv.notify_all();
  return e.failed ? nullptr : e.fn;
}

// Whether jit_function has a compiled kernel for this shape already.
bool jit_compiled(dfr_context *ctx, const DevPlan &P, int LDSM, int NC, int U, int which) {
  const std::string key =
      std::to_string(ctx->device) + "\n" + shape_source(P, LDSM, NC, U, ctx->tune.scan_prefetch, which);
  std::lock_guard<std::mutex> g(g_mu);
  auto it = g_cache.find(key);
  return it != g_cache.end() && !it->second.failed && !it->second.compiling;
}

// The generated source of a plan's pass A (for tests and inspection).
std::string jit_scan_source(const DevPlan &P, int LDSM, int NC, int U, int PF) {
  return shape_source(P, LDSM, NC, U, PF, JIT_SCAN);
}
