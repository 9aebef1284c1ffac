// Value programs: an aggregate's expression (checked arithmetic, CASE)
// evaluated over the rows a query keeps into a new device column, which the
// query then aggregates like any other column (dfr_column_compute).
//
// One thread per row, 64 consecutive rows per wave, so a wave writes one
// validity word of the output. The program is a postfix list of at most 32
// nodes, each with its own type (Int32, Int64 or Float64; operands convert
// to it), evaluated on a stack of at most 8 entries. An entry carries its
// value, its validity and a "poison" mask: the nodes whose checked integer
// arithmetic overflowed on valid operands while producing it. CASE keeps
// only the taken branch's poison, so an overflow in an untaken branch never
// faults; a kept row's final poison is OR'd into the call's fault mask, and
// the host reports the first faulting node in program order (the CPU
// evaluates node by node: kernels/binary.mojo).
//
// Included at the end of groupby.hip (shares its column loads, leaf
// evaluation and pool).

namespace {

constexpr int kValThreads = 256;

struct DevValueOp {
  int32_t op, arg, dtype, pad;
  int64_t lit;
};

struct DevCondition {
  int32_t n_leaves;
  uint32_t truth[8];
  int32_t col[kMaxLeaves];  // input of each leaf
  DevLeaf leaf[kMaxLeaves];
};

struct DevValueProgram {
  int32_t n_ops, n_conds, n_inputs, out_dtype;
  int64_t rows;
  DevValueOp ops[DFR_MAX_VALUE_OPS];
  DevCondition conds[DFR_MAX_CONDITIONS + 1];  // the last one is `keep`
  DevCol in[DFR_MAX_VALUE_INPUTS];
};

__device__ __forceinline__ uint64_t val_load(const DevCol &c, int64_t row, bool *valid) {
  uint64_t v;
  if (c.dtype == T_INT32)
    v = (uint64_t)(int64_t)((const int32_t *)c.values)[row];
  else if (c.dtype == T_DICT)
    v = (uint64_t)((const uint32_t *)c.values)[row];
  else
    v = ((const uint64_t *)c.values)[row];
  if (c.valid) {
    const uint64_t b = (uint64_t)(c.vbit + row);
    *valid = ((c.valid[b >> 6] >> (b & 63)) & 1ull) != 0;
  } else {
    *valid = true;
  }
  return v;
}

// A condition's Kleene is-true for one row (a null takes the false side).
__device__ __forceinline__ bool cond_true(const DevValueProgram &P, const DevCondition &C, int64_t row) {
  uint32_t m = 0;
  for (int l = 0; l < C.n_leaves; ++l) {
    const DevCol &a = P.in[C.col[l]];
    bool ok = false, ok2 = false;
    const uint64_t v = val_load(a, row, &ok);
    uint64_t v2 = 0;
    if (C.leaf[l].col2 >= 0) v2 = val_load(P.in[C.leaf[l].col2], row, &ok2);
    // (leaf_true applies DFR_LEAF_NOT and keeps a null "not true")
    m |= (uint32_t)leaf_true(C.leaf[l], a.is_float != 0, v, ok, v2, ok2) << l;
  }
  return ((C.truth[m >> 5] >> (m & 31)) & 1u) != 0;
}

__global__ void __launch_bounds__(kValThreads)
    k_value_program(const DevValueProgram *__restrict__ prog, void *out_values, uint64_t *out_valid,
                    uint32_t *fault) {
  const DevValueProgram &P = *prog;
  const int64_t row = (int64_t)blockIdx.x * kValThreads + threadIdx.x;
  const bool live = row < P.rows;
  bool ok = false;
  uint32_t poison = 0;
  if (live && cond_true(P, P.conds[P.n_conds], row)) {
    uint64_t sv[DFR_MAX_VALUE_STACK];
    bool sok[DFR_MAX_VALUE_STACK];
    uint32_t sp_[DFR_MAX_VALUE_STACK];
    int8_t st[DFR_MAX_VALUE_STACK];  // each entry's type
    int sp = 0;
    for (int i = 0; i < P.n_ops; ++i) {
      const DevValueOp o = P.ops[i];
      const int T = o.dtype;
      auto as = [&](int k) -> uint64_t {  // entry k converted to T
        if (T == T_FLOAT64 && st[k] != T_FLOAT64) return f64_bits((double)(int64_t)sv[k]);
        return sv[k];
      };
      switch (o.op) {
        case DFR_VAL_COLUMN: {
          const DevCol &c = P.in[o.arg];
          bool v_ok;
          sv[sp] = val_load(c, row, &v_ok);
          sok[sp] = v_ok;
          sp_[sp] = 0;
          st[sp] = (int8_t)(c.dtype == T_FLOAT64 ? T_FLOAT64 : (c.dtype == T_INT32 ? T_INT32 : T_INT64));
          ++sp;
          break;
        }
        case DFR_VAL_LITERAL:
        case DFR_VAL_NULL:
          sv[sp] = (uint64_t)o.lit;
          sok[sp] = o.op == DFR_VAL_LITERAL;
          sp_[sp] = 0;
          st[sp] = (int8_t)T;
          ++sp;
          break;
        case DFR_VAL_NEG:
        case DFR_VAL_ABS: {
          const int k = sp - 1;
          const uint64_t x = as(k);
          bool over = false;
          uint64_t r;
          if (T == T_FLOAT64) {
            const double d = bits_f64(x);
            r = f64_bits(o.op == DFR_VAL_NEG ? -d : fabs(d));
          } else {
            const int64_t lo = T == T_INT32 ? (int64_t)INT32_MIN : INT64_MIN;
            const int64_t a = (int64_t)x;
            over = a == lo;
            r = (uint64_t)((o.op == DFR_VAL_NEG || a < 0) ? (int64_t)(0ull - (uint64_t)a) : a);
          }
          if (over && sok[k]) sp_[k] |= 1u << i;
          sv[k] = r;
          st[k] = (int8_t)T;
          break;
        }
        case DFR_VAL_CASE: {
          // THEN below ELSE on the stack
          const int e = sp - 1, t = sp - 2;
          const bool c = cond_true(P, P.conds[o.arg], row);
          const int k = c ? t : e;
          const uint64_t x = as(k);
          sv[t] = x;
          sok[t] = sok[k];
          sp_[t] = sp_[k];
          st[t] = (int8_t)T;
          --sp;
          break;
        }
        default: {  // binary arithmetic
          const int b = sp - 1, a = sp - 2;
          const uint64_t x = as(a), y = as(b);
          const bool both = sok[a] && sok[b];
          bool over = false;
          uint64_t r;
          if (T == T_FLOAT64) {
            const double p = bits_f64(x), q = bits_f64(y);
            double z;
            switch (o.op) {
              case DFR_VAL_ADD: z = p + q; break;
              case DFR_VAL_SUB: z = p - q; break;
              case DFR_VAL_MUL: z = p * q; break;
              default: z = p / q; break;
            }
            r = f64_bits(z);
          } else {
            const int64_t p = (int64_t)x, q = (int64_t)y;
            long long z;
            switch (o.op) {
              case DFR_VAL_ADD: over = __builtin_add_overflow((long long)p, (long long)q, &z); break;
              case DFR_VAL_SUB: over = __builtin_sub_overflow((long long)p, (long long)q, &z); break;
              default: over = __builtin_mul_overflow((long long)p, (long long)q, &z); break;
            }
            if (T == T_INT32 && !over) over = z < (long long)INT32_MIN || z > (long long)INT32_MAX;
            r = (uint64_t)z;
          }
          sv[a] = r;
          sok[a] = both;
          sp_[a] = sp_[a] | sp_[b] | ((over && both) ? (1u << i) : 0u);
          st[a] = (int8_t)T;
          --sp;
          break;
        }
      }
    }
    ok = sok[0];
    poison = sp_[0];
    if (ok) {
      if (P.out_dtype == T_INT32)
        ((int32_t *)out_values)[row] = (int32_t)(int64_t)sv[0];
      else
        ((uint64_t *)out_values)[row] = sv[0];
    }
  }
  if (poison) atomicOr(fault, poison);
  const uint64_t ball = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && row < P.rows) out_valid[row >> 6] = ball;
}

const char *value_fault_text(int op, int dtype) {
  const bool i64 = dtype == DFR_INT64;
  switch (op) {
    case DFR_VAL_ADD: return i64 ? "Int64 sum overflow" : "int32 addition overflow";
    case DFR_VAL_SUB: return i64 ? "Int64 subtraction overflow" : "int32 subtraction overflow";
    case DFR_VAL_MUL: return i64 ? "Int64 multiplication overflow" : "int32 multiplication overflow";
    case DFR_VAL_NEG: return i64 ? "Int64 negation overflow" : "int32 negation overflow";
    default: return i64 ? "Int64 abs overflow" : "int32 abs overflow";
  }
}

int compile_condition(const dfr_predicate &pr, int n_inputs, const int32_t *in_dtype, DevCondition &C,
                      const char *what, char **error) {
  if (pr.n_leaves < 0 || pr.n_leaves > DFR_MAX_LEAVES)
    return dfr_fail(error, DFR_INVALID, "%s has %d leaves", what, pr.n_leaves);
  C.n_leaves = pr.n_leaves;
  for (int i = 0; i < 8; ++i) C.truth[i] = pr.truth[i];
  for (int l = 0; l < pr.n_leaves; ++l) {
    const dfr_leaf &L = pr.leaves[l];
    if (L.column < 0 || L.column >= n_inputs)
      return dfr_fail(error, DFR_INVALID, "%s leaf %d names input %d", what, l, L.column);
    if (L.op < DFR_EQ || L.op > DFR_IS_NOT_NULL)
      return dfr_fail(error, DFR_INVALID, "%s leaf %d: unknown comparison %d", what, l, L.op);
    if ((L.flags & DFR_LEAF_MOD) && in_dtype[L.column] == DFR_FLOAT64)
      return dfr_fail(error, DFR_UNSUPPORTED, "%s leaf %d: Float64 modulo", what, l);
    int col2 = -1;
    if (L.flags & DFR_LEAF_COLUMN) {
      if (L.literal < 0 || L.literal >= n_inputs)
        return dfr_fail(error, DFR_INVALID, "%s leaf %d compares with input %lld", what, l, (long long)L.literal);
      const int32_t a = in_dtype[L.column], b = in_dtype[L.literal];
      const bool ints = (a == DFR_INT32 || a == DFR_INT64) && (b == DFR_INT32 || b == DFR_INT64);
      if ((L.flags & DFR_LEAF_MOD) || L.op > DFR_GE || !(ints || (a == DFR_FLOAT64 && b == DFR_FLOAT64)))
        return dfr_fail(error, DFR_UNSUPPORTED, "%s leaf %d: column comparisons take two integer or two Float64 "
                        "columns", what, l);
      col2 = (int)L.literal;
    }
    C.col[l] = L.column;
    DevLeaf &D = C.leaf[l];
    D.op = L.op;
    D.flags = L.flags;
    D.bit = l;
    D.col2 = col2;
    D.lit = L.literal;
    D.mod = L.modulus;
  }
  return DFR_OK;
}

}  // namespace

extern "C" int dfr_column_compute(dfr_context *ctx, const dfr_value_program *program, dfr_column *const *inputs,
                                  int32_t n_inputs, dfr_column **out, char **error) {
  if (!ctx || !program || !out || (n_inputs > 0 && !inputs)) return dfr_fail(error, DFR_INVALID, "null argument");
  if (n_inputs < 1 || n_inputs > DFR_MAX_VALUE_INPUTS)
    return dfr_fail(error, DFR_INVALID, "a value program reads 1 to %d inputs, not %d", DFR_MAX_VALUE_INPUTS, n_inputs);
  const dfr_value_program &G = *program;
  if (G.n_ops < 1 || G.n_ops > DFR_MAX_VALUE_OPS)
    return dfr_fail(error, DFR_INVALID, "a value program has 1 to %d nodes, not %d", DFR_MAX_VALUE_OPS, G.n_ops);
  if (G.n_conditions < 0 || G.n_conditions > DFR_MAX_CONDITIONS)
    return dfr_fail(error, DFR_INVALID, "a value program has at most %d conditions", DFR_MAX_CONDITIONS);
  static_assert(sizeof(DevValueProgram) + 64 <= kStagingBytes, "staging holds a value program");
  DevValueProgram P;
  memset(&P, 0, sizeof P);
  int32_t in_dtype[DFR_MAX_VALUE_INPUTS];
  const int64_t rows = inputs[0] ? inputs[0]->length : 0;
  for (int i = 0; i < n_inputs; ++i) {
    const dfr_column *c = inputs[i];
    if (!c) return dfr_fail(error, DFR_INVALID, "input %d is null", i);
    if (c->ctx->device != ctx->device) return dfr_fail(error, DFR_INVALID, "input %d is on another device", i);
    if (c->length != rows) return dfr_fail(error, DFR_INVALID, "inputs have different lengths");
    in_dtype[i] = c->dtype;
    P.in[i] = c->dev();
  }
  // the stack, simulated: operand counts, depth and types (an operand is
  // never wider than its node: Int32 < Int64 < Float64)
  int depth = 0;
  int tstack[DFR_MAX_VALUE_OPS + 1];
  auto rank = [](int32_t d) { return d == DFR_INT32 ? 0 : (d == DFR_INT64 ? 1 : 2); };
  for (int i = 0; i < G.n_ops; ++i) {
    const dfr_value_op &o = G.ops[i];
    if (o.dtype != DFR_INT32 && o.dtype != DFR_INT64 && o.dtype != DFR_FLOAT64)
      return dfr_fail(error, DFR_INVALID, "node %d has type %d", i, o.dtype);
    int pops = 0;
    switch (o.op) {
      case DFR_VAL_COLUMN:
        if (o.arg < 0 || o.arg >= n_inputs) return dfr_fail(error, DFR_INVALID, "node %d reads input %d", i, o.arg);
        if (in_dtype[o.arg] == DFR_DICT)
          return dfr_fail(error, DFR_UNSUPPORTED, "node %d: dictionary columns are not values", i);
        if ((in_dtype[o.arg] == DFR_FLOAT64) != (o.dtype == DFR_FLOAT64) ||
            (in_dtype[o.arg] == DFR_INT64 && o.dtype == DFR_INT32))
          return dfr_fail(error, DFR_INVALID, "node %d: input %d is not of the node's type", i, o.arg);
        break;
      case DFR_VAL_LITERAL:
      case DFR_VAL_NULL: break;
      case DFR_VAL_NEG:
      case DFR_VAL_ABS: pops = 1; break;
      case DFR_VAL_DIV:
        if (o.dtype != DFR_FLOAT64) return dfr_fail(error, DFR_INVALID, "node %d: division is Float64", i);
        pops = 2;
        break;
      case DFR_VAL_ADD:
      case DFR_VAL_SUB:
      case DFR_VAL_MUL: pops = 2; break;
      case DFR_VAL_CASE:
        if (o.arg < 0 || o.arg >= G.n_conditions)
          return dfr_fail(error, DFR_INVALID, "node %d takes condition %d", i, o.arg);
        pops = 2;
        break;
      default: return dfr_fail(error, DFR_INVALID, "node %d: unknown operation %d", i, o.op);
    }
    if (depth < pops) return dfr_fail(error, DFR_INVALID, "node %d pops an empty stack", i);
    for (int k = depth - pops; k < depth; ++k)
      if (rank(tstack[k]) > rank(o.dtype))
        return dfr_fail(error, DFR_INVALID, "node %d: an operand is wider than the node's type", i);
    depth += 1 - pops;
    if (depth <= DFR_MAX_VALUE_STACK) tstack[depth - 1] = o.op == DFR_VAL_COLUMN ? in_dtype[o.arg] : o.dtype;
    if (depth > DFR_MAX_VALUE_STACK)
      return dfr_fail(error, DFR_INVALID, "the program needs more than %d stack entries", DFR_MAX_VALUE_STACK);
    P.ops[i] = DevValueOp{o.op, o.arg, o.dtype == DFR_INT32 ? T_INT32 : (o.dtype == DFR_INT64 ? T_INT64 : T_FLOAT64),
                          0, o.literal};
  }
  if (depth != 1) return dfr_fail(error, DFR_INVALID, "the program leaves %d values", depth);
  int st;
  for (int c = 0; c < G.n_conditions; ++c)
    if ((st = compile_condition(G.conditions[c], n_inputs, in_dtype, P.conds[c], "condition", error))) return st;
  if ((st = compile_condition(G.keep, n_inputs, in_dtype, P.conds[G.n_conditions], "keep", error))) return st;
  P.n_ops = G.n_ops;
  P.n_conds = G.n_conditions;
  P.n_inputs = n_inputs;
  const int32_t out_dtype = G.ops[G.n_ops - 1].dtype;
  P.out_dtype = out_dtype == DFR_INT32 ? T_INT32 : (out_dtype == DFR_INT64 ? T_INT64 : T_FLOAT64);
  P.rows = rows;
  HIP_TRY(hipSetDevice(ctx->device));
  dfr_column *col = new dfr_column();
  col->ctx = ctx;
  col->dtype = out_dtype;
  col->length = rows;
  const size_t w = out_dtype == DFR_INT32 ? 4 : 8;
  const uint64_t nwords = ((uint64_t)rows + 63) / 64 + 1;
  void *dprog = nullptr;
  struct Guard {
    dfr_context *ctx;
    dfr_column *&col;
    void *&dprog;
    ~Guard() {
      if (dprog) ctx->pool.release(dprog);
      if (col) dfr_column_release(col);
    }
  } guard{ctx, col, dprog};
  if ((st = ctx->pool.alloc(std::max<size_t>(16, (size_t)rows * w), &col->owned_values, error))) return st;
  if ((st = ctx->pool.alloc(nwords * 8, &col->owned_valid, error))) return st;
  if ((st = ctx->pool.alloc(sizeof(DevValueProgram) + 64, &dprog, error))) return st;
  col->values = col->owned_values;
  col->valid = (const uint64_t *)col->owned_valid;
  col->vbit = 0;
  uint32_t *dfault = (uint32_t *)((char *)dprog + sizeof(DevValueProgram));
  // (the program and the fault word travel in one copy; the host's staging
  // bytes are not rewritten before the synchronisation below)
  memcpy(ctx->staging, &P, sizeof P);
  memset(ctx->staging + sizeof P, 0, 64);
  HIP_TRY(hipMemcpyAsync(dprog, ctx->staging, sizeof P + 64, hipMemcpyHostToDevice, ctx->stream));
  if ((st = fill(ctx, (uint64_t *)col->owned_valid, nwords, 0))) return st;
  if (rows > 0) {
    hipLaunchKernelGGL(k_value_program, dim3((unsigned)((rows + kValThreads - 1) / kValThreads)), dim3(kValThreads),
                       0, ctx->stream, (const DevValueProgram *)dprog, col->owned_values, (uint64_t *)col->owned_valid,
                       dfault);
    ctx->obs.launches++;
  }
  uint32_t *hfault = (uint32_t *)(ctx->staging + kReadback);
  HIP_TRY(hipMemcpyAsync(hfault, dfault, 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->obs.waits++;
  if (*hfault) {
    const int first = __builtin_ctz(*hfault);
    return dfr_fail(error, DFR_DATA_FAULT, "%s", value_fault_text(G.ops[first].op, G.ops[first].dtype));
  }
  if ((st = column_stats(ctx, col, error))) return st;
  *out = col;
  col = nullptr;
  return DFR_OK;
}
