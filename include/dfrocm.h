/*
 * dfrocm.h — C ABI of libdfrocm, the MI355X filtered group-by provider for
 * dataframe_mojo.
 *
 * The reference's accelerator boundary is thirteen Mojo functions
 * (providers/cpu/dataframe_provider/__init__.mojo:15-70, imported only by
 * dataframe/_accel_provider.mojo:10-24). A provider package written in the
 * host language implements those functions and reaches the device through
 * this library. Every entry point below is plain C: pointers, sizes and
 * integer status codes, no C++ or torch types.
 *
 * Conventions follow the reference's one native library
 * (native/dfparquet/dfparquet.cc:51-56, dataframe/io/native_library.mojo):
 *   - every fallible call returns an int status (DFR_OK on success);
 *   - on failure, when `error` is non-null, *error receives a malloc'd,
 *     NUL-terminated message that the caller frees with dfr_free();
 *   - only dfr_* symbols are exported (version script libdfrocm.map).
 *
 * Threading: one dfr_context serves one caller thread at a time. Concurrent
 * callers use distinct contexts (each owns its own HIP stream).
 */
#ifndef DFROCM_H
#define DFROCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFR_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define DFR_OK 0
/* The query is outside what the device path supports; reported before any
 * device work is submitted (accelerator contract: unsupported work raises
 * before submission, docs/design/accelerator-contract.md:95-103). */
#define DFR_UNSUPPORTED 1
/* A data fault: an aggregate's final value is out of its result type's
 * range. The message text is the reference's own error text
 * ("Int64 expression sum overflow", function.mojo:2786-2798;
 * "int32 expression sum overflow", function.mojo:2906-2914). */
#define DFR_DATA_FAULT 2
/* The query does not fit the context's HBM budget. */
#define DFR_CAPACITY 3
/* A HIP runtime error. */
#define DFR_DEVICE 4
/* Invalid arguments (null pointers, mismatched lengths, bad indexes). */
#define DFR_INVALID 5
/* dfr_server_wait: the ticket's query has not finished within the timeout. */
#define DFR_PENDING 6

/* ---- column types ------------------------------------------------------ */
#define DFR_INT32 1
#define DFR_INT64 2
#define DFR_FLOAT64 3
/* A dictionary-encoded String column: uint32 codes into `dict_size`
 * canonical labels (storage/string.mojo:35-89 StringCodes; equal labels must
 * share one code, kernels/canonical_labels.mojo:13-47). Labels stay on the
 * host; results carry codes. */
#define DFR_DICT 4

/* ---- devices (no HIP context is created) -------------------------------- */
typedef struct dfr_device_info {
    int32_t ordinal;
    int32_t compute_units;
    int64_t memory_bytes;  /* HBM size of the device */
    char name[64];         /* marketing name, e.g. "AMD Instinct MI355X" */
    char arch[32];         /* e.g. "gfx950" */
} dfr_device_info;

/* ABI version of the loaded library (DFR_ABI_VERSION it was built with). */
int dfr_abi_version(void);

/* Frees an error string or any other buffer this library malloc'd. */
void dfr_free(void *p);

/* Number of GPU agents, read from the KFD topology in sysfs. Opens no HIP
 * runtime and creates no context, as the provider's installed()/devices()
 * may not (accelerator-contract.md:27; devices.mojo:24-40). */
int dfr_device_count(int32_t *count, char **error);
int dfr_device_info_get(int32_t ordinal, dfr_device_info *info, char **error);

/* ---- context: one device, one stream, one HBM budget -------------------- */
typedef struct dfr_context dfr_context;

/* budget_bytes <= 0 means "the device's free memory". The budget bounds
 * what the context allocates (DATAFRAME_ACCEL_MEMORY_LIMIT semantics,
 * dataframe_metal/metal.mojo:36-46; query_scope.mojo:17-29). */
int dfr_context_create(int32_t device, int64_t budget_bytes, dfr_context **out,
                       char **error);
void dfr_context_destroy(dfr_context *ctx);

/* Execution observation, as dataframe/execution_observation.mojo:23-94
 * records it for a device route. */
typedef struct dfr_observation {
    int64_t launches;          /* kernels launched */
    int64_t completed;         /* launches known complete (after a sync) */
    int64_t upload_bytes;      /* host -> device */
    int64_t download_bytes;    /* device -> host */
    int64_t reserved_bytes;    /* device bytes currently held */
    int64_t peak_reserved_bytes;
    int64_t waits;             /* host waits on the stream */
    int64_t queries;           /* group-by queries executed */
    int64_t jit_compiles;      /* pass A kernels compiled at run time for a query shape */
    int64_t jit_launches;      /* pass A launches of such kernels */
    int64_t hot_keys;          /* heavy-hitter keys seeded into pass A tables, summed over queries */
} dfr_observation;
int dfr_context_observe(const dfr_context *ctx, dfr_observation *out);

/* Frees the device blocks the context keeps for reuse (released results'
 * and queries' scratch); blocks in use are not touched. */
void dfr_context_trim(dfr_context *ctx);

/* The identity of the sources this library was built from (a hash of its
 * kernels, runtime and this header). A calibration profile records the
 * build it was measured with and is applied only to that build (no
 * counterpart in the reference's C ABI; the Metal provider's profiles carry
 * the same kind of identity, dataframe_metal/calibration.mojo:29-41). */
const char *dfr_build_id(void);

/* What became of the device's calibration profile when the context was
 * created: "applied: <path> (<n> settings)" or "none: <reason>". Profiles
 * (JSON, written by the autotuner) are read from
 * $DATAFRAME_ROCM_CALIBRATION ("off": none) or from
 * calibration/<arch>-<CUs>cu.json beside the library; one is applied only
 * when its identity (device name, arch, compute units, build id) matches,
 * as the reference's identity-matched profiles
 * (dataframe_metal/calibration.mojo:1-12). *text is freed with dfr_free. */
int dfr_context_calibration(const dfr_context *ctx, char **text, char **error);

/* Sets one launch or table parameter of the context (no counterpart in the
 * reference: the autotuner's knob). Keys: "block", "wg_per_cu",
 * "lds_bytes" (threads per workgroup, persistent workgroups per CU and LDS
 * table bytes of the one-pass atomic scan and of a batch), "scan_block",
 * "scan_wg_per_cu", "scan_lds_bytes" (the same for the partitioned path's
 * pass A), "lds_probe", "lds_mode" (-1 auto, 0 none, 1 direct,
 * 2 hashed), "direct_max_slots", "hash_probe", "hash_slots", "agg_mode"
 * (1 partitioned: rows that miss the LDS table are spilled and folded per
 * key partition; 0 atomic: they fold into HBM with device atomics),
 * "part_lds_bytes" (LDS of a partition table), "scan_unroll" (row pairs a
 * scan thread loads at once: 1, 2 or 4), "two_level" (-1 auto: a second
 * partition level when the groups may outgrow one level's LDS tables; 0
 * never; 1 always), "max_parts" (first-level key partitions at most, a
 * power of two up to 4096), "sub_lds_bytes" and "sub_block" (LDS bytes and
 * threads of a second-level partition table's workgroup), "sub_fill"
 * (records per sub-partition on average, in percent of its table's entries,
 * 10 to 90; a sub-partition whose groups outgrow its table doubles the
 * sub-partitions and runs the second level again), "sub_wg_per_cu"
 * (those workgroups per CU, each looping over sub-partitions: -1 as many as
 * fit on a CU at once, 0 one workgroup per sub-partition), "batch_tile"
 * (rows of a batch work item, at most), "batch_fill" (a batch of fewer rows
 * than batch_tile x this x the grid's workgroups is cut into shorter items,
 * this many per workgroup, none shorter than a query's LDS table; 0: items of
 * batch_tile rows only), "batch_streams" (a batch's partitioned queries run on this many
 * side streams beside the batch kernel, 0 to 8; 0 and 1: on the context's
 * stream), "batch_graph" (1: a batch's launch sequence
 * is captured and run as one HIP graph), "batch_part_rows" (a batch query
 * with at least this many rows whose key domain outgrows the batch kernel's
 * LDS table runs the partitioned passes instead, after the batch kernel;
 * 0: never; under graph capture their runtime-compiled kernels are built
 * before the capture begins), "jit" (pass A compiled at run time
 * for the query's shape: 0 never, 1 when the input has at least "jit_rows"
 * rows, 2 always; the precompiled kernel runs when it cannot be built),
 * "jit_rows", "jit_wait" (0: a runtime-compiled kernel is used only when
 * it was compiled already, so no call waits seconds for a compile, and a
 * batch query is routed to the partitioned passes only then; a server's
 * lanes run with 0), "scan_prefetch" (1: the runtime-compiled pass A issues the
 * next tile's loads before it folds the current tile; 0: load, then fold),
 * "lds_odd_stride" (1: pass A's LDS table entries are an odd number of
 * 8-byte words apart; 0: exactly the state's words), "part_odd_stride" (the
 * same for the pass C partition tables; 1, the default, spreads an
 * instruction's 64-bit accesses of random entries over every LDS bank pair),
 * "part_block" (threads of a first-level partition table's workgroup),
 * "two_level_parts" (first-level partitions of a two-level query, a power
 * of two up to "max_parts"), "scatter_slices" (the partitioned pass A's
 * workgroups count their spilled records in this many histograms, one per
 * run of their waves, and each is moved by its own scatter workgroup),
 * "scatter_chunk" (records a scatter workgroup sorts at once: 8192 or
 * 16384), "scatter_stage" (1, the default: the scatters load records of 2
 * to 4 words whole, sort them through LDS and stream them out; 0: they sort
 * indices and gather the records again), "hot_keys" (the partitioned pass A
 * seeds every workgroup's hashed table with the heavy-hitter keys of a
 * sample of the kept rows, whose states then merge in one dense reduction
 * instead of through the partitions: -1, the default, from "hot_rows" rows
 * (default 2^24) when the packed key domain is at most "hot_domain"
 * (default 2^24) and the query has one partition level; 1 whenever that
 * domain allows; 0 never), "hot_sample" (rows sampled, default 2^19),
 * "hot_share" (percent of the sampled kept rows the seeded keys must carry,
 * else nothing is seeded; default 30). Seeding changes no result.
 * "rec_pack" (1, the default: integer values of spilled row records whose
 * column range fits ride in the record's key word, which shortens the
 * records the partitions move; 0: every value has its own word). */
int dfr_context_tune(dfr_context *ctx, const char *key, int64_t value, char **error);

/* The current value of a dfr_context_tune key (the built-in default, the
 * calibration profile's, or the last dfr_context_tune); no counterpart in
 * the reference: the autotuner reads each knob before it varies it. */
int dfr_context_tune_get(const dfr_context *ctx, const char *key, int64_t *value, char **error);

/* ---- device columns (the column hand-off) ------------------------------ */
typedef struct dfr_column dfr_column;

/* Copies `length` values starting at element `offset` of `values` to the
 * device. `validity` is an LSB-first bitmap (Arrow order, as
 * storage/column.mojo:1-34 and storage/bitmap.mojo:6-48), where bit
 * `bit_offset + i` is row i; NULL means every row is valid. `dict_size` is
 * the number of labels of a DFR_DICT column (ignored otherwise). Column
 * statistics (min/max, float exponent range, null count) are computed on the
 * device at upload. */
int dfr_column_upload(dfr_context *ctx, int32_t dtype, const void *values,
                      int64_t offset, int64_t length, const uint8_t *validity,
                      int64_t bit_offset, int64_t dict_size, dfr_column **out,
                      char **error);

/* Wraps device memory the caller owns (zero copy): `device_values` and
 * `device_validity` (64-bit LSB-first words, may be NULL) must stay valid
 * until dfr_column_release. Statistics are computed on the device once. */
int dfr_column_wrap(dfr_context *ctx, int32_t dtype, const void *device_values,
                    int64_t offset, int64_t length,
                    const uint64_t *device_validity, int64_t bit_offset,
                    int64_t dict_size, dfr_column **out, char **error);

/* ---- chunked and selected inputs (streaming hand-off) -------------------
 * A column that reaches the device as several pieces: the chunks of a
 * chunked Series (series.mojo chunks()), or the batches a CPU pipeline
 * hands over with a selection instead of a copy (SelectionBatch,
 * vector/selection_batches.mojo:13-75: each batch keeps its source column,
 * a start and an optional row list). Batch b contributes `count` rows,
 * after the rows of batches 0..b-1; its row i is row start + rows[i] of its
 * source window (rows NULL: start + i). The sources are host memory. Only
 * the span of source rows a batch selects is copied, in one transfer for
 * the whole column, and a device kernel gathers the rows and their
 * validity bits into one device column (which then has statistics as an
 * upload does). A row outside its source window is DFR_INVALID before any
 * transfer, as SelectionBatch.validate rejects it
 * (vector/selection_batches.mojo:34-58). */
typedef struct dfr_source {
    const void *values;      /* host: the source column's value buffer */
    int64_t offset;          /* element offset of its window */
    int64_t length;          /* rows of its window */
    const uint8_t *validity; /* LSB-first bitmap, bit bit_offset + i is window row i; NULL: all valid */
    int64_t bit_offset;
} dfr_source;
typedef struct dfr_selection {
    dfr_source source;
    int64_t start;           /* first window row the batch reads */
    const uint32_t *rows;    /* count row numbers relative to start; NULL: contiguous */
    int64_t count;           /* rows the batch contributes */
} dfr_selection;
int dfr_column_upload_selection(dfr_context *ctx, int32_t dtype, const dfr_selection *batches,
                                int32_t n_batches, int64_t dict_size, dfr_column **out,
                                char **error);

/* A window [offset, offset + length) of a column, without device work: it
 * borrows the parent's buffers (the parent must outlive it) and takes the
 * parent's statistics as bounds, so it costs no launch; the reference's
 * columns are windows the same way (storage/column.mojo:18-30). A query
 * result over a slice is the same whatever other slices run with it. */
int dfr_column_slice(const dfr_column *parent, int64_t offset, int64_t length, dfr_column **out,
                     char **error);
void dfr_column_release(dfr_column *column);
int64_t dfr_column_length(const dfr_column *column);
int64_t dfr_column_null_count(const dfr_column *column);
/* An upper estimate of the column's distinct non-null values, from a
 * HyperLogLog sketch made at hand-off (a slice reports its parent's); it
 * bounds the groups a key column can form, which sizes hash tables. */
int64_t dfr_column_distinct(const dfr_column *column);

/* ---- the query ---------------------------------------------------------- */
/* Predicate leaves: `column <op> literal`, or `floor_mod(column, modulus)
 * <op> literal` when DFR_LEAF_MOD is set, or `column <op> other column`
 * when DFR_LEAF_COLUMN is set (`literal` is then the other column's index
 * into dfr_query.columns: both Int32/Int64, compared by value, or both
 * Float64; no modulus), or a null test. A comparison with
 * a null value is not true (a null is neither true nor false, so it stays
 * not true under DFR_LEAF_NOT as well). Float comparisons are IEEE (NaN
 * compares false, NaN != NaN is true; docs/expressions.md:93-94). The
 * modulo takes the divisor's sign and modulo zero is null
 * (docs/expressions.md:91). `literal` holds an int64 value, or the bits of
 * a double when the column is DFR_FLOAT64. */
#define DFR_EQ 0
#define DFR_NE 1
#define DFR_LT 2
#define DFR_LE 3
#define DFR_GT 4
#define DFR_GE 5
#define DFR_IS_NULL 6
#define DFR_IS_NOT_NULL 7
#define DFR_LEAF_MOD 1 /* leaf flag: compare floor_mod(column, modulus) */
#define DFR_LEAF_NOT 2 /* leaf flag: Kleene NOT of the comparison */
#define DFR_LEAF_COLUMN 4 /* leaf flag: the right-hand side is column `literal` */
#define DFR_MAX_LEAVES 8

typedef struct dfr_leaf {
    int32_t column;  /* index into dfr_query.columns */
    int32_t op;      /* DFR_EQ .. DFR_IS_NOT_NULL */
    int32_t flags;   /* DFR_LEAF_MOD, DFR_LEAF_NOT, DFR_LEAF_COLUMN */
    int32_t pad;
    int64_t literal; /* int64 value or double bits; a column index under DFR_LEAF_COLUMN */
    int64_t modulus; /* for DFR_LEAF_MOD; modulo by zero is null */
} dfr_leaf;

/* A row is kept when bit m of the truth table is set, m being the leaves'
 * is-true bits (leaf i is bit i). With no leaves every row is kept. NOT is
 * pushed into the leaves by the host compiler, so AND/OR of is-true bits
 * is SQL filter semantics even with nulls
 * (dataframe_metal/predicate_program.mojo:1-71). */
typedef struct dfr_predicate {
    int32_t n_leaves;
    int32_t pad;
    uint32_t truth[8]; /* 256 bits */
    dfr_leaf leaves[DFR_MAX_LEAVES];
} dfr_predicate;

/* Aggregates (expr.mojo:2869-3006). */
#define DFR_AGG_LEN 0   /* rows in the group, Int64 */
#define DFR_AGG_COUNT 1 /* non-null values, Int64 */
#define DFR_AGG_SUM 2   /* Int64 for Int64, Int32 for Int32, Float64 */
#define DFR_AGG_MIN 3   /* input type */
#define DFR_AGG_MAX 4   /* input type */
#define DFR_AGG_MEAN 5  /* Float64 */
#define DFR_AGG_VAR 6   /* Float64, ddof */
#define DFR_AGG_STD 7   /* Float64, ddof */

typedef struct dfr_agg {
    int32_t op;
    int32_t column;    /* index into dfr_query.columns (ignored for LEN) */
    int32_t ddof;      /* VAR/STD; the reference's default is 1 */
    int32_t pad;
    int64_t min_count; /* SUM: null when fewer valid values (default 0) */
} dfr_agg;

#define DFR_MAX_KEYS 4
#define DFR_MAX_COLUMNS 12

/* filter(predicate).group_by(keys, maintain_order).agg(aggs) over columns of
 * one frame (lazy_api.mojo:418-470,3223). Keys and columns all have the same
 * length. A key may also appear among the columns. */
typedef struct dfr_query {
    int32_t n_keys;
    int32_t n_columns;
    dfr_column *const *keys;
    dfr_column *const *columns;
    dfr_predicate predicate;
    int32_t n_aggs;
    int32_t maintain_order; /* groups in order of their first kept row */
    const dfr_agg *aggs;
} dfr_query;

/* ---- value programs: arithmetic and CASE inside aggregates --------------
 * An aggregate may reduce an expression instead of a column, e.g.
 * (col("price") * col("quantity")).sum() or
 * when(col("q") > 0).then(col("p") * 2).otherwise(0).mean()
 * (docs/apple-metal.md:104-120, 635-650, 711-725). dfr_column_compute
 * evaluates such an expression, a postfix program, for every row the `keep`
 * predicate keeps (the query's filter: other rows are null and never
 * evaluated) into a new device column that a query then aggregates.
 * Semantics (docs/expressions.md:50-100): nulls propagate through
 * arithmetic; integer + - * and unary - / abs are checked at the node's
 * width (Int32 nodes compute wider and range-check) and an overflow in a
 * kept row is a data fault with the reference's text ("Int64 sum
 * overflow", "Int64 subtraction overflow", "Int64 multiplication
 * overflow", "int32 addition overflow", ..., kernels/binary.mojo:108-122,
 * 415-421, 1620-1635), also when a later operand makes the result null;
 * when several nodes overflow, the first in program order is reported, as
 * the CPU evaluates node by node. `/` is IEEE Float64 division. CASE takes
 * its THEN value where its condition is true and its ELSE value otherwise
 * (a null condition takes ELSE); only the taken branch's overflow counts. */
#define DFR_VAL_COLUMN 0   /* push input `arg` */
#define DFR_VAL_LITERAL 1  /* push `literal` (int64 value or double bits) */
#define DFR_VAL_NULL 2     /* push a null */
#define DFR_VAL_ADD 3
#define DFR_VAL_SUB 4
#define DFR_VAL_MUL 5
#define DFR_VAL_DIV 6      /* Float64 only */
#define DFR_VAL_NEG 7
#define DFR_VAL_ABS 8
#define DFR_VAL_CASE 9     /* pops ELSE, then THEN; pushes condition `arg` ? THEN : ELSE */
#define DFR_MAX_VALUE_OPS 32
#define DFR_MAX_VALUE_INPUTS 8
#define DFR_MAX_CONDITIONS 4
#define DFR_MAX_VALUE_STACK 8

typedef struct dfr_value_op {
    int32_t op;
    int32_t arg;      /* COLUMN: input index; CASE: condition index */
    int32_t dtype;    /* the node's type: DFR_INT32, DFR_INT64 or DFR_FLOAT64; operands convert to it */
    int32_t pad;
    int64_t literal;  /* LITERAL: an int64 value or the bits of a double */
} dfr_value_op;

typedef struct dfr_value_program {
    int32_t n_ops;
    int32_t n_conditions;
    dfr_value_op ops[DFR_MAX_VALUE_OPS];
    dfr_predicate conditions[DFR_MAX_CONDITIONS];  /* leaf columns index the inputs */
    dfr_predicate keep;                            /* rows evaluated; the others are null */
} dfr_value_program;

/* Computes `program` over `inputs` (equal lengths; Int32, Int64, Float64, and
 * dictionary columns in conditions only) into a new column of the last
 * node's type. DFR_DATA_FAULT (reference text) on a checked overflow in a
 * kept row; DFR_INVALID for a malformed program (checked before any device
 * work: stack depth, operand indexes, types). */
int dfr_column_compute(dfr_context *ctx, const dfr_value_program *program, dfr_column *const *inputs,
                       int32_t n_inputs, dfr_column **out, char **error);

/* ---- results (device resident until copied) ---------------------------- */
typedef struct dfr_result dfr_result;

/* Runs one query. On DFR_OK *out holds the result, whose buffers stay on
 * the device (storage/residence.mojo:33-80) until copied out. */
int dfr_groupby(dfr_context *ctx, const dfr_query *query, dfr_result **out,
                char **error);

/* Runs n independent queries as one batch (the reference has no batch
 * entry point; this serves callers with many small queries, for which one
 * launch sequence per query costs more than the work). One persistent
 * kernel drains a device work queue of (query, row tile) items, and the
 * batch is compacted and finished together, with one host synchronisation.
 * results[i] holds what dfr_groupby(queries[i]) returns: the same groups and
 * values, bit for bit (group order may differ: without maintain_order it is
 * unspecified). statuses[i] is that query's status: a query that fails to
 * compile (DFR_UNSUPPORTED, DFR_INVALID, maintain_order queries included)
 * or faults (DFR_DATA_FAULT) gets no result and, when `errors` is non-null,
 * a message in errors[i] to free with dfr_free. The call's own status is
 * DFR_OK when the batch ran. */
int dfr_groupby_batch(dfr_context *ctx, const dfr_query *queries, int32_t n, dfr_result **results,
                      int32_t *statuses, char **errors, char **error);

/* ---- serving ------------------------------------------------------------
 * A request loop over one device (the reference has no server; this is for
 * callers that answer many concurrent queries over device-resident
 * columns). `lanes` worker threads, each with its own context and stream,
 * take waiting queries in arrival order, as many as fit `max_batch` and the
 * device bytes (`budget_bytes`, 0: no limit) not held by queries in flight,
 * and run them as one batch (dfr_groupby_batch). Queries that arrive while
 * batches run wait only until a lane frees, so with two or more lanes their
 * work overlaps on the device. A result equals what dfr_groupby gives for
 * the query alone. The query's columns may belong to any context on the
 * same device and must outlive the wait; results come from the lanes'
 * contexts, and each keeps its lane's context alive: it may be released
 * before or after dfr_server_destroy. */
typedef struct dfr_server dfr_server;

typedef struct dfr_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t batches;
    int64_t max_batch;       /* largest batch run */
    int64_t concurrent_max;  /* most lanes running at once */
    int64_t queued;          /* waiting for a lane now */
    int64_t in_flight_bytes; /* admission estimate held by running batches */
} dfr_server_stats;

int dfr_server_create(dfr_context *ctx, int32_t lanes, int32_t max_batch, int64_t budget_bytes,
                      dfr_server **out, char **error);
/* Queues a query (copied; its columns are not) and returns its ticket. */
int dfr_server_submit(dfr_server *server, const dfr_query *query, int64_t *ticket, char **error);
/* Waits up to timeout_ms (< 0: no limit) for a ticket; returns the query's
 * status (DFR_PENDING on time-out) and hands over its result and, when
 * `latency_ns` is non-null, the time from its submission to the end of its
 * batch. A ticket is answered once. */
int dfr_server_wait(dfr_server *server, int64_t ticket, int64_t timeout_ms, dfr_result **out,
                    int64_t *latency_ns, char **error);
int dfr_server_stats_get(dfr_server *server, dfr_server_stats *out);
/* Runs every queued query, then stops the lanes. */
void dfr_server_destroy(dfr_server *server);

int64_t dfr_result_groups(const dfr_result *result);
/* DFR_INT32 / DFR_INT64 / DFR_FLOAT64 / DFR_DICT of key k or aggregate a. */
int32_t dfr_result_key_dtype(const dfr_result *result, int32_t k);
int32_t dfr_result_agg_dtype(const dfr_result *result, int32_t a);

/* Copies key k (or aggregate a) to host memory: `values` receives groups
 * values of the column's width (4 bytes for DFR_INT32/DFR_DICT, else 8);
 * `validity` (may be NULL) receives ceil(groups/8) bytes of LSB-first bits.
 * Row order is the result's group order. */
int dfr_result_copy_key(const dfr_result *result, int32_t k, void *values,
                        uint8_t *validity, char **error);
int dfr_result_copy_agg(const dfr_result *result, int32_t a, void *values,
                        uint8_t *validity, char **error);
/* Each group's first kept row (its ordering key under maintain_order). */
int dfr_result_copy_first_rows(const dfr_result *result, int64_t *rows,
                               char **error);
/* Device bytes the result holds. */
int64_t dfr_result_device_bytes(const dfr_result *result);
/* Releases the caller's handle; the device memory goes once no dfr_buffer
 * of the result is held either. */
void dfr_result_release(dfr_result *result);

/* ---- result buffers as device residences --------------------------------
 * The reference keeps a buffer that lives in accelerator memory as a
 * _DeviceResidence: an opaque context and two C callbacks,
 * _CopyToHost(context, host address, bytes) -> 0 on success and
 * _ReleaseDevice(context) (dataframe/storage/residence.mojo:20-22, 33-80).
 * A dfr_buffer is that context for one output buffer of a result; it keeps
 * its result alive until released, so the result handle may be released
 * first. dfr_buffer_copy_to_host and dfr_buffer_release have exactly the
 * callbacks' shapes (Mojo Int is 64 bits), so their addresses are the
 * residence's `to_host` and `release`. */
#define DFR_BUFFER_KEY 0        /* kind: key column `index` */
#define DFR_BUFFER_AGG 1        /* kind: aggregate `index` */
#define DFR_BUFFER_FIRST_ROWS 2 /* kind: each group's first kept row (int64; maintain_order) */
#define DFR_BUFFER_VALUES 0     /* part: groups values of the column's width */
#define DFR_BUFFER_VALIDITY 1   /* part: ceil(groups / 8) bytes of LSB-first bits */
typedef struct dfr_buffer dfr_buffer;
/* A residence handle for one buffer of `result`; *bytes receives its size. */
int dfr_result_buffer(dfr_result *result, int32_t kind, int32_t index, int32_t part, dfr_buffer **out,
                      int64_t *bytes, char **error);
/* _CopyToHost: copies the first `bytes` bytes (at most the buffer's size) of
 * the buffer to host_dst; returns 0 (DFR_OK) on success, else a status. */
int64_t dfr_buffer_copy_to_host(void *buffer, void *host_dst, int64_t bytes);
/* _ReleaseDevice. */
void dfr_buffer_release(void *buffer);

/* ---- a group-by that allocates nothing ---------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid it).
 * dfr_groupby_into runs a query like dfr_groupby, but writes the result
 * into the caller's device buffers and calls no hipMalloc or hipFree (nor
 * any allocating library routine). Its other device memory (plan, group
 * table, compaction arrays, under maintain_order the row bitmap) comes from
 * a fixed per-GPU scratch array of 64 MiB held until the process ends, so a
 * query whose table and arrays exceed it fails with DFR_CAPACITY before any
 * work, as does one with more groups than `capacity`. Every call on one GPU
 * shares that GPU's scratch: calls that overlap on one GPU race, so callers
 * serialise them. It runs the one-pass aggregation (no partition buffers),
 * and its columns' own statistics must settle the Float64 accumulators
 * (DFR_UNSUPPORTED otherwise). On DFR_OK, *groups holds the group count:
 * rows [0, groups) of every output, in the order dfr_groupby would give
 * under maintain_order, unspecified otherwise. */
typedef struct dfr_output_column {
    void *values;       /* device: capacity values of the column's width (4 bytes for
                           DFR_INT32/DFR_DICT, else 8) */
    uint64_t *validity; /* device: capacity / 64 + 1 words of LSB-first bits */
} dfr_output_column;
typedef struct dfr_output {
    int64_t capacity;          /* groups the buffers hold */
    dfr_output_column *keys;   /* query->n_keys of them (host array) */
    dfr_output_column *aggs;   /* query->n_aggs of them (host array) */
    int64_t *first_rows;       /* device, capacity entries; required under maintain_order */
} dfr_output;
int dfr_groupby_into(dfr_context *ctx, const dfr_query *query, const dfr_output *out, int64_t *groups,
                     char **error);

/* Milliseconds the device spent on the query's kernels (HIP events on the
 * context's stream). */
double dfr_result_device_ms(const dfr_result *result);

/* Milliseconds of one phase of the query, from HIP events recorded around
 * it on the context's stream: DFR_PHASE_INIT (group-table set-up),
 * DFR_PHASE_SCAN (the fused filter + aggregate scan), DFR_PHASE_ORDER
 * (compaction and ordering), DFR_PHASE_FINISH (keys and results). */
#define DFR_PHASE_INIT 0
#define DFR_PHASE_SCAN 1
#define DFR_PHASE_ORDER 2
#define DFR_PHASE_FINISH 3
double dfr_result_phase_ms(const dfr_result *result, int32_t phase);

#ifdef __cplusplus
}
#endif
#endif /* DFROCM_H */
