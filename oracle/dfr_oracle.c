/*
 * dfr_oracle.c — TEST INFRASTRUCTURE ONLY. A CPU restatement of the
 * reference CPU provider's grouped aggregation, used as the checker of
 * libdfrocm. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline load it; the product never does.
 *
 * The reference (Mojo) cannot be built here, so this file restates its
 * algorithm from the source:
 *   - group rows: null keys form one group; groups in order of their first
 *     kept row (dataframe/aggregation/hash_table.mojo:1-49, 44-47;
 *     state.mojo:267 _merge_ascending_runs)
 *   - per-row state updates, kind by kind (function.mojo:187-256 kinds,
 *     :611-673 _slot_kind, :1143-1235 _scatter): int sums in 128 bits
 *     (_K_INT_SUM) or 64 bits for <=32-bit inputs (_K_NARROW_SUM), float sums
 *     added in row order (_K_FLOAT_SUM), Welford moments (_K_MOMENTS,
 *     scalar_states.mojo:103-144), min/max with a best value and flags
 *     (1 seen, 2 NaN seen), strict comparisons so ties keep the earlier
 *     value
 *   - finish (function.mojo:2516-2804 _Finalizer.scan, :2891-2917
 *     _sum_typed; reducer.mojo:3391-3599): count/len Int64; Int64 sum
 *     raises "Int64 expression sum overflow" only on an out-of-range final
 *     total of a valid group; Int32 sums stay Int32 and raise
 *     "int32 expression sum overflow"; sums are null below min_count; mean
 *     is total / count (null when count is 0); var is m2 / (n - ddof), null
 *     when n - ddof <= 0; std its square root; min/max null when nothing
 *     was seen, NaN only when every valid value was NaN.
 * This is one worker's sequential order; the reference's multi-worker
 * merges (combine, function.mojo:1559-1779) change float rounding only.
 *
 * Also: double-double per-group sums and two-pass variances, the
 * compensated reference against which both the GPU and this oracle report
 * per-group float error.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define DFO_INT32 1
#define DFO_INT64 2
#define DFO_FLOAT64 3
#define DFO_DICT 4

#define AGG_LEN 0
#define AGG_COUNT 1
#define AGG_SUM 2
#define AGG_MIN 3
#define AGG_MAX 4
#define AGG_MEAN 5
#define AGG_VAR 6
#define AGG_STD 7

typedef struct {
    int32_t dtype;
    int32_t pad;
    const void *values;
    const uint8_t *valid; /* one byte per row (0/1), or NULL = all valid */
} dfo_col;

typedef struct {
    int32_t op;
    int32_t col;
    int32_t ddof;
    int32_t pad;
    int64_t min_count;
} dfo_agg;

typedef __int128 i128;

typedef struct {
    i128 isum;
    double fsum;
    int64_t count;
    int64_t mcount;
    double mean, m2;
    int64_t ibest;
    double fbest;
    int flags; /* 1 seen, 2 NaN seen */
} state_t;

static int col_valid(const dfo_col *c, int64_t r) { return c->valid ? c->valid[r] != 0 : 1; }

static int64_t col_int(const dfo_col *c, int64_t r) {
    switch (c->dtype) {
        case DFO_INT32: return ((const int32_t *)c->values)[r];
        case DFO_DICT: return (int64_t)((const uint32_t *)c->values)[r];
        default: return ((const int64_t *)c->values)[r];
    }
}

/* A Float64 key's canonical word (group_by.mojo:319-329, nested_keys.mojo
 * _canonical): -0.0 and 0.0 are one key, every NaN is one key. */
static int64_t key_word(const dfo_col *c, int64_t r) {
    if (c->dtype != DFO_FLOAT64) return col_int(c, r);
    double x = ((const double *)c->values)[r];
    if (x != x) return (int64_t)0x7FF8000000000000ull;
    if (x == 0.0) return 0;
    int64_t w;
    memcpy(&w, &x, 8);
    return w;
}

static double col_f64(const dfo_col *c, int64_t r) {
    if (c->dtype == DFO_FLOAT64) return ((const double *)c->values)[r];
    return (double)col_int(c, r);
}

/* --- key hash table: key tuple (values + null bits) -> group index ----- */
typedef struct {
    int n_keys;
    int64_t cap;
    int64_t *slot_group; /* -1 empty */
    int64_t *keyvals;    /* per group: n_keys values */
    uint8_t *keynull;    /* per group: n_keys flags */
    int64_t groups, gcap;
} table_t;

static uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint64_t key_hash(int n, const int64_t *v, const uint8_t *nul) {
    uint64_t h = 0x12345678;
    for (int k = 0; k < n; ++k) h = mix(h ^ (nul[k] ? 0xABCDEFull : (uint64_t)v[k] * 31 + (uint64_t)k));
    return h;
}

static int key_eq(int n, const int64_t *a, const uint8_t *an, const int64_t *b, const uint8_t *bn) {
    for (int k = 0; k < n; ++k) {
        if (an[k] != bn[k]) return 0;
        if (!an[k] && a[k] != b[k]) return 0;
    }
    return 1;
}

static void table_grow(table_t *t);

static int64_t table_find(table_t *t, const int64_t *v, const uint8_t *nul, int *created) {
    if ((t->groups + 1) * 2 > t->cap) table_grow(t);
    uint64_t mask = (uint64_t)t->cap - 1;
    uint64_t i = key_hash(t->n_keys, v, nul) & mask;
    for (;;) {
        int64_t g = t->slot_group[i];
        if (g < 0) break;
        if (key_eq(t->n_keys, t->keyvals + g * t->n_keys, t->keynull + g * t->n_keys, v, nul)) {
            *created = 0;
            return g;
        }
        i = (i + 1) & mask;
    }
    if (t->groups == t->gcap) {
        t->gcap = t->gcap ? t->gcap * 2 : 1024;
        t->keyvals = realloc(t->keyvals, (size_t)t->gcap * t->n_keys * sizeof(int64_t));
        t->keynull = realloc(t->keynull, (size_t)t->gcap * t->n_keys);
    }
    int64_t g = t->groups++;
    memcpy(t->keyvals + g * t->n_keys, v, t->n_keys * sizeof(int64_t));
    memcpy(t->keynull + g * t->n_keys, nul, t->n_keys);
    t->slot_group[i] = g;
    *created = 1;
    return g;
}

static void table_grow(table_t *t) {
    int64_t ncap = t->cap ? t->cap * 2 : 1024;
    int64_t *ns = malloc((size_t)ncap * sizeof(int64_t));
    for (int64_t i = 0; i < ncap; ++i) ns[i] = -1;
    uint64_t mask = (uint64_t)ncap - 1;
    for (int64_t g = 0; g < t->groups; ++g) {
        uint64_t i = key_hash(t->n_keys, t->keyvals + g * t->n_keys, t->keynull + g * t->n_keys) & mask;
        while (ns[i] >= 0) i = (i + 1) & mask;
        ns[i] = g;
    }
    free(t->slot_group);
    t->slot_group = ns;
    t->cap = ncap;
}

static void set_err(char *err, int errlen, const char *msg) {
    if (err && errlen > 0) {
        strncpy(err, msg, (size_t)errlen - 1);
        err[errlen - 1] = 0;
    }
}

/* Folds row r into group state s[0..n_aggs) (LEN is counted by the caller). */
static void fold_row(state_t *s0, int n_aggs, const dfo_agg *aggs, const dfo_col *cols, int64_t r) {
    for (int a = 0; a < n_aggs; ++a) {
        const dfo_agg *A = &aggs[a];
        state_t *s = &s0[a];
        if (A->op == AGG_LEN) continue;
        const dfo_col *c = &cols[A->col];
        if (!col_valid(c, r)) continue;
        s->count++;
        const int fl = c->dtype == DFO_FLOAT64;
        switch (A->op) {
            case AGG_COUNT: break;
            case AGG_SUM:
            case AGG_MEAN:
                if (fl)
                    s->fsum += col_f64(c, r);
                else
                    s->isum += (i128)col_int(c, r);
                break;
            case AGG_MIN:
            case AGG_MAX: {
                const int is_max = A->op == AGG_MAX;
                if (fl) {
                    double x = col_f64(c, r);
                    if (isnan(x)) {
                        s->flags |= 2;
                        break;
                    }
                    int better = is_max ? (x > s->fbest) : (x < s->fbest);
                    if (!(s->flags & 1) || better) {
                        s->fbest = x;
                        s->flags |= 1;
                    }
                } else {
                    int64_t x = col_int(c, r);
                    int better = is_max ? (x > s->ibest) : (x < s->ibest);
                    if (!(s->flags & 1) || better) {
                        s->ibest = x;
                        s->flags |= 1;
                    }
                }
                break;
            }
            default: { /* VAR / STD: Welford, as VarState.add */
                double x = col_f64(c, r);
                s->mcount++;
                double delta = x - s->mean;
                s->mean += delta / (double)s->mcount;
                s->m2 += delta * (x - s->mean);
                break;
            }
        }
    }
}

/* Finishes aggregate a of one group into agg_bits/agg_valid[a * n + g]. */
static int finish_agg(const state_t *s, int64_t rows, const dfo_agg *A, const dfo_col *cols, int a, int64_t n,
                      int64_t g, uint64_t *agg_bits, uint8_t *agg_valid, const char **errmsg) {
    const dfo_col *c = A->op == AGG_LEN ? NULL : &cols[A->col];
    uint64_t bits = 0;
    int valid = 1, status = 0;
    switch (A->op) {
        case AGG_LEN: bits = (uint64_t)rows; break;
        case AGG_COUNT: bits = (uint64_t)s->count; break;
        case AGG_SUM:
            valid = s->count >= A->min_count;
            if (c->dtype == DFO_FLOAT64) {
                double v = valid ? s->fsum : 0.0;
                memcpy(&bits, &v, 8);
            } else {
                i128 tot = valid ? s->isum : 0;
                if (c->dtype == DFO_INT32) {
                    if (tot > 2147483647 || tot < -2147483647 - 1) {
                        *errmsg = "int32 expression sum overflow";
                        status = 2;
                    }
                } else if (tot > (i128)INT64_MAX || tot < (i128)INT64_MIN) {
                    *errmsg = "Int64 expression sum overflow";
                    status = 2;
                }
                bits = (uint64_t)(int64_t)tot;
            }
            break;
        case AGG_MIN:
        case AGG_MAX:
            if (c->dtype == DFO_FLOAT64) {
                valid = (s->flags & 3) != 0;
                double v = (s->flags & 1) ? s->fbest : ((s->flags & 2) ? NAN : 0.0);
                memcpy(&bits, &v, 8);
            } else {
                valid = (s->flags & 1) != 0;
                bits = valid ? (uint64_t)s->ibest : 0;
            }
            break;
        case AGG_MEAN: {
            valid = s->count > 0;
            double tot = c->dtype == DFO_FLOAT64 ? s->fsum : (double)s->isum;
            double v = valid ? tot / (double)s->count : 0.0;
            memcpy(&bits, &v, 8);
            break;
        }
        default: {
            valid = s->mcount - (int64_t)A->ddof > 0;
            double v = valid ? s->m2 / (double)(s->mcount - A->ddof) : 0.0;
            if (A->op == AGG_STD && valid) v = sqrt(v);
            memcpy(&bits, &v, 8);
            break;
        }
    }
    agg_bits[(int64_t)a * n + g] = bits;
    agg_valid[(int64_t)a * n + g] = (uint8_t)valid;
    return status;
}

/* Writes key k of group g (table slot tg) to key_vals/key_valid[k * n + g]. */
static void emit_key(const table_t *t, int64_t tg, const dfo_col *keys, int k, int64_t first_row, int64_t n, int64_t g,
                     int64_t *key_vals, uint8_t *key_valid) {
    const int nk = t->n_keys;
    key_valid[k * n + g] = !t->keynull[tg * nk + k];
    key_vals[k * n + g] = t->keyvals[tg * nk + k];
    /* a Float64 key reports the group's first row's own value, as the
     * reference's first() of the key (group_by.mojo:319-329) */
    if (keys[k].dtype == DFO_FLOAT64 && !t->keynull[tg * nk + k])
        memcpy(&key_vals[k * n + g], &((const double *)keys[k].values)[first_row], 8);
}

/* The key tuple of row r. */
static void row_key(int n_keys, const dfo_col *keys, int64_t r, int64_t *kv, uint8_t *kn) {
    for (int k = 0; k < n_keys; ++k) {
        kn[k] = !col_valid(&keys[k], r);
        kv[k] = kn[k] ? 0 : key_word(&keys[k], r);
    }
}

/* The groups of one run of rows, folded in the given (ascending) order. */
typedef struct {
    table_t t;
    state_t *st;
    int64_t *rows_of, *first_rows;
    int64_t scap;
} groups_t;

static void groups_init(groups_t *G, int n_keys, int n_aggs) {
    memset(G, 0, sizeof *G);
    G->t.n_keys = n_keys;
    table_grow(&G->t);
    G->scap = 1024;
    G->st = calloc((size_t)G->scap * (n_aggs ? n_aggs : 1), sizeof(state_t));
    G->rows_of = malloc((size_t)G->scap * sizeof(int64_t));
    G->first_rows = malloc((size_t)G->scap * sizeof(int64_t));
}

static void groups_add(groups_t *G, int n_keys, const dfo_col *keys, int n_aggs, const dfo_agg *aggs,
                       const dfo_col *cols, int64_t r) {
    int64_t kv[8];
    uint8_t kn[8];
    row_key(n_keys, keys, r, kv, kn);
    int created;
    const int na = n_aggs ? n_aggs : 1;
    int64_t g = table_find(&G->t, kv, kn, &created);
    if (created) {
        if (g >= G->scap) {
            int64_t ncap = G->scap * 2;
            G->st = realloc(G->st, (size_t)ncap * na * sizeof(state_t));
            memset(G->st + G->scap * na, 0, (size_t)(ncap - G->scap) * na * sizeof(state_t));
            G->rows_of = realloc(G->rows_of, (size_t)ncap * sizeof(int64_t));
            G->first_rows = realloc(G->first_rows, (size_t)ncap * sizeof(int64_t));
            G->scap = ncap;
        }
        G->first_rows[g] = r;
        G->rows_of[g] = 0;
    }
    G->rows_of[g]++;
    fold_row(&G->st[g * n_aggs], n_aggs, aggs, cols, r);
}

static void groups_free(groups_t *G) {
    free(G->st);
    free(G->rows_of);
    free(G->first_rows);
    free(G->t.slot_group);
    free(G->t.keyvals);
    free(G->t.keynull);
}

/* Writes the groups of G at output rows off.. (keys, first rows, aggregates). */
static int groups_emit(const groups_t *G, int n_keys, const dfo_col *keys, int n_aggs, const dfo_agg *aggs,
                       const dfo_col *cols, int64_t n, int64_t off, int64_t *first_rows, int64_t *key_vals,
                       uint8_t *key_valid, uint64_t *agg_bits, uint8_t *agg_valid, const char **errmsg) {
    int status = 0;
    for (int64_t g = 0; g < G->t.groups; ++g) {
        first_rows[off + g] = G->first_rows[g];
        for (int k = 0; k < n_keys; ++k)
            emit_key(&G->t, g, keys, k, G->first_rows[g], n, off + g, key_vals, key_valid);
        for (int a = 0; a < n_aggs; ++a) {
            int st = finish_agg(&G->st[g * n_aggs + a], G->rows_of[g], &aggs[a], cols, a, n, off + g, agg_bits,
                                agg_valid, errmsg);
            if (st) status = st;
        }
    }
    return status;
}

/*
 * Groups the kept rows (kept[r] != 0) by the key columns and finishes each
 * aggregate. Outputs are sized for `n` groups by the caller:
 *   first_rows[g], key_vals[k * n + g], key_valid[k * n + g],
 *   agg_bits[a * n + g] (int64 value or double bits), agg_valid[a * n + g].
 * Groups come in order of their first kept row.
 * Returns 0, or 2 with `err` set to the reference's error text.
 */
int dfo_groupby(int64_t n, int n_keys, const dfo_col *keys, const uint8_t *kept, int n_cols,
                const dfo_col *cols, int n_aggs, const dfo_agg *aggs, int64_t *out_groups,
                int64_t *first_rows, int64_t *key_vals, uint8_t *key_valid, uint64_t *agg_bits,
                uint8_t *agg_valid, char *err, int errlen) {
    (void)n_cols;
    groups_t G;
    groups_init(&G, n_keys, n_aggs);
    for (int64_t r = 0; r < n; ++r)
        if (kept[r]) groups_add(&G, n_keys, keys, n_aggs, aggs, cols, r);
    *out_groups = G.t.groups;
    const char *msg = NULL;
    int status = groups_emit(&G, n_keys, keys, n_aggs, aggs, cols, n, 0, first_rows, key_vals, key_valid, agg_bits,
                             agg_valid, &msg);
    if (status) set_err(err, errlen, msg);
    groups_free(&G);
    return status;
}

/*
 * The same grouping by `threads` workers, as the reference CPU engine
 * organises it (morsel-parallel workers, per-worker radix partitions by key
 * hash, then partition-parallel aggregation and finish; hash_table.mojo:1-49,
 * 2850 _FinishJob). Worker t takes the contiguous rows
 * [n * t / threads, n * (t + 1) / threads) and appends each kept row to its
 * list for the row's key partition; each partition is then folded by one
 * worker over the lists in worker order, i.e. in ascending row order, so
 * every group sees its rows in the sequential order and every state (float
 * sums and moments included) is bit-identical to dfo_groupby's. Groups come
 * partition by partition (not in first-row order); first_rows tells the
 * caller the order dfo_groupby gives.
 */
typedef struct {
    int64_t *v;
    int64_t n, cap;
} rowvec_t;

int dfo_groupby_parallel(int threads, int64_t n, int n_keys, const dfo_col *keys, const uint8_t *kept, int n_cols,
                         const dfo_col *cols, int n_aggs, const dfo_agg *aggs, int64_t *out_groups,
                         int64_t *first_rows, int64_t *key_vals, uint8_t *key_valid, uint64_t *agg_bits,
                         uint8_t *agg_valid, char *err, int errlen) {
    (void)n_cols;
    enum { LOG_P = 8, NP = 1 << LOG_P };
    if (threads < 1) threads = 1;
    const int T = threads;
    rowvec_t *lists = calloc((size_t)T * NP, sizeof(rowvec_t));
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const int64_t r0 = n * t / T, r1 = n * (t + 1) / T;
        int64_t kv[8];
        uint8_t kn[8];
        for (int64_t r = r0; r < r1; ++r) {
            if (!kept[r]) continue;
            row_key(n_keys, keys, r, kv, kn);
            const int p = (int)(key_hash(n_keys, kv, kn) >> (64 - LOG_P));
            rowvec_t *L = &lists[(size_t)t * NP + p];
            if (L->n == L->cap) {
                L->cap = L->cap ? L->cap * 2 : 256;
                L->v = realloc(L->v, (size_t)L->cap * sizeof(int64_t));
            }
            L->v[L->n++] = r;
        }
    }
    groups_t *parts = calloc(NP, sizeof(groups_t));
#pragma omp parallel for schedule(dynamic) num_threads(T)
    for (int p = 0; p < NP; ++p) {
        groups_init(&parts[p], n_keys, n_aggs);
        for (int t = 0; t < T; ++t) {
            const rowvec_t *L = &lists[(size_t)t * NP + p];
            for (int64_t i = 0; i < L->n; ++i) groups_add(&parts[p], n_keys, keys, n_aggs, aggs, cols, L->v[i]);
        }
    }
    int64_t off[NP + 1];
    off[0] = 0;
    for (int p = 0; p < NP; ++p) off[p + 1] = off[p] + parts[p].t.groups;
    *out_groups = off[NP];
    int status = 0;
    const char *msg = NULL;
#pragma omp parallel for schedule(dynamic) num_threads(T)
    for (int p = 0; p < NP; ++p) {
        const char *m = NULL;
        int st = groups_emit(&parts[p], n_keys, keys, n_aggs, aggs, cols, n, off[p], first_rows, key_vals, key_valid,
                             agg_bits, agg_valid, &m);
        if (st) {
#pragma omp critical
            {
                status = st;
                msg = m;
            }
        }
    }
    if (status) set_err(err, errlen, msg);
    for (int p = 0; p < NP; ++p) groups_free(&parts[p]);
    free(parts);
    for (int i = 0; i < T * NP; ++i) free(lists[i].v);
    free(lists);
    return status;
}

/* --- double-double per-group references ------------------------------ */

static void two_sum(double a, double b, double *s, double *e) {
    *s = a + b;
    double bb = *s - a;
    *e = (a - (*s - bb)) + (b - bb);
}

static void dd_add(double *hi, double *lo, double x) {
    double s, e;
    two_sum(*hi, x, &s, &e);
    e += *lo;
    two_sum(s, e, hi, lo);
}

static void two_prod(double a, double b, double *p, double *e) {
    *p = a * b;
    *e = fma(a, b, -*p);
}

/*
 * For each group g (0..G-1) over rows with gid[r] >= 0 and valid[r]:
 * sum_hi/lo = double-double sum, var = two-pass variance with ddof
 * computed in double-double (mean in double-double, squared deviations
 * accumulated in double-double).
 */
void dfo_dd_moments(int64_t n, const int64_t *gid, const double *x, const uint8_t *valid, int64_t G,
                    int ddof, double *sum_hi, double *sum_lo, int64_t *count, double *var) {
    for (int64_t g = 0; g < G; ++g) {
        sum_hi[g] = sum_lo[g] = 0.0;
        count[g] = 0;
        var[g] = NAN;
    }
    for (int64_t r = 0; r < n; ++r) {
        int64_t g = gid[r];
        if (g < 0 || (valid && !valid[r])) continue;
        dd_add(&sum_hi[g], &sum_lo[g], x[r]);
        count[g]++;
    }
    double *mh = malloc((size_t)(G ? G : 1) * sizeof(double));
    double *ml = malloc((size_t)(G ? G : 1) * sizeof(double));
    double *qh = calloc((size_t)(G ? G : 1), sizeof(double));
    double *ql = calloc((size_t)(G ? G : 1), sizeof(double));
    for (int64_t g = 0; g < G; ++g) {
        /* mean = sum / n in double-double: one Newton correction */
        double n_ = (double)count[g];
        double m = count[g] ? sum_hi[g] / n_ : 0.0;
        double p, e;
        two_prod(m, n_, &p, &e);
        double rem = ((sum_hi[g] - p) - e + sum_lo[g]);
        mh[g] = m;
        ml[g] = count[g] ? rem / n_ : 0.0;
    }
    for (int64_t r = 0; r < n; ++r) {
        int64_t g = gid[r];
        if (g < 0 || (valid && !valid[r])) continue;
        /* d = x - mean in double-double, then d*d accumulated */
        double s, e;
        two_sum(x[r], -mh[g], &s, &e);
        e -= ml[g];
        double dh, dl;
        two_sum(s, e, &dh, &dl);
        double p, pe;
        two_prod(dh, dh, &p, &pe);
        pe += 2.0 * dh * dl;
        dd_add(&qh[g], &ql[g], p);
        dd_add(&qh[g], &ql[g], pe);
    }
    for (int64_t g = 0; g < G; ++g) {
        if (count[g] - ddof > 0) var[g] = (qh[g] + ql[g]) / (double)(count[g] - ddof);
    }
    free(mh);
    free(ml);
    free(qh);
    free(ql);
}
