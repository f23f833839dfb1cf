/*
 * tdd.h -- the device layer: a thin C ABI between the C17 host engine and the
 * hand-written HIP kernels for gfx950 (MI355X), implemented in the .hip
 * files of teide_amd/csrc/dev.
 *
 * Plain pointers and sizes only.  All "device" pointers are HBM addresses
 * (normally blocks of the HBM arena, see td_dev.h); every call runs on the
 * library's single HIP stream and returns a td_err_t code (TD_OK on
 * success).  A call that produces host-visible scalars synchronises the
 * stream before returning.
 *
 * Each kernel entry names the reference CPU routine whose semantics it
 * implements (file:line in the reference tree).
 */
#ifndef TDD_H
#define TDD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------- runtime -- */

typedef struct {
    char     name[64];
    char     arch[32];          /* e.g. "gfx950" */
    int32_t  n_cu;
    int32_t  n_xcd_hint;        /* 8 on MI355X */
    uint64_t hbm_bytes;
    uint64_t lds_bytes_per_cu;
    int32_t  clock_mhz;
} tdd_info_t;

int  tdd_init(void);                         /* device 0, one stream; idempotent */
int  tdd_available(void);                    /* 1 when a GPU is usable */
int  tdd_info(tdd_info_t* out);
int  tdd_malloc(void** dptr, size_t bytes);  /* raw hipMalloc (arena backing) */
int  tdd_free(void* dptr);
int  tdd_h2d(void* dst, const void* src, size_t bytes);
int  tdd_d2h(void* dst, const void* src, size_t bytes);
int  tdd_d2d(void* dst, const void* src, size_t bytes);
int  tdd_fill(void* dst, uint64_t pattern, size_t bytes);  /* kernel fill, 8-byte pattern */
/* bytes of open file fd from offset off into device memory dst: chunks read
 * into two pinned staging buffers in turn (several reader threads per
 * chunk), each copied asynchronously while the next is read.  The
 * reference reads column files into host memory (store/col.c:449-560). */
int  tdd_file_h2d(int fd, int64_t off, void* dst, size_t bytes);
int  tdd_sync(void);
int  tdd_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* Kernel timing with HIP events on the library stream.  While enabled,
 * every launch is bracketed by two events; tdd_prof_read() synchronises and
 * returns per-kernel totals (name, launches, total ms). */
typedef struct { char name[48]; int64_t launches; double total_ms; } tdd_prof_row_t;
int  tdd_prof_enable(int on);
int  tdd_prof_reset(void);
int  tdd_prof_read(tdd_prof_row_t* rows, int cap, int* n_rows);

/* ------------------------------------------------------------- columns -- */

typedef struct {
    const void* data;   /* device address of element 0 (NULL = no column) */
    int8_t   type;      /* td.h type tag */
    uint8_t  attrs;     /* TD_SYM index width in bits 0-1 */
    uint8_t  esz;       /* element bytes */
    uint8_t  ranged;    /* 1: every element (widened to int64) lies in [lo, hi] */
    uint8_t  pad[4];
    int64_t  lo, hi;    /* known bounds of an integer column (column statistics) */
} tdd_col_t;

/* A device TD_SEL: pointers into one device copy of the td.h TD_SEL payload
 * (meta | seg_flags | seg_popcnt | bits, td.h:546-589). */
typedef struct {
    void*     base;
    uint8_t*  flags;
    uint16_t* popcnt;
    uint64_t* bits;
    int64_t   nrows;
    uint32_t  n_segs;
    uint32_t  pad;
} tdd_sel_t;

/* ----------------------------------------------------------- predicates -- */
/* A filter predicate compiled to postfix form over up to 16 columns:
 * comparisons of a column with a constant or another column, combined by
 * AND / OR / NOT.  Comparisons run in int64 when both sides are integer
 * typed and in f64 when either is F64, the reference's typing rule
 * (exec.c:760-918, 982-1051). */
#define TDD_PRED_MAX_INS  32
#define TDD_PRED_MAX_COLS 16
enum {
    TDD_P_CMP_CONST = 1,   /* push col[a] CMP constant */
    TDD_P_CMP_COL   = 2,   /* push col[a] CMP col[b] */
    TDD_P_AND       = 3,
    TDD_P_OR        = 4,
    TDD_P_NOT       = 5,
    TDD_P_CONST     = 6,   /* push ci != 0 */
    TDD_P_BOOLCOL   = 7    /* push col[a] != 0 (a BOOL / U8 column) */
};
typedef struct {
    uint8_t kind, cmp, a, b;      /* cmp: OP_EQ..OP_GE */
    uint8_t as_f64, pad[3];
    int64_t ci;
    double  cf;
} tdd_pred_ins_t;
typedef struct {
    int32_t        n_ins, n_cols;
    tdd_col_t      cols[TDD_PRED_MAX_COLS];
    tdd_pred_ins_t ins[TDD_PRED_MAX_INS];
} tdd_pred_t;

/* Evaluate `pred` over rows [0, out->nrows) into a device TD_SEL, ANDed with
 * `and_with` when given: wave64 ballots form the bit words directly, popcounts
 * give segment flags and counts.  Replaces expr_eval_full + td_sel_from_pred
 * (+ td_sel_and for chained filters): exec.c:1254-1273, sel.c:128-161,
 * sel.c:167-189, exec.c:11706-11731. */
int tdd_sel_eval(const tdd_pred_t* pred, const tdd_sel_t* and_with, tdd_sel_t* out);
/* BOOL bytes -> TD_SEL (td_sel_from_pred, sel.c:128-161). */
int tdd_sel_from_bool(const uint8_t* bools, tdd_sel_t* out);
/* a AND b (td_sel_and, sel.c:167-189). */
int tdd_sel_and(const tdd_sel_t* a, const tdd_sel_t* b, tdd_sel_t* out);
/* Rebuild flags / popcounts / total from the bits (td_sel_recompute, sel.c:77-122). */
int tdd_sel_recompute(tdd_sel_t* s);
/* total_pass read back to the host. */
int tdd_sel_total(const tdd_sel_t* s, int64_t* total);

/* ------------------------------------------------- compaction / gather -- */

/* Row ids of the selected rows, ascending, into idx (total_pass entries):
 * the match_idx build of sel_compact (exec.c:2362-2388). */
int tdd_sel_index(const tdd_sel_t* s, int64_t* idx);
/* Stream compaction of up to 16 columns through a selection, fused with the
 * index build: dst[c][j] = src[c][row of j-th selected bit].  Replaces
 * sel_compact + multi_gather_fn (exec.c:2328-2480, 1852-1899). */
int tdd_sel_compact(const tdd_sel_t* s, int ncols, const tdd_col_t* src, void* const* dst);
/* dst[i] = src[idx[i]]; idx < 0 writes zeros when nullable (gather_fn,
 * exec.c:1910-1939). */
int tdd_gather(const tdd_col_t* src, const int64_t* idx, int64_t n, int nullable, void* dst);

/* -------------------------------------------------------------- group-by -- */

#define TDD_MAX_KEYS 8
#define TDD_MAX_AGGS 8

typedef struct {
    int64_t           nrows;
    const tdd_sel_t*  mask;               /* NULL: every row */
    int32_t           n_keys, n_aggs;
    tdd_col_t         keys[TDD_MAX_KEYS];
    uint16_t          ops[TDD_MAX_AGGS];  /* OP_SUM ... OP_VAR_POP */
    tdd_col_t         aggs[TDD_MAX_AGGS]; /* data may be NULL for OP_COUNT */
    int32_t           force_path;         /* 0 auto, 1 direct array, 2 hash, 3 radix-partitioned direct array, 4 hash-partitioned hash, 5 hash-partitioned hash after hot-key pre-aggregation at any size, 6 as 4 with a first try of 1,024-entry tables (the regrow path) */
    int32_t           pad;
} tdd_group_spec_t;

typedef struct {
    int64_t  n_groups;
    void*    keys[TDD_MAX_KEYS];   /* arena blocks; element type = key column type */
    void*    vals[TDD_MAX_AGGS];   /* arena blocks; element type = out_types[a] */
    int8_t   out_types[TDD_MAX_AGGS];
    int32_t  path;                 /* 1 LDS direct array, 2 global direct array, 3 hash, 4 partitioned direct array, 5 partitioned hash, 6 partitioned hash with hot keys folded first; 20 + p: key tuples wider than 64 bits (an F64 among several keys, or integer ranges multiplying past 2^63) numbered by a hash of their raw key words, then grouped by that number on path p */
    int32_t  tries;                /* hash paths: table sizes tried (1 when the first fitted) */
    int64_t  part_rows;            /* partitioned hash: rows the partition passes carried (the
                                      rows of hot keys are folded before them) */
} tdd_group_result_t;

/* Masked min/max of an integer key column (minmax_scan_fn, exec.c:4544-4604);
 * *mn > *mx when no row is selected. */
int tdd_minmax(const tdd_col_t* col, int64_t nrows, const tdd_sel_t* mask,
               int64_t* mn, int64_t* mx);
/* tdd_minmax with no device allocation at all (no hipMalloc / hipFree, no
 * arena block): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation.  No
 * counterpart in the reference.  Its partial results live in a static
 * device array of 2 + 2 * 2048 int64 words per GPU; TD_ERR_RANGE when a
 * call would need more than 2048 partial pairs (more than 512 CUs).  Calls
 * that overlap on one GPU race on that array. */
int tdd_minmax_noalloc(const tdd_col_t* col, int64_t nrows, const tdd_sel_t* mask,
                       int64_t* mn, int64_t* mx);
/* Group-by with COUNT / SUM / AVG / MIN / MAX / FIRST / LAST / VAR family.
 * Integer keys whose masked ranges multiply to a modest slot count use a
 * direct array (LDS-private per workgroup when it fits, else global atomics),
 * the device form of the reference's DA path (exec.c:5948-6404); other keys
 * use a global open-addressing table with LDS pre-aggregation, the device
 * form of its hash paths (exec.c:3868-4526, 6406-6856).  Output types and
 * AVG / VAR arithmetic follow emit_agg_columns (exec.c:4632-4773).  Groups
 * come out in ascending key order on the direct-array paths.  n_keys == 0:
 * one group of every selected row, none when no row is selected (the
 * scalar aggregate, exec.c:5766-5940). */
int tdd_group(const tdd_group_spec_t* spec, tdd_group_result_t* out);

/* Product of rows [0, n) of col into out8 (8 bytes): a double for an F64
 * column, else the wrapping int64 product of the values widened to int64;
 * 1 for n == 0.  A fixed-shape tree, so the f64 bits depend only on the
 * input.  Replaces exec_reduction's OP_PROD (exec.c:1589-1617, 1779, 1812). */
int tdd_reduce_prod(const tdd_col_t* col, int64_t n, void* out8);
/* out[i] = in[i]'s bits with NaNs made one canonical NaN and -0.0 made
 * +0.0: the value normalisation of exec_count_distinct (exec.c:1697-1701). */
int tdd_f64_canon(const double* in, int64_t n, int64_t* out);

/* ------------------------------------------------------------------ join -- */

typedef struct {
    int64_t           nl, nr;
    int32_t           n_keys, join_type;  /* 0 inner, 1 left, 2 full */
    tdd_col_t         lkeys[TDD_MAX_KEYS];
    tdd_col_t         rkeys[TDD_MAX_KEYS];
    const tdd_sel_t*  lsel;               /* left rows that take part; NULL: all */
    int32_t           force_path;         /* 0 auto, 1 direct map, 2 radix, 3 partitioned fused join-group, 4 XCD-partitioned fused join-group */
    int32_t           radix_bits;         /* 0: chosen from the build size */
} tdd_join_spec_t;

typedef struct {
    int64_t  n_pairs;
    int64_t* l_idx;       /* arena blocks, n_pairs each; -1 = no partner */
    int64_t* r_idx;
    int32_t  path;        /* 1 direct map, 2 radix partitioned */
    int32_t  radix_bits;
} tdd_join_result_t;

/* Equi-join producing row-index pairs.  Dense unique integer build keys use
 * a direct-indexed chained map; everything else a two-pass radix partition
 * (LDS histograms, LDS-staged coalesced scatter) followed by per-partition
 * build + probe with the hash table in LDS.  Replaces exec_join's radix and
 * chained paths (exec.c:7431-8392); pairs of a left row appear in one
 * contiguous run, unmatched left rows (left / full) get r = -1 and unmatched
 * right rows (full) come last with l = -1. */
int tdd_join(const tdd_join_spec_t* spec, tdd_join_result_t* out);

/* The materialising radix join: INNER or LEFT equi-join that writes the
 * output columns directly -- lcols[0..n_lcols) of every output row's left
 * row, then rcols[0..n_rcols) of its right row (zeros for a LEFT row with no
 * partner) -- with no pair arrays and no random gathers.  Both sides are
 * radix partitioned once by key (the left side carrying its output columns,
 * so its partitioned buffers are the output); each partition's build keys
 * go into an LDS hash table and its probe rows stream through it.  Rows come
 * out in partition order, as the reference's radix join emits them
 * (exec.c:7431-7950, output columns exec.c:8607-8728).  A single integer
 * build key dense in its range uses a map instead of the partitions: the
 * output rows are then the selected probe rows in row order, left columns
 * compacted, right columns read through the map (path 4).  Returns
 * TD_ERR_NYI, with nothing allocated, when the build side repeats a key, a
 * build partition outgrows an LDS table or the join is FULL: run tdd_join
 * then. */
typedef struct {
    tdd_join_spec_t   join;
    int32_t           n_lcols, n_rcols;
    tdd_col_t         lcols[16];
    tdd_col_t         rcols[16];
} tdd_join_mat_spec_t;

typedef struct {
    int64_t  n_rows;
    void*    lcols[16];           /* arena blocks, element bytes of lcols[c] */
    void*    rcols[16];
    int32_t  path;                /* 3 materialising radix, 4 materialising map */
    int32_t  radix_bits;
} tdd_join_mat_result_t;

int tdd_join_mat(const tdd_join_mat_spec_t* spec, tdd_join_mat_result_t* out);

/* Fused filter -> join -> group-by: the probe side is filtered by `lpred`
 * (and/or `join.lsel`), probed against the build side and aggregated into
 * the group table without materialising the filtered table or the join
 * output.  Group keys and aggregate inputs name a side (0 left, 1 right) and
 * a column.  Has no single reference counterpart: it computes what the
 * reference computes with JOIN(FILTER(..)) followed by GROUP on the join
 * result in a second td_execute (exec.c:11840-11857 then 11759-11838). */
typedef struct {
    tdd_join_spec_t   join;               /* inner join */
    const tdd_pred_t* lpred;              /* NULL: no filter */
    int32_t           n_keys, n_aggs;
    tdd_col_t         keys[TDD_MAX_KEYS];
    uint8_t           key_side[TDD_MAX_KEYS];
    uint16_t          ops[TDD_MAX_AGGS];
    tdd_col_t         aggs[TDD_MAX_AGGS];
    uint8_t           agg_side[TDD_MAX_AGGS];
} tdd_join_group_spec_t;

/* out->path is the group-by's path, plus 10 when the join ran as the
 * materialising radix join of only the grouped columns before grouping. */
int tdd_join_group(const tdd_join_group_spec_t* spec, tdd_group_result_t* out);

/* Batched fused filter -> join -> group-by: many queries of one output
 * shape per launch sequence (see tdd_join_group for a single query).  No
 * reference counterpart: the reference runs each query's two td_execute
 * calls on its own.  prepare() plans the batch, checks the build keys of
 * every query for uniqueness and (tune != 0) times candidate probe launch
 * shapes once per batch-size class; run() executes it (use_graph: capture
 * the launch sequence as a hipGraph on the first run and replay it after),
 * writing every batched query's groups into one output block per column,
 * query q's rows at [offsets[q], offsets[q+1]).  status(q) != 0 marks a
 * query the batch could not take (run it with tdd_join_group).  Queries
 * whose build keys are not a dense unique range join through radix
 * partitions of both sides with LDS hash tables (the reference's radix
 * join, exec.c:7431-7950, for every such query of the batch at once);
 * radix() returns their number, the autotuned build rows per partition
 * and the largest radix bits of any of them. */
typedef struct tdd_jg_batch tdd_jg_batch_t;
int  tdd_jg_batch_prepare(const tdd_join_group_spec_t* const* specs, int n, int tune,
                          tdd_jg_batch_t** out);
int  tdd_jg_batch_status(const tdd_jg_batch_t* b, int q);
int  tdd_jg_batch_shape(const tdd_jg_batch_t* b, int64_t* tile_rows, int* block);
int  tdd_jg_batch_radix(const tdd_jg_batch_t* b, int64_t* part_rows, int* max_bits);
int  tdd_jg_batch_run(tdd_jg_batch_t* b, int use_graph, tdd_group_result_t* out,
                      int64_t* offsets);
void tdd_jg_batch_free(tdd_jg_batch_t* b);

/* ---------------------------------------------------------- expressions -- */
/* An element-wise expression compiled to registers and instructions, the
 * reference's expression compiler form (exec.c:696-918): registers are
 * columns, constants or temporaries, each computed as TD_I64, TD_F64 or
 * TD_BOOL; an instruction applies one opcode (OP_NEG .. OP_CAST, OP_ADD ..
 * OP_MAX2) to one or two registers; promotions are explicit OP_CASTs. */
#define TDD_EXPR_MAX_REGS 16
#define TDD_EXPR_MAX_INS  48
enum { TDD_R_COL = 0, TDD_R_CONST = 1, TDD_R_TMP = 2 };
typedef struct { uint8_t op, dst, src1, src2; } tdd_expr_ins_t;   /* src2 0xFF: unary */
typedef struct {
    uint8_t   kind;
    int8_t    type;
    uint8_t   pad[6];
    tdd_col_t col;
    double    cf;
    int64_t   ci;
} tdd_expr_reg_t;
typedef struct {
    int32_t        n_regs, n_ins, out_reg, pad;
    tdd_expr_reg_t regs[TDD_EXPR_MAX_REGS];
    tdd_expr_ins_t ins[TDD_EXPR_MAX_INS];
} tdd_expr_t;

/* out[0..n) = the expression over rows 0..n: 8-byte i64 / f64 or 1-byte
 * bool elements by the output register's type.  Arithmetic as the
 * reference evaluates it (exec.c:960-1100): i64 wrapping, DIV / MOD by zero
 * (and INT64_MIN / -1) give 0, f64 DIV / MOD by zero give 0, CAST f64 ->
 * i64 clamps. */
int tdd_expr_eval(const tdd_expr_t* e, int64_t n, void* out);

/* The bias of an affine aggregate input, SUM / AVG(col +- c), applied to
 * the group results of SUM / AVG(col) as the reference's emit does
 * (exec.c:4660-4705): f64 SUM += bias_f * count, i64 SUM += bias_i * count
 * (wrapping), AVG += bias_f.  counts: per-group row counts (SUM only). */
int tdd_affine_fix(void* vals, int8_t type, int avg, const int64_t* counts, int64_t n,
                   int64_t bias_i, double bias_f);

/* The per-row column of a MAPCOMMON (partition key) column: out[r] = kv[p]
 * for rows r in [offs[p], offs[p+1]) (device arrays; offs has n_parts + 1
 * entries, offs[0] = 0), elements of esz bytes (materialize_mapcommon,
 * exec.c:201-241). */
int tdd_mapcommon_fill(void* out, int esz, const void* kv, const int64_t* offs, int n_parts,
                       int64_t n);

/* ----------------------------------------------------------------- sort -- */
/* One sort key: a column, its direction and NaN placement, and for SYM
 * columns the rank of every symbol id by its string (device array). */
typedef struct {
    tdd_col_t       col;
    int32_t         desc, nulls_first;
    const uint32_t* rank;
} tdd_sort_key_t;

/* perm[0..n) = rows in sorted order: keys compared in order, each by the
 * order-preserving word of the reference's radix encoding
 * (radix_encode_fn, exec.c:2739-2870), ties kept in row order.  Replaces
 * exec_sort's radix / merge sorts (exec.c:3370-3862); top-N is a prefix of
 * this permutation (the reference's heap top-N, exec.c:3176-3340, leaves
 * the order of equal keys at the boundary unspecified). */
int tdd_sort_perm(const tdd_sort_key_t* keys, int nk, int64_t n, int64_t* perm);

/* --------------------------------------------------- shared-scan serving -- */
/* Requests of the family
 *   GROUP(dim.key; SUM / AVG / COUNT of fact columns)
 *     over JOIN(FILTER(fact, fact.pcol CMP t), fact.fk = dim.pk)
 * with a per-request constant t, comparison and group key share one scan of
 * the resident fact table.  The scan is cut into chunks; a step folds one
 * chunk into a cube per group key, cell (pcol value, group): row count and
 * one sum per value column.  Prefix / suffix sums over the pcol axis then
 * answer every active request's predicate for that chunk, and each request
 * adds them into its own direct-array group table.  A request that starts
 * at any step sees every chunk once after n_chunks steps.  COUNT and
 * integer sums are exact; f64 sums add the same values in another order.
 * A request with any other filter (req.pred: an f64 or wide-domain column,
 * several comparisons) shares the same chunks: each step evaluates its
 * predicate on the chunk's rows and folds the passing rows into its table
 * (through LDS copies of the tables of up to 16 requests at a time).
 * The dim join keys must be unique integers (a direct map per group key,
 * built at open and kept resident).  No reference counterpart: the
 * reference runs each request as its own two td_execute calls. */
#define TDD_SRV_MAX_KEYS 4
#define TDD_SRV_MAX_VALS 4

typedef struct {
    int64_t   nf, nd;                     /* fact rows, dim rows */
    tdd_col_t fk, pk;                     /* join key columns (integer) */
    tdd_col_t pcol;                       /* filtered fact column: integer, ranged */
    int32_t   n_vals, n_keys;
    tdd_col_t vals[TDD_SRV_MAX_VALS];     /* fact value columns (f64 or integer) */
    tdd_col_t keys[TDD_SRV_MAX_KEYS];     /* dim group-key columns (integer, ranged) */
    int64_t   chunk_rows;                 /* 0: an eighth of the fact table */
} tdd_srv_spec_t;

typedef struct {
    int32_t  key;                         /* index into spec.keys */
    int32_t  cmp;                         /* OP_EQ, OP_LT, OP_LE, OP_GT or OP_GE */
    int64_t  t;                           /* pcol CMP t */
    int32_t  n_aggs, pad;
    uint16_t ops[TDD_MAX_AGGS];           /* OP_SUM, OP_AVG or OP_COUNT */
    int8_t   src[TDD_MAX_AGGS];           /* value column index, -2: pcol, -1: none (COUNT) */
    /* non-NULL: the request's own filter over fact rows (columns of nf
     * rows), evaluated on every chunk of the shared scan instead of cmp / t;
     * copied at tdd_srv_begin */
    const tdd_pred_t* pred;
} tdd_srv_req_t;

typedef struct tdd_srv tdd_srv_t;
typedef struct tdd_srv_acc tdd_srv_acc_t;

/* Builds the per-key direct maps and the chunk cubes.  TD_ERR_NYI when the
 * build keys are not unique integers or pcol is not a small integer range. */
int     tdd_srv_open(const tdd_srv_spec_t* spec, tdd_srv_t** out);
int64_t tdd_srv_chunks(const tdd_srv_t* s);
int64_t tdd_srv_bytes(const tdd_srv_t* s);          /* device bytes held by the server */
/* Device bytes a request holds while active (accumulator + output bound);
 * TD_ERR_NYI when the server cannot answer it. */
int     tdd_srv_req_bytes(const tdd_srv_t* s, const tdd_srv_req_t* r, int64_t* bytes);
/* Allocates and clears a request's accumulator. */
int     tdd_srv_begin(tdd_srv_t* s, const tdd_srv_req_t* r, tdd_srv_acc_t** out);
/* One step: scan chunk (step number mod n_chunks) for the keys the active
 * requests use, then fold it into each of them. */
int     tdd_srv_step(tdd_srv_t* s, tdd_srv_acc_t* const* active, int n);
/* Groups of finished requests in ascending key order: out[i] for acc[i]. */
int     tdd_srv_finish(tdd_srv_t* s, tdd_srv_acc_t* const* acc, int n, tdd_group_result_t* out);
void    tdd_srv_acc_free(tdd_srv_acc_t* a);
void    tdd_srv_close(tdd_srv_t* s);

/* ------------------------------------------------- host services (C side) -- */
/* Implemented by the host engine (arena.c); the device layer allocates its
 * scratch and result blocks from the HBM arena through these. */
void* td__arena_alloc(size_t bytes);
void  td__arena_free(void* dptr);

#ifdef __cplusplus
}
#endif
#endif
