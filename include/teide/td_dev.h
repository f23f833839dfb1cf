/*
 * td_dev.h -- MI355X extensions of the td.h API.  None of these has a
 * counterpart in the reference engine; td.h programs do not need them.
 *
 * Residency.  A vector is either host-resident (the reference's layout,
 * mmod 0) or device-resident: a host td_t proxy with mmod 3 whose elements
 * live in the HBM arena, a buddy allocator over device memory.  td_execute
 * accepts both.  When every column a query touches is device-resident its
 * results stay device-resident (proxies); otherwise host columns are copied
 * in for the call and results come back as ordinary host vectors, so an
 * unchanged td.h program behaves exactly as with the reference.
 *
 * The elements of a device-resident vector are not host-addressable:
 * td_data() of a proxy holds the device address, not the values.
 */
#ifndef TD_DEV_H
#define TD_DEV_H

#include <teide/td.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when the GPU and the library's kernels are usable (initialises them). */
int      td_dev_available(void);

/* Device-resident copy of a vector, TD_SEL or table (tables: a new table of
 * device-resident copies of every column).  Resident input: a device-side
 * copy.  Returns TD_ERR_PTR(TD_ERR_OOM) when the arena budget is exceeded. */
td_t*    td_dev_put(td_t* v);
/* Host copy of a vector, TD_SEL or table (host input: retained as is). */
td_t*    td_dev_get(td_t* v);
/* 1 when v (or every column of table v) is device-resident; a parted
 * column is when all its segments are, a MAPCOMMON column always is. */
int      td_dev_is_resident(td_t* v);
/* Column files and splayed tables (the reference's on-disk format,
 * store/col.c, store/splay.c) loaded straight into the HBM arena, through
 * pinned staging buffers, with the statistics td_dev_put computes.  The
 * splayed form loads the symbol table first when sym_path is given. */
td_t*    td_dev_col_load(const char* path);
td_t*    td_dev_splay_load(const char* dir, const char* sym_path);
/* A partitioned database's table (store/part.c: td_read_parted's layout,
 * a MAPCOMMON partition-key column then parted columns) with every parted
 * column's partitions back to back in one arena block, its segments views
 * into that block; queries read such a column as one flat column in place.
 * The MAPCOMMON column (partition keys and row counts) stays on the host. */
td_t*    td_dev_parted_load(const char* db_root, const char* table_name);
/* Wait for all queued device work. */
td_err_t td_dev_sync(void);

/* Arena accounting and budget (admission control uses these). */
typedef struct {
    uint64_t in_use;         /* bytes handed out */
    uint64_t peak;
    uint64_t reserved;       /* device memory held by the arena */
    uint64_t budget;         /* cap on reserved; 0 = 90% of device memory */
    uint64_t device_total;
} td_dev_mem_t;
void     td_dev_mem_stats(td_dev_mem_t* out);
void     td_dev_set_budget(uint64_t bytes);

/* Kernel timing: while enabled every kernel launch is timed with HIP events
 * on the library stream; td_dev_prof_read fills up to cap rows of
 * (kernel name, launches, total milliseconds). */
typedef struct { char name[48]; int64_t launches; double total_ms; } td_dev_prof_t;
void     td_dev_prof_enable(int on);
void     td_dev_prof_reset(void);
int      td_dev_prof_read(td_dev_prof_t* rows, int cap);

/* Batched execution: many DAGs of the fused shape GROUP over the graph's
 * inner JOIN(FILTER*(fact), dim), with resident inputs, run as one batch of
 * launches (same results as td_execute on each; results are resident).
 * DAGs the batch cannot take run through td_execute.  results[i] receives
 * an owned table or an error pointer.  Returns TD_OK or the first error.
 * No reference counterpart (the reference runs queries one at a time). */
td_err_t td_dev_execute_batch(td_graph_t* const* graphs, td_op_t* const* roots, int64_t n,
                              td_t** results);
/* The same, prepared once and run many times: the first run captures the
 * launch sequence as a hipGraph, later runs replay it.  The graphs and
 * their tables must outlive the handle. */
typedef struct td_dev_batch td_dev_batch_t;
td_dev_batch_t* td_dev_batch_prepare(td_graph_t* const* graphs, td_op_t* const* roots,
                                     int64_t n);
td_err_t td_dev_batch_run(td_dev_batch_t* b, td_t** results);
/* batched queries, autotuned probe tile rows and workgroup size */
int64_t  td_dev_batch_info(const td_dev_batch_t* b, int64_t* tile_rows, int* block);
/* batched queries on the radix hash path (build keys not a dense unique
 * range), their autotuned build rows per partition and largest radix bits */
int64_t  td_dev_batch_radix(const td_dev_batch_t* b, int64_t* part_rows, int* max_bits);
void     td_dev_batch_free(td_dev_batch_t* b);

/* Serving: a request loop over a resident fact / dim pair.  Requests are
 * td.h DAGs; those of the form
 *   GROUP(dim key; SUM / AVG / COUNT of fact columns)
 *     over the graph's inner JOIN(FILTER(fact, pcol CMP constant), fk = pk)
 * with the configured columns share one scan of the fact table (continuous
 * batching: a request submitted while others run joins the scan at the next
 * step and finishes after one pass over every chunk); any other DAG runs
 * through td_execute when admitted.  Admission is in submission order while
 * request slots (max_active) and the device-memory budget (resident tables +
 * server tables + every active request's accumulator) allow; a request that
 * could never fit finishes with TD_ERR_OOM.  Results equal td_execute's on
 * the same DAG (integers exact, f64 sums in another order).  Graphs must
 * outlive their results.  No reference counterpart. */
typedef struct td_dev_server td_dev_server_t;
typedef struct {
    td_t*    fact;             /* resident tables (td_dev_put) */
    td_t*    dim;
    int64_t  fk, pk;           /* join columns (symbol ids); pk unique integers */
    int64_t  pcol;             /* fact column filtered by requests: small integer range */
    int64_t  vals[4];          /* fact value columns requests aggregate */
    int64_t  keys[4];          /* dim group-key columns */
    int32_t  n_vals, n_keys;
    uint64_t budget;           /* device bytes; 0: no cap beyond the arena's */
    int64_t  chunk_rows;       /* fact rows per step; 0: an eighth of the table */
    int32_t  max_active;       /* request slots; 0: 4096 */
    int32_t  pad;
} td_dev_serve_cfg_t;
typedef struct {
    int64_t  queued, active, ready;    /* waiting, running, finished not yet polled */
    int64_t  steps, chunks;            /* steps run; steps per request */
    int64_t  shared, alone;            /* requests admitted to the scan / run alone */
    uint64_t held, base, budget;       /* bytes held now, of which fixed; budget */
} td_dev_serve_info_t;
td_dev_server_t* td_dev_serve_open(const td_dev_serve_cfg_t* cfg);
/* Queue a request; returns its id (>= 0) or -error. */
int64_t  td_dev_serve_submit(td_dev_server_t* s, td_graph_t* g, td_op_t* root);
/* Admit what fits, run one scan step, finish requests that have seen every
 * chunk; returns the number of requests finished (or -error). */
int      td_dev_serve_step(td_dev_server_t* s);
/* Pop a finished request: 1 with *id / *result (owned table or error
 * pointer), 0 when none is ready. */
int      td_dev_serve_poll(td_dev_server_t* s, int64_t* id, td_t** result);
void     td_dev_serve_info(const td_dev_server_t* s, td_dev_serve_info_t* out);
void     td_dev_serve_close(td_dev_server_t* s);

/* Cancellation points: td_execute polls td_cancel() between DAG nodes and
 * the device layer polls it between the launches of multi-pass operators
 * (partition passes, retries, join phases); a cancelled query returns
 * TD_ERR_CANCEL with its device scratch freed.  Test hook: the k-th poll
 * from now acts as a td_cancel() (0 disables); td_dev_cancel_polls counts
 * the polls made so far. */
void     td_dev_cancel_after(int64_t polls);
int64_t  td_dev_cancel_polls(void);

/* Out of device memory: a join or group-by whose partitioned plan cannot
 * get its buffers falls back to a plan that needs less (the pair-producing
 * join; the global hash table or direct array), as the reference falls
 * back to its chained join and sequential group table (exec.c:7677-7683,
 * 8219-8231, 6486-6498); with a forced path it returns TD_ERR_OOM.  Test
 * hook: the next `count` arena requests of at least `min_bytes` fail. */
void     td_dev_fail_allocs(uint64_t min_bytes, int64_t count);

/* Physical-plan hints for tests and measurement: force the join path
 * (0 auto, 1 direct map, 2 radix, 3 partitioned fused join-group,
 * 4 XCD-partitioned fused join-group) / group path (0 auto, 1 direct
 * array, 2 hash, 3 radix-partitioned direct array with small leaf tables,
 * 4 hash-partitioned hash with small LDS tables, 5 hash-partitioned hash
 * after hot-key pre-aggregation at any input size, 6 as 4 with a first try
 * of 1,024-entry global table and list, which then regrow) / radix bits
 * (0 auto) for subsequent td_execute calls. */
void     td_dev_force_paths(int join_path, int group_path, int radix_bits);
/* The physical plan of the last td_execute as a short text. */
const char* td_dev_last_plan(void);

#ifdef __cplusplus
}
#endif
#endif
