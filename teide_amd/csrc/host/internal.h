/*
 * internal.h -- declarations shared by the C17 host modules of libteide_amd.
 */
#ifndef TD_INTERNAL_H
#define TD_INTERNAL_H

#include <teide/td.h>
#include <teide/td_dev.h>
#include <teide/tdd.h>

#include <stdatomic.h>
#include <string.h>

/* mmod values.  0 heap and 1 file mapping follow td.h:234; 3 marks a
 * device-resident block whose payload lives in the HBM arena (td_dev.h);
 * 0xFE tags a free heap block for buddy coalescing. */
#define TD_MMOD_HEAP  0
#define TD_MMOD_FILE  1
#define TD_MMOD_HBM   3
#define TD_MMOD_FREE  0xFE

/* Payload of a host proxy for a device-resident object (mmod == 3). */
typedef struct {
    void*    dptr;     /* device address of the elements (or TD_SEL payload) */
    uint64_t bytes;    /* allocated device bytes */
    int64_t  lo, hi;   /* integer columns loaded by td_dev_put: min / max */
    int32_t  ranged;   /* lo / hi valid (resident columns are immutable) */
    int32_t  pad;
    td_t*    owner;    /* non-NULL: a view into owner's block (held) */
} td_hbm_ref_t;

static inline td_hbm_ref_t* td_hbm_ref(td_t* v) { return (td_hbm_ref_t*)td_data(v); }
static inline int td_is_resident(const td_t* v) {
    return v && !TD_IS_ERR(v) && v->mmod == TD_MMOD_HBM;
}

/* ---- heap.c ---- */
void   td__heap_lock(void);
void   td__heap_unlock(void);

/* ---- arena.c: HBM arena (buddy allocator over device memory) ---- */
void*  td__arena_alloc(size_t bytes);          /* NULL on exhaustion */
void   td__arena_free(void* dptr);
size_t td__arena_in_use(void);
size_t td__arena_peak(void);
size_t td__arena_reserved(void);
int    td__dev_ready(void);                    /* lazy device + arena init */
void   td__arena_release_cached(void);
void   td__dev_forget(td_t* v);                /* td_free hook: drop cached upload */

/* Device-resident vector: host proxy with mmod == 3 and the elements in the
 * arena.  bytes >= len * elem size (TD_SEL: full payload). */
td_t*  td__hbm_vec_new(int8_t type, uint8_t attrs, int64_t len, size_t bytes);
int    td__dev_stats(td_t* d);                  /* min / max of a resident integer column */
td_t*  td__hbm_wrap(int8_t type, uint8_t attrs, int64_t len, void* dptr, size_t bytes);
void   td__hbm_free(td_t* v);
/* A resident vector viewing len elements of `owner`'s block from byte
 * offset `off`; holds a reference on owner. */
td_t*  td__hbm_view(td_t* owner, int8_t type, uint8_t attrs, int64_t len, size_t off,
                    size_t bytes);

/* ---- dev_api.c ---- */
td_t*  td__to_host(td_t* v);          /* consumes v: host copy of a proxy / table */
int    td__force_join_path(void);
int    td__force_group_path(void);
int    td__force_radix_bits(void);
void   td__set_plan(const char* fmt, ...);

size_t td__arena_budget(void);
void   td__arena_set_budget(size_t bytes);

/* ---- sel.c ---- */
size_t td__sel_payload_bytes(int64_t nrows);
void   td__sel_layout(int64_t nrows, size_t* off_flags, size_t* off_pop, size_t* off_bits);
tdd_sel_t td__sel_view(void* base, int64_t nrows);
td_t*  td__sel_new_dev(int64_t nrows);
void*  td__sel_dev_payload(td_t* s, void** tmp);

/* ---- core.c ---- */
int    td__cancel_requested(void);
int    td__cancel_poll(void);
void   td__cancel_reset(void);

/* ---- sym / table helpers ---- */
int64_t td__intern_cstr(const char* s);
const char* td__sym_cstr(int64_t id, size_t* len);

/* ---- exec helpers shared by exec.c and dev_api.c ---- */
td_op_ext_t* td__find_ext(td_graph_t* g, uint32_t node_id);
td_op_t*     td__graph_new_node(td_graph_t* g);

/* Element size honouring TD_SYM width. */
static inline uint8_t td__esz(const td_t* v) {
    return td_sym_elem_size(v->type, v->attrs);
}

#endif
