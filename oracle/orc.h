/*
 * orc.h -- the oracle: a sequential plain-C restatement of the reference
 * engine's filter -> join -> group-by semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as a checker.  The
 * product (libteide_amd.so) never links, loads or calls it.
 *
 * Each function cites the reference code it restates.  Values are exchanged
 * as raw arrays tagged with td.h type codes; no td_t objects are involved.
 */
#ifndef ORC_H
#define ORC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A column: raw data, td.h type tag, attrs (TD_SYM width in bits 0-1). */
typedef struct {
    const void* data;
    int8_t      type;
    uint8_t     attrs;
} orc_col_t;

/* ------------------------------------------------------------ selection -- */

/* Bytes of a TD_SEL payload for nrows (src/vec/sel.c:37-47). */
int64_t orc_sel_payload_bytes(int64_t nrows);
/* BOOL bytes -> TD_SEL payload: LSB-first bit words, per-1024-row segment
 * flags and popcounts, total_pass (sel.c:77-161). */
void    orc_sel_from_bool(const uint8_t* pred, int64_t nrows, uint8_t* payload);
/* AND of two selections of equal length (sel.c:167-189). */
void    orc_sel_and(const uint8_t* a, const uint8_t* b, int64_t nrows, uint8_t* out);
/* Row ids of the set bits, ascending (exec.c:2362-2388).  Returns count. */
int64_t orc_sel_to_idx(const uint8_t* payload, int64_t nrows, int64_t* idx);

/* ------------------------------------------------------------ predicates -- */

/* out[i] = (col[i] OP constant), evaluated in i64 when both sides are
 * integer-typed and in f64 when either is F64 (exec.c:760-918, 982-1051).
 * op is one of OP_EQ..OP_GE. */
void orc_cmp_const(orc_col_t col, int64_t n, int op, int const_is_f64,
                   int64_t ci, double cf, uint8_t* out);
/* out[i] = (a[i] OP b[i]) with the same promotion rule. */
void orc_cmp_cols(orc_col_t a, orc_col_t b, int64_t n, int op, uint8_t* out);
/* Boolean combinators on BOOL bytes: OP_AND, OP_OR, OP_NOT (b ignored). */
void orc_bool_op(int op, const uint8_t* a, const uint8_t* b, int64_t n, uint8_t* out);

/* ---------------------------------------------------------------- gather -- */

/* dst[i] = src[idx[i]] for esz-byte elements; idx < 0 writes zeros when
 * nullable (exec.c:1852-1939). */
void orc_gather(const void* src, int esz, const int64_t* idx, int64_t m,
                void* dst, int nullable);

/* ------------------------------------------------------------------ join -- */

/* Equi-join on nk key columns.  join_type 0 inner, 1 left, 2 full
 * (exec.c:7431-8741).  Keys compare by value: F64 with ==, every other type
 * read as int64 (exec.c:7754-7768).  A NULL key column never matches.
 * Pairs are emitted for left rows in ascending order, each left row's
 * matches in ascending right order; unmatched left rows (left/full) get
 * r = -1; unmatched right rows (full) are appended with l = -1, ascending.
 * Returns the pair count; writes at most cap pairs. */
int64_t orc_join(const orc_col_t* lkeys, const orc_col_t* rkeys, int nk,
                 int64_t nl, int64_t nr, int join_type,
                 int64_t* l_idx, int64_t* r_idx, int64_t cap);

/* ---------------------------------------------------------------- group -- */

/* Result of orc_group, caller-allocated with room for cap groups:
 *   keys[k * cap + g]  key k of group g as int64 (F64 keys as raw bits)
 *   vals[a * cap + g]  aggregate a: double bits when out_types[a] == F64,
 *                      else int64
 *   sum_abs[a*cap+g]   sum of |x| over the group (f64 tolerance checks)
 *   comp[a*cap+g]      compensated (Neumaier) sum for SUM/AVG inputs
 * order: 1 = ascending composite key (the reference's direct-array path,
 * exec.c:5948-6404), 0 = first appearance (its hash paths, compared after a
 * canonical sort).  Returns the group count, or -1 on an unsupported op. */
typedef struct {
    int64_t* keys;
    uint64_t* vals;
    double*  sum_abs;
    double*  comp;
    int8_t   out_types[8];
    int      order;
} orc_group_out_t;

int64_t orc_group(const orc_col_t* keys, int nk, const uint16_t* ops,
                  const orc_col_t* aggs, int na, int64_t nrows,
                  const uint64_t* mask_bits, int64_t cap, orc_group_out_t* out);

#ifdef __cplusplus
}
#endif
#endif
