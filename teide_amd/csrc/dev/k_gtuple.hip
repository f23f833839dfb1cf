/*
 * k_gtuple.hip -- group-by on key tuples that no single 64-bit word holds:
 * two or more keys with an F64 among them, or integer keys whose ranges
 * multiply past 2^63.
 *
 * Each selected row finds its tuple in a global open-addressing table of
 * row ids (a 64-bit hash of the keys' raw 8-byte values picks the slot, a
 * compare of every key column confirms it), and takes the id of the row
 * that first claimed the tuple's slot.  The group-by then runs on that one
 * int64 key, and each output key column is gathered at the groups' ids.
 * Keys compare by their raw 8-byte values, as the reference's group table
 * does with memcmp over its key words (exec.c:4131-4165): -0.0 and 0.0 are
 * different keys, a NaN equals only a NaN with the same bits.
 */
#include "dev_common.h"
#include "join_keys.h"

struct TupleKeys {
    tdd_col_t cols[TDD_MAX_KEYS];
    int32_t   n, pad;
};

__device__ __forceinline__ uint64_t tuple_hash(const TupleKeys& K, int64_t r) {
    uint64_t h = 0x9e3779b97f4a7c15ull;
    for (int k = 0; k < K.n; k++) h = jmix(h ^ (uint64_t)ld_i64(K.cols[k], r)) + (uint64_t)k;
    return h;
}

__device__ __forceinline__ bool tuple_eq(const TupleKeys& K, int64_t a, int64_t b) {
    for (int k = 0; k < K.n; k++)
        if (ld_i64(K.cols[k], a) != ld_i64(K.cols[k], b)) return false;
    return true;
}

/* The table holds fewer tuples than half its slots, so every probe
 * sequence meets an empty slot or its tuple. */
__global__ void k_tuple_ids(TupleKeys K, const uint64_t* mask, int64_t nrows,
                            unsigned long long* table, uint64_t tmask, int64_t* gid) {
    const unsigned long long EMPTY = ~0ull;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
         r += (int64_t)gridDim.x * blockDim.x) {
        if (!row_selected(mask, r)) continue;
        uint64_t s = tuple_hash(K, r) & tmask;
        for (;;) {
            unsigned long long cur = __hip_atomic_load(&table[s], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
            if (cur == EMPTY) {
                cur = atomicCAS(&table[s], EMPTY, (unsigned long long)r);
                if (cur == EMPTY) { gid[r] = r; break; }
            }
            if (tuple_eq(K, (int64_t)cur, r)) { gid[r] = (int64_t)cur; break; }
            s = (s + 1) & tmask;
        }
    }
}

int group_tuples(const tdd_group_spec_t* spec, int64_t nsel, tdd_group_result_t* out) {
    memset(out, 0, sizeof(*out));
    POLL_CANCEL();
    const int64_t n = spec->nrows;
    uint64_t cap = 1024;
    while (cap < 2 * (uint64_t)(nsel > 0 ? nsel : 0) + 1024) cap <<= 1;
    SCRATCH(table, (size_t)cap * 8);
    SCRATCH(gid, (size_t)(n > 0 ? n : 1) * 8);
    int rc = tdd_fill(table.p, ~0ull, (size_t)cap * 8);
    if (rc) return rc;
    TupleKeys K;
    memset(&K, 0, sizeof(K));
    K.n = spec->n_keys;
    for (int k = 0; k < K.n; k++) K.cols[k] = spec->keys[k];
    if (n > 0)
        LAUNCH("group_tuple_ids", k_tuple_ids, stream_grid(ceil_div(n, 256 * 4)), 256, 0, K,
               spec->mask ? (const uint64_t*)spec->mask->bits : nullptr, n,
               table.as<unsigned long long>(), (uint64_t)(cap - 1), gid.as<int64_t>());
    tdd_group_spec_t G = *spec;
    G.n_keys = 1;
    memset(G.keys, 0, sizeof(G.keys));
    G.keys[0].data = gid.p;
    G.keys[0].type = TD_I64;
    G.keys[0].esz = 8;
    tdd_group_result_t R;
    rc = tdd_group(&G, &R);
    if (rc) return rc;
    const int64_t ng = R.n_groups;
    bool ok = true;
    for (int k = 0; k < spec->n_keys && ok; k++) {
        out->keys[k] = td__arena_alloc((size_t)(ng > 0 ? ng : 1) * spec->keys[k].esz);
        ok = out->keys[k] && tdd_gather(&spec->keys[k], (const int64_t*)R.keys[0], ng, 0,
                                        out->keys[k]) == TD_OK;
    }
    td__arena_free(R.keys[0]);
    if (!ok) {
        for (int k = 0; k < spec->n_keys; k++) { td__arena_free(out->keys[k]); out->keys[k] = nullptr; }
        for (int a = 0; a < spec->n_aggs; a++) td__arena_free(R.vals[a]);
        return TD_ERR_OOM;
    }
    for (int a = 0; a < spec->n_aggs; a++) {
        out->vals[a] = R.vals[a];
        out->out_types[a] = R.out_types[a];
    }
    out->n_groups = ng;
    out->path = 20 + R.path;
    out->tries = R.tries;
    out->part_rows = R.part_rows;
    HC(hipStreamSynchronize(tdd::g_stream));
    return TD_OK;
}
