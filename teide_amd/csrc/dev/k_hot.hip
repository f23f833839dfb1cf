/*
 * k_hot.hip -- hot-key pre-aggregation for the partitioned hash group-by.
 *
 * On skewed keys (Zipf data: a few hundred keys can hold most of the rows)
 * the partitioned hash group-by would carry every row of every hot key
 * through its scatter passes and LDS leaf tables.  Here the hot keys are
 * found first and their rows folded in one pass over the rows, so that the
 * scatter passes only carry the cold rows:
 *
 *  1. sample: every step-th selected row's key is counted, per workgroup
 *     in LDS and then in a global open-addressing table (about 2 M samples);
 *  2. a histogram of the sampled keys' counts by power of two picks the
 *     count threshold that keeps at most HOT_MAX keys; the pre-pass is
 *     taken only when those keys hold at least HOT_MIN_SHARE of the sample;
 *  3. the keys at or over the threshold are compacted into the hot list;
 *  4. fold (k_pg_hotfold, with the partitioned group-by's row reader and
 *     record fold): every workgroup loads the hot list into an LDS hash set
 *     with an LDS accumulator block per hot key, walks its 64-row words,
 *     folds the rows of hot keys into those blocks, and writes the words'
 *     cold rows (selected, key not hot) as selection words; the blocks
 *     merge into the global hot cells once per workgroup;
 *  5. the group-by then runs on the cold selection, and the hot keys'
 *     cells merge into its global hash table before the groups are emitted.
 *
 * The hot keys are the group-by's own 64-bit keys (ht_key), so a hot key
 * stands for exactly the rows the hash group-by would have grouped with it.
 */
#include "group_plan.h"


#define HOT_MAX       512                /* hot keys at most */
#define HOT_SAMPLES   ((int64_t)1 << 21)
#define HOT_MIN_SHARE 0.25               /* the hot keys' share of the sample */

/* Count one key into the global table (claim its slot, add n). */
__device__ __forceinline__ void hot_count_add(unsigned long long* tkeys, unsigned int* tcnt,
                                              uint64_t tmask, uint64_t key, unsigned int n) {
    uint64_t i = mix64(key) & tmask;
    for (;;) {
        const unsigned long long k = tkeys[i];
        if (k == key) break;
        if (k == EMPTY_KEY) {
            const unsigned long long prev = atomicCAS(&tkeys[i], EMPTY_KEY, (unsigned long long)key);
            if (prev == EMPTY_KEY || prev == key) break;
        }
        i = (i + 1) & tmask;
    }
    atomicAdd(&tcnt[i], n);
}

#define HOT_LSET 2048                    /* per-workgroup LDS counts of the sample */

/* Every step-th selected row's key counted: first in this workgroup's LDS
 * table (so that the hottest keys do not all hit one global counter), then
 * the LDS counts added to the global table once per workgroup. */
__global__ __launch_bounds__(256) void k_hot_sample(GroupPlan P, const uint64_t* mask,
                                                    int64_t nrows, int64_t step,
                                                    unsigned long long* tkeys, unsigned int* tcnt,
                                                    uint64_t tmask) {
    __shared__ unsigned long long lk[HOT_LSET];
    __shared__ unsigned int lc[HOT_LSET];
    for (int i = threadIdx.x; i < HOT_LSET; i += blockDim.x) { lk[i] = EMPTY_KEY; lc[i] = 0; }
    __syncthreads();
    const int64_t ns = (nrows + step - 1) / step;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < ns;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = s * step;
        if (mask && !((mask[row >> 6] >> (row & 63)) & 1)) continue;
        const uint64_t key = ht_key(P, row);
        if (key == EMPTY_KEY) continue;               /* the all-ones key stays cold */
        uint32_t x = (uint32_t)mix64(key) & (HOT_LSET - 1);
        bool done = false;
        for (int p = 0; p < 16 && !done; p++) {
            const unsigned long long k = lk[x];
            if (k == EMPTY_KEY) {
                const unsigned long long prev = atomicCAS(&lk[x], EMPTY_KEY, (unsigned long long)key);
                if (prev == EMPTY_KEY || prev == key) { atomicAdd(&lc[x], 1u); done = true; }
            } else if (k == key) {
                atomicAdd(&lc[x], 1u);
                done = true;
            }
            x = (x + 1) & (HOT_LSET - 1);
        }
        if (!done) hot_count_add(tkeys, tcnt, tmask, key, 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HOT_LSET; i += blockDim.x)
        if (lk[i] != EMPTY_KEY && lc[i]) hot_count_add(tkeys, tcnt, tmask, lk[i], lc[i]);
}

/* hist[2 * b]: sampled keys counted in [2^b, 2^(b+1)); hist[2 * b + 1]:
 * their samples.  Binned in LDS, added to hist once per workgroup (most
 * keys fall in bin 0: one global counter per row would serialise). */
__global__ __launch_bounds__(256) void k_hot_hist(const unsigned long long* tkeys,
                                                  const unsigned int* tcnt, int64_t n,
                                                  unsigned long long* hist) {
    __shared__ unsigned long long lh[64];
    if (threadIdx.x < 64) lh[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (tkeys[i] == EMPTY_KEY || tcnt[i] == 0) continue;
        const int b = 31 - __clz(tcnt[i]);
        atomicAdd(&lh[2 * b], 1ull);
        atomicAdd(&lh[2 * b + 1], (unsigned long long)tcnt[i]);
    }
    __syncthreads();
    if (threadIdx.x < 64 && lh[threadIdx.x]) atomicAdd(&hist[threadIdx.x], lh[threadIdx.x]);
}

__global__ void k_hot_pick(const unsigned long long* tkeys, const unsigned int* tcnt, int64_t n,
                           unsigned int thr, unsigned long long* hot, unsigned int* nhot) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        if (tkeys[i] == EMPTY_KEY || tcnt[i] < thr) continue;
        const unsigned int at = atomicAdd(nhot, 1u);
        if (at < HOT_MAX) hot[at] = tkeys[i];
    }
}

/* Identity values of n accumulator blocks. */
__global__ void k_hot_init(uint64_t* cells, int64_t n, GroupPlan P) {
    const int64_t total = n * P.ncells;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
        cells[i] = cell_identity(P.kind[i % P.ncells]);
}

/* The hot keys' cells into the group-by's global hash table. */
__global__ void k_hot_merge(GroupPlan P, const unsigned long long* hot, int nhot,
                            const uint64_t* hcells, unsigned long long* gkeys, uint64_t gmask,
                            uint64_t* gcells, int* overflow) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhot || hcells[(int64_t)h * P.ncells] == 0) return;
    const int64_t g = ht_find(gkeys, gmask, hot[h]);
    if (g < 0) { *overflow = 1; return; }
    cells_merge(&gcells[g * P.ncells], &hcells[(int64_t)h * P.ncells], P);
}

/* ------------------------------------------------------------------ host -- */

struct tdd_hot {
    void*  keys = nullptr;                /* hot list (device) */
    void*  cells = nullptr;               /* nhot x ncells, identities (device) */
    int    nhot = 0;
    ~tdd_hot() {
        (void)hipStreamSynchronize(tdd::g_stream);
        if (keys) td__arena_free(keys);
        if (cells) td__arena_free(cells);
    }
};

/* Steps 1-3 over rows [0, nrows) selected by mask.  *out stays NULL when
 * the keys are not skewed enough, there are fewer than 4 M rows (unless
 * forced) or the accumulator blocks would not fit LDS; otherwise it holds
 * the hot list and its cells, ready for the fold (step 4, k_pg_hotfold). */
int hot_split(const GroupPlan& P, const uint64_t* mask, int64_t nrows, bool forced,
              tdd_hot** out) {
    *out = nullptr;
    const int nc = P.ncells;
    int cap = HOT_MAX;                    /* the fold's LDS blocks within 64 KiB */
    while (cap > 0 && (size_t)2048 * 10 + (size_t)cap * nc * 8 > 64 * 1024) cap /= 2;
    if (cap < 64 || nrows <= 0 || (!forced && nrows < ((int64_t)1 << 22))) return TD_OK;
    const int64_t step = nrows / HOT_SAMPLES > 1 ? nrows / HOT_SAMPLES : 1;
    const int64_t ns = (nrows + step - 1) / step;
    uint64_t tcap = 1024;
    while ((int64_t)tcap < 2 * ns) tcap <<= 1;
    SCRATCH(tkeys, tcap * 8);
    SCRATCH(tcnt, tcap * 4);
    SCRATCH(hist, 64 * 8 + 16);
    int rc = tdd_fill(tkeys.p, EMPTY_KEY, tcap * 8);
    if (!rc) rc = tdd_fill(tcnt.p, 0, tcap * 4);
    if (!rc) rc = tdd_fill(hist.p, 0, 64 * 8 + 16);
    if (rc) return rc;
    LAUNCH("group_hot", k_hot_sample, stream_grid(ceil_div(ns, 256)), 256, 0, P, mask, nrows, step,
           tkeys.as<unsigned long long>(), tcnt.as<unsigned int>(), (uint64_t)(tcap - 1));
    LAUNCH("group_hot", k_hot_hist, stream_grid(ceil_div((int64_t)tcap, 256)), 256, 0,
           (const unsigned long long*)tkeys.as<unsigned long long>(),
           (const unsigned int*)tcnt.as<unsigned int>(), (int64_t)tcap,
           hist.as<unsigned long long>());
    unsigned long long hh[64];
    HC(hipMemcpyAsync(hh, hist.p, sizeof(hh), hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    /* the lowest power-of-two threshold (>= 4 samples) keeping <= cap keys */
    unsigned long long total = 0, nk = 0, share = 0;
    for (int b = 0; b < 32; b++) total += hh[2 * b + 1];
    int tb = -1;
    for (int b = 31; b >= 2; b--) {
        if (nk + hh[2 * b] > (unsigned long long)cap) break;
        nk += hh[2 * b];
        share += hh[2 * b + 1];
        tb = b;
    }
    if (tb < 0 || nk == 0 || total == 0 || (double)share < HOT_MIN_SHARE * (double)total)
        return TD_OK;
    tdd_hot* h = new tdd_hot();
    h->keys = td__arena_alloc((size_t)HOT_MAX * 8);
    h->cells = td__arena_alloc((size_t)cap * nc * 8);
    if (!h->keys || !h->cells) { delete h; return TD_ERR_OOM; }
    unsigned int* dn = (unsigned int*)(hist.as<unsigned long long>() + 64);
    LAUNCH("group_hot", k_hot_pick, stream_grid(ceil_div((int64_t)tcap, 256)), 256, 0,
           (const unsigned long long*)tkeys.as<unsigned long long>(),
           (const unsigned int*)tcnt.as<unsigned int>(), (int64_t)tcap, (unsigned int)(1u << tb),
           (unsigned long long*)h->keys, dn);
    unsigned int nhot = 0;
    HC(hipMemcpyAsync(&nhot, dn, 4, hipMemcpyDeviceToHost, tdd::g_stream));
    HC(hipStreamSynchronize(tdd::g_stream));
    h->nhot = (int)(nhot < (unsigned)cap ? nhot : (unsigned)cap);
    const int ib = stream_grid(ceil_div((int64_t)h->nhot * nc, 256 * 8));
    LAUNCH("group_hot", k_hot_init, ib, 256, 0, (uint64_t*)h->cells, (int64_t)h->nhot, P);
    *out = h;
    return TD_OK;
}

int hot_count(const tdd_hot* h) { return h->nhot; }
const unsigned long long* hot_keys(const tdd_hot* h) { return (const unsigned long long*)h->keys; }
uint64_t* hot_cells(const tdd_hot* h) { return (uint64_t*)h->cells; }

int hot_merge(const GroupPlan& P, const tdd_hot* h, unsigned long long* gkeys, uint64_t gmask,
              uint64_t* gcells, int* overflow) {
    if (h->nhot <= 0) return TD_OK;
    LAUNCH("group_hot", k_hot_merge, (int)ceil_div(h->nhot, 256), 256, 0, P,
           (const unsigned long long*)h->keys, h->nhot, (const uint64_t*)h->cells, gkeys, gmask,
           gcells, overflow);
    return TD_OK;
}

void hot_free(tdd_hot* h) { delete h; }
