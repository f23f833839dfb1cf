/*
 * core.c -- error strings, element sizes, platform and thread helpers.
 */
#define _GNU_SOURCE
#include "internal.h"

#include <fcntl.h>
#include <pthread.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

/* Element sizes by positive type tag (the reference's src/core/types.c:27-49). */
const uint8_t td_type_sizes[TD_TYPE_COUNT] = {
    8, 1, 1, 1, 2, 4, 8, 8, 0, 4, 4, 8, 16, 8, 0, 0, 0, 0, 0, 0, 8,
};

const char* td_err_str(td_err_t e) {
    static const char* const names[] = {
        "ok", "out of memory", "type error", "range error", "length error",
        "rank error", "domain error", "not yet implemented", "io error",
        "schema error", "corrupt", "cancelled",
    };
    if ((unsigned)e < sizeof(names) / sizeof(names[0])) return names[e];
    return "unknown error";
}

/* ------------------------------------------------------------- platform -- */

void* td_vm_alloc(size_t size) {
    void* p = mmap(NULL, size, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    return p == MAP_FAILED ? NULL : p;
}
void td_vm_free(void* ptr, size_t size) { if (ptr) munmap(ptr, size); }

void* td_vm_map_file(const char* path, size_t* out_size) {
    int fd = open(path, O_RDONLY);
    if (fd < 0) return NULL;
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size == 0) { close(fd); return NULL; }
    void* p = mmap(NULL, (size_t)st.st_size, PROT_READ | PROT_WRITE, MAP_PRIVATE, fd, 0);
    close(fd);
    if (p == MAP_FAILED) return NULL;
    if (out_size) *out_size = (size_t)st.st_size;
    return p;
}
void td_vm_unmap_file(void* ptr, size_t size) { if (ptr) munmap(ptr, size); }
void td_vm_advise_seq(void* ptr, size_t size) { madvise(ptr, size, MADV_SEQUENTIAL); }
void td_vm_advise_willneed(void* ptr, size_t size) { madvise(ptr, size, MADV_WILLNEED); }
void td_vm_release(void* ptr, size_t size) { madvise(ptr, size, MADV_DONTNEED); }
void* td_vm_alloc_aligned(size_t size, size_t alignment) {
    void* p = NULL;
    if (posix_memalign(&p, alignment < sizeof(void*) ? sizeof(void*) : alignment, size))
        return NULL;
    return p;
}

typedef struct { td_thread_fn fn; void* arg; } thr_boot_t;
static void* thr_main(void* a) {
    thr_boot_t b = *(thr_boot_t*)a;
    free(a);
    b.fn(b.arg);
    return NULL;
}

td_err_t td_thread_create(td_thread_t* t, td_thread_fn fn, void* arg) {
    thr_boot_t* b = malloc(sizeof(*b));
    if (!b) return TD_ERR_OOM;
    b->fn = fn;
    b->arg = arg;
    pthread_t th;
    if (pthread_create(&th, NULL, thr_main, b) != 0) { free(b); return TD_ERR_OOM; }
    *t = (td_thread_t)th;
    return TD_OK;
}

td_err_t td_thread_join(td_thread_t t) {
    return pthread_join((pthread_t)t, NULL) == 0 ? TD_OK : TD_ERR_IO;
}

uint32_t td_thread_count(void) {
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (uint32_t)n : 1;
}

_Atomic(uint32_t) td_parallel_flag;
void td_parallel_begin(void) { atomic_store(&td_parallel_flag, 1); }
void td_parallel_end(void) { atomic_store(&td_parallel_flag, 0); }

/* The device replaces the reference's CPU worker pool (src/ops/pool.c): the
 * pool entry points only manage the cooperative cancel flag. */
static _Atomic(int) g_cancel;
td_err_t td_pool_init(uint32_t n_workers) { (void)n_workers; return TD_OK; }
void td_pool_destroy(void) {}
void td_cancel(void) { atomic_store(&g_cancel, 1); }
int  td__cancel_requested(void) { return atomic_load(&g_cancel); }
void td__cancel_reset(void) { atomic_store(&g_cancel, 0); }

/* Cancellation points: td_execute polls between DAG nodes and the device
 * layer between the launches of its multi-pass operators.  A poll sees a
 * td_cancel() from any thread; the test hook td_dev_cancel_after(k) makes
 * the k-th poll from now act as one. */
static _Atomic(int64_t) g_cancel_after, g_polls;
int td__cancel_poll(void) {
    atomic_fetch_add(&g_polls, 1);
    int64_t k = atomic_load(&g_cancel_after);
    if (k > 0 && atomic_fetch_sub(&g_cancel_after, 1) == 1) atomic_store(&g_cancel, 1);
    return atomic_load(&g_cancel);
}
void    td_dev_cancel_after(int64_t polls) { atomic_store(&g_cancel_after, polls > 0 ? polls : 0); }
int64_t td_dev_cancel_polls(void) { return atomic_load(&g_polls); }
