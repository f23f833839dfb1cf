/*
 * stubs.c -- td.h entry points outside the filter -> join -> group-by path
 * and its loaders (parted tables, CSV, graph relations).  They exist so that
 * programs written against the reference link unchanged; each reports
 * TD_ERR_NYI.  Column files and splayed tables live in storage.c.
 */
#include "internal.h"

td_err_t td_meta_save_d(td_t* schema, const char* path) { (void)schema; (void)path; return TD_ERR_NYI; }
td_t* td_meta_load_d(const char* path) { (void)path; return TD_ERR_PTR(TD_ERR_NYI); }
td_t* td_read_csv(const char* path) { (void)path; return TD_ERR_PTR(TD_ERR_NYI); }
td_t* td_read_csv_opts(const char* path, char d, bool h, const int8_t* t, int32_t n) {
    (void)path; (void)d; (void)h; (void)t; (void)n; return TD_ERR_PTR(TD_ERR_NYI);
}
td_err_t td_write_csv(td_t* table, const char* path) { (void)table; (void)path; return TD_ERR_NYI; }

td_rel_t* td_rel_build(td_t* t, const char* fk, int64_t n, bool s) {
    (void)t; (void)fk; (void)n; (void)s; return NULL;
}
td_rel_t* td_rel_from_edges(td_t* t, const char* a, const char* b, int64_t n1, int64_t n2, bool s) {
    (void)t; (void)a; (void)b; (void)n1; (void)n2; (void)s; return NULL;
}
td_err_t td_rel_save(td_rel_t* rel, const char* dir) { (void)rel; (void)dir; return TD_ERR_NYI; }
td_rel_t* td_rel_load(const char* dir) { (void)dir; return NULL; }
td_rel_t* td_rel_mmap(const char* dir) { (void)dir; return NULL; }
void td_rel_free(td_rel_t* rel) { (void)rel; }
