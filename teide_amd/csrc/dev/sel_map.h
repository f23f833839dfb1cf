/*
 * sel_map.h -- stream compaction that also reads right columns through a
 * dense key map: the materialising join's output when its build keys are
 * unique and dense in their range.
 */
#pragma once

#include "dev_common.h"

#define TDD_MAP_MAXC 16

/* Right columns laid out by key: src[c][key - kmin] holds the value of the
 * build row with that key, present bit (key - kmin) says one exists
 * (present == nullptr: every key of the range does). */
struct TddMapGather {
    tdd_col_t       key;             /* probe key column (source rows) */
    int64_t         kmin, range;
    const uint32_t* present;
    int32_t         n, pad;
    const void*     src[TDD_MAP_MAXC];
    void*           dst[TDD_MAP_MAXC];
    uint8_t         esz[TDD_MAP_MAXC];
};

/* tdd_sel_compact of ncols columns, plus M->n right columns written for
 * the same output rows: the value at the row's key, 0 without one. */
int tdd__sel_compact_map(const tdd_sel_t* s, int ncols, const tdd_col_t* src, void* const* dst,
                         const TddMapGather* M);
