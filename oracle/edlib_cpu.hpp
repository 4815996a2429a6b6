// ORACLE (test infrastructure only; never linked into the product).
//
// CPU restatement of the pairwise step racon takes from edlib
// (martinsos/edlib >= 1.2, 5-argument edlibNewAlignConfig; call sites racon
// src/overlap.cpp:205-224 and test/racon_test.cpp:16-25). edlib is an
// un-vendored submodule (vendor/edlib is empty, no commit pinned), so this
// restates its published algorithm: Myers 1999 bit-vector edit distance in
// 64-bit blocks with Hyyro's global (NW) boundary, Ukkonen banding with k
// starting at 64 and doubling until the distance fits, and the path found by
// direct traceback when (2*8+4) * blocks * |t| < 1 MiB, else by Hirschberg
// splitting the target in half (Hirschberg 1975, Sosic & Sikic 2017).
// Traceback preference at a cell: up (insertion), then left (deletion), then
// diagonal (match/mismatch).
//
// Two versions are given:
//   edlib_align_nw        - banded, block-by-block like edlib;
//   edlib_align_nw_exact  - the same recursion on exact (unbanded) DP values.
// They must agree (tests check it); the second is the definition the device
// kernel implements.
#pragma once

#include <cstdint>
#include <vector>

namespace oracle {

enum : uint8_t { kEdMatch = 0, kEdInsert = 1, kEdDelete = 2, kEdMismatch = 3 };

struct EdlibResult {
  int32_t edit_distance = -1;
  std::vector<uint8_t> ops;  // alignment path, kEd* codes
  int status = 0;            // 0 ok, 1 error
};

// EDLIB_MODE_NW + EDLIB_TASK_PATH with k = -1 (racon's configuration).
EdlibResult edlib_align_nw(const char* query, int32_t query_length,
                           const char* target, int32_t target_length,
                           bool find_path = true);

// Same result computed from exact DP values (full matrix, no band).
EdlibResult edlib_align_nw_exact(const char* query, int32_t query_length,
                                 const char* target, int32_t target_length);

// Plain Levenshtein distance (EDLIB_TASK_DISTANCE); used by the accuracy
// metric of racon's tests (racon_test.cpp:16-25).
int32_t edit_distance(const char* query, int32_t query_length,
                      const char* target, int32_t target_length);

}  // namespace oracle
