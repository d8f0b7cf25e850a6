// Launch wrappers for the bordered (multiply connected) system.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rs_device.hpp"

namespace iegpu {

void launch_mark_skel(const DevState& s, int* mask, int64_t ndof,
                      hipStream_t st);
void launch_copy(double* dst, const double* src, int64_t n, hipStream_t st);
// Row gather: dst row r = src row map[r] (rows of row_words 32-bit words;
// map[r] < 0 zeroes the row).  Every map entry indexes a row of src.
void launch_gather_rows(void* dst, const void* src, const int* map, int n_rows, int row_words,
                        hipStream_t st);
int64_t border_partial_blocks(int64_t ndof);
void launch_border_gram(const double* P, const double* C, const int* mask,
                        int64_t ndof, int h, double* partial, double* out,
                        hipStream_t st);
void launch_border_gemv(const double* P, const double* x, const int* mask,
                        int64_t ndof, int h, double* partial, double* out,
                        double sign, hipStream_t st);
template <typename R>
void launch_border_factor(const DevState& s, const double* mU,
                          const double* mPsiT, const double* gram,
                          int64_t ndof, int h, double* Smat, int* Spiv,
                          hipStream_t st);
void launch_border_solve(const DevState& s, const double* Smat,
                         const int* Spiv, int h, double* x, const double* tail,
                         const double* g, double* alpha, double* ybuf,
                         hipStream_t st);
void launch_red_update(double* x, const double* C, const int* mask,
                       const double* alpha, int64_t ndof, int h,
                       hipStream_t st);

}  // namespace iegpu
