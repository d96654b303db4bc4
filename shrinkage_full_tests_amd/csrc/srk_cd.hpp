// srk_cd.hpp -- critical-dimension measurement kernels (see srk_cd.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace srk {

// Largest dynamic LDS a CD block may use, and whether a signal length N with
// upsampling factor L and sinc half-width A fits it.
size_t cd_max_lds_bytes();
bool cd_shape_ok(int N, int L, int A);

// CD of B row-major signals [B][N] at threshold thr; per signal the nearest
// crossings left of / at-or-right of the centre, in pixels, and their
// difference.  Returns a hipError_t.
int cd_measure_rows(const float* in, int B, int N, int L, int A, float thr, float* left_px,
                    float* right_px, float* cd_px, hipStream_t stream);

// Same for the centre row (center_row != 0) or centre column of
// E_crop + c R, E [B][N][N] and R [B][W][W] column-major, crop offset e.
int cd_measure_shrink(const float* E, const float* R, int N, int W, int e, int B, float c,
                      int center_row, int L, int A, float thr, float* left_px, float* right_px,
                      float* cd_px, hipStream_t stream);

// Polyphase table coeff [L][2A+1] (rows normalised to sum 1), and the batched
// upsampling out [B][N L] of row-major signals with a given table.
int cd_build_table(int L, int A, float* coeff, hipStream_t stream);
int cd_upsample(const float* in, const float* coeff, int B, int N, int L, int A, float* out,
                hipStream_t stream);

// Zero-crossing CD from the caller's table (coeff, or null to build it) and /
// or upsampled signals (up [B][N L], or null to upsample `in` on the fly).
int cd_measure_ex(const float* in, const float* coeff, const float* up, int B, int N, int L,
                  int A, float* left_px, float* right_px, float* cd_px, hipStream_t stream);

}  // namespace srk
