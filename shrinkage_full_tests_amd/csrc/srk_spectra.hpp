// srk_spectra.hpp -- device generation of the SRSRK kernel spectra (FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace srk {

// Device scratch needed for kernel width kw at tile size M.
size_t spectra_scratch_bytes(int kw, int M);

// Gx, Gy spectra [M][M/2+1] (reference layout, FP32) and optionally the
// spatial kernels [kw][kw] (FP32), computed in FP64 on the device; scratch
// holds spectra_scratch_bytes(kw, M) bytes.  Returns a hipError_t.
int spectra_srsrk_device(int kw, double sigma, double dx, double cw, int M, void* scratch,
                         float2* gx, float2* gy, float* kx32, float* ky32, hipStream_t stream);

}  // namespace srk
