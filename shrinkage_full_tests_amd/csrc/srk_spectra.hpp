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

// Spectra handed over at size Ns (reference layout) -> the spatial kernel on
// the erosion box |x|, |y| <= e (FP64 inverse DFT restricted to the box) ->
// its spectra at size M (reference layout, FP32; skipped when gx_dst or
// gy_dst is null).  *ratio is set to a device double that receives the
// relative RMS of the Ns-periodic kernel outside the box (Parseval).  scratch
// holds resample_scratch_bytes(Ns_max, e, M) bytes, Ns <= Ns_max.
size_t resample_scratch_bytes(int Ns_max, int e, int M);
int resample_spectra_device(int Ns, int e, int M, const float2* gx_src, const float2* gy_src,
                            void* scratch, float2* gx_dst, float2* gy_dst, double** ratio,
                            int Ns_max, hipStream_t stream);

}  // namespace srk
