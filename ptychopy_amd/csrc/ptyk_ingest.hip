// Diffraction-frame preprocessing of on-disk data (DiffractionLoader::run, DiffractionLoader.cpp:55-131)
// on the GPU: one gather per output pixel over a batch of raw frames composes
//   1. the square crop / zero pad of S x S centred at (center_row, center_col)
//      (d_symmetrizeDiffraction with its host offsets, DiffractionsKernels.cu:100-107); a frame that
//      is already S x S is taken as is, whatever the centre;
//   2. the 90-degree rotation out[i][j] = in[j][S-1-i] (d_rot90, utilitiesKernels.cu:841-851), which
//      the reference applies only when rotate_90_times % 4 == 1 (utilitiesKernels.cu:2031-2049);
//   3. the FFT shift out[i][j] = in[(i + S/2) % S][(j + S/2) % S] (h_shiftFFT: d_shiftX then
//      d_shiftY by round(S/2), utilitiesKernels.cu:737-781, 1945-1963);
//   4. optionally the threshold and square root of d_preprocessDiffractions
//      (DiffractionsKernels.cu:109-118): non-finite -> 0, v < t -> 0 else v - t, sqrt.  The reference's
//      load path leaves this step commented out (Diffractions.cpp:226-229).
// A crop window reaching past the raw frame reads zeros; the reference's copy instead writes past
// its S x S output when the centre is closer than S/2 to the first row or column but the window
// ends inside the frame (DiffractionLoader.cpp:93-106): here the window is clipped to S x S.
// Bound: HBM (4 B read + 4 B written per output pixel); index arithmetic only.
#include <hip/hip_runtime.h>

#include "ptyk_internal.h"

namespace ptyk {

namespace {

__global__ void k_ingest(const float* __restrict__ raw, float* __restrict__ out, long total, IngestArgs a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int S = a.size;
  const long ss = (long)S * S;
  const long b = i / ss;
  const int p = (int)(i - b * ss);
  int r = p / S, c = p % S;
  if (a.fft_shift) {
    r = (r + S / 2) % S;
    c = (c + S / 2) % S;
  }
  if (a.rotate) {
    const int t = r;
    r = c;
    c = S - 1 - t;
  }
  int rr = r, cc = c;
  if (!(a.raw_rows == S && a.raw_cols == S)) {
    rr = r + a.center_row - S / 2;
    cc = c + a.center_col - S / 2;
  }
  float v = 0.f;
  if (rr >= 0 && rr < a.raw_rows && cc >= 0 && cc < a.raw_cols)
    v = raw[b * (long)a.raw_rows * a.raw_cols + (long)rr * a.raw_cols + cc];
  if (a.threshold_sqrt) {
    if (!isfinite(v)) v = 0.f;
    v = v < a.threshold ? 0.f : v - a.threshold;
    // the correctly rounded float square root: sqrt in FP64 (IEEE) then one rounding to FP32, which
    // cannot land on a tie for a square root (53 >= 2 x 24 + 2 bits)
    if (a.square_root) v = (float)sqrt((double)v);
  }
  out[i] = v;
}

}  // namespace

hipError_t ingest_frames(const float* raw, float* out, long count, const IngestArgs& a, hipStream_t s) {
  const long total = count * (long)a.size * a.size;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ingest, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, raw, out, total, a);
  return hipGetLastError();
}

}  // namespace ptyk
