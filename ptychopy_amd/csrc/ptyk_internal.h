// Internal host-side declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

namespace ptyk {

// Supported transform lengths (square frames M x M).
inline bool frame_size_supported(int m) {
  return m == 16 || m == 32 || m == 64 || m == 128 || m == 256 || m == 512;
}

// ---- batched 2D FFT of n frames of m x m complex64, in place (rows then columns).
hipError_t fft2_batch(float2* data, long n_frames, int m, bool inverse, const float2* tw, hipStream_t s);

// ---- DM (difference map) per-iteration launches; see ptyk_dm.hip for the equations.
struct DmArgs {
  float2* psi;            // N x M x M exit-wave state
  float2* obj;            // object canvas rows x cols
  float2* probe;          // mode 0, M x M
  const __half* amp;      // N x M x M amplitudes (FP16)
  const float* amp_scale; // N per-frame scales
  const int2* roi;        // N rounded ROI origins (row, col)
  const int* row_ptr;     // obj_rows + 1 : CSR of frames covering each object row
  const int* row_frames;  // N*M frame ids, ascending per object row
  double2* probe_partial; // probe_chunks x M x M
  double* obj_partial;    // obj_rows * obj_col_chunks
  double* probe_partial_sum;  // M (one per probe row block)
  float* factors;         // device: [0] = 1/objMax cached, [1] = 1/probeMax cached
  double* scal;           // device: [0] objSum, [1] probeSum
  const float2* tw;
  long n_frames;
  int m, obj_rows, obj_cols;
  int probe_chunks, frames_per_chunk;
  int phase_constraint, update_probe;
};
hipError_t dm_iteration(const DmArgs& a, hipStream_t s, hipEvent_t* ev /* nullable: 8 events for per-kernel timing */);
int dm_obj_col_chunks(int obj_cols);

// ---- ePIE: one persistent workgroup runs all positions of one iteration in order.
struct EpieArgs {
  float2* obj; float* obj_int;          // canvas and |O|^2 cache
  float2* probe; float* probe_int;      // K x M x M
  float2* psi_old; float2* psi;         // K x M x M scratch
  const __half* amp; const float* amp_scale;
  const float* amp_sum;                 // per frame sum of amplitudes (RMS normalisation)
  const float2* pos;                    // float positions (row, col)
  const int* order;                     // position order for this iteration
  double* scal;                         // [0] objMax, [1] probeMax, [2] error accumulator
  const float2* tw;
  int n_pos, m, k_use, k_total, obj_rows, obj_cols;
  int phase_constraint, update_probe, modal, rms, tex_quantize;
};
hipError_t epie_iteration(const EpieArgs& a, hipStream_t s);

// ---- misc kernels (ptyk_misc.hip)
hipError_t quantize_frames(const float* src_dev, __half* dst, float* scales, long n, int m, hipStream_t s);
hipError_t dequantize_frames(const __half* src, const float* scales, float* dst, long n, int m, hipStream_t s);
hipError_t frame_sums(const __half* amp, const float* scales, float* sums, long n, int m, hipStream_t s);
hipError_t mean_amplitude(const __half* amp, const float* scales, float* avg, long n, int m, hipStream_t s);
hipError_t probe_sinc_gauss(float2* probe, int m, float window, float factor, hipStream_t s);
hipError_t modulus_project_single(float2* f, const float* amp, int m, hipStream_t s);
hipError_t probe_normalize(float2* probe, float* probe_int, int k_total, int m, double* scal_out, hipStream_t s);
hipError_t sum_abs2(const float2* x, long count, double* out, hipStream_t s);
hipError_t max_abs2(const float2* x, long count, double* out, hipStream_t s);
hipError_t intensities(const float2* x, float* out, long count, hipStream_t s);
hipError_t simulate_exit_waves(const float2* obj, int obj_rows, int obj_cols, const float2* probe, const float2* pos,
                               float2* psi, long n, int m, int tex_quantize, hipStream_t s);
hipError_t abs_frames(const float2* psi, float* out, long count, int squared, hipStream_t s);
hipError_t modes_from_base(float2* probe, float* probe_int, int k_total, int m, float factor, hipStream_t s);
hipError_t orthogonalize_modes(float2* probe, float2* tmp, int k_total, int m, hipStream_t s);

}  // namespace ptyk
