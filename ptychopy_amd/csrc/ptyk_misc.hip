// Frame codec, probe initialisation, simulator and small reductions.
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

// ------------------------------------------------------------- FP16 frame codec
// Per frame: scale = 2^ceil(log2(max/32768)) so that A/scale is exact in FP32 and the
// largest amplitude maps into (16384, 32768]; stored value = half(A/scale).
__global__ __launch_bounds__(PTYK_NT) void k_frame_scale(const float* __restrict__ src, float* __restrict__ scales,
                                                         long mm) {
  __shared__ float red[PTYK_NT / 64];
  const float* f = src + (long)blockIdx.x * mm;
  float mx = 0.f;
  for (long i = threadIdx.x; i < mm; i += PTYK_NT) mx = fmaxf(mx, fabsf(f[i]));
  mx = block_max<PTYK_NT>(mx, red);
  if (threadIdx.x == 0) {
    float sc = 1.f;
    if (mx > 0.f) {
      int e;
      frexpf(mx / 32768.f, &e);  // mx/32768 = m * 2^e, m in [0.5, 1)
      sc = ldexpf(1.f, e);
    }
    scales[blockIdx.x] = sc;
  }
}

__global__ void k_frame_quantize(const float* __restrict__ src, const float* __restrict__ scales,
                                 __half* __restrict__ dst, long mm, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) dst[i] = __float2half_rn(src[i] / scales[i / mm]);
}

__global__ void k_frame_dequantize(const __half* __restrict__ src, const float* __restrict__ scales,
                                   float* __restrict__ dst, long mm, long total) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) dst[i] = __half2float(src[i]) * scales[i / mm];
}

hipError_t quantize_frames(const float* src, __half* dst, float* scales, long n, int m, hipStream_t s) {
  const long mm = (long)m * m, total = n * mm;
  hipLaunchKernelGGL(k_frame_scale, dim3((unsigned)n), dim3(PTYK_NT), 0, s, src, scales, mm);
  hipLaunchKernelGGL(k_frame_quantize, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, scales, dst, mm,
                     total);
  return hipGetLastError();
}

hipError_t dequantize_frames(const __half* src, const float* scales, float* dst, long n, int m, hipStream_t s) {
  const long mm = (long)m * m, total = n * mm;
  hipLaunchKernelGGL(k_frame_dequantize, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, scales, dst, mm,
                     total);
  return hipGetLastError();
}

// Per-frame sum of amplitudes (Diffractions::fillSquaredSums, Diffractions.cpp:234-242).
__global__ __launch_bounds__(PTYK_NT) void k_frame_sums(const __half* __restrict__ amp, const float* __restrict__ sc,
                                                        float* __restrict__ sums, long mm) {
  __shared__ double red[PTYK_NT / 64];
  const __half* f = amp + (long)blockIdx.x * mm;
  double acc = 0.0;
  for (long i = threadIdx.x; i < mm; i += PTYK_NT) acc += (double)load_amp(f, i, sc[blockIdx.x]);
  acc = block_sum<PTYK_NT>(acc, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = (float)acc;
}

hipError_t frame_sums(const __half* amp, const float* scales, float* sums, long n, int m, hipStream_t s) {
  hipLaunchKernelGGL(k_frame_sums, dim3((unsigned)n), dim3(PTYK_NT), 0, s, amp, scales, sums, (long)m * m);
  return hipGetLastError();
}

// Mean amplitude over frames per detector pixel (d_normalizeDiffractionIntensities,
// ProbeKernels.cu:388-407). One thread per pixel, FP64 accumulation in frame order.
__global__ void k_mean_amp(const __half* __restrict__ amp, const float* __restrict__ sc, float* __restrict__ avg,
                           long n, long mm) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= mm) return;
  double acc = 0.0;
  for (long j = 0; j < n; ++j) acc += (double)load_amp(amp, j * mm + p, sc[j]);
  avg[p] = (float)(acc / (double)n);
}

hipError_t mean_amplitude(const __half* amp, const float* scales, float* avg, long n, int m, hipStream_t s) {
  const long mm = (long)m * m;
  hipLaunchKernelGGL(k_mean_amp, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, s, amp, scales, avg, n, mm);
  return hipGetLastError();
}

// ------------------------------------------------------------- probe initialisation
// P0 = |sinc(r * dx_s/beamSize) * exp(-r^2/w^2)|, r measured from (M/2, M/2)
// (d_initProbe / d_hypotOgrid / d_calculateGaussian, ProbeKernels.cu:52-63, 207-218).
__global__ void k_probe_sinc_gauss(float2* __restrict__ p, int m, float window, float factor) {
  const int i = blockIdx.x, j = threadIdx.x;
  if (j >= m) return;
  const float a = (float)((double)i - m * 0.5), b = (float)((double)j - m * 0.5);
  const float r = sqrtf(a * a + b * b);
  float v = r * factor;
  v = (v != 0.f) ? (float)((double)sinpif(v) / (3.141592653589793 * (double)v)) : 1.f;
  const float g = expf(powf(window, -2.f) * r * r * -1.f);
  p[i * m + j] = make_float2(fabsf(v * g), 0.f);
}

hipError_t probe_sinc_gauss(float2* probe, int m, float window, float factor, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_sinc_gauss, dim3(m), dim3(m < 64 ? 64 : m), 0, s, probe, m, window, factor);
  return hipGetLastError();
}

// Single-mode modulus projection in the Fourier domain: A exp(i arg F) / M^2 (d_adjustFFT).
__global__ void k_adjust_single(float2* __restrict__ f, const float* __restrict__ amp, long mm, float nrm) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= mm) return;
  f[p] = modulus_phase(f[p], amp[p], nrm);
}

hipError_t modulus_project_single(float2* f, const float* amp, int m, hipStream_t s) {
  const long mm = (long)m * m;
  hipLaunchKernelGGL(k_adjust_single, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, s, f, amp, mm,
                     1.f / (float)mm);
  return hipGetLastError();
}

// Probe::normalize (Probe.cpp:303-318): P <- P / max|P| over all modes, then |P|^2 and
// probeMax = max |P|^2. One workgroup (the probe is at most 5 x 512^2).
__global__ __launch_bounds__(1024) void k_probe_normalize(float2* __restrict__ p, float* __restrict__ pint, long count,
                                                          double* __restrict__ out_max) {
  __shared__ float red[16];
  float mx = 0.f;
  for (long i = threadIdx.x; i < count; i += 1024) {
    const float2 v = p[i];
    mx = fmaxf(mx, sqrtf(v.x * v.x + v.y * v.y));
  }
  mx = block_max<1024>(mx, red);
  const float f = (float)(1.0 / (double)mx);
  float m2 = 0.f;
  for (long i = threadIdx.x; i < count; i += 1024) {
    float2 v = p[i];
    v = make_float2(v.x * f, v.y * f);
    p[i] = v;
    const float a = sqrtf(v.x * v.x + v.y * v.y);
    pint[i] = a * a;
    m2 = fmaxf(m2, a * a);
  }
  m2 = block_max<1024>(m2, red);
  if (threadIdx.x == 0) *out_max = m2;
}

hipError_t probe_normalize(float2* probe, float* probe_int, int k_total, int m, double* scal_out, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_normalize, dim3(1), dim3(1024), 0, s, probe, probe_int, (long)k_total * m * m, scal_out);
  return hipGetLastError();
}

// Probe::initProbeModes (Probe.cpp:320-330, d_initModesFromBase ProbeKernels.cu:82-98):
// P_k = factor * P_0 for k >= 1, intensities |P_k|^2.
__global__ void k_modes_from_base(float2* __restrict__ p, float* __restrict__ pint, long mm, int k_total, float factor) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= mm) return;
  const float2 b = cscale(p[q], factor);
  const float a = sqrtf(b.x * b.x + b.y * b.y);
  for (int k = 1; k < k_total; ++k) {
    p[(long)k * mm + q] = b;
    pint[(long)k * mm + q] = a * a;
  }
}

hipError_t modes_from_base(float2* probe, float* probe_int, int k_total, int m, float factor, hipStream_t s) {
  const long mm = (long)m * m;
  hipLaunchKernelGGL(k_modes_from_base, dim3((unsigned)((mm + 255) / 256)), dim3(256), 0, s, probe, probe_int, mm,
                     k_total, factor);
  return hipGetLastError();
}

// Probe::orthogonalize (Probe.cpp:229-242) with h_projectUtoV (ProbeKernels.cu:509-515):
// for i >= 1, for j = i-1 .. 0:  P_i -= (<copy_i, P_j> / <P_j, P_j>) P_j, where
// <a, b> = sum conj(a) b and copy_i is mode i before the sweep. One workgroup, FP64 sums.
__global__ __launch_bounds__(1024) void k_orthogonalize(float2* __restrict__ p, float2* __restrict__ copy, int k_total,
                                                        long mm) {
  __shared__ double red[16];
  for (long q = threadIdx.x; q < (long)k_total * mm; q += 1024) copy[q] = p[q];
  __syncthreads();
  for (int i = 1; i < k_total; ++i) {
    for (int j = i - 1; j >= 0; --j) {
      double nx = 0.0, ny = 0.0, dd = 0.0;
      for (long q = threadIdx.x; q < mm; q += 1024) {
        const float2 u = p[(long)j * mm + q], v = copy[(long)i * mm + q];
        nx += (double)v.x * u.x + (double)v.y * u.y;  // conj(v) * u
        ny += (double)v.x * u.y - (double)v.y * u.x;
        dd += (double)u.x * u.x + (double)u.y * u.y;
      }
      nx = block_sum<1024>(nx, red);
      ny = block_sum<1024>(ny, red);
      dd = block_sum<1024>(dd, red);
      const float fx = (float)(nx / dd), fy = (float)(ny / dd);
      for (long q = threadIdx.x; q < mm; q += 1024) {
        const float2 u = p[(long)j * mm + q];
        const float2 t = make_float2(u.x * fx - u.y * fy, u.x * fy + u.y * fx);
        float2 w = p[(long)i * mm + q];
        p[(long)i * mm + q] = make_float2(w.x - t.x, w.y - t.y);
      }
      __syncthreads();
    }
  }
}

hipError_t orthogonalize_modes(float2* probe, float2* tmp, int k_total, int m, hipStream_t s) {
  hipLaunchKernelGGL(k_orthogonalize, dim3(1), dim3(1024), 0, s, probe, tmp, k_total, (long)m * m);
  return hipGetLastError();
}

// ------------------------------------------------------------- reductions
__global__ __launch_bounds__(1024) void k_sum_abs2(const float2* __restrict__ x, long count, double* __restrict__ out) {
  __shared__ double red[16];
  double acc = 0.0;
  for (long i = threadIdx.x; i < count; i += 1024) acc += (double)x[i].x * x[i].x + (double)x[i].y * x[i].y;
  acc = block_sum<1024>(acc, red);
  if (threadIdx.x == 0) *out = acc;
}
hipError_t sum_abs2(const float2* x, long count, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_abs2, dim3(1), dim3(1024), 0, s, x, count, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(1024) void k_max_abs2(const float2* __restrict__ x, long count, double* __restrict__ out) {
  __shared__ float red[16];
  float mx = 0.f;
  for (long i = threadIdx.x; i < count; i += 1024) {
    const float a = sqrtf(x[i].x * x[i].x + x[i].y * x[i].y);
    mx = fmaxf(mx, a * a);
  }
  mx = block_max<1024>(mx, red);
  if (threadIdx.x == 0) *out = mx;
}
hipError_t max_abs2(const float2* x, long count, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max_abs2, dim3(1), dim3(1024), 0, s, x, count, out);
  return hipGetLastError();
}

__global__ void k_intensities(const float2* __restrict__ x, float* __restrict__ out, long count) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const float a = sqrtf(x[i].x * x[i].x + x[i].y * x[i].y);
    out[i] = a * a;
  }
}
hipError_t intensities(const float2* x, float* out, long count, hipStream_t s) {
  hipLaunchKernelGGL(k_intensities, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, x, out, count);
  return hipGetLastError();
}

// ------------------------------------------------------------- simulator
// psi_n = bilinear(O, pos_n) * P0 (d_simulatePSI, SampleKernels.cu:113-122).
__device__ __forceinline__ float2 bilinear_sample(const float2* __restrict__ o, int rows, int cols, float u, float v,
                                                  int quantize) {
  const float xb = u - 0.5f, yb = v - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb);
  float al = xb - fx, be = yb - fy;
  if (quantize) {
    al = rintf(al * 256.f) * (1.f / 256.f);
    be = rintf(be * 256.f) * (1.f / 256.f);
  }
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = min(max(i0 + 1, 0), cols - 1), j1 = min(max(j0 + 1, 0), rows - 1);
  i0 = min(max(i0, 0), cols - 1);
  j0 = min(max(j0, 0), rows - 1);
  const float2 t00 = o[(long)j0 * cols + i0], t01 = o[(long)j0 * cols + i1];
  const float2 t10 = o[(long)j1 * cols + i0], t11 = o[(long)j1 * cols + i1];
  const float w00 = (1.f - al) * (1.f - be), w01 = al * (1.f - be), w10 = (1.f - al) * be, w11 = al * be;
  return make_float2(w00 * t00.x + w01 * t01.x + w10 * t10.x + w11 * t11.x,
                     w00 * t00.y + w01 * t01.y + w10 * t10.y + w11 * t11.y);
}

__global__ void k_sim_exit(const float2* __restrict__ obj, int rows, int cols, const float2* __restrict__ probe,
                           const float2* __restrict__ pos, float2* __restrict__ psi, int m, long total, int quantize) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long mm = (long)m * m;
  const long n = i / mm;
  const int px = (int)(i - n * mm);
  const int ii = px / m, jj = px % m;
  const float2 ps = pos[n];
  const float2 roi = bilinear_sample(obj, rows, cols, (float)jj + ps.y + 0.5f, (float)ii + ps.x + 0.5f, quantize);
  psi[i] = cmul(roi, probe[px]);
}

hipError_t simulate_exit_waves(const float2* obj, int rows, int cols, const float2* probe, const float2* pos,
                               float2* psi, long n, int m, int quantize, hipStream_t s) {
  const long total = n * m * m;
  hipLaunchKernelGGL(k_sim_exit, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, obj, rows, cols, probe, pos,
                     psi, m, total, quantize);
  return hipGetLastError();
}

__global__ void k_abs_frames(const float2* __restrict__ x, float* __restrict__ out, long count, int squared) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    const float a2 = x[i].x * x[i].x + x[i].y * x[i].y;
    out[i] = squared ? a2 : sqrtf(a2);
  }
}
hipError_t abs_frames(const float2* psi, float* out, long count, int squared, hipStream_t s) {
  hipLaunchKernelGGL(k_abs_frames, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, psi, out, count, squared);
  return hipGetLastError();
}

}  // namespace ptyk
