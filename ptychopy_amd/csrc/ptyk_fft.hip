// Batched power-of-two 2D FFT over n frames of m x m complex64, staged in LDS.
//
// Replaces cufftPlanMany(rank 2, {H,W}) C2C plans (CudaFFTPlan.cpp:91-101) for the
// untangled call sites: probe initialisation (utilities.cpp:138-154), the simulator
// (Diffractions.cpp:105-107) and the FFT entry point used by the tests.
// Pass 1: each workgroup transforms a batch of whole rows (contiguous) in LDS.
// Pass 2: each workgroup transforms strips of 16 columns (128 B per row segment).
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

template <int M>
struct FftCfg {
  static constexpr int RNB = plan_min_batch<M, PTYK_NT>();
  static constexpr int CNB = (plan_min_batch<M, PTYK_NT>() + 15) / 16 * 16;
};

template <int M, bool INV>
__global__ __launch_bounds__(PTYK_NT) void k_fft_rows(float2* __restrict__ data, long total_rows,
                                                      const float2* __restrict__ tw) {
  constexpr int NB = FftCfg<M>::RNB;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  const long row0 = (long)blockIdx.x * NB;
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const long g = row0 + e / M;
    lds[e] = g < total_rows ? data[g * M + (e % M)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<M, INV, NB, PTYK_NT, RowLayout<M, M>, false>(lds, tw, tid);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const long g = row0 + e / M;
    if (g < total_rows) data[g * M + (e % M)] = lds[e];
  }
}

template <int M, bool INV>
__global__ __launch_bounds__(PTYK_NT) void k_fft_cols(float2* __restrict__ data, long total_strips,
                                                      const float2* __restrict__ tw) {
  constexpr int NB = FftCfg<M>::CNB;
  constexpr int G = NB / 16;
  constexpr int SPF = M / 16;
  __shared__ float2 lds[NB * M];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    float2 v = make_float2(0.f, 0.f);
    if (s < total_strips) {
      const long f = s / SPF;
      const int col = (int)(s % SPF) * 16 + c;
      v = data[(f * M + i) * M + col];
    }
    lds[e] = v;
  }
  __syncthreads();
  fft_lds<M, INV, NB, PTYK_NT, StripLayout<M>, true>(lds, tw, tid);
  for (int e = tid; e < NB * M; e += PTYK_NT) {
    const int g = e / (M * 16), i = (e >> 4) % M, c = e & 15;
    const long s = (long)blockIdx.x * G + g;
    if (s < total_strips) {
      const long f = s / SPF;
      const int col = (int)(s % SPF) * 16 + c;
      data[(f * M + i) * M + col] = lds[e];
    }
  }
}

template <int M, bool INV>
static hipError_t fft2_impl(float2* data, long n, const float2* tw, hipStream_t s) {
  const long rows = n * M;
  const long strips = n * (M / 16);
  const long gr = (rows + FftCfg<M>::RNB - 1) / FftCfg<M>::RNB;
  const long gc = (strips + FftCfg<M>::CNB / 16 - 1) / (FftCfg<M>::CNB / 16);
  if (gr > 0x7fffffffL || gc > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_fft_rows<M, INV>), dim3((unsigned)gr), dim3(PTYK_NT), 0, s, data, rows, tw);
  hipLaunchKernelGGL((k_fft_cols<M, INV>), dim3((unsigned)gc), dim3(PTYK_NT), 0, s, data, strips, tw);
  return hipGetLastError();
}

hipError_t fft2_batch(float2* data, long n, int m, bool inv, const float2* tw, hipStream_t s) {
  if (n <= 0) return hipSuccess;
#define PTYK_FFT_CASE(MM) \
  case MM: return inv ? fft2_impl<MM, true>(data, n, tw, s) : fft2_impl<MM, false>(data, n, tw, s);
  switch (m) {
    PTYK_FFT_CASE(16)
    PTYK_FFT_CASE(32)
    PTYK_FFT_CASE(64)
    PTYK_FFT_CASE(128)
    PTYK_FFT_CASE(256)
    PTYK_FFT_CASE(512)
    default: return hipErrorInvalidValue;
  }
#undef PTYK_FFT_CASE
}

}  // namespace ptyk
