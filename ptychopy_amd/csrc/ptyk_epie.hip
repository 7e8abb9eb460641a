// ePIE iteration on gfx950 (replaces ePIE.cpp:65-129 and the kernels it drives:
// d_extractObjectArray (SampleKernels.cu:167-178), d_complexMultiply, the cuFFT pair of
// applyModulusConstraint (utilities.cpp:138-154) with d_adjustFFT / d_adjustModalFFT
// (utilitiesKernels.cu:633-682), d_updateObjectArray(M) (SampleKernels.cu:244-294),
// d_updateProbe(M) (ProbeKernels.cu:425-458), the thrust maxima behind
// Sample/Probe::updateMaxIntensity and d_calculateER (utilitiesKernels.cu:684-712)).
//
// ePIE is a chain of dependent position updates, so one persistent workgroup (16 waves on
// one CU) walks every position of an iteration in the given order; the exit waves live in
// an L2-resident scratch and each 2D transform is staged through LDS in row and column
// chunks. There is one launch per iteration and no host round trip per position.
#include "ptyk_device.h"
#include "ptyk_internal.h"

namespace ptyk {

#define PTYK_EPIE_NT 1024

template <int M, int NT>
struct WgFft {
  static constexpr int RNB = plan_min_batch<M, NT>();
  static constexpr int CNB = (plan_min_batch<M, NT>() + 15) / 16 * 16;
  static constexpr int LDS = (RNB > CNB ? RNB : CNB) * M;  // float2 elements
};

// In-place 2D transform of one M x M frame held in global memory by one workgroup.
template <int M, bool INV, int NT>
__device__ void wg_fft2(float2* __restrict__ g, float2* lds, const float2* __restrict__ tw) {
  constexpr int RNB = WgFft<M, NT>::RNB;
  constexpr int CNB = WgFft<M, NT>::CNB;
  const int tid = threadIdx.x;
  for (int r0 = 0; r0 < M; r0 += RNB) {
    for (int e = tid; e < RNB * M; e += NT) {
      const int row = r0 + e / M;
      lds[e] = row < M ? g[row * M + (e % M)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<M, INV, RNB, NT, RowLayout<M, M>, false>(lds, tw, tid);
    for (int e = tid; e < RNB * M; e += NT) {
      const int row = r0 + e / M;
      if (row < M) g[row * M + (e % M)] = lds[e];
    }
    __syncthreads();
  }
  for (int c0 = 0; c0 < M; c0 += CNB) {
    for (int e = tid; e < CNB * M; e += NT) {
      const int col = c0 + e % CNB;
      lds[e] = col < M ? g[(e / CNB) * M + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<M, INV, CNB, NT, ColLayout<CNB>, true>(lds, tw, tid);
    for (int e = tid; e < CNB * M; e += NT) {
      const int col = c0 + e % CNB;
      if (col < M) g[(e / CNB) * M + col] = lds[e];
    }
    __syncthreads();
  }
}

// Bilinear sample with clamp addressing, matching a linear-filtered, unnormalised texture
// fetch at (u, v) = (col + offY + 0.5, row + offX + 0.5) (SampleKernels.cu:167-178, 493-508).
// quantize: round the weights to 1/256 (8 fractional bits, as texture units store them).
__device__ __forceinline__ float2 bilinear(const float2* __restrict__ o, int rows, int cols, float u, float v,
                                           int quantize) {
  const float xb = u - 0.5f, yb = v - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb);
  float al = xb - fx, be = yb - fy;
  if (quantize) {
    al = rintf(al * 256.f) * (1.f / 256.f);
    be = rintf(be * 256.f) * (1.f / 256.f);
  }
  int i0 = (int)fx, j0 = (int)fy;
  int i1 = i0 + 1, j1 = j0 + 1;
  i0 = min(max(i0, 0), cols - 1);
  i1 = min(max(i1, 0), cols - 1);
  j0 = min(max(j0, 0), rows - 1);
  j1 = min(max(j1, 0), rows - 1);
  const float2 t00 = o[(long)j0 * cols + i0], t01 = o[(long)j0 * cols + i1];
  const float2 t10 = o[(long)j1 * cols + i0], t11 = o[(long)j1 * cols + i1];
  const float w00 = (1.f - al) * (1.f - be), w01 = al * (1.f - be), w10 = (1.f - al) * be, w11 = al * be;
  return make_float2(w00 * t00.x + w01 * t01.x + w10 * t10.x + w11 * t11.x,
                     w00 * t00.y + w01 * t01.y + w10 * t10.y + w11 * t11.y);
}

template <int M>
__global__ __launch_bounds__(PTYK_EPIE_NT) void k_epie(EpieArgs a) {
  extern __shared__ float2 lds[];
  __shared__ float redf[PTYK_EPIE_NT / 64];
  __shared__ double redd[PTYK_EPIE_NT / 64];
  constexpr int NT = PTYK_EPIE_NT;
  constexpr int MM = M * M;
  const int tid = threadIdx.x;
  const int ku = a.k_use;
  // The reference keeps these maxima as float (real_t); factors are 1.0/max rounded to float.
  float obj_max = (float)a.scal[0];
  float probe_max = (float)a.scal[1];
  double err = 0.0;
  const long canvas = (long)a.obj_rows * a.obj_cols;

  for (int s = 0; s < a.n_pos; ++s) {
    const int fi = a.order[s];
    const float2 pos = a.pos[fi];
    const float ascale = a.amp_scale[fi];
    const __half* amp = a.amp + (long)fi * MM;

    // psi_old_k = ROI * P_k  (bilinear ROI at the fractional position)
    for (int px = tid; px < MM; px += NT) {
      const int ii = px / M, jj = px % M;
      const float2 roi = bilinear(a.obj, a.obj_rows, a.obj_cols, (float)jj + pos.y + 0.5f, (float)ii + pos.x + 0.5f,
                                  a.tex_quantize);
      for (int k = 0; k < ku; ++k) {
        const float2 v = cmul(roi, a.probe[(long)k * MM + px]);
        a.psi_old[(long)k * MM + px] = v;
        a.psi[(long)k * MM + px] = v;
      }
    }
    __syncthreads();
    for (int k = 0; k < ku; ++k) wg_fft2<M, false, NT>(a.psi + (long)k * MM, lds, a.tw);

    // Fourier modulus constraint (d_adjustFFT / d_adjustModalFFT)
    const float nrm = 1.f / (float)MM;
    for (int px = tid; px < MM; px += NT) {
      const float A = load_amp(amp, px, ascale);
      if (ku == 1) {
        a.psi[px] = modulus_phase(a.psi[px], A, nrm);
      } else {
        float ms = 0.f;
        for (int k = 0; k < ku; ++k) ms += cabs2(a.psi[(long)k * MM + px]);
        const float f = A * rsqrtf(ms) * nrm;
        for (int k = 0; k < ku; ++k) a.psi[(long)k * MM + px] = cscale(a.psi[(long)k * MM + px], f);
      }
    }
    __syncthreads();
    for (int k = 0; k < ku; ++k) wg_fft2<M, true, NT>(a.psi + (long)k * MM, lds, a.tw);

    // Object update at the rounded position (d_updateObjectArray(M))
    const int rx = (int)roundf(pos.x), ry = (int)roundf(pos.y);
    const float co = (float)(1.0 / (double)probe_max);
    for (int px = tid; px < MM; px += NT) {
      const int ii = px / M, jj = px % M;
      float2 u = make_float2(0.f, 0.f);
      for (int k = 0; k < ku; ++k) {
        const long q = (long)k * MM + px;
        u = cadd(u, cmulcj(a.probe[q], csub(a.psi[q], a.psi_old[q])));
      }
      const long oi = (long)(rx + ii) * a.obj_cols + (ry + jj);
      float2 o = cadd(a.obj[oi], cscale(u, co));
      const float mag = sqrtf(o.x * o.x + o.y * o.y);
      if (a.phase_constraint) o = clamp_phase(o);
      a.obj[oi] = o;
      a.obj_int[oi] = mag * mag;
    }
    __syncthreads();
    {
      float mx = 0.f;
      for (long q = tid; q < canvas; q += NT) mx = fmaxf(mx, a.obj_int[q]);
      obj_max = block_max<NT>(mx, redf);
    }

    // Probe update with the updated object (d_updateProbe(M)); Probe::updateMaxIntensity
    if (a.update_probe) {
      const float cp = (float)(1.0 / (double)obj_max);
      for (int px = tid; px < MM; px += NT) {
        const int ii = px / M, jj = px % M;
        const float2 oc = a.obj[(long)(rx + ii) * a.obj_cols + (ry + jj)];
        for (int k = 0; k < ku; ++k) {
          const long q = (long)k * MM + px;
          float2 p = cadd(cscale(cmulcj(oc, csub(a.psi[q], a.psi_old[q])), cp), a.probe[q]);
          a.probe[q] = p;
          const float ia = sqrtf(p.x * p.x + p.y * p.y);
          a.probe_int[q] = ia * ia;
        }
      }
      __syncthreads();
      float mx = 0.f;
      if (a.modal) {
        for (int px = tid; px < MM; px += NT) {
          float sm = 0.f;
          for (int k = 0; k < a.k_total; ++k) sm += a.probe_int[(long)k * MM + px];
          mx = fmaxf(mx, sm);
        }
      } else {
        for (long q = tid; q < (long)a.k_total * MM; q += NT) mx = fmaxf(mx, a.probe_int[q]);
      }
      probe_max = block_max<NT>(mx, redf);
    }

    // Optional Fourier error (PhaserUtil::calculateER; ePIE.cpp:124-125)
    if (a.rms) {
      for (int k = 0; k < ku; ++k) wg_fft2<M, false, NT>(a.psi_old + (long)k * MM, lds, a.tw);
      double acc = 0.0;
      for (int px = tid; px < MM; px += NT) {
        float2 sm = make_float2(0.f, 0.f);
        for (int k = 0; k < ku; ++k) sm = cadd(sm, a.psi_old[(long)k * MM + px]);
        acc += fabs((double)load_amp(amp, px, ascale) - sqrt((double)sm.x * sm.x + (double)sm.y * sm.y));
      }
      acc = block_sum<NT>(acc, redd);
      err += acc / ku / (double)a.amp_sum[fi];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.scal[0] = obj_max;
    a.scal[1] = probe_max;
    a.scal[2] = err;
  }
}

template <int M>
static hipError_t epie_impl(const EpieArgs& a, hipStream_t s) {
  const size_t lds = (size_t)WgFft<M, PTYK_EPIE_NT>::LDS * sizeof(float2);
  hipError_t e = hipFuncSetAttribute((const void*)k_epie<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_epie<M>), dim3(1), dim3(PTYK_EPIE_NT), lds, s, a);
  return hipGetLastError();
}

hipError_t epie_iteration(const EpieArgs& a, hipStream_t s) {
  switch (a.m) {
    case 16: return epie_impl<16>(a, s);
    case 32: return epie_impl<32>(a, s);
    case 64: return epie_impl<64>(a, s);
    case 128: return epie_impl<128>(a, s);
    case 256: return epie_impl<256>(a, s);
    case 512: return epie_impl<512>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ptyk
