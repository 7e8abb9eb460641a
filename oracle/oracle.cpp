// FP64 CPU oracle of the reference's hot-path equations -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this library,
// and only as the checker / the timed CPU baseline.  The engine never calls it.
//
// Each function restates the reference algorithm in double precision, in the order the
// reference applies it (not in the fused form the GPU engine uses), citing the reference
// file:line it follows.  Complex arrays are interleaved (re, im) doubles, row-major.
//
// Pins (see tests/test_oracle_*.py): orc_fft2 against numpy.fft and closed forms;
// the modulus projection against its closed form and idempotence; ePIE/DM updates against
// fixed points (exact data => zero update) and hand-computed 1-frame cases; bilinear
// extraction against integer gathers and separable linear interpolation; the mesh against
// float32 closed forms.  Functions without such a check are marked "parity unpinned".
#include "oracle_common.h"

using namespace orc;

namespace {

// Fourier modulus constraint on the K spectra Phi_k at psi + k mm (PhaserUtil::applyModulusConstraint,
// utilities.cpp:138-154): K = 1, Phi' = A exp(i atan2 Phi) / M^2 (d_adjustFFT,
// utilitiesKernels.cu:619-653); K > 1, Phi'_k = Phi_k A / sqrt(sum_k |Phi_k|^2) / M^2
// (d_adjustModalFFT, utilitiesKernels.cu:655-682).
void modulus_spectra(cd* psi, const double* A, int k, int m) {
  const long mm = (long)m * m;
  const double nrm = 1.0 / (double)mm;
  for (long q = 0; q < mm; ++q) {
    if (k == 1) {
      psi[q] = A[q] * unit_phase(psi[q]) * nrm;
    } else {
      double s2 = 0;
      for (int kk = 0; kk < k; ++kk) s2 += std::norm(psi[(size_t)kk * mm + q]);
      const double f = A[q] / std::sqrt(s2) * nrm;
      for (int kk = 0; kk < k; ++kk) psi[(size_t)kk * mm + q] *= f;
    }
  }
}

}  // namespace

extern "C" {

// The modulus constraint on K exit waves (k x m x m, in place): FFT, modulus_spectra, unnormalised
// inverse FFT.
void orc_modulus_modes(double* psi, const double* amp, int k, int m) {
  Fft1 p(m);
  const long mm = (long)m * m;
  for (int kk = 0; kk < k; ++kk) fft2(C(psi) + kk * mm, m, false, p);
  modulus_spectra(C(psi), amp, k, m);
  for (int kk = 0; kk < k; ++kk) fft2(C(psi) + kk * mm, m, true, p);
}

// Probe::orthogonalize (Probe.cpp:228-242): for i = 1..K-1 and j = i-1 down to 0,
// P_i -= P_j (sum conj(v_i) P_j / sum |P_j|^2), with v_i a copy of mode i taken before the sweep
// (h_projectUtoV with h_innerProduct(a, b) = sum conj(a) b, ProbeKernels.cu:509-515 and
// utilitiesKernels.cu:854-867).  The coefficient conjugates the textbook projection
// sum conj(P_j) v_i; for real modes the two agree.
void orc_orthogonalize_modes(double* probe, int k, int m) {
  const long mm = (long)m * m;
  cd* P = C(probe);
  const std::vector<cd> copy(P, P + (long)k * mm);
  for (int i = 1; i < k; ++i)
    for (int j = i - 1; j >= 0; --j) {
      cd num = 0;
      double den = 0;
      for (long q = 0; q < mm; ++q) {
        num += std::conj(copy[(long)i * mm + q]) * P[(long)j * mm + q];
        den += std::norm(P[(long)j * mm + q]);
      }
      const cd f = num / den;
      for (long q = 0; q < mm; ++q) P[(long)i * mm + q] -= P[(long)j * mm + q] * f;
    }
}

// Probe::initProbeModes (Probe.cpp:320-330): modes 1..K-1 <- 0.05 P0 with intensities
// |0.05 P0|^2 (d_initModesFromBase, ProbeKernels.cu:82-97), then orthogonalize.
void orc_init_probe_modes(double* probe, double* probe_int, int k, int m) {
  const long mm = (long)m * m;
  cd* P = C(probe);
  for (int kk = 1; kk < k; ++kk)
    for (long q = 0; q < mm; ++q) {
      P[(long)kk * mm + q] = P[q] * 0.05;
      const double a = std::abs(P[(long)kk * mm + q]);
      probe_int[(long)kk * mm + q] = a * a;
    }
  orc_orthogonalize_modes(probe, k, m);
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Batched 2D FFT, n frames of m x m.
void orc_fft2(double* data, int64_t n, int m, int inverse) {
  Fft1 p(m);
#pragma omp parallel for schedule(dynamic)
  for (int64_t i = 0; i < n; ++i) fft2(C(data) + i * (long)m * m, m, inverse != 0, p);
}

// ROI extraction: bilinear fetch at (row + pos.x, col + pos.y) (d_extractObjectArray).
void orc_extract_bilinear(const double* obj, int rows, int cols, float px, float py, int m, int quantize,
                          double* out) {
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      C(out)[(long)i * m + j] =
          tex_bilinear(C(obj), rows, cols, (float)j + py + 0.5f, (float)i + px + 0.5f, quantize);
}

// Single-mode Fourier modulus projection (PhaserUtil::applyModulusConstraint with
// d_adjustFFT, utilities.cpp:138-154, utilitiesKernels.cu:633-653):
// psi <- IFFT_u( A exp(i atan2(Phi)) / M^2 ),  Phi = FFT(psi).
void orc_modulus_projection(double* psi, const double* amp, int m) {
  Fft1 p(m);
  cd* f = C(psi);
  fft2(f, m, false, p);
  const double nrm = 1.0 / ((double)m * m);
  for (long q = 0; q < (long)m * m; ++q) f[q] = amp[q] * unit_phase(f[q]) * nrm;
  fft2(f, m, true, p);
}

// Initial probe (Probe::init without a guess, Probe.cpp:360-372; d_initProbe
// ProbeKernels.cu:207-218; d_normalizeDiffractionIntensities ProbeKernels.cu:388-407;
// Probe::normalize Probe.cpp:303-318).  probe: k_total x m x m (modes >= 1 zeroed).
// Returns max |P|^2 after normalisation.
double orc_probe_init(const double* amp, int64_t n, int m, int k_total, double beam_size, double dx_s,
                      double* probe) {
  const long mm = (long)m * m;
  std::vector<double> avg(mm, 0.0);
  for (int64_t j = 0; j < n; ++j)
    for (long q = 0; q < mm; ++q) avg[q] += amp[j * mm + q];
  for (long q = 0; q < mm; ++q) avg[q] /= (double)n;
  const double window = (double)m / (2 * 2.35), factor = dx_s / beam_size;
  std::fill(probe, probe + 2 * mm * k_total, 0.0);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const double a = i - m * 0.5, b = j - m * 0.5;
      const double r = std::sqrt(a * a + b * b);
      double v = r * factor;
      v = v != 0 ? std::sin(M_PI * v) / (M_PI * v) : 1.0;
      const double g = std::exp(-r * r / (window * window));
      C(probe)[(long)i * m + j] = std::abs(v * g);
    }
  orc_modulus_projection(probe, avg.data(), m);
  double mx = 0;
  for (long q = 0; q < mm * k_total; ++q) mx = std::max(mx, std::abs(C(probe)[q]));
  double m2 = 0;
  for (long q = 0; q < mm * k_total; ++q) {
    C(probe)[q] /= mx;
    m2 = std::max(m2, std::norm(C(probe)[q]));
  }
  return m2;
}

// Simulator (Diffractions::simulate, Diffractions.cpp:80-116; d_simulatePSI
// SampleKernels.cu:113-122): out_j = |FFT2(bilinear(O, pos_j) * P)| (or squared).
void orc_simulate(const double* obj, int rows, int cols, const double* probe, const float* pos, int64_t n, int m,
                  int squared, int quantize, double* out) {
  Fft1 p(m);
  const long mm = (long)m * m;
#pragma omp parallel for schedule(dynamic)
  for (int64_t j = 0; j < n; ++j) {
    std::vector<cd> f(mm);
    for (int i = 0; i < m; ++i)
      for (int c = 0; c < m; ++c)
        f[(long)i * m + c] = tex_bilinear(C(obj), rows, cols, (float)c + pos[2 * j + 1] + 0.5f,
                                          (float)i + pos[2 * j] + 0.5f, quantize) *
                             C(probe)[(long)i * m + c];
    fft2(f.data(), m, false, p);
    for (long q = 0; q < mm; ++q) out[j * mm + q] = squared ? std::norm(f[q]) : std::abs(f[q]);
  }
}

// One ePIE iteration (ePIE::iteration, ePIE.cpp:65-129).
//  obj rows x cols, obj_int rows x cols (|O|^2 cache), probe / probe_int k_total x m x m,
//  amp n x m x m amplitudes, amp_sum n (sum of amplitudes per frame), pos n x (row, col)
//  float32, order: n indices.  scal: [0] objMax, [1] probeMax (in/out), [2] error sum (out).
//  k_use modes take part (K when modal reconstruction has begun, else 1); modal selects
//  d_adjustModalFFT and the modal probe maximum.
void orc_epie_iteration(double* obj, double* obj_int, int rows, int cols, double* probe, double* probe_int,
                        int k_total, int k_use, int modal, const double* amp, const double* amp_sum,
                        const float* pos, const int32_t* order, int64_t n, int m, int phase_constraint,
                        int update_probe, int rms, int quantize, double* scal) {
  Fft1 p(m);
  const long mm = (long)m * m;
  cd* O = C(obj);
  cd* P = C(probe);
  std::vector<cd> psi_old((size_t)k_use * mm), psi((size_t)k_use * mm);
  double obj_max = scal[0], probe_max = scal[1], err = 0.0;
  for (int64_t s = 0; s < n; ++s) {
    const int fi = order[s];
    const float px = pos[2 * fi], py = pos[2 * fi + 1];
    const double* A = amp + (long)fi * mm;
    // psi_old_k = ROI * P_k (CXMath::multiply with firstAElementOnly, CXMath.h:49-58)
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        const cd roi = tex_bilinear(O, rows, cols, (float)j + py + 0.5f, (float)i + px + 0.5f, quantize);
        for (int k = 0; k < k_use; ++k) psi_old[(size_t)k * mm + i * m + j] = roi * P[(size_t)k * mm + i * m + j];
      }
    psi = psi_old;
    for (int k = 0; k < k_use; ++k) fft2(psi.data() + (size_t)k * mm, m, false, p);
    modulus_spectra(psi.data(), A, k_use, m);
    for (int k = 0; k < k_use; ++k) fft2(psi.data() + (size_t)k * mm, m, true, p);
    // object update at the rounded position (d_updateObjectArray(M), SampleKernels.cu:244-294)
    const int rx = (int)std::round(px), ry = (int)std::round(py);  // roundf: half away from zero
    const double co = 1.0 / probe_max;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) {
        cd u = 0;
        for (int k = 0; k < k_use; ++k) {
          const size_t q = (size_t)k * mm + i * m + j;
          u += std::conj(P[q]) * (psi[q] - psi_old[q]);
        }
        const long oi = (long)(rx + i) * cols + (ry + j);
        cd o = O[oi] + u * co;
        const double mag = std::abs(o);
        if (phase_constraint) o = clamp_phase(o);
        O[oi] = o;
        obj_int[oi] = mag * mag;
      }
    obj_max = 0;
    for (long q = 0; q < (long)rows * cols; ++q) obj_max = std::max(obj_max, obj_int[q]);
    // probe update with the updated object (d_updateProbe(M), ProbeKernels.cu:425-458)
    if (update_probe) {
      const double cp = 1.0 / obj_max;
      for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
          const cd oc = std::conj(O[(long)(rx + i) * cols + (ry + j)]);
          for (int k = 0; k < k_use; ++k) {
            const size_t q = (size_t)k * mm + i * m + j;
            P[q] += oc * (psi[q] - psi_old[q]) * cp;
            probe_int[q] = std::norm(P[q]);
          }
        }
      probe_max = 0;
      if (modal) {
        for (long q = 0; q < mm; ++q) {
          double sm = 0;
          for (int k = 0; k < k_total; ++k) sm += probe_int[(size_t)k * mm + q];
          probe_max = std::max(probe_max, sm);
        }
      } else {
        for (long q = 0; q < mm * k_total; ++q) probe_max = std::max(probe_max, probe_int[q]);
      }
    }
    // Fourier error (PhaserUtil::calculateER + d_calculateER; utilities.cpp:188-195)
    if (rms) {
      for (int k = 0; k < k_use; ++k) fft2(psi_old.data() + (size_t)k * mm, m, false, p);
      double acc = 0;
      for (long q = 0; q < mm; ++q) {
        cd sm = 0;
        for (int k = 0; k < k_use; ++k) sm += psi_old[(size_t)k * mm + q];
        acc += std::fabs(A[q] - std::abs(sm));
      }
      err += acc / k_use / amp_sum[fi];
    }
  }
  scal[0] = obj_max;
  scal[1] = probe_max;
  scal[2] = err;
}

// One DM iteration (DM::iteration, DM.cu:80-128, with the callbacks d_calculatePsi,
// d_modulusConstraint and d_updateObject, SampleKernels.cu:418-488).  Object and probe
// reads use the values at the start of the iteration (the reference reads them while
// other threads add to them); all increments are summed.
//  psi n x m x m (in/out), probe m x m (mode 0), factors[2] = cached {1/objMax, 1/probeMax}
//  (in/out, DM.cu:46-73), sums[2] = {sum |O|^2, sum |P|^2} after the update (out).
void orc_dm_iteration(double* obj, int rows, int cols, double* probe, double* psi, const double* amp,
                      const float* pos, int64_t n, int m, int phase_constraint, int update_probe, double* factors,
                      double* sums) {
  Fft1 p(m);
  const long mm = (long)m * m;
  cd* O = C(obj);
  cd* P = C(probe);
  const std::vector<cd> O0(O, O + (long)rows * cols), P0(P, P + mm);
  const double s_o = factors[1], s_p = factors[0];
  const double nrm = 1.0 / (double)mm;
#pragma omp parallel for schedule(dynamic)
  for (int64_t j = 0; j < n; ++j) {
    cd* ps = C(psi) + j * mm;
    const int rx = (int)std::round(pos[2 * j]), ry = (int)std::round(pos[2 * j + 1]);
    std::vector<cd> f(mm), po(mm);
    for (int i = 0; i < m; ++i)
      for (int c = 0; c < m; ++c) {
        po[i * m + c] = P0[i * m + c] * O0[(long)(rx + i) * cols + ry + c];
        f[i * m + c] = 2.0 * po[i * m + c] - ps[i * m + c];
      }
    fft2(f.data(), m, false, p);
    for (long q = 0; q < mm; ++q)
      f[q] = amp[j * mm + q] * unit_phase(f[q]) * nrm;
    fft2(f.data(), m, true, p);
    for (long q = 0; q < mm; ++q) ps[q] = ps[q] + f[q] - 2.0 * po[q];
  }
  // increments, summed per object / probe pixel in frame order
  std::vector<cd> dO((long)rows * cols, 0.0), dP(mm, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    const cd* ps = C(psi) + j * mm;
    const int rx = (int)std::round(pos[2 * j]), ry = (int)std::round(pos[2 * j + 1]);
    for (int i = 0; i < m; ++i)
      for (int c = 0; c < m; ++c) {
        const long oi = (long)(rx + i) * cols + ry + c;
        cd u = std::conj(P0[i * m + c]) * ps[i * m + c] * s_o;
        if (phase_constraint) u = clamp_phase(u);
        dO[oi] += u;
        if (update_probe) dP[i * m + c] += std::conj(O0[oi]) * ps[i * m + c] * s_p;
      }
  }
  double so = 0, sp = 0;
  for (long q = 0; q < (long)rows * cols; ++q) {
    O[q] += dO[q];
    so += std::norm(O[q]);
  }
  for (long q = 0; q < mm; ++q) {
    P[q] += dP[q];
    sp += std::norm(P[q]);
  }
  sums[0] = so;
  sums[1] = sp;
  // DM::checkCache at the start of the next iteration (DM.cu:54-73)
  const double passed[2] = {1.0 / so, 1.0 / sp};
  for (int i = 0; i < 2; ++i)
    if (std::fabs(factors[i] - passed[i]) > 1e-3) factors[i] = passed[i];
}

// Normalised object error against a known object T over the illuminated region (the engine's
// metric; no reference counterpart): coverage c = sum_j sum_k |P_k|^2 at ROI j; pixels with
// c > frac max(c); gamma = <O,T>/<O,O>; e = ||gamma O - T|| / ||T||.  out = {e, Re g, Im g, pixels}.
void orc_object_error(const double* obj, const double* truth, const double* probe, int k, int m, const int32_t* roi,
                      int64_t n, int rows, int cols, double frac, double* out) {
  const long canvas = (long)rows * cols, mm = (long)m * m;
  std::vector<double> pw(mm, 0.0), cov(canvas, 0.0);
  for (int i = 0; i < k; ++i)
    for (long q = 0; q < mm; ++q) pw[q] += std::norm(C(probe)[i * mm + q]);
  for (int64_t j = 0; j < n; ++j)
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < m; ++c) cov[(long)(roi[2 * j] + r) * cols + roi[2 * j + 1] + c] += pw[(long)r * m + c];
  double mx = 0;
  for (long q = 0; q < canvas; ++q) mx = std::max(mx, cov[q]);
  const cd* O = C(obj);
  const cd* T = C(truth);
  double so = 0, st = 0, cnt = 0;
  cd ot = 0;
  for (long q = 0; q < canvas; ++q)
    if (cov[q] > frac * mx) {
      so += std::norm(O[q]);
      st += std::norm(T[q]);
      ot += std::conj(O[q]) * T[q];
      cnt += 1;
    }
  const cd g = so > 0 ? ot / so : cd(0);
  double sd = 0;
  for (long q = 0; q < canvas; ++q)
    if (cov[q] > frac * mx) sd += std::norm(g * O[q] - T[q]);
  out[0] = st > 0 ? std::sqrt(sd / st) : 0.0;
  out[1] = g.real();
  out[2] = g.imag();
  out[3] = cnt;
}

}  // extern "C"
