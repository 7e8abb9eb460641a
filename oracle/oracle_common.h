// Shared pieces of the FP64 oracle (test infrastructure only): complex type, a radix-2 FFT
// with cuFFT's sign and scaling conventions, and small helpers.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef std::complex<double> cd;

// ------------------------------------------------------------------ FFT
// Iterative radix-2 decimation-in-time with bit reversal; forward exp(-2 pi i kn/N),
// inverse unnormalised (cuFFT conventions, CudaFFTPlan.cpp:91-101).
struct Fft1 {
  int n = 0, logn = 0;
  std::vector<int> rev;
  std::vector<cd> w;  // exp(-2 pi i k / n), k < n/2
  explicit Fft1(int n_) : n(n_) {
    while ((1 << logn) < n) ++logn;
    rev.resize(n);
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < logn; ++b)
        if (i & (1 << b)) r |= 1 << (logn - 1 - b);
      rev[i] = r;
    }
    w.resize(n / 2 > 0 ? n / 2 : 1);
    for (int k = 0; k < n / 2; ++k) w[k] = std::polar(1.0, -2.0 * M_PI * k / n);
  }
  void run(cd* x, long stride, bool inv) const {
    std::vector<cd> a(n);
    for (int i = 0; i < n; ++i) a[rev[i]] = x[i * stride];
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len / 2, step = n / len;
      for (int s = 0; s < n; s += len)
        for (int k = 0; k < half; ++k) {
          cd t = w[k * step];
          if (inv) t = std::conj(t);
          const cd u = a[s + k], v = a[s + k + half] * t;
          a[s + k] = u + v;
          a[s + k + half] = u - v;
        }
    }
    for (int i = 0; i < n; ++i) x[i * stride] = a[i];
  }
};

inline void fft2(cd* f, int m, bool inv, const Fft1& p) {
  for (int r = 0; r < m; ++r) p.run(f + (long)r * m, 1, inv);
  for (int c = 0; c < m; ++c) p.run(f + c, m, inv);
}

// exp(i arg f) with arg 0 for f == 0 whatever the signs of the zeros.  The reference computes
// atan2f(f.y, f.x) on cuFFT output (utilitiesKernels.cu:626-631, SampleKernels.cu:441-444);
// for an all-zero input (first DM/ePIE iteration from O = 0) the signs of cuFFT's zeros, and
// so atan2's +-pi versus 0, are not specified -- this value is the one for +0 inputs.
inline cd unit_phase(cd f) {
  if (f.real() == 0.0 && f.imag() == 0.0) return cd(1.0, 0.0);
  return std::polar(1.0, std::atan2(f.imag(), f.real()));
}

inline cd clamp_phase(cd u) {  // SampleKernels.cu:209-215
  double arg = std::atan2(u.imag(), u.real());
  arg = arg <= 0 ? arg : 0;
  return std::polar(std::abs(u), arg);
}

// Linear-filtered, clamp-addressed, unnormalised texture fetch at (u, v)
// (SampleKernels.cu:167-178, 493-508).  quantize: weights rounded to 1/256.
inline cd tex_bilinear(const cd* o, int rows, int cols, float u, float v, int quantize) {
  const float xb = u - 0.5f, yb = v - 0.5f;
  const float fx = std::floor(xb), fy = std::floor(yb);
  double al = (double)(xb - fx), be = (double)(yb - fy);
  if (quantize) {
    al = std::nearbyint(al * 256.0) / 256.0;
    be = std::nearbyint(be * 256.0) / 256.0;
  }
  auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const int i0 = cl((int)fx, cols - 1), i1 = cl((int)fx + 1, cols - 1);
  const int j0 = cl((int)fy, rows - 1), j1 = cl((int)fy + 1, rows - 1);
  return (1 - al) * (1 - be) * o[(long)j0 * cols + i0] + al * (1 - be) * o[(long)j0 * cols + i1] +
         (1 - al) * be * o[(long)j1 * cols + i0] + al * be * o[(long)j1 * cols + i1];
}

inline cd* C(double* p) { return reinterpret_cast<cd*>(p); }
inline const cd* C(const double* p) { return reinterpret_cast<const cd*>(p); }

}  // namespace orc
