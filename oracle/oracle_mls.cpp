// FP64 CPU oracle of the reference's MLs (maximum-likelihood, least-squares step) solver --
// TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker, never by the engine.
//
// Restates, in double precision and in the reference's order of operations:
//   MLS::iteration                      MLS.cpp:74-490
//   IPtychoScanMesh::find_reconstruction_ROI, updateScanPositionsDiff
//                                       IPtychoScanMesh.cpp:123-172, 340-369
//   Probe::initEvo, remove_extra_degree, calc_object_norm, updated_cached_illumination,
//         get_illumination_probe, shift_probe, update_probe, update_variable_probe,
//         gradient_position_solver, ortho_modes
//                                       Probe.cpp:507-729, 887-1118, 1261-1517, 244-294
//   Diffractions::get_fourier_error, modulus_constraint
//                                       Diffractions.cpp:261-310
//   Sample::update_object               Sample.cpp:277-331
//   element kernels                     utilitiesKernels.cu:171-189, 349-466, 973-987
//
// Arithmetic that decides integer ROI origins (positions, their mean correction, the rounding)
// is kept in float exactly as the reference does it, so ROI origins are bit-exact; everything
// else is FP64.  Atomic scatters are summed in frame order.
//
// Where this restatement departs from the reference on purpose (each is also a departure of the
// engine):
//  * a group of one frame gathers at the same origin as larger groups (the reference's host
//    gather for such groups reads one pixel up and left of its GPU gather and scatter,
//    Probe.cpp:792-797 vs utilitiesKernels.cu:1064-1084);
//  * with variable_probe == variable_intensity == 0 the illumination probe is P0 (the reference
//    leaves that buffer unset, Probe.cpp:896-919);
//  * the LSQ 2x2 system is solved in FP64 by Cramer's rule (the reference uses Eigen's
//    colPivHouseholderQr in complex float, MLS.cpp:384-398; same solution for a regular system).
//
// Pins (tests/test_oracle_mls.py): shift by an integer number of pixels equals np.roll; the
// position-solver derivative of a plane wave equals its closed form; a consistent data set
// (A = |FFT(P O)|) is a fixed point of the whole iteration; LSQ on hand-built 2x2 systems;
// ortho_modes output orthogonal with powers in the documented order; ROI origins against a
// float32 numpy restatement; initial normalisation against closed forms.
#include "oracle_common.h"

using namespace orc;

namespace {

struct MlsPrm {  // layout mirrored by oracle/__init__.py MlsParams
  int32_t iter, pps, probe_reconstruct, object_reconstruct;
  int32_t variable_probe, variable_intensity, variable_probe_modes, apply_subpix_shift;
  int32_t apply_multimodal_update, n_objects, n_rec, reserved;
  int32_t box[4];  // calc_object_norm box, 1-based inclusive: x0, x1, y0, y1
  double delta_p, beta_lsq, beta_object, beta_probe;
};

// fftshifted frequency grid (i/M - 0.5 shifted by M/2, Probe.cpp:92-123 with d_shiftY)
inline double freq(int i, int m) { return (double)((i + m / 2) % m) / m - 0.5; }

// Probe::shift_probe (Probe.cpp:1261-1337): x <- IFFT(FFT(x) e^{-2 pi i (s.x f[col] + s.y f[row])}) / M^2
void shift_frame(cd* x, int m, float sx, float sy, const Fft1& p) {
  fft2(x, m, false, p);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      const double ph = -2.0 * M_PI * ((double)sx * freq(c, m) + (double)sy * freq(r, m));
      x[(long)r * m + c] *= std::polar(1.0, ph);
    }
  fft2(x, m, true, p);
  const double s = 1.0 / ((double)m * m);
  for (long q = 0; q < (long)m * m; ++q) x[q] *= s;
}

// Float ROI arithmetic of find_reconstruction_ROI (IPtychoScanMesh.cpp:123-172).
void roi_of(const float* pos, int n, int m, int rows, int cols, int32_t* roi, float* sub) {
  const float cx = (float)std::ceil(rows * 1.0 / 2 - m * 1.0 / 2);
  const float cy = (float)std::ceil(cols * 1.0 / 2 - m * 1.0 / 2);
  for (int i = 0; i < n; ++i) {
    const float p1 = pos[2 * i] + cx, p2 = pos[2 * i + 1] + cy;
    const unsigned tx = (unsigned)std::round(p1), ty = (unsigned)std::round(p2);
    sub[2 * i] = p2 - (float)ty;  // note the order: .x is the column fraction
    sub[2 * i + 1] = p1 - (float)tx;
    roi[2 * i] = (int32_t)tx - 1;
    roi[2 * i + 1] = (int32_t)ty - 1;
  }
}

inline double sgn(double v) { return (double)((0.0 < v) - (v < 0.0)); }

// Gradient position solver for one frame (Probe.cpp:956-1118): dX = IFFT(FFT(O) 2 pi i f[col])/M^2,
// dY = IFFT(FFT(O) 2 pi i f[row])/M^2, step = sum Re(chi conj(d vp)) / sum |d vp|^2 per axis.
void position_step(const cd* op, const cd* vp, const cd* chi, int m, const Fft1& p, double* dx, double* dy,
                   cd* dX_out = nullptr, cd* dY_out = nullptr) {
  const long mm = (long)m * m;
  std::vector<cd> F(op, op + mm), dX(mm), dY(mm);
  fft2(F.data(), m, false, p);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      const cd t = F[(long)r * m + c] * cd(0.0, 2.0 * M_PI);
      dX[(long)r * m + c] = t * freq(c, m);
      dY[(long)r * m + c] = t * freq(r, m);
    }
  fft2(dX.data(), m, true, p);
  fft2(dY.data(), m, true, p);
  const double nrm = 1.0 / (double)mm;
  double nx = 0, dxs = 0, ny = 0, dys = 0;
  for (long q = 0; q < mm; ++q) {
    dX[q] *= nrm;
    dY[q] *= nrm;
    if (dX_out) dX_out[q] = dX[q];
    if (dY_out) dY_out[q] = dY[q];
    if (!vp) continue;
    const cd ax = dX[q] * vp[q], ay = dY[q] * vp[q];
    dxs += std::norm(ax);
    nx += (chi[q] * std::conj(ax)).real();
    dys += std::norm(ay);
    ny += (chi[q] * std::conj(ay)).real();
  }
  *dx = vp ? nx / dxs : 0.0;
  *dy = vp ? ny / dys : 0.0;
}

// Optimal object/probe steps of one group (d_get_optimal_step_lsq, utilitiesKernels.cu:445-466,
// then MLS.cpp:370-398): per frame j with dOP = dO P and dPO = dP O,
//   [[sum|dOP|^2 + l + 0.1 mean_j(..), sum dOP conj(dPO)], [conj(.), sum|dPO|^2 + l + 0.1 mean_j(..)]] x
//     = [sum Re(conj(dOP) chi), sum Re(conj(dPO) chi)],   l = M^2 * (float)(FLT_EPSILON / M^2).
// x: G x 2 (object step, probe step).
void lsq_steps(const cd* oup, const cd* vp, const cd* pum, const cd* op, const cd* chi, int G, int m, cd* x) {
  const long mm = (long)m * m;
  const double lam = (float)(1.1920928955078125e-07 / mm);
  std::vector<double> a1(G), a4(G), b1(G), b2(G);
  std::vector<cd> a2(G);
  for (int j = 0; j < G; ++j) {
    double s1 = 0, s4 = 0, t1 = 0, t2 = 0;
    cd s2 = 0;
    for (long q = 0; q < mm; ++q) {
      const cd dop = oup[j * mm + q] * vp[j * mm + q];
      const cd dpo = pum[q] * op[j * mm + q];
      s1 += std::norm(dop);
      s2 += dop * std::conj(dpo);
      s4 += std::norm(dpo);
      t1 += (std::conj(dop) * chi[j * mm + q]).real();
      t2 += (std::conj(dpo) * chi[j * mm + q]).real();
    }
    a1[j] = s1 + lam * mm;
    a2[j] = s2;
    a4[j] = s4 + lam * mm;
    b1[j] = t1;
    b2[j] = t2;
  }
  double m1 = 0, m4 = 0;
  for (int j = 0; j < G; ++j) {
    m1 += a1[j];
    m4 += a4[j];
  }
  m1 = 0.1 * m1 / G;
  m4 = 0.1 * m4 / G;
  for (int j = 0; j < G; ++j) {
    const cd A11 = a1[j] + m1, A12 = a2[j], A21 = std::conj(a2[j]), A22 = a4[j] + m4;
    const cd det = A11 * A22 - A12 * A21;
    x[2 * j] = (b1[j] * A22 - A12 * b2[j]) / det;
    x[2 * j + 1] = (A11 * b2[j] - A21 * b1[j]) / det;
  }
}


// Diffractions::get_fourier_error and modulus_constraint for one frame (Diffractions.cpp:261-310;
// d_modulus_amplitude, DiffractionsKernels.cu:127-137).  psi + kk * stride holds the spectrum of
// mode kk; ap = sqrt(sum_k |Psi_k|^2); the error is sqrt(mean((A - ap)^2 / 0.25)) (the real
// h_normalize divides by its factor, utilitiesKernels.cu:1672-1678); chi_k = IFFT_u((A / (ap +
// 1e-9) - 1) Psi_k) / M^2 replaces each spectrum.  Returns the error.
double frame_modulus(cd* psi, long stride, const double* A, int k, int m, const Fft1& fp) {
  const long mm = (long)m * m;
  double e2 = 0;
  for (long q = 0; q < mm; ++q) {
    double s2 = 0;
    for (int kk = 0; kk < k; ++kk) s2 += std::norm(psi[kk * stride + q]);
    const double ap = std::sqrt(s2);
    e2 += (A[q] - ap) * (A[q] - ap) / 0.25;
    const double R = A[q] / (ap + 1e-9) - 1.0;
    for (int kk = 0; kk < k; ++kk) psi[kk * stride + q] *= R;
  }
  const double nrm = 1.0 / (double)mm;
  for (int kk = 0; kk < k; ++kk) {
    cd* ps = psi + kk * stride;
    fft2(ps, m, true, fp);
    for (long q = 0; q < mm; ++q) ps[q] *= nrm;
  }
  return std::sqrt(e2 / mm);
}

// Probe::update_variable_probe, variable-mode part (Probe.cpp:1366-1477): with w = p / max(p) at
// each frame's ROI, img_j = w (pu_j - pum), c_j = (mean Re(V conj img_j) + e1_j) / sum_j e1_j^2,
// U = sum_j c_j img_j / M^2, V <- V + U relax_U / max(1, rms U), relax_U = G / N, V <- V / rms V;
// then e1_j += mean Re(chi_j conj(op_j V)) / (mean |op_j V|^2 + 0.1 mean_j mean |op_j V|^2).
void variable_probe_update(int G, int n, int m, const int32_t* gi, const int32_t* roi, const double* illum, int rows,
                           int cols, const cd* pu, const cd* pum, const cd* op, const cd* chi0, cd* V, double* evo1) {
  const long mm = (long)m * m;
  const double nrm = 1.0 / (double)mm;
  const double relax_u = (double)G / n;
  double evn = 0;
  for (int j = 0; j < G; ++j) evn += evo1[gi[j]] * evo1[gi[j]];
  double pmax = 0;
  for (long q = 0; q < (long)rows * cols; ++q) pmax = std::max(pmax, illum[q]);
  std::vector<cd> img((size_t)G * mm), acc(mm, cd(0));
  for (int j = 0; j < G; ++j) {
    const int f = gi[j];
    double sre = 0;
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < m; ++c) {
        const long q = (long)r * m + c;
        const double w = illum[(long)(roi[2 * f] + r) * cols + roi[2 * f + 1] + c] / pmax;
        img[j * mm + q] = w * (pu[j * mm + q] - pum[q]);
        sre += (V[q] * std::conj(img[j * mm + q])).real();
      }
    const double cj = (sre / mm + evo1[f]) / evn;
    for (long q = 0; q < mm; ++q) acc[q] += img[j * mm + q] * cj;
  }
  double s2 = 0;
  for (long q = 0; q < mm; ++q) {
    acc[q] *= nrm;
    s2 += std::norm(acc[q]);
  }
  const double sc = relax_u / std::max(1.0, std::sqrt(s2 / mm));
  double sv = 0;
  for (long q = 0; q < mm; ++q) {
    V[q] += acc[q] * sc;
    sv += std::norm(V[q]);
  }
  const double vn = std::sqrt(sv / mm);
  for (long q = 0; q < mm; ++q) V[q] *= 1.0 / vn;
  std::vector<double> den(G), num(G);
  double dsum = 0;
  for (int j = 0; j < G; ++j) {
    double d = 0, u = 0;
    for (long q = 0; q < mm; ++q) {
      const cd ov = op[j * mm + q] * V[q];
      d += std::norm(ov);
      u += (chi0[j * mm + q] * std::conj(ov)).real();
    }
    den[j] = d / mm;
    num[j] = u / mm;
    dsum += den[j];
  }
  const double m3 = dsum / G;
  for (int j = 0; j < G; ++j) evo1[gi[j]] += num[j] / (den[j] + 0.1 * m3);
}

// The probe-update and object-update directions of mode ll for one group (MLS.cpp:251-297):
// pu_j = chi_j conj(op_j) (shifted back by the frame's sub-pixel shift for ll = 1 with
// apply_subpix_shift), pum = mean_j (p / (p + 0.01 MAX_ILLUM))_roi_j pu_j (d_addFactorDivide);
// the object direction oup_j = chi_j conj(vp_j) for llp = 1 and P_llp chi_j (no conjugate,
// MLS.cpp:294-297) for llp > 1.  pu/pum are formed when probe_rec, oup when obj_rec.
void mode_directions(int ll, int llp, int G, int m, int subpix, const int32_t* gi, const int32_t* roi,
                     const float* sub, const double* illum, int cols, double max_illum, const cd* chi, const cd* op,
                     const cd* vp, const cd* P, bool probe_rec, bool obj_rec, cd* pu, cd* pum, cd* oup,
                     const Fft1& fp) {
  const long mm = (long)m * m;
  std::fill(pum, pum + mm, cd(0));
  if (probe_rec) {
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < G; ++j) {
      const int f = gi[j];
      cd* u = pu + j * mm;
      for (long q = 0; q < mm; ++q) u[q] = chi[j * mm + q] * std::conj(op[j * mm + q]);
      if (ll == 1 && subpix) shift_frame(u, m, -sub[2 * f], -sub[2 * f + 1], fp);
    }
    const double fw = max_illum * 0.01;
    for (int j = 0; j < G; ++j) {
      const int f = gi[j];
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c) {
          const double pv = illum[(long)(roi[2 * f] + r) * cols + roi[2 * f + 1] + c];
          pum[(long)r * m + c] += pv / (pv + fw) * pu[j * mm + (long)r * m + c];
        }
    }
    for (long q = 0; q < mm; ++q) pum[q] /= (double)G;
  }
  if (obj_rec)
    for (int j = 0; j < G; ++j)
      for (long q = 0; q < mm; ++q)
        oup[j * mm + q] = (llp == 1) ? chi[j * mm + q] * std::conj(vp[j * mm + q])
                                     : P[(llp - 1) * mm + q] * chi[j * mm + q];
  else
    std::fill(oup, oup + (long)G * mm, cd(0));
}
}  // namespace

extern "C" {

void orc_mls_lsq(const double* oup, const double* vp, const double* pum, const double* op, const double* chi, int G,
                 int m, double* x) {
  lsq_steps(C(oup), C(vp), C(pum), C(op), C(chi), G, m, C(x));
}

void orc_mls_roi(const float* pos, int n, int m, int rows, int cols, int32_t* roi, float* sub) {
  roi_of(pos, n, m, rows, cols, roi, sub);
}

// Pieces exported for pinning: the sub-pixel shift and the position-solver derivatives/step.
void orc_mls_shift(double* x, int m, float sx, float sy) {
  Fft1 p(m);
  shift_frame(C(x), m, sx, sy, p);
}
void orc_mls_position_step(const double* op, const double* vp, const double* chi, int m, double* out,
                           double* dX, double* dY) {
  Fft1 p(m);
  position_step(C(op), vp ? C(vp) : nullptr, chi ? C(chi) : nullptr, m, p, out, out + 1, dX ? C(dX) : nullptr,
                dY ? C(dY) : nullptr);
}

// Probe::calc_object_norm (Probe.cpp:671-699) over the precalculated 1-based inclusive box
// (x0, x1, y0, y1): object_norm = rms(O W) / rms(W) with W the illumination; the probe modes and
// the variable mode are multiplied by it and the object by its inverse (the complex h_normalize
// multiplies, utilitiesKernels.cu:1646-1653).
void orc_mls_object_norm(const int32_t* box, int rows, int cols, double* obj, const double* illum, int k, int m,
                         double* probe, double* vmode) {
  const long mm = (long)m * m, canvas = (long)rows * cols;
  cd* O = C(obj);
  const int x0 = box[0] - 1, y0 = box[2] - 1;
  const int bx = box[1] - box[0] + 1, by = box[3] - box[2] + 1;
  double sw = 0, so = 0;
  for (int r = 0; r < bx; ++r)
    for (int c = 0; c < by; ++c) {
      const long q = (long)(x0 + r) * cols + y0 + c;
      sw += illum[q] * illum[q];
      so += std::norm(O[q] * illum[q]);
    }
  const double cnt = (double)bx * by;
  const double onorm = std::sqrt(so / cnt) / std::sqrt(sw / cnt);
  for (long q = 0; q < (long)k * mm; ++q) C(probe)[q] *= onorm;
  for (long q = 0; q < mm; ++q) C(vmode)[q] *= onorm;
  for (long q = 0; q < canvas; ++q) O[q] *= 1.0 / onorm;
}

// Probe::remove_extra_degree (Probe.cpp:615-669): V <- V / rms V, e1 <- e1 rms V; then over
// len = number of evolution rows (2), i.e. the first two positions only: e0 <- (e0 - mean e0)
// 0.99 + 1, e1 <- e1 - mean e1 (means kept in a float2), e0 <- e0 / rms(e0), P0 <- P0 rms(e0).
// evo: 2 x n.
void orc_mls_remove_extra(int n, int m, double* probe, double* vmode, double* evo) {
  const long mm = (long)m * m;
  cd* P = C(probe);
  cd* V = C(vmode);
  double* evo0 = evo;
  double* evo1 = evo + n;
  double s = 0;
  for (long q = 0; q < mm; ++q) s += std::norm(V[q]);
  const double vn = std::sqrt(s / mm);
  for (long q = 0; q < mm; ++q) V[q] /= vn;
  for (int i = 0; i < n; ++i) evo1[i] *= vn;
  const int len = std::min(2, n);
  double a = 0, b = 0;
  for (int i = 0; i < len; ++i) {
    a += evo0[i];
    b += evo1[i];
  }
  const double ma = (float)(a / len), mb = (float)(b / len);  // stored in a float2
  double s0 = 0;
  for (int i = 0; i < len; ++i) {
    evo0[i] = (evo0[i] - ma) * 0.99 + 1;
    evo1[i] = evo1[i] - mb;
    s0 += evo0[i] * evo0[i];
  }
  const double cn = std::sqrt(s0 / len);
  for (int i = 0; i < len; ++i) evo0[i] /= cn;
  for (long q = 0; q < mm; ++q) P[q] *= cn;
}

// frame_modulus on one frame's K spectra (k x m x m, replaced by chi); returns the error.
double orc_mls_frame_modulus(double* psi, const double* amp, int k, int m) {
  Fft1 p(m);
  return frame_modulus(C(psi), (long)m * m, amp, k, m, p);
}

// variable_probe_update for one group; pu G x m x m, pum m x m, op / chi0 G x m x m, vmode m x m
// (in/out), evo1 n (in/out), roi n x 2, illum rows x cols.
void orc_mls_variable_probe(int G, int n, int m, const int32_t* gi, const int32_t* roi, const double* illum, int rows,
                            int cols, const double* pu, const double* pum, const double* op, const double* chi0,
                            double* vmode, double* evo1) {
  variable_probe_update(G, n, m, gi, roi, illum, rows, cols, C(pu), C(pum), C(op), C(chi0), C(vmode), evo1);
}

// mode_directions for one group (pu, pum, oup out).
void orc_mls_mode_directions(int ll, int llp, int G, int m, int subpix, const int32_t* gi, const int32_t* roi,
                             const float* sub, const double* illum, int cols, double max_illum, const double* chi,
                             const double* op, const double* vp, const double* probe, int probe_rec, int obj_rec,
                             double* pu, double* pum, double* oup) {
  Fft1 p(m);
  mode_directions(ll, llp, G, m, subpix, gi, roi, sub, illum, cols, max_illum, C(chi), C(op), C(vp), C(probe),
                  probe_rec != 0, obj_rec != 0, C(pu), C(pum), C(oup), p);
}

// Probe::initEvo (Probe.cpp:507-613) with h_preCalillum (ProbeKernels.cu:691-730).
//  evo: 2 x n (row 0 = intensity coefficients, row 1 = variable-mode coefficients; row 1 comes
//  in as the raw 1e-6 N(0,1) draws); frame_power: sum of each frame's intensities (before the
//  square root, Diffractions::fillSquaredSums).  illum: rows x cols (out); *max_illum (out).
void orc_mls_init(int n, int m, int k, int rows, int cols, double* obj, double* probe, double* vmode, double* evo,
                  int variable_probe_modes, const int32_t* roi, const double* frame_power, double* illum,
                  double* max_illum) {
  const long mm = (long)m * m, canvas = (long)rows * cols;
  cd* O = C(obj);
  cd* P = C(probe);
  cd* V = C(vmode);
  for (int i = 0; i < n; ++i) evo[i] = 1.0;
  if (variable_probe_modes > 0) {
    double s = 0;
    for (long q = 0; q < mm; ++q) s += std::norm(V[q]);
    const double rv = std::sqrt(s / mm);
    for (long q = 0; q < mm; ++q) V[q] /= rv;
    for (int i = 0; i < n; ++i) evo[n + i] *= rv;
  }
  std::fill(illum, illum + canvas, 0.0);
  for (int j = 0; j < n; ++j)
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < m; ++c)
        illum[(long)(roi[2 * j] + r) * cols + roi[2 * j + 1] + c] += std::norm(P[(long)r * m + c]);
  double si = 0, soi = 0;
  for (long q = 0; q < canvas; ++q) {
    si += illum[q];
    soi += illum[q] * std::norm(O[q]);
  }
  const double onorm = std::sqrt(soi / si);
  for (long q = 0; q < canvas; ++q) O[q] /= onorm;
  double pw = 0;
  for (long q = 0; q < (long)k * mm; ++q) pw += std::norm(P[q]);
  const double meanillum = pw / mm;
  std::vector<double> corr(n);
  for (int i = 0; i < n; ++i) corr[i] = meanillum / (frame_power[i] / mm);
  std::sort(corr.begin(), corr.end());
  double med = (n % 2 == 0) ? (corr[n / 2 - 1] + corr[n / 2]) / 2 : corr[n / 2];
  const double cr = std::sqrt(med * mm);
  for (long q = 0; q < (long)k * mm; ++q) P[q] /= cr;
  for (long q = 0; q < mm; ++q) V[q] /= cr;
  double mx = 0;
  for (long q = 0; q < canvas; ++q) {
    illum[q] /= cr * cr;
    mx = std::max(mx, illum[q]);
  }
  *max_illum = mx;
}

// Probe::ortho_modes (Probe.cpp:244-294): modified Gram-Schmidt in mode order, then modes
// 1..K-1 (not mode 0) take the modes in order of decreasing pre-orthogonalisation power.
void orc_mls_ortho_modes(double* probe, int k, int m) {
  const long mm = (long)m * m;
  cd* P = C(probe);
  std::vector<std::pair<double, int>> pw;
  for (int i = 0; i < k; ++i) {
    double s = 0;
    for (long q = 0; q < mm; ++q) s += std::norm(P[i * mm + q]);
    pw.push_back({-s, i});
    for (int j = 0; j < i; ++j) {
      double sr = 0;
      cd sl = 0;
      for (long q = 0; q < mm; ++q) {
        sr += std::norm(P[j * mm + q]);
        sl += P[i * mm + q] * std::conj(P[j * mm + q]);
      }
      const cd proj = sl / sr;
      for (long q = 0; q < mm; ++q) P[i * mm + q] -= proj * P[j * mm + q];
    }
  }
  std::sort(pw.begin(), pw.end());
  const std::vector<cd> copy(P, P + k * mm);
  for (int i = 1; i < k; ++i)
    std::copy(copy.begin() + pw[i].second * mm, copy.begin() + (pw[i].second + 1) * mm, P + i * mm);
}

// One MLs iteration.  Groups are given in processing order: frames gframes[goffs[g]..goffs[g+1]).
//  obj rows x cols, probe k x m x m, vmode m x m, evo 2 x n, illum rows x cols, pos / pos_o n x 2
//  (float, row then column coordinate), roi n x 2 (0-based origins, in/out), sub n x 2 (in/out),
//  beta_obj / beta_probe n complex, amp n x m x m amplitudes, ferr n (out; written when iter < 10).
void orc_mls_iteration(const MlsPrm* prm, int n, int m, int k, int rows, int cols, double* obj, double* probe,
                       double* vmode, double* evo, double* illum, double max_illum, float* pos, const float* pos_o,
                       int32_t* roi, float* sub, double* beta_obj, double* beta_probe, const double* amp,
                       const int32_t* gframes, const int32_t* goffs, int n_groups, double* ferr) {
  const int iter = prm->iter;
  const long mm = (long)m * m, canvas = (long)rows * cols;
  cd* O = C(obj);
  cd* P = C(probe);
  cd* V = C(vmode);
  cd* BO = C(beta_obj);
  cd* BP = C(beta_probe);
  double* evo0 = evo;
  double* evo1 = evo + n;
  Fft1 fp(m);

  // MLS.cpp:86-89
  if (iter == 1 || iter >= prm->pps) roi_of(pos, n, m, rows, cols, roi, sub);

  orc_mls_object_norm(prm->box, rows, cols, obj, illum, k, m, probe, vmode);

  // IPtychoScanMesh::updateScanPositionsDiff (float sums, IPtychoScanMesh.cpp:340-369)
  {
    float s1x = 0, s1y = 0, s2x = 0, s2y = 0;
    for (int i = 0; i < n; ++i) {
      s1x += pos[2 * i];
      s1y += pos[2 * i + 1];
      s2x += pos_o[2 * i];
      s2y += pos_o[2 * i + 1];
    }
    const float dx = s1x / n - s2x / n, dy = s1y / n - s2y / n;
    for (int i = 0; i < n; ++i) {
      pos[2 * i] -= dx;
      pos[2 * i + 1] -= dy;
    }
  }

  if (prm->variable_probe_modes && iter > prm->probe_reconstruct) orc_mls_remove_extra(n, m, probe, vmode, evo);

  // Probe::updated_cached_illumination (Probe.cpp:701-729): running average with the new
  // |P0|^2 projections.  MAX_ILLUM stays the initEvo value (this writes m_MAX_ILLUM).
  {
    std::vector<double> add(canvas, 0.0);
    for (int j = 0; j < n; ++j)
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
          add[(long)(roi[2 * j] + r) * cols + roi[2 * j + 1] + c] += std::norm(P[(long)r * m + c]);
    for (long q = 0; q < canvas; ++q) illum[q] = (illum[q] + add[q]) / 2.0;
  }

  const bool var_any = prm->variable_probe || prm->variable_intensity;
  for (int g = 0; g < n_groups; ++g) {
    const int32_t* gi = gframes + goffs[g];
    const int G = goffs[g + 1] - goffs[g];
    std::vector<cd> op((size_t)G * mm), vp((size_t)G * mm), psi((size_t)k * G * mm), pu((size_t)G * mm),
        oup((size_t)G * mm), pum(mm);
    // get_projections at origin t-1 (utilitiesKernels.cu:1064-1084)
    for (int j = 0; j < G; ++j) {
      const int f = gi[j];
      for (int r = 0; r < m; ++r)
        for (int c = 0; c < m; ++c)
          op[j * mm + (long)r * m + c] = O[(long)(roi[2 * f] + r) * cols + roi[2 * f + 1] + c];
    }
    // Probe::get_illumination_probe (Probe.cpp:887-954)
#pragma omp parallel for schedule(dynamic)
    for (int j = 0; j < G; ++j) {
      const int f = gi[j];
      cd* v = vp.data() + j * mm;
      for (long q = 0; q < mm; ++q) v[q] = var_any ? evo0[f] * P[q] + evo1[f] * V[q] : P[q];
      if (prm->apply_subpix_shift) shift_frame(v, m, sub[2 * f], sub[2 * f + 1], fp);
      for (int kk = 0; kk < k; ++kk) {
        cd* ps = psi.data() + ((size_t)kk * G + j) * mm;
        for (long q = 0; q < mm; ++q)
          ps[q] = (kk == 0 && prm->apply_subpix_shift) ? op[j * mm + q] * v[q] : P[kk * mm + q] * op[j * mm + q];
        fft2(ps, m, false, fp);
      }
      // Fourier error (Diffractions.cpp:261-279; MLS.cpp:234-237: only for iter < 10, since
      // iter / min(20, 2^floor(2 + iter/50)) is a floating-point quotient) and
      // modulus constraint (Diffractions.cpp:281-310; DiffractionsKernels.cu:127-137)
      const double ev = frame_modulus(psi.data() + (size_t)j * mm, (long)G * mm, amp + (long)f * mm, k, m, fp);
      if (iter < 10) ferr[f] = ev;
    }
    const cd* chi0 = psi.data();

    const int ll_range = std::max(k, prm->n_objects);
    for (int ll = 1; ll <= ll_range; ++ll) {
      const bool obj_rec = iter >= prm->object_reconstruct && (prm->apply_multimodal_update || ll <= prm->n_objects);
      const bool probe_rec = iter >= prm->probe_reconstruct;
      const int llp = std::min(k, ll);
      const cd* chi = psi.data() + (size_t)(ll - 1) * G * mm;
      mode_directions(ll, llp, G, m, prm->apply_subpix_shift, gi, roi, sub, illum, cols, max_illum, chi, op.data(),
                      vp.data(), P, probe_rec, obj_rec, pu.data(), pum.data(), oup.data(), fp);
      std::vector<cd> upd;  // object_upd_sum after the preconditioner
      if (obj_rec) {
        upd.assign(canvas, cd(0.0, (double)(float)2.220446049250313e-16));  // (0, DBL_EPSILON) as float
        if (prm->delta_p != 0) {
          for (int j = 0; j < G; ++j) {
            const int f = gi[j];
            for (int r = 0; r < m; ++r)
              for (int c = 0; c < m; ++c)
                upd[(long)(roi[2 * f] + r) * cols + roi[2 * f + 1] + c] += oup[j * mm + (long)r * m + c];
          }
          const double fd = max_illum * prm->delta_p;
          for (long q = 0; q < canvas; ++q) upd[q] /= (illum[q] + fd);  // d_object_sum_update_Gfun
          if (prm->beta_lsq > 0)
            for (int j = 0; j < G; ++j) {
              const int f = gi[j];
              for (int r = 0; r < m; ++r)
                for (int c = 0; c < m; ++c)
                  oup[j * mm + (long)r * m + c] = upd[(long)(roi[2 * f] + r) * cols + roi[2 * f + 1] + c];
            }
        }
      }
      // LSQ step (MLS.cpp:349-416)
      if (prm->beta_lsq > 0 && ll == 1 && obj_rec && probe_rec) {
        std::vector<cd> x((size_t)2 * G);
        lsq_steps(oup.data(), vp.data(), pum.data(), op.data(), chi0, G, m, x.data());
        for (int j = 0; j < G; ++j) {
          BO[gi[j]] = prm->beta_object * prm->beta_lsq * x[2 * j];
          BP[gi[j]] = prm->beta_probe * prm->beta_lsq * x[2 * j + 1];
        }
      }
      // Sample::update_object (scan ids all 1: one group mean) and Probe::update_probe
      if (obj_rec) {
        cd bo = 0;
        for (int j = 0; j < G; ++j) bo += BO[gi[j]];
        bo /= (double)G;
        for (long q = 0; q < canvas; ++q) O[q] += upd[q] * bo;
      }
      if (probe_rec && ll <= k) {
        cd bp = 0;
        for (int j = 0; j < G; ++j) bp += BP[gi[j]];
        bp /= (double)G;
        for (long q = 0; q < mm; ++q) P[(ll - 1) * mm + q] += pum[q] * bp;
      }
      // Probe::gradient_position_solver (Probe.cpp:956-1118): the column-derivative step is
      // added to the row coordinate and the row-derivative step to the column coordinate.
      if (iter >= prm->pps && ll == 1) {
#pragma omp parallel for schedule(dynamic)
        for (int j = 0; j < G; ++j) {
          double dx, dy;
          position_step(op.data() + j * mm, vp.data() + j * mm, chi0 + j * mm, m, fp, &dx, &dy);
          const int f = gi[j];
          pos[2 * f] = (float)(pos[2 * f] + std::min(std::fabs(dx), 0.2) * sgn(dx));
          pos[2 * f + 1] = (float)(pos[2 * f + 1] + std::min(std::fabs(dy), 0.2) * sgn(dy));
        }
      }
      // Probe::update_variable_probe (Probe.cpp:1366-1517)
      if (var_any && iter > prm->probe_reconstruct + 1 && ll == 1) {
        if (prm->variable_probe) {
          variable_probe_update(G, n, m, gi, roi, illum, rows, cols, pu.data(), pum.data(), op.data(), chi0, V, evo1);
        }
        if (prm->variable_intensity) {
          for (int j = 0; j < G; ++j) {
            double d = 0, u = 0;
            for (long q = 0; q < mm; ++q) {
              const cd ov = op[j * mm + q] * P[q];
              d += std::norm(ov);
              u += (chi0[j * mm + q] * std::conj(ov)).real();
            }
            evo0[gi[j]] += 0.1 * u / d;
          }
        }
      }
    }
  }
  if (k > prm->n_rec) orc_mls_ortho_modes(probe, k, m);
}

}  // extern "C"
