// FP64 CPU oracle of the warm-start network -- TEST INFRASTRUCTURE ONLY (loaded by tests/ as
// the checker).  Same architecture and parameter layout as the engine (include/ptyk.h,
// ptyk_net_*): 3x3 same convolutions with ReLU in four levels (c, 2c, 4c, 8c channels), 2x2
// max pooling, nearest x2 upsampling with the skip concatenated after the upsampled channels,
// final 1x1 convolution to two linear outputs, mean squared error.
//
// emulate_bf16 = 1 rounds to bfloat16 exactly where the engine stores or feeds BF16: the
// weights entering a convolution, every stored activation (after ReLU), and the gradient
// operand of the two gradient products of each convolution (not the bias sums, the ReLU masks,
// pooling or upsampling).  emulate_bf16 = 0 is plain FP64, pinned by finite differences.
#include <cstring>

#include "oracle_common.h"

using namespace orc;

namespace {

inline double bf(double v) {  // round to nearest even bfloat16 (through float)
  float f = (float)v;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct T4 {  // batch x c x h x w, FP64
  int n = 0, c = 0, h = 0;
  std::vector<double> v;
  T4() {}
  T4(int n_, int c_, int h_) : n(n_), c(c_), h(h_), v((size_t)n_ * c_ * h_ * h_, 0.0) {}
  double& at(int b, int ch, int y, int x) { return v[(((size_t)b * c + ch) * h + y) * h + x]; }
  double at(int b, int ch, int y, int x) const { return v[(((size_t)b * c + ch) * h + y) * h + x]; }
};

struct Layer {
  int cin, cout, ks, relu;
  long woff, boff;
};

std::vector<Layer> layers_for(int c, long* np) {
  const int tab[15][4] = {{1, c, 3, 1},          {c, c, 3, 1},         {c, 2 * c, 3, 1},     {2 * c, 2 * c, 3, 1},
                          {2 * c, 4 * c, 3, 1},  {4 * c, 4 * c, 3, 1}, {4 * c, 8 * c, 3, 1}, {8 * c, 8 * c, 3, 1},
                          {12 * c, 4 * c, 3, 1}, {4 * c, 4 * c, 3, 1}, {6 * c, 2 * c, 3, 1}, {2 * c, 2 * c, 3, 1},
                          {3 * c, c, 3, 1},      {c, c, 3, 1},         {c, 2, 1, 0}};
  std::vector<Layer> L;
  long off = 0;
  for (auto& t : tab) {
    Layer l{t[0], t[1], t[2], t[3], 0, 0};
    l.woff = off;
    off += (long)l.cout * l.cin * l.ks * l.ks;
    l.boff = off;
    off += l.cout;
    L.push_back(l);
  }
  *np = off;
  return L;
}

struct Net {
  std::vector<Layer> L;
  std::vector<double> w;  // parameters (bf16-rounded weights when emulating)
  int emu = 0;
};

T4 conv(const Net& net, int li, const T4& x) {
  const Layer& l = net.L[li];
  T4 y(x.n, l.cout, x.h);
  const int r = l.ks / 2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < x.n; ++b)
    for (int co = 0; co < l.cout; ++co)
      for (int yy = 0; yy < x.h; ++yy)
        for (int xx = 0; xx < x.h; ++xx) {
          double s = 0;
          for (int ci = 0; ci < l.cin; ++ci)
            for (int dy = 0; dy < l.ks; ++dy)
              for (int dx = 0; dx < l.ks; ++dx) {
                const int sy = yy + dy - r, sx = xx + dx - r;
                if (sy < 0 || sy >= x.h || sx < 0 || sx >= x.h) continue;
                s += net.w[l.woff + ((long)co * l.cin + ci) * l.ks * l.ks + dy * l.ks + dx] * x.at(b, ci, sy, sx);
              }
          s += net.w[l.boff + co];
          if (l.relu) s = std::max(s, 0.0);
          y.at(b, co, yy, xx) = (net.emu && l.relu) ? bf(s) : s;
        }
  return y;
}

// dz: gradient at the layer output (after the ReLU mask); accumulates dW, db; returns dx
T4 conv_back(const Net& net, int li, const T4& x, const T4& dz, std::vector<double>& g, bool want_dx) {
  const Layer& l = net.L[li];
  const int r = l.ks / 2;
  T4 dzr = dz;
  if (net.emu)
    for (auto& v : dzr.v) v = bf(v);
  for (int co = 0; co < l.cout; ++co) {
    double sb = 0;
    for (int b = 0; b < dz.n; ++b)
      for (int yy = 0; yy < dz.h; ++yy)
        for (int xx = 0; xx < dz.h; ++xx) sb += dz.at(b, co, yy, xx);
    g[l.boff + co] += sb;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int co = 0; co < l.cout; ++co)
    for (int ci = 0; ci < l.cin; ++ci)
      for (int dy = 0; dy < l.ks; ++dy)
        for (int dx = 0; dx < l.ks; ++dx) {
          double s = 0;
          for (int b = 0; b < x.n; ++b)
            for (int yy = 0; yy < x.h; ++yy)
              for (int xx = 0; xx < x.h; ++xx) {
                const int sy = yy + dy - r, sx = xx + dx - r;
                if (sy < 0 || sy >= x.h || sx < 0 || sx >= x.h) continue;
                s += dzr.at(b, co, yy, xx) * x.at(b, ci, sy, sx);
              }
          g[l.woff + ((long)co * l.cin + ci) * l.ks * l.ks + dy * l.ks + dx] += s;
        }
  T4 dx_(x.n, l.cin, x.h);
  if (!want_dx) return dx_;
#pragma omp parallel for collapse(2) schedule(static)
  for (int b = 0; b < x.n; ++b)
    for (int ci = 0; ci < l.cin; ++ci)
      for (int yy = 0; yy < x.h; ++yy)
        for (int xx = 0; xx < x.h; ++xx) {
          double s = 0;
          for (int co = 0; co < l.cout; ++co)
            for (int dy = 0; dy < l.ks; ++dy)
              for (int dx = 0; dx < l.ks; ++dx) {
                const int qy = yy - (dy - r), qx = xx - (dx - r);
                if (qy < 0 || qy >= x.h || qx < 0 || qx >= x.h) continue;
                s += net.w[l.woff + ((long)co * l.cin + ci) * l.ks * l.ks + dy * l.ks + dx] * dzr.at(b, co, qy, qx);
              }
          dx_.at(b, ci, yy, xx) = s;
        }
  return dx_;
}

void mask(T4& g, const T4& a) {
  for (size_t i = 0; i < g.v.size(); ++i)
    if (!(a.v[i] > 0)) g.v[i] = 0;
}

T4 pool(const T4& x, std::vector<int>& arg) {
  T4 y(x.n, x.c, x.h / 2);
  arg.assign(y.v.size(), 0);
  for (int b = 0; b < x.n; ++b)
    for (int c = 0; c < x.c; ++c)
      for (int yy = 0; yy < y.h; ++yy)
        for (int xx = 0; xx < y.h; ++xx) {
          double best = -INFINITY;
          int bi = 0;
          for (int q = 0; q < 4; ++q) {
            const double v = x.at(b, c, 2 * yy + (q >> 1), 2 * xx + (q & 1));
            if (v > best) {
              best = v;
              bi = q;
            }
          }
          y.at(b, c, yy, xx) = best;
          arg[(((size_t)b * y.c + c) * y.h + yy) * y.h + xx] = bi;
        }
  return y;
}

T4 up_cat(const T4& lo, const T4& skip) {  // [up(lo), skip]
  T4 y(lo.n, lo.c + skip.c, skip.h);
  for (int b = 0; b < lo.n; ++b)
    for (int yy = 0; yy < y.h; ++yy)
      for (int xx = 0; xx < y.h; ++xx) {
        for (int c = 0; c < lo.c; ++c) y.at(b, c, yy, xx) = lo.at(b, c, yy / 2, xx / 2);
        for (int c = 0; c < skip.c; ++c) y.at(b, lo.c + c, yy, xx) = skip.at(b, c, yy, xx);
      }
  return y;
}

}  // namespace

extern "C" {

int64_t orc_nn_num_params(int c) {
  long np;
  layers_for(c, &np);
  return np;
}

// Forward (out: batch x 2 x s x s), loss, and (grad != NULL) the parameter gradient.  The loss is
// the weighted mean squared error sum w(p) (y - t)^2 / (batch * 2 * sum_p w(p)), w an s x s
// per-pixel weight shared by both output channels and every sample (NULL: w = 1).
void orc_nn_forward_backward(const float* params, int s, int c, int batch, const float* xin, const float* tgt,
                             const float* weight, int emulate_bf16, double* out, double* loss, double* grad) {
  Net net;
  long np;
  net.L = layers_for(c, &np);
  net.emu = emulate_bf16;
  net.w.assign(params, params + np);
  if (net.emu)
    for (const Layer& l : net.L)
      for (long i = 0; i < (long)l.cout * l.cin * l.ks * l.ks; ++i) net.w[l.woff + i] = bf(net.w[l.woff + i]);
  T4 x(batch, 1, s);
  for (size_t i = 0; i < x.v.size(); ++i) x.v[i] = net.emu ? bf(xin[i]) : xin[i];
  std::vector<int> g0, g1, g2;
  T4 a0 = conv(net, 0, x), e0 = conv(net, 1, a0), p0 = pool(e0, g0);
  T4 a1 = conv(net, 2, p0), e1 = conv(net, 3, a1), p1 = pool(e1, g1);
  T4 a2 = conv(net, 4, p1), e2 = conv(net, 5, a2), p2 = pool(e2, g2);
  T4 a3 = conv(net, 6, p2), bn = conv(net, 7, a3);
  T4 cat2 = up_cat(bn, e2);
  T4 a4 = conv(net, 8, cat2), d2 = conv(net, 9, a4);
  T4 cat1 = up_cat(d2, e1);
  T4 a5 = conv(net, 10, cat1), d1 = conv(net, 11, a5);
  T4 cat0 = up_cat(d1, e0);
  T4 a6 = conv(net, 12, cat0), d0 = conv(net, 13, a6);
  T4 y = conv(net, 14, d0);
  const size_t cnt = y.v.size(), ss = (size_t)s * s;
  double wsum = 0;
  for (size_t p = 0; p < ss; ++p) wsum += weight ? weight[p] : 1.0;
  const double norm = (double)batch * 2 * wsum;
  double l = 0;
  T4 gy(batch, 2, s);
  for (size_t i = 0; i < cnt; ++i) {
    const double t = net.emu ? bf(tgt[i]) : tgt[i];
    const double w = weight ? weight[i % ss] : 1.0;
    const double d = y.v[i] - t;
    l += w * d * d;
    gy.v[i] = 2.0 * w * d / norm;
    if (out) out[i] = y.v[i];
  }
  *loss = l / norm;
  if (!grad) return;
  std::vector<double> g(np, 0.0);
  auto split = [](const T4& cat, int c_lo, T4* lo, T4* skip) {
    *lo = T4(cat.n, c_lo, cat.h);
    *skip = T4(cat.n, cat.c - c_lo, cat.h);
    for (int b = 0; b < cat.n; ++b)
      for (int ch = 0; ch < cat.c; ++ch)
        for (int yy = 0; yy < cat.h; ++yy)
          for (int xx = 0; xx < cat.h; ++xx) {
            if (ch < c_lo) lo->at(b, ch, yy, xx) = cat.at(b, ch, yy, xx);
            else skip->at(b, ch - c_lo, yy, xx) = cat.at(b, ch, yy, xx);
          }
  };
  auto up_back = [](const T4& gup) {
    T4 r(gup.n, gup.c, gup.h / 2);
    for (int b = 0; b < gup.n; ++b)
      for (int ch = 0; ch < gup.c; ++ch)
        for (int yy = 0; yy < gup.h; ++yy)
          for (int xx = 0; xx < gup.h; ++xx) r.at(b, ch, yy / 2, xx / 2) += gup.at(b, ch, yy, xx);
    return r;
  };
  auto pool_back = [](const T4& gp, const std::vector<int>& arg, T4& gin) {
    for (int b = 0; b < gp.n; ++b)
      for (int ch = 0; ch < gp.c; ++ch)
        for (int yy = 0; yy < gp.h; ++yy)
          for (int xx = 0; xx < gp.h; ++xx) {
            const int q = arg[(((size_t)b * gp.c + ch) * gp.h + yy) * gp.h + xx];
            gin.at(b, ch, 2 * yy + (q >> 1), 2 * xx + (q & 1)) += gp.at(b, ch, yy, xx);
          }
  };
  T4 gd0 = conv_back(net, 14, d0, gy, g, true);
  mask(gd0, d0);
  T4 ga6 = conv_back(net, 13, a6, gd0, g, true);
  mask(ga6, a6);
  T4 gcat0 = conv_back(net, 12, cat0, ga6, g, true);
  T4 gup0, ge0;
  split(gcat0, 2 * c, &gup0, &ge0);
  T4 gd1 = up_back(gup0);
  mask(gd1, d1);
  T4 ga5 = conv_back(net, 11, a5, gd1, g, true);
  mask(ga5, a5);
  T4 gcat1 = conv_back(net, 10, cat1, ga5, g, true);
  T4 gup1, ge1;
  split(gcat1, 4 * c, &gup1, &ge1);
  T4 gd2 = up_back(gup1);
  mask(gd2, d2);
  T4 ga4 = conv_back(net, 9, a4, gd2, g, true);
  mask(ga4, a4);
  T4 gcat2 = conv_back(net, 8, cat2, ga4, g, true);
  T4 gup2, ge2;
  split(gcat2, 8 * c, &gup2, &ge2);
  T4 gbn = up_back(gup2);
  mask(gbn, bn);
  T4 ga3 = conv_back(net, 7, a3, gbn, g, true);
  mask(ga3, a3);
  T4 gp2 = conv_back(net, 6, p2, ga3, g, true);
  pool_back(gp2, g2, ge2);
  mask(ge2, e2);
  T4 ga2 = conv_back(net, 5, a2, ge2, g, true);
  mask(ga2, a2);
  T4 gp1 = conv_back(net, 4, p1, ga2, g, true);
  pool_back(gp1, g1, ge1);
  mask(ge1, e1);
  T4 ga1 = conv_back(net, 3, a1, ge1, g, true);
  mask(ga1, a1);
  T4 gp0 = conv_back(net, 2, p0, ga1, g, true);
  pool_back(gp0, g0, ge0);
  mask(ge0, e0);
  T4 ga0 = conv_back(net, 1, a0, ge0, g, true);
  mask(ga0, a0);
  conv_back(net, 0, x, ga0, g, false);
  std::copy(g.begin(), g.end(), grad);
}

// One 'same' convolution with bias and optional ReLU, for pinning against a library routine.
void orc_nn_conv2d(const double* x, const double* w, const double* b, int n, int cin, int cout, int ks, int h,
                   int relu, double* y) {
  Net net;
  net.L = {Layer{cin, cout, ks, relu, 0, (long)cout * cin * ks * ks}};
  net.w.assign(w, w + (long)cout * cin * ks * ks);
  net.w.insert(net.w.end(), b, b + cout);
  T4 xt(n, cin, h);
  std::copy(x, x + xt.v.size(), xt.v.begin());
  T4 yt = conv(net, 0, xt);
  std::copy(yt.v.begin(), yt.v.end(), y);
}

}  // extern "C"
