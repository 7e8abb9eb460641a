// Host-side driver implementation (see phaser.h).  Compiled as C++ with contraction off; the
// mesh arithmetic reproduces nvcc's fused multiply-adds explicitly.
#include "phaser.h"

#include <sys/stat.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "ptycho.h"

namespace ptycho {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
inline float uniform_open0(std::mt19937_64& g) {
  return (float)(((double)(g() >> 11) + 1.0) * (1.0 / 9007199254740992.0));
}
bool parse_pair(const std::string& v, double* a, double* b, bool* pair) {
  const size_t c = v.find(',');
  *pair = c != std::string::npos;
  *a = atof(v.substr(0, c).c_str());
  *b = *pair ? atof(v.substr(c + 1).c_str()) : *a;
  return true;
}
}  // namespace

// ---------------------------------------------------------------- parameters
void Parameters::finalize() {
  p.symmetric_array_size = r.desiredShape;
  e.dx_s = ptycho_sample_pixel_size(e.lambda, e.z_d, (int32_t)r.desiredShape, e.dx_d);
  e.dx = e.dx_s;
}

bool Parameters::parseCommandLine(const std::vector<std::string>& args, std::string* err) {
  for (const std::string& tok : args) {
    std::string t = tok;
    while (!t.empty() && t[0] == '-') t.erase(0, 1);
    if (t.empty()) continue;
    const size_t eq = t.find('=');
    const std::string key = t.substr(0, eq), val = eq == std::string::npos ? "" : t.substr(eq + 1);
    const int iv = atoi(val.c_str());
    double a = 0, b = 0;
    bool pair = false;
    if (key == "sqrtData") p.flags |= 1;
    else if (key == "fftShiftData") p.flags |= 4;
    else if (key == "beamstopMask") p.beamstopMask = iv > 0;
    else if (key == "blind") r.blind = iv > 0;
    else if (key == "simulate") r.simulated = iv > 0;
    else if (key == "spiralScan") e.spiralScan = iv > 0;
    else if (key == "bitDepth") e.bit_depth = iv;
    else if (key == "flipScanAxis") r.flipScanAxis = iv > 0;
    else if (key == "mirror1stScanAxis") r.mirrorX = iv > 0;
    else if (key == "mirror2ndScanAxis") r.mirrorY = iv > 0;
    else if (key == "RMS") r.calculateRMS = iv > 0;
    else if (key == "binaryOutput") r.binaryOutput = iv > 0;
    else if (key == "threads") p.io_threads = iv;
    else if (key == "i") r.iterations = iv;
    else if (key == "save") r.save = iv;
    else if (key == "T") r.time = atof(val.c_str());
    else if (key == "size") r.desiredShape = iv;
    else if (key == "PPS") r.probe_pos_search = iv;
    else if (key == "probeModes") r.probeModes = r.nProbes = iv;
    else if (key == "fs") p.fileStartIndex = iv;
    else if (key == "dpf") p.diffsPerFile = iv;
    else if (key == "angles") p.projectionsNum = iv;
    else if (key == "threshold") p.threshold_raw_data = (float)iv;
    else if (key == "jitterRadius") r.jitterRadius = iv;
    else if (key == "rotate90") p.rotate_90_times = iv;
    else if (key == "overlap") r.halo = iv;
    else if (key == "shareFrequency") r.shareFrequency = iv;
    else if (key == "updateProbe") r.updateProbe = iv;
    else if (key == "updateModes") r.updateProbeModes = iv;
    else if (key == "updateVis") r.updateVis = iv;
    else if (key == "phaseConstraint") r.phaseConstraint = iv;
    else if (key == "fp") p.dataFilePattern = val;
    else if (key == "hdf5path") p.hdf5DatasetName = val;
    else if (key == "scanDims") {
      parse_pair(val, &a, &b, &pair);
      e.scanDims_x = (unsigned)a;
      e.scanDims_y = pair ? (unsigned)b : 1;
    } else if (key == "drift") {
      parse_pair(val, &a, &b, &pair);
      e.drift_x = (float)a;
      e.drift_y = (float)b;
    } else if (key == "qxy") {
      parse_pair(val, &a, &b, &pair);
      p.symmetric_array_center_x = (int)a;
      p.symmetric_array_center_y = (int)b;
    } else if (key == "nxy") {
      parse_pair(val, &a, &b, &pair);
      p.raw_file_size_x = (int)a;
      p.raw_file_size_y = (int)b;
    } else if (key == "step") {
      parse_pair(val, &a, &b, &pair);
      e.stepSize_x = (float)a;
      e.stepSize_y = (float)b;
    } else if (key == "beamSize") e.beamSize = atof(val.c_str());
    else if (key == "lambda") e.lambda = atof(val.c_str());
    else if (key == "dx_d") e.dx_d = atof(val.c_str());
    else if (key == "z") e.z_d = atof(val.c_str());
    else if (key == "jobID") r.reconstructionID = val;
    else if (key == "algorithm") r.algorithm = val;
    else if (key == "lf") r.positionsFilename = val;
    else if (key == "probeGuess") r.probeGuess = val;
    else if (key == "objectGuess") r.objectGuess = val;
    else if (key == "wsServerUrl") r.wsServerUrl = val;
    else if (key == "delta_p") r.delta_p = atof(val.c_str());
    else if (key == "seed") r.seed = strtoull(val.c_str(), nullptr, 10);
    else {
      if (err) *err = "unknown option -" + key;
      return false;
    }
  }
  finalize();
  return true;
}

bool Parameters::setKeyword(const std::string& key, const std::string& val, std::string* err) {
  const long iv = atol(val.c_str());
  const double dv = atof(val.c_str());
  // guards as in Parameters::parseFromCPython (Parameters.cpp:380-653)
  if (key == "jobID") { if (!val.empty()) r.reconstructionID = val; }
  else if (key == "fp") { if (!val.empty()) p.dataFilePattern = val; }
  else if (key == "fs") { if (iv >= 0) p.fileStartIndex = (unsigned)iv; }
  else if (key == "hdf5path") { if (!val.empty()) p.hdf5DatasetName = val; }
  else if (key == "dpf") { if (iv > 1) p.diffsPerFile = (unsigned)iv; }
  else if (key == "beamSize") { if (dv > 0) e.beamSize = dv; }
  else if (key == "probeGuess") { if (!val.empty()) r.probeGuess = val; }
  else if (key == "objectGuess") { if (!val.empty()) r.objectGuess = val; }
  else if (key == "size") {
    if (iv > 16) r.desiredShape = (unsigned)iv;
    else { if (err) *err = "the size must be large than 16"; return false; }
  } else if (key == "qx") p.symmetric_array_center_x = (int)iv;
  else if (key == "qy") p.symmetric_array_center_y = (int)iv;
  else if (key == "nx") p.raw_file_size_x = (int)iv;
  else if (key == "ny") p.raw_file_size_y = (int)iv;
  else if (key == "scanDimsx") { if (iv > 0) e.scanDims_x = (unsigned)iv; }
  else if (key == "scanDimsy") { if (iv > 0) e.scanDims_y = (unsigned)iv; }
  else if (key == "spiralScan") { if (iv == 1) e.spiralScan = true; }
  else if (key == "flipScanAxis") { if (iv == 1) r.flipScanAxis = true; }
  else if (key == "mirror1stScanAxis") { if (iv == 1) r.mirrorX = true; }
  else if (key == "mirror2ndScanAxis") { if (iv == 1) r.mirrorY = true; }
  else if (key == "stepx") { if (dv > 0) e.stepSize_x = (float)dv; }
  else if (key == "stepy") { if (dv > 0) e.stepSize_y = (float)dv; }
  else if (key == "probeModes") { if (iv >= 1) r.probeModes = r.nProbes = (unsigned)iv; }
  else if (key == "lambd") { if (dv > 0) e.lambda = dv; }
  else if (key == "dx_d") { if (dv > 0) e.dx_d = dv; }
  else if (key == "z") { if (dv > 0) e.z_d = dv; }
  else if (key == "iter") { if (iv >= 0) r.iterations = (unsigned)iv; }
  else if (key == "T") { if (iv > 0) r.time = (double)iv; }
  else if (key == "jitterRadius") { if (iv > 0) r.jitterRadius = (unsigned)iv; }
  else if (key == "delta_p") { r.delta_p = dv; }  // the reference accepts and drops it; honoured here
  else if (key == "threshold") { if (iv > 0) p.threshold_raw_data = (float)iv; }
  else if (key == "rotate90") { if (iv >= 0) p.rotate_90_times = (unsigned)iv; }
  else if (key == "sqrtData") { if (iv == 1) p.flags |= 1; }
  else if (key == "fftShiftData") { if (iv == 1) p.flags |= 4; }
  else if (key == "binaryOutput") { if (iv == 1) r.binaryOutput = true; }
  else if (key == "simulate") { if (iv == 1) r.simulated = true; }
  else if (key == "phaseConstraint") { if (iv >= 0) r.phaseConstraint = (unsigned)iv; }
  else if (key == "updateProbe") { if (iv >= 0) r.updateProbe = (unsigned)iv; }
  else if (key == "updateModes") { if (iv >= 0) r.updateProbeModes = (unsigned)iv; }
  else if (key == "beamstopMask") { if (iv == 1) p.beamstopMask = true; }
  else if (key == "lf") { if (!val.empty()) r.positionsFilename = val; }
  else if (key == "PPS") { if (iv >= 1) r.probe_pos_search = (int)iv; }
  else if (key == "algorithm") { if (!val.empty()) r.algorithm = val; }
  else if (key == "RMS") { r.calculateRMS = iv > 0; }
  else if (key == "seed") { r.seed = strtoull(val.c_str(), nullptr, 10); }
  else {
    if (err) *err = "unknown keyword " + key;
    return false;
  }
  finalize();
  return true;
}

// ---------------------------------------------------------------- meshes
void ScanMesh::finish(const std::vector<F2>& raw, unsigned rows, unsigned cols, unsigned pitch, unsigned probe,
                      bool flip) {
  // extents over the pitched buffer with FLT_MIN / FLT_MAX identities (utilitiesKernels.cu:1291-1301);
  // raw holds rows x pitch entries, padding included
  F2 mx{FLT_MIN, FLT_MIN}, mn{FLT_MAX, FLT_MAX};
  for (const F2& v : raw) {
    mx.x = std::max(mx.x, v.x);
    mx.y = std::max(mx.y, v.y);
    mn.x = std::min(mn.x, v.x);
    mn.y = std::min(mn.y, v.y);
  }
  maxima = mx;
  minima = mn;
  canvas_x = (int)std::ceil((mx.x - mn.x) + (float)probe);
  canvas_y = (int)std::ceil((mx.y - mn.y) + (float)probe);
  positions.clear();
  indices.clear();
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < cols; ++j) {
      const F2 v = raw[(size_t)i * pitch + j];
      positions.push_back(flip ? F2{v.y - mn.y, v.x - mn.x} : F2{v.x - mn.x, v.y - mn.y});
      indices.push_back(i * cols + j);
    }
}

bool CartesianScanMesh::generate(const Parameters& prm, std::string* err) {
  ptycho_mesh_params mp{};
  mp.nx = (int32_t)prm.e.scanDims_x;
  mp.ny = (int32_t)prm.e.scanDims_y;
  mp.step_x = prm.e.stepSize_x;
  mp.step_y = prm.e.stepSize_y;
  mp.dx_s = prm.e.dx_s;
  mp.drift_x = prm.e.drift_x;
  mp.drift_y = prm.e.drift_y;
  mp.jitter_radius = (int32_t)prm.r.jitterRadius;
  mp.mirror_x = prm.r.mirrorX;
  mp.mirror_y = prm.r.mirrorY;
  mp.flip = prm.r.flipScanAxis;
  mp.probe_size = (int32_t)prm.r.desiredShape;
  mp.seed = prm.r.seed;
  const size_t n = (size_t)mp.nx * mp.ny;
  std::vector<float> pos(2 * n);
  int32_t canvas[2];
  float ext[4];
  if (ptycho_mesh_cartesian(&mp, pos.data(), canvas, ext) != 0) {
    if (err) *err = "bad Cartesian mesh parameters";
    return false;
  }
  positions.resize(n);
  indices.resize(n);
  for (size_t q = 0; q < n; ++q) {
    positions[q] = F2{pos[2 * q], pos[2 * q + 1]};
    indices[q] = (unsigned)q;
  }
  minima = F2{ext[0], ext[1]};
  maxima = F2{ext[2], ext[3]};
  canvas_x = canvas[0];
  canvas_y = canvas[1];
  return true;
}

// list and spiral meshes live in a (ceil(n/1024) x 1024) pitched buffer in the reference, and
// every entry of it -- padding included -- goes through the position kernel and the extents
static const unsigned kListPitch = 1024;

bool ListScanMesh::generate(const Parameters& prm, std::string* err) {
  const size_t n = metres.size() / 2;
  if (n == 0) {
    if (err) *err = "empty position list";
    return false;
  }
  const unsigned rows = (unsigned)((n + kListPitch - 1) / kListPitch);
  const float o = (float)(1.0 / prm.e.dx_s);
  std::mt19937_64 gen(prm.r.seed);
  std::vector<F2> raw((size_t)rows * kListPitch);
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < kListPitch; ++j) {
      const size_t q = (size_t)i * kListPitch + j;
      float jx = 0.f, jy = 0.f;
      if (prm.r.jitterRadius > 0) {
        const float r = (float)prm.r.jitterRadius;
        jx = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
        jy = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
      }
      float vx = q < n ? (float)metres[2 * q] : 0.f, vy = q < n ? (float)metres[2 * q + 1] : 0.f;
      vx = std::fmaf((float)i, prm.e.drift_x, vx);
      vy = std::fmaf((float)j, prm.e.drift_y, vy);
      raw[q] = F2{std::fmaf(vx, o, jx), std::fmaf(vy, o, jy)};
    }
  std::vector<F2> first(raw.begin(), raw.end());
  finish(first, 1, (unsigned)n, (unsigned)raw.size(), prm.r.desiredShape, prm.r.flipScanAxis);
  return true;
}

bool SpiralScanMesh::generate(const Parameters& prm, std::string* err) {
  const size_t n = (size_t)prm.e.scanDims_x * prm.e.scanDims_y;
  if (n == 0) {
    if (err) *err = "empty scan";
    return false;
  }
  const unsigned rows = (unsigned)((n + kListPitch - 1) / kListPitch);
  const float o = (float)(1.0 / prm.e.dx_s);
  std::mt19937_64 gen(prm.r.seed);
  std::vector<F2> raw((size_t)rows * kListPitch);
  for (unsigned i = 0; i < rows; ++i)
    for (unsigned j = 0; j < kListPitch; ++j) {
      const size_t q = (size_t)i * kListPitch + j;
      float jx = 0.f, jy = 0.f;
      if (prm.r.jitterRadius > 0) {
        const float r = (float)prm.r.jitterRadius;
        jx = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
        jy = std::fmaf(-(uniform_open0(gen) * 2.0f), r, r);
      }
      const float s = std::sqrt((float)q);
      float vx = s * std::cos(4.0f * s) * prm.e.stepSize_x;
      float vy = s * std::sin(4.0f * s) * prm.e.stepSize_y;
      vx = std::fmaf((float)i, prm.e.drift_x, vx);
      vy = std::fmaf((float)j, prm.e.drift_y, vy);
      raw[q] = F2{std::fmaf(vx, o, jx), std::fmaf(vy, o, jy)};
    }
  finish(raw, 1, (unsigned)n, (unsigned)raw.size(), prm.r.desiredShape, prm.r.flipScanAxis);
  return true;
}

// ---------------------------------------------------------------- methods
int EPIE::iteration(Engine& eng, const ScanMesh&, bool phaseConstraint, bool updateProbe, bool updateProbeModes,
                    unsigned, bool rms, double* err) {
  // random visiting order: repeatedly take a random one of the remaining positions
  // (ePIE.cpp:80-97, rand() % remaining; seeded here)
  std::vector<int32_t> remaining(eng.n);
  for (int i = 0; i < eng.n; ++i) remaining[i] = i;
  order_.resize(eng.n);
  for (int s = 0; s < eng.n; ++s) {
    const size_t k = (size_t)(rng_() % remaining.size());
    order_[s] = remaining[k];
    remaining.erase(remaining.begin() + k);
  }
  uint32_t flags = 0;
  if (phaseConstraint) flags |= PTYK_FLAG_PHASE_CONSTRAINT;
  if (updateProbe) flags |= PTYK_FLAG_UPDATE_PROBE;
  if (updateProbeModes) flags |= PTYK_FLAG_UPDATE_MODES;
  if (rms) flags |= PTYK_FLAG_RMS;
  return ptyk_epie_iteration(eng.ctx, order_.data(), flags, err);
}

int DMMethod::iteration(Engine& eng, const ScanMesh&, bool phaseConstraint, bool updateProbe, bool, unsigned,
                        bool, double* err) {
  uint32_t flags = 0;
  if (phaseConstraint) flags |= PTYK_FLAG_PHASE_CONSTRAINT;
  if (updateProbe) flags |= PTYK_FLAG_UPDATE_PROBE;
  if (err) *err = 0.0;  // DM::iteration returns 0 (DM.cu:127)
  return ptyk_dm_iteration(eng.ctx, flags);
}

// ---------------------------------------------------------------- IPhaser
bool IPhaser::simulate_frames(std::vector<float>* amp) {
  // Diffractions::simulate (Diffractions.cpp:80-116) over a seeded random-phase object of the
  // canvas size with the simulated probe (the reference reads lena/baboon PGM files here)
  const int n = eng.n, m = eng.m;
  std::mt19937_64 g(prm.r.seed + 17);
  std::vector<float> obj((size_t)2 * eng.rows * eng.cols);
  for (size_t q = 0; q < obj.size() / 2; ++q) {
    const double a = 0.9 + 0.1 * ((double)(g() >> 11) / 9007199254740992.0);
    const double ph = 0.5 * (2.0 * ((double)(g() >> 11) / 9007199254740992.0) - 1.0);
    obj[2 * q] = (float)(a * std::cos(ph));
    obj[2 * q + 1] = (float)(a * std::sin(ph));
  }
  std::vector<float> probe((size_t)2 * m * m);
  const double w = (double)m / (2 * 2.35);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const double a = i - m * 0.5, b = j - m * 0.5, r = std::sqrt(a * a + b * b);
      double v = r * prm.e.dx_s / prm.e.beamSize;
      v = v != 0 ? std::sin(M_PI * v) / (M_PI * v) : 1.0;
      probe[2 * (i * m + j)] = (float)std::fabs(v * std::exp(-r * r / (w * w)));
    }
  std::vector<float> pos(2 * (size_t)n);
  for (int q = 0; q < n; ++q) {
    pos[2 * q] = mesh->positions[q].x;
    pos[2 * q + 1] = mesh->positions[q].y;
  }
  amp->resize((size_t)n * m * m);
  return ptyk_simulate(eng.ctx, obj.data(), eng.rows, eng.cols, probe.data(), pos.data(), n, 0, amp->data()) == 0;
}

bool IPhaser::init() {
  prm.finalize();
  const bool mls = prm.r.algorithm == "MLs";
  if (mls) {
    err = "the MLs method is not available in this build";
    return false;
  }
  // initScanMesh (IPhaser.cpp:128-171)
  if (!positions_m.empty()) {
    auto* lm = new ListScanMesh();
    lm->metres = positions_m;
    mesh.reset(lm);
  } else if (!prm.r.positionsFilename.empty()) {
    std::ifstream in(prm.r.positionsFilename);
    if (!in.is_open()) {
      err = "Failed to load raster positions file (" + prm.r.positionsFilename + ")";
      return false;
    }
    auto* lm = new ListScanMesh();
    const size_t n = (size_t)prm.e.scanDims_x * prm.e.scanDims_y;
    for (size_t i = 0; i < n; ++i) {
      double x, y;
      char comma;
      if (!(in >> x >> comma >> y) || comma != ',') {
        err = "Invalid raster positions file format [" + prm.r.positionsFilename + "]";
        return false;
      }
      lm->metres.push_back(x);
      lm->metres.push_back(y);
    }
    mesh.reset(lm);
  } else if (prm.e.spiralScan) {
    mesh.reset(new SpiralScanMesh());
  } else {
    mesh.reset(new CartesianScanMesh());
  }
  if (!mesh->generate(prm, &err)) return false;
  const int n = (int)mesh->positions.size(), m = (int)prm.r.desiredShape;
  if (!prm.r.simulated && diffraction.empty()) {
    err = prm.p.dataFilePattern.empty()
              ? "Could not find data path parameter! Did you forget to provide `-fp` argument?"
              : "file input (" + prm.p.dataFilePattern + ") is not supported by this build; pass diffractionNP";
    return false;
  }
  if (!diffraction.empty() && (diff_n != n || diff_rows != m || diff_cols != m)) {
    err = "diffraction array must be N x size x size with N = number of scan positions";
    return false;
  }
  // initSample / initProbe / initDiffractions: one engine handle
  ptyk_config cfg{};
  cfg.n_frames = n;
  cfg.frame_size = m;
  cfg.n_modes = (int32_t)prm.r.probeModes;
  cfg.obj_rows = mesh->canvas_x;
  cfg.obj_cols = mesh->canvas_y;
  cfg.device = 0;
  if (eng.ctx) {
    ptyk_destroy(eng.ctx);
    eng.ctx = nullptr;
  }
  if (ptyk_create(&cfg, &eng.ctx) != PTYK_OK) {
    err = ptyk_last_error(nullptr);
    return false;
  }
  eng.n = n;
  eng.m = m;
  eng.k = cfg.n_modes;
  eng.rows = cfg.obj_rows;
  eng.cols = cfg.obj_cols;
  std::vector<float> pos(2 * (size_t)n);
  for (int q = 0; q < n; ++q) {
    pos[2 * q] = mesh->positions[q].x;
    pos[2 * q + 1] = mesh->positions[q].y;
  }
  if (ptyk_set_positions(eng.ctx, pos.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  std::vector<float> amp;
  if (!diffraction.empty()) {
    amp = diffraction;
  } else if (!simulate_frames(&amp)) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (ptyk_set_frames_f32(eng.ctx, amp.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  return true;
}

bool IPhaser::addPhasingMethod(const std::string& name, unsigned iterations) {
  if (name == "ePIE")
    methods.push_back(CXPhasing{name, std::unique_ptr<IPhasingMethod>(new EPIE(prm.r.seed)), iterations});
  else if (name == "DM")
    methods.push_back(CXPhasing{name, std::unique_ptr<IPhasingMethod>(new DMMethod()), iterations});
  else {
    err = "Reconstruction algorithm [" + name + "] not implemented";
    return false;
  }
  return true;
}

bool IPhaser::prePhase() {
  // IPhaser::prePhase, non-MLs branch (IPhaser.cpp:296-333)
  if (!eng.ctx) {
    err = "init() must be called before phase()";
    return false;
  }
  const int m = eng.m;
  if (!prm.r.probeGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.probeGuess, &g, &r, &c) || r != m || c != m) {
      err = "Probe guess file (" + prm.r.probeGuess + ") not found or not " + std::to_string(m) + "x" +
            std::to_string(m);
      return false;
    }
    probe_guess = g;
    probe_guess_rows = r;
    probe_guess_cols = c;
  }
  if (!probe_guess.empty()) {
    if (probe_guess_rows != m || probe_guess_cols != m) {
      err = "probe guess must be size x size";
      return false;
    }
    std::vector<float> all((size_t)2 * eng.k * m * m, 0.f);
    std::copy(probe_guess.begin(), probe_guess.end(), all.begin());
    if (ptyk_set_probe(eng.ctx, all.data()) != PTYK_OK || ptyk_refresh_maxima(eng.ctx) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  } else if (ptyk_init_probe(eng.ctx, prm.e.beamSize, prm.e.dx_s) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (ptyk_clear_object(eng.ctx) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  if (!prm.r.objectGuess.empty()) {
    std::vector<float> g;
    int r = 0, c = 0;
    if (!load_complex_csv(prm.r.objectGuess, &g, &r, &c)) {
      err = "Object guess file (" + prm.r.objectGuess + ") not found";
      return false;
    }
    object_guess = g;
    obj_guess_rows = r;
    obj_guess_cols = c;
  }
  if (!object_guess.empty()) {
    // Sample::loadarr / loadGuess: the guess fills the canvas from its origin
    std::vector<float> canvas((size_t)2 * eng.rows * eng.cols, 0.f);
    for (int i = 0; i < std::min(eng.rows, obj_guess_rows); ++i)
      for (int j = 0; j < std::min(eng.cols, obj_guess_cols); ++j) {
        canvas[2 * ((size_t)i * eng.cols + j)] = object_guess[2 * ((size_t)i * obj_guess_cols + j)];
        canvas[2 * ((size_t)i * eng.cols + j) + 1] = object_guess[2 * ((size_t)i * obj_guess_cols + j) + 1];
      }
    if (ptyk_set_object(eng.ctx, canvas.data()) != PTYK_OK) {
      err = ptyk_last_error(eng.ctx);
      return false;
    }
  }
  errors.assign(prm.r.iterations, 0.0);
  for (auto& pm : methods) pm.method->initMem(*mesh, m);
  return true;
}

bool IPhaser::phaseStep(IPhasingMethod* m, unsigned i, double* e) {
  // IPhaser::phaseStep (IPhaser.cpp:391-398)
  const int rc = m->iteration(eng, *mesh, i <= prm.r.phaseConstraint, i >= prm.r.updateProbe,
                              i >= prm.r.updateProbeModes, i, prm.r.calculateRMS, e);
  if (rc != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  return true;
}

void IPhaser::postPhase() {
  if (eng.ctx) ptyk_end_modal(eng.ctx);
  for (auto& pm : methods) pm.method->endPhasing();
}

bool IPhaser::phase() {
  if (!eng.ctx || !mesh) {
    err = "init() must be called before phase()";
    return false;
  }
  t0_ = now_s();
  if (!prePhase()) return false;
  // IPhaser::phaseLoop (IPhaser.cpp:367-389)
  for (auto& pm : methods)
    for (unsigned i = 1; i <= pm.iterations; ++i) {
      double e = 0.0;
      if (!phaseStep(pm.method.get(), i, &e)) return false;
      if (i - 1 < errors.size()) errors[i - 1] += e;
      if (i % prm.r.save == 0) writeResultsToDisk((int)(i / prm.r.save));
      if (prm.r.time >= 0 && now_s() - t0_ >= prm.r.time) break;
    }
  postPhase();
  phase_seconds = now_s() - t0_;
  return true;
}

bool IPhaser::phaseinit() {
  if (!eng.ctx || !mesh) {
    err = "init() must be called before phaseinit()";
    return false;
  }
  t0_ = now_s();
  return prePhase();
}

bool IPhaser::phasestepvis(unsigned i) {
  for (auto& pm : methods) {
    double e = 0.0;
    if (!phaseStep(pm.method.get(), i, &e)) return false;
    if (i >= 1 && i - 1 < errors.size()) errors[i - 1] += e;
    if (prm.r.time >= 0 && now_s() - t0_ >= prm.r.time) break;
  }
  return true;
}

bool IPhaser::phasepost() {
  postPhase();
  phase_seconds = now_s() - t0_;
  return true;
}

bool IPhaser::writeResultsToDisk(int r) {
  if (!eng.ctx) return false;
  mkdir("data", 0755);
  const bool bin = prm.r.binaryOutput;
  std::vector<float> probe((size_t)2 * eng.k * eng.m * eng.m), obj((size_t)2 * eng.rows * eng.cols);
  if (ptyk_get_probe(eng.ctx, probe.data()) != PTYK_OK || ptyk_get_object(eng.ctx, obj.data()) != PTYK_OK) {
    err = ptyk_last_error(eng.ctx);
    return false;
  }
  char fn[512];
  snprintf(fn, sizeof fn, "data/%s_probes_%d.%s", prm.r.reconstructionID.c_str(), r, bin ? "bin" : "csv");
  bool ok = save_complex(fn, probe.data(), eng.k * eng.m, eng.m, bin);
  snprintf(fn, sizeof fn, "data/%s_object_%d.%s", prm.r.reconstructionID.c_str(), r, bin ? "bin" : "csv");
  ok = save_complex(fn, obj.data(), eng.rows, eng.cols, bin) && ok;
  return ok;
}

// ---------------------------------------------------------------- CSV
bool save_complex(const std::string& path, const float* c, int rows, int cols, bool binary) {
  std::ofstream out(path, binary ? std::ios::out | std::ios::binary : std::ios::out);
  if (!out.is_open()) return false;
  if (binary) {
    out.write((const char*)c, (std::streamsize)rows * cols * 2 * sizeof(float));
    return true;
  }
  for (int i = 0; i < rows; ++i) {
    for (int j = 0; j < cols; ++j) {
      const float re = c[2 * ((size_t)i * cols + j)], im = c[2 * ((size_t)i * cols + j) + 1];
      out << re;
      if (im >= 0) out << "+";
      out << im << "j";
      if (j < cols - 1) out << ", ";
    }
    out << std::endl;
  }
  return true;
}

// Complex token "a+bj" / "a-bj" / "(a+bj)" / "a", as written by save_complex and numpy.
// The reference's operator>> (Cuda2DArray.cpp:54-102) splits at the first 'e' and assumes a
// two-digit exponent on the real part, which misreads e.g. "0.5+2.3e-06j"; here the split is
// the last sign that does not follow an exponent marker.
static void parse_complex_token(std::string t, float* re, float* im) {
  std::string u;
  for (char ch : t)
    if (ch != ',' && ch != '(' && ch != ')' && ch != ' ') u += ch;
  const bool imag = !u.empty() && (u.back() == 'j' || u.back() == 'i');
  if (imag) u.pop_back();
  size_t split = std::string::npos;
  for (size_t k = u.size(); k-- > 1;)
    if ((u[k] == '+' || u[k] == '-') && u[k - 1] != 'e' && u[k - 1] != 'E') {
      split = k;
      break;
    }
  if (split == std::string::npos) {
    if (imag) {
      *re = 0.f;
      *im = (float)atof(u.c_str());
    } else {
      *re = (float)atof(u.c_str());
      *im = 0.f;
    }
    return;
  }
  *re = (float)atof(u.substr(0, split).c_str());
  *im = (float)atof(u.substr(split).c_str());
}

bool load_complex_csv(const std::string& path, std::vector<float>* c, int* rows, int* cols) {
  std::ifstream in(path);
  if (!in.is_open()) return false;
  c->clear();
  std::string line;
  int r = 0, cc = -1;
  while (std::getline(in, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream ls(line);
    std::string tok;
    int k = 0;
    while (ls >> tok) {
      float re, im;
      parse_complex_token(tok, &re, &im);
      c->push_back(re);
      c->push_back(im);
      ++k;
    }
    if (cc < 0) cc = k;
    if (k != cc) return false;
    ++r;
  }
  *rows = r;
  *cols = cc < 0 ? 0 : cc;
  return r > 0;
}

}  // namespace ptycho
