// Host-side driver: configuration, scan meshes, the IPhasingMethod seam and IPhaser.
//
// Mirrors the reference's C++ layer (IPhaser.h:135-184, IPhasingMethod.h:53-69,
// Parameters.h:55-203, IPtychoScanMesh.h) with the GPU state behind a ptyk_ctx handle
// (include/ptyk.h) instead of process-wide singletons.
#pragma once

#include <cstdint>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "ptyk.h"

namespace ptycho {

struct F2 {
  float x, y;
};

// ---------------------------------------------------------------- parameters
// Field names and defaults follow Parameters.h:55-166 (C++ defaults); the Python entry
// points apply their own defaults first (ptychopy.c:199-240).
struct PreprocessingParams {
  int flags = 0;  // 1 square root, 4 fft shift (Parameters.h:48-53)
  int symmetric_array_center_x = 128, symmetric_array_center_y = 128;
  int raw_file_size_x = 256, raw_file_size_y = 256;
  std::string dataFilePattern, hdf5DatasetName = "/entry/data/data";
  unsigned symmetric_array_size = 256, fileStartIndex = 0, diffsPerFile = 1, projectionsNum = 0;
  unsigned rotate_90_times = 0, bin_factor = 1, io_threads = 1;
  float threshold_raw_data = 0.f;
  bool beamstopMask = false;
};

struct ExperimentParams {
  double lambda = 2.48e-10, beamSize = 1e-7, dx_d = 172e-6, dx = 1.4089e-8, dx_s = 5.631549e-09, z_d = 1.0;
  unsigned scanDims_x = 12, scanDims_y = 12;
  float stepSize_x = 40e-9f, stepSize_y = 40e-9f;
  float drift_x = 0.f, drift_y = 0.f;
  bool spiralScan = false;
  int bit_depth = 32;
  // 2^bit_depth - 1; the reference's default shifts by the type width (Parameters.h:97-98),
  // which never saturates amplitudes -- kept as 2^32 - 1.
  double pixelSaturation = 4294967295.0;
};

struct ReconstructionParams {
  std::string reconstructionID, algorithm = "ePIE";
  unsigned iterations = 100, save = 999999;
  double time = -1.0;
  unsigned desiredShape = 256, probeModes = 1, jitterRadius = 0, halo = 0, shareFrequency = 10;
  unsigned updateProbe = 10, updateProbeModes = 20, updateVis = 10, phaseConstraint = 1;
  unsigned nProbes = 1;
  int Niter = 100, variable_probe_modes = 1, variable_intensity = 1, variable_probe = 1;
  double beta_LSQ = 0.9;
  int beta_object = 1, beta_probe = 1, probe_pos_search = 5, probe_reconstruct = 1, apply_subpix_shift = 1;
  int apply_multimodal_update = 0, object_reconstruct = 1;
  unsigned Nobjects = 1;
  double delta_p = 0.1;
  int Nrec = 1;
  bool blind = true, simulated = false, flipScanAxis = false, mirrorX = false, mirrorY = false;
  bool calculateRMS = false, binaryOutput = false;
  std::string positionsFilename, probeGuess, objectGuess, wsServerUrl;
  uint64_t seed = 0;  // replaces the reference's time-seeded srand / curand streams
  bool compactGroups = false;  // MLs groups from get_close_indices ("MLC") instead of the random partitions
};

class Parameters {
 public:
  PreprocessingParams p;
  ExperimentParams e;
  ReconstructionParams r;
  // Parameters::parseFromCommandLine (Parameters.cpp:66-378): "-key=value" tokens.
  bool parseCommandLine(const std::vector<std::string>& args, std::string* err);
  // One keyword of Parameters::parseFromCPython (Parameters.cpp:380-653), value as text.
  bool setKeyword(const std::string& key, const std::string& value, std::string* err);
  // Derived quantities computed at the end of both parsers (Parameters.cpp:374-376, 650-652).
  void finalize();
};

// ---------------------------------------------------------------- scan meshes
class ScanMesh {
 public:
  // MLs geometry (generateMeshML): positions stay unshifted and unflipped (addScanPositionMLs,
  // IPtychoScanMesh.cpp:111-120), the canvas is the MLs object size and box the
  // calc_object_norm region (precalculate_ROI).
  bool ml = false;
  int box[4] = {0, 0, 0, 0};
  std::vector<F2> positions;  // pixels, shifted to the canvas (addScanPosition) unless ml
  std::vector<unsigned> indices;
  F2 minima{0, 0}, maxima{0, 0};
  int canvas_x = 0, canvas_y = 0;
  virtual ~ScanMesh() = default;
  virtual bool generate(const Parameters& prm, std::string* err) = 0;
  // calculateMeshDimensions + addScanPosition over raw positions laid out in a pitched
  // buffer of `rows` x `cols` (zero padding included in the extents).
  void finish(const std::vector<F2>& raw, unsigned rows, unsigned cols, unsigned pitch, unsigned probe, bool flip);
};

class CartesianScanMesh : public ScanMesh {
 public:
  bool generate(const Parameters& prm, std::string* err) override;
};

// Positions given in metres (ListScanMesh::loadarr / loadList, ListScanMesh.cpp:55-103;
// d_generateMesh<2>, ScanMeshKernels.cu:46-93).
class ListScanMesh : public ScanMesh {
 public:
  std::vector<double> metres;  // n x 2
  bool generate(const Parameters& prm, std::string* err) override;
};

// sqrt(i) (cos, sin)(4 sqrt(i)) * step (d_generateMesh<3>; SpiralScanMesh.cpp:53-78).
class SpiralScanMesh : public ScanMesh {
 public:
  bool generate(const Parameters& prm, std::string* err) override;
};

// ---------------------------------------------------------------- physics objects
// The seam's value types: datatypes.h's real_t (float) and CUDA's float2 / uint2.
typedef float real_t;
typedef F2 float2;
struct uint2 {
  unsigned x, y;
};
typedef ScanMesh IPtychoScanMesh;

// Views with the reference's class roles (Diffractions.h, Probe.h, Sample.h) over one engine
// handle: the frame store, the probe modes and the object canvas are device arrays of the
// ptyk_ctx, and a method reaches them through ctx.  The views own nothing.
class Diffractions {
 public:
  ptyk_ctx* ctx = nullptr;
  int n = 0, m = 0;  // N frames of M x M
};
class Probe {
 public:
  ptyk_ctx* ctx = nullptr;
  int m = 0, modes = 1;
};
class Sample {
 public:
  ptyk_ctx* ctx = nullptr;
  int rows = 0, cols = 0;
};

struct Engine {
  ptyk_ctx* ctx = nullptr;
  int n = 0, m = 0, k = 1, rows = 0, cols = 0;
  Diffractions diffractions;
  Probe probe;
  Sample sample;
  void bind() {
    diffractions.ctx = ctx, diffractions.n = n, diffractions.m = m;
    probe.ctx = ctx, probe.m = m, probe.modes = k;
    sample.ctx = ctx, sample.rows = rows, sample.cols = cols;
  }
  ~Engine() {
    if (ctx) ptyk_destroy(ctx);
  }
};

// IPhasingMethod (IPhasingMethod.h:53-69), pure virtuals with the reference's signatures.
// initMem sees the scan mesh and the probe size once before the first iteration; iteration
// returns what the reference's method returns (ePIE: the averaged Fourier error when RMS is
// set, else 0; DM: 0; MLs: 1) and MLs writes fourierErrors[iter - 1] (MLS.cpp:84, 236).
// The reference exit()s on a failed CUDA call; here a method records the engine's status
// code instead (status(), PTYK_OK on success), which IPhaser::phaseStep checks.
class IPhasingMethod {
 public:
  IPhasingMethod() {}
  virtual ~IPhasingMethod() {}
  virtual void initMem(IPtychoScanMesh* scanMesh, uint2 probeSize) = 0;
  virtual real_t iteration(Diffractions*, Probe*, Sample*, IPtychoScanMesh*,
                           std::vector<std::vector<real_t> >& fourierErrors, const std::vector<float2>& scanPositions,
                           bool phaseConstraint, bool updateProbe, bool updateProbeModes, unsigned int iter,
                           bool RMS = false) = 0;
  virtual void endPhasing() = 0;
  int status() const { return status_; }

 protected:
  int status_ = PTYK_OK;
};

// ePIE (ePIE.h:55-70)
class EPIE : public IPhasingMethod {
 public:
  explicit EPIE(uint64_t seed) : rng_(seed) {}
  void initMem(IPtychoScanMesh*, uint2) override {}
  real_t iteration(Diffractions*, Probe*, Sample*, IPtychoScanMesh*, std::vector<std::vector<real_t> >& fourierErrors,
                   const std::vector<float2>& scanPositions, bool phaseConstraint, bool updateProbe,
                   bool updateProbeModes, unsigned int iter, bool RMS = false) override;
  void endPhasing() override {}
  const std::vector<int32_t>& lastOrder() const { return order_; }

 private:
  std::mt19937_64 rng_;
  std::vector<int32_t> order_;
};

// DM (DM.h:50-77)
class DMMethod : public IPhasingMethod {
 public:
  void initMem(IPtychoScanMesh*, uint2) override {}
  real_t iteration(Diffractions*, Probe*, Sample*, IPtychoScanMesh*, std::vector<std::vector<real_t> >& fourierErrors,
                   const std::vector<float2>& scanPositions, bool phaseConstraint, bool updateProbe,
                   bool updateProbeModes, unsigned int iter, bool RMS = false) override;
  void endPhasing() override {}
};

// MLS (MLS.h:48-69): the frame partitions of get_nonoverlapping_indices are drawn once; each
// iteration takes a random one and a shuffled group order (MLS.cpp:106-113), and the engine runs
// the whole iteration (ptyk_mls_iteration).
class MLSMethod : public IPhasingMethod {
 public:
  MLSMethod(uint64_t seed, int niter, bool compact = false)
      : rng_(seed + 202), seed_(seed), niter_(niter), compact_(compact) {}
  void initMem(IPtychoScanMesh* scanMesh, uint2 probeSize) override;
  real_t iteration(Diffractions*, Probe*, Sample*, IPtychoScanMesh*, std::vector<std::vector<real_t> >& fourierErrors,
                   const std::vector<float2>& scanPositions, bool phaseConstraint, bool updateProbe,
                   bool updateProbeModes, unsigned int iter, bool RMS = false) override;
  void endPhasing() override {}
  std::vector<int32_t> lastGroups, lastOffsets;

 private:
  std::mt19937_64 rng_;
  uint64_t seed_;
  int niter_;
  bool compact_;
  std::vector<int32_t> frames_, offsets_;
  int32_t groups_ = 0, parts_ = 0;
  int64_t n_ = 0;
};

struct CXPhasing {
  std::string name;
  std::unique_ptr<IPhasingMethod> method;
  unsigned iterations;
};

// ---------------------------------------------------------------- driver
class IPhaser {
 public:
  Parameters prm;
  // inputs of the numpy entry points (ptychopy.c:284-441)
  std::vector<float> diffraction;  // n x m x m, as given (amplitudes; intensities for MLs)
  int diff_n = 0, diff_rows = 0, diff_cols = 0;
  std::vector<double> positions_m;  // n x 2 metres
  std::vector<float> object_guess;  // complex64
  int obj_guess_rows = 0, obj_guess_cols = 0;
  bool obj_guess_centred = false;  // read from an objectGuess file: placed at the canvas centre
  std::vector<float> probe_guess;  // complex64 m x m
  int probe_guess_rows = 0, probe_guess_cols = 0;

  std::unique_ptr<ScanMesh> mesh;
  Engine eng;
  std::vector<CXPhasing> methods;
  std::vector<double> errors;
  std::vector<std::vector<real_t> > fourierErrors;  // IPhaser::m_fourierErrors (IPhaser.cpp:341)
  std::string err;
  double phase_seconds = 0.0;

  bool init();                                         // IPhaser::init (IPhaser.cpp:224-262)
  bool addPhasingMethod(const std::string& name, unsigned iterations);  // IPhaser.cpp:264-274
  bool phase();                                        // prePhase + phaseLoop + postPhase
  bool phaseinit();                                    // step mode (IPhaser.cpp:458-471)
  bool phasestepvis(unsigned i);                       // IPhaser.cpp:479-494
  bool phasepost();
  bool writeResultsToDisk(int r = 0);                  // IPhaser.cpp:417-428

 private:
  bool prePhase();
  bool phaseStep(IPhasingMethod* m, unsigned i, real_t* e);
  void postPhase();
  bool simulate_frames(std::vector<float>* amp);
  bool read_frames(int n, int m, std::vector<float>* amp);  // -fp= frame files
  bool mls_prephase();
  bool mls_probe_modes(std::vector<float>* modes);
  std::vector<double> frame_power_;  // MLs: sum of each frame's intensities
  double t0_ = 0.0;
};

// Probe::initMLH with explicit draws (2 per extra mode): modes (K x M x M complex64, mode 0 given
// when non-zero, else the zone-plate model) are replaced by the initial MLs modes.
int init_mlh_modes(ptyk_ctx* ctx, int m, int k, double lambda, double dx_s, double beam_size, const float* draws,
                   bool tilt, std::vector<float>* modes, std::string* err);

// CSV "a+bj" writer / reader of Cuda2DArray (Cuda2DArray.cpp:46-108, Cuda2DArray.hpp:439-464).
bool save_complex(const std::string& path, const float* c64, int rows, int cols, bool binary);
bool load_complex_csv(const std::string& path, std::vector<float>* c64, int* rows, int* cols);

}  // namespace ptycho
