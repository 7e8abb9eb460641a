/*
 * ptyk.h -- C ABI of the gfx950 ptychography engine.
 *
 * One handle owns the device state of one reconstruction: the FP16 frame store, the
 * scan positions, the object canvas, the probe modes and the per-method solver state.
 * Handles replace the reference's process-wide singletons (CXParams, FFTPlanner,
 * PhaserUtil; Parameters.h:203, CudaFFTPlan.h:125, utilities.h:639), so several
 * reconstructions can live in one process.
 *
 * All pointers are HOST pointers unless stated otherwise; sizes are element counts.
 * Complex arrays are interleaved float32 (re, im), row-major, no pitch padding.
 * Every function returns PTYK_OK (0) or a negative PTYK_E* code and never exits the
 * process (the reference calls exit() on errors: utilities.h:73-92, ePIE.cpp:100-105).
 */
#ifndef PTYK_H
#define PTYK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptyk_ctx ptyk_ctx;

enum {
  PTYK_OK = 0,
  PTYK_EINVAL = -1,  /* bad argument (shape, size, null pointer, unsupported M) */
  PTYK_ENOMEM = -2,  /* device allocation failed */
  PTYK_EDEVICE = -3, /* HIP runtime error */
  PTYK_ESTATE = -4,  /* call order: frames/positions/probe not set yet */
  PTYK_ERANGE = -5   /* a scan position puts a ROI outside the object canvas */
};

/* Iteration flags, the booleans IPhaser::phaseStep passes to IPhasingMethod::iteration
 * (IPhaser.cpp:391-398; IPhasingMethod.h:53-69). */
enum {
  PTYK_FLAG_PHASE_CONSTRAINT = 1u, /* i <= phaseConstraint */
  PTYK_FLAG_UPDATE_PROBE = 2u,     /* i >= updateProbe */
  PTYK_FLAG_UPDATE_MODES = 4u,     /* i >= updateProbeModes */
  PTYK_FLAG_RMS = 8u               /* accumulate the Fourier error (ePIE) */
};

/* Scalar state slots readable with ptyk_get_scalars. */
enum {
  PTYK_SCAL_OBJ_MAX = 0,   /* Sample::m_maxObjectIntensity (max |O|^2, or sum for DM) */
  PTYK_SCAL_PROBE_MAX = 1, /* Probe::m_maxIntensity (max |P|^2, or sum for DM) */
  PTYK_SCAL_ERROR = 2,     /* last iteration's error accumulator (ePIE: sum over positions) */
  PTYK_SCAL_DM_FACTOR_OBJ = 4,   /* DM cached 1/objMax (probe-update factor; DM.cu:46-73) */
  PTYK_SCAL_DM_FACTOR_PROBE = 5, /* DM cached 1/probeMax (object-update factor) */
  PTYK_SCAL_COUNT = 8
};

typedef struct {
  int32_t n_frames;    /* N diffraction patterns */
  int32_t frame_size;  /* M: square frames M x M, power of two in [16, 512] */
  int32_t n_modes;     /* K probe modes (Probe(size, modes), Probe.cpp:54-68) */
  int32_t obj_rows;    /* object canvas (Sample(arrayShapeX, arrayShapeY), Sample.cpp:55-62) */
  int32_t obj_cols;
  int32_t device;      /* HIP device ordinal */
  int32_t tex_quantize;/* 1: bilinear weights rounded to 1/256 as a texture unit does; 0: exact */
  int32_t reserved;
} ptyk_config;

/* Create / destroy.  Replaces the device allocations of Sample, Probe, Diffractions
 * (Sample.cpp:55-62, Probe.cpp:54-68, Diffractions.cpp:159-186). */
int ptyk_create(const ptyk_config* cfg, ptyk_ctx** out);
void ptyk_destroy(ptyk_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed ptyk_create when ctx is NULL). */
const char* ptyk_last_error(const ptyk_ctx* ctx);

/* Frame store.  Amplitudes (not intensities), DC at [0,0], N x M x M
 * (Diffractions::loadarr, Diffractions.cpp:159-186).  Stored as FP16 with a per-frame
 * power-of-two FP32 scale; ptyk_get_frames_f32 returns the dequantised values the
 * solvers actually use. */
int ptyk_set_frames_f32(ptyk_ctx* ctx, const float* amplitudes);
int ptyk_set_frames_f16(ptyk_ctx* ctx, const uint16_t* halves, const float* scales);
int ptyk_get_frames_f32(ptyk_ctx* ctx, float* out);
int ptyk_get_frame_scales(ptyk_ctx* ctx, float* out);

/* Scan positions in pixels, N x (row, col) float32, already shifted to the canvas
 * (IPtychoScanMesh::addScanPosition, IPtychoScanMesh.cpp:102-109).  ROI origins are
 * roundf(pos) (ePIE.cpp:114-115, SampleKernels.cu:426); every ROI must lie inside
 * the canvas, else PTYK_ERANGE. */
int ptyk_set_positions(ptyk_ctx* ctx, const float* pos_rc);
/* The rounded ROI origins the engine uses, N x (row, col) int32. */
int ptyk_get_roi(ptyk_ctx* ctx, int32_t* out);

/* Object canvas rows x cols complex64 and probe modes K x M x M complex64
 * (Sample::loadarr / Probe::init with probearr; Sample.cpp:184-188, Probe.cpp:342-359). */
int ptyk_set_object(ptyk_ctx* ctx, const float* obj);
int ptyk_get_object(ptyk_ctx* ctx, float* obj);
int ptyk_clear_object(ptyk_ctx* ctx); /* Sample::clearObjectArray (Sample.cpp:216-220) */
int ptyk_set_probe(ptyk_ctx* ctx, const float* probe);
int ptyk_get_probe(ptyk_ctx* ctx, float* probe);
/* DM exit-wave state, N x M x M complex64 (DM::m_psi, DM.cu:95-100). */
int ptyk_set_psi(ptyk_ctx* ctx, const float* psi);
int ptyk_get_psi(ptyk_ctx* ctx, float* psi);

/* Scalar state; setting the DM factor slots also marks the DM cache as initialised. */
int ptyk_get_scalars(ptyk_ctx* ctx, double* out /* PTYK_SCAL_COUNT */);
int ptyk_set_scalars(ptyk_ctx* ctx, const double* in /* PTYK_SCAL_COUNT */);
/* Recompute objMax = max |O|^2 and probeMax = max |P|^2 from the current arrays
 * (Sample/Probe::updateIntensities with useSum = false). */
int ptyk_refresh_maxima(ptyk_ctx* ctx);

/* Initial probe for ePIE/DM (Probe::init without a guess, Probe.cpp:360-372):
 * P0 = |sinc(r dx_s/beamSize) exp(-r^2/w^2)|, w = M/(2*2.35); P0 <- IFFT(mean_j(A_j)
 * exp(i arg FFT(P0)))/M^2; all modes normalised to max|P| = 1 (Probe::normalize). */
int ptyk_init_probe(ptyk_ctx* ctx, double beam_size, double dx_s);

/* ePIE modal reconstruction start / end (Probe::beginModalReconstruction and
 * endModalReconstruction, Probe.cpp:1768-1782; initProbeModes, Probe.cpp:320-330). */
int ptyk_begin_modal(ptyk_ctx* ctx);
int ptyk_end_modal(ptyk_ctx* ctx);
/* Probe::orthogonalize (Probe.cpp:228-242) on the handle's probe modes: for i >= 1 and
 * j = i-1 .. 0, P_i -= P_j sum(conj(v_i) P_j) / sum |P_j|^2 with v_i mode i before the sweep (the
 * reference's h_projectUtoV coefficient, ProbeKernels.cu:509-515). */
int ptyk_orthogonalize_modes(ptyk_ctx* ctx);

/* One ePIE iteration over all positions in the given order (ePIE::iteration,
 * ePIE.cpp:65-129).  order: N frame indices (a permutation).  err (nullable)
 * receives the averaged Fourier error when PTYK_FLAG_RMS is set, else 0. */
int ptyk_epie_iteration(ptyk_ctx* ctx, const int32_t* order, uint32_t flags, double* err);

/* One DM iteration (DM::iteration, DM.cu:80-128).  The first call takes the cached
 * normalisation factors from the current objMax/probeMax (DM.cu:88-92). */
int ptyk_dm_iteration(ptyk_ctx* ctx, uint32_t flags);
/* Same, timing the launches with HIP events on the engine stream, summed per kernel kind:
 * kernel_ms[8] = {row pass, column pass, inverse row pass, probe gather, object gather,
 *                 probe apply, object apply, finish}. */
int ptyk_dm_iteration_timed(ptyk_ctx* ctx, uint32_t flags, float* kernel_ms);

/* Tuning knobs (no reference counterpart; chosen by the autotuner).  They change launch shapes
 * only: every value computes the same iteration.
 * PTYK_TUNE_DM_WAVE: frames per DM wave for M in {128, 256, 512}; 0 = automatic (about
 * 128 MiB of exit waves per wave), -1 = one pass over all frames.  get returns the value in
 * effect (0 when the single-pass form runs).
 * PTYK_TUNE_MLS_MOD_CB: column width (8, 16 or 32; default 16) of the 64-row blocks of the MLs
 * modulus pass for M in {128, 256, 512}.
 * PTYK_TUNE_MLS_PUM_FPC: frames per chunk (2..64; default 4) of the MLs probe-mode update.
 * PTYK_TUNE_MLS_GRAPHS: 1 replays each MLs group from a HIP graph captured once per group shape
 * and flag set (one launch per group); 0 (default) launches every kernel from the host.
 * PTYK_TUNE_DM_GRAPH: 1 replays every DM iteration after the first from a HIP graph captured once
 * per flag set on the handle's stream (re-captured when the schedule changes); 0 (default)
 * launches the iteration's kernels from the host.
 * PTYK_TUNE_MLS_ROW_WAVES: 1 (default) runs the MLs exit-wave and inverse row passes and the
 * probe-mode update for M in {128, 256} as wave tiles (one wave per tile of whole row groups,
 * no workgroup barrier between its transform stages); 0 as workgroup tiles.  The row passes give
 * bitwise the same values either way; the probe-mode sums add their frames in another order. */
enum { PTYK_TUNE_DM_WAVE = 1, PTYK_TUNE_MLS_MOD_CB = 2, PTYK_TUNE_MLS_PUM_FPC = 3, PTYK_TUNE_MLS_GRAPHS = 4,
       PTYK_TUNE_DM_GRAPH = 5, PTYK_TUNE_MLS_ROW_WAVES = 6 };
int ptyk_set_tuning(ptyk_ctx* ctx, int32_t key, int32_t value);
int ptyk_get_tuning(ptyk_ctx* ctx, int32_t key, int32_t* value);

/* ---------------------------------------------------------------- MLs
 * Maximum-likelihood reconstruction with least-squares steps (MLS::iteration, MLS.cpp:74-490),
 * the variable probe (Probe::update_variable_probe), sub-pixel probe shifts and gradient
 * position refinement.  Positions are in pixels and NOT shifted to the canvas (MLs keeps them
 * centred, CartesianScanMesh.cpp:94-177); ROI origins are round(pos + ceil(rows/2 - M/2)) - 1
 * per axis (IPtychoScanMesh.cpp:123-172).  The object canvas and the K probe modes are the
 * handle's (ptyk_set_object / ptyk_set_probe before ptyk_mls_begin). */
typedef struct {
  int32_t probe_pos_search;        /* PPS: position refinement from this iteration on (ptychopy: 20) */
  int32_t probe_reconstruct;       /* probe updates from this iteration on (1) */
  int32_t object_reconstruct;      /* object updates from this iteration on (1) */
  int32_t variable_probe;          /* 1 */
  int32_t variable_intensity;      /* 1 */
  int32_t variable_probe_modes;    /* 1 (one extra mode) */
  int32_t apply_subpix_shift;      /* 1 */
  int32_t apply_multimodal_update; /* 0 */
  int32_t n_objects;               /* 1 (the only supported value) */
  int32_t n_rec;                   /* 1: modes are re-orthogonalised after an iteration when K > n_rec */
  int32_t object_box[4];           /* calc_object_norm box, 1-based inclusive: row0, row1, col0, col1
                                      (IPtychoScanMesh::precalculate_ROI, IPtychoScanMesh.cpp:179-184) */
  double delta_p;                  /* preconditioner regulariser (0.1) */
  double beta_lsq;                 /* 0.9 */
  double beta_object;              /* 1 */
  double beta_probe;               /* 1 */
  int32_t reserved[4];
} ptyk_mls_config;

/* Start an MLs reconstruction (Probe::initEvo, Probe.cpp:507-613): positions n x 2 (px),
 * variable_mode M x M complex64 (Probe::initVarModes draws), evolution_draws n values for the
 * variable-mode coefficients (1e-6 N(0,1) draws in the reference), frame_power n sums of each
 * frame's intensities (NULL: sums of the squared stored amplitudes).  Normalises the object,
 * the probe modes and the variable mode, builds the illumination canvas; LSQ steps start at 1.
 * PTYK_ERANGE if a ROI falls outside the canvas. */
int ptyk_mls_begin(ptyk_ctx* ctx, const ptyk_mls_config* cfg, const float* positions, const float* variable_mode,
                   const double* evolution_draws, const double* frame_power);
/* One MLs iteration (iter is 1-based).  The frames are processed in groups: group g is
 * group_frames[group_offsets[g] .. group_offsets[g+1]), in the given group order (the
 * reference's random partition and shuffled order, MLS.cpp:106-113; IPtychoScanMesh.cpp:826-913).
 * A frame may appear at most once.  fourier_errors (nullable, n values) receives the per-frame
 * Fourier errors when iter < 10 (MLS.cpp:234-237).  PTYK_ERANGE if refined positions moved a
 * ROI outside the canvas (the ROI is clamped; the state stays usable). */
int ptyk_mls_iteration(ptyk_ctx* ctx, int32_t iter, const int32_t* group_frames, const int32_t* group_offsets,
                       int32_t n_groups, float* fourier_errors);
/* Same, timing the main-stream launches with HIP events, summed per kind over the iteration:
 * kernel_ms[15] = {per-iteration steps (ROIs, object norm, illumination, mode
 * re-orthogonalisation), illumination probes formed in line (0 when the previous group formed
 * them on the second stream), exit-wave row pass, modulus column pass, inverse row pass, probe
 * update and its shift, preconditioned object direction, LSQ sums and solve, step application,
 * probe modes 2..K, waits for the second stream; then the own kernel time of the exit-wave row
 * pass, the modulus column pass and the inverse row pass (start/stop events recorded by their
 * dispatches: the intervals above also carry the events' own processing on the stream); then
 * the host time to enqueue the iteration}. */
int ptyk_mls_iteration_timed(ptyk_ctx* ctx, int32_t iter, const int32_t* group_frames, const int32_t* group_offsets,
                             int32_t n_groups, float* fourier_errors, float* kernel_ms);
/* MLs state; every non-null buffer is copied out (get) or in (set). */
typedef struct {
  float* positions;        /* n x 2 current positions (px) */
  float* positions_start;  /* n x 2 positions at ptyk_mls_begin */
  int32_t* roi;            /* n x 2 ROI origins (0-based) */
  float* subpixel;         /* n x 2 sub-pixel shifts (column fraction, row fraction) */
  double* evolution;       /* 2 x n probe-evolution coefficients */
  float* variable_mode;    /* M x M complex64 */
  float* illumination;     /* rows x cols cached illumination */
  float* beta_object;      /* n complex64 LSQ object steps */
  float* beta_probe;       /* n complex64 LSQ probe steps */
  float* fourier_errors;   /* n */
  double max_illumination; /* Probe::MAX_ILLUM (get: out; set: used when > 0) */
} ptyk_mls_state;
int ptyk_mls_get_state(ptyk_ctx* ctx, ptyk_mls_state* st);
int ptyk_mls_set_state(ptyk_ctx* ctx, const ptyk_mls_state* st);

/* ---------------------------------------------------------------- batched scans
 * Several independent reconstructions (handles on one device, any sizes) advanced together
 * (no reference counterpart: the reference runs one reconstruction per process).  One DM
 * iteration of every scan is captured once per flag set into a HIP graph with one branch per
 * scan and replayed afterwards; each scan's result is bitwise that of ptyk_dm_iteration on it
 * alone.  The first call (and any call after a scan's buffers changed) runs eagerly or
 * re-captures.  Calls return after the launch; ptyk_batch_synchronize waits. */
typedef struct ptyk_batch ptyk_batch;
int ptyk_batch_create(ptyk_ctx* const* ctxs, int32_t n, ptyk_batch** out);
void ptyk_batch_destroy(ptyk_batch* batch);
const char* ptyk_batch_last_error(const ptyk_batch* batch);
int ptyk_batch_dm_iteration(ptyk_batch* batch, uint32_t flags);
int ptyk_batch_synchronize(ptyk_batch* batch);
int ptyk_batch_graph_count(const ptyk_batch* batch); /* graphs captured so far */

/* ---------------------------------------------------------------- object error
 * Normalised object error against a known object (no reference counterpart; the reference's
 * d_realSpaceER compares magnitudes only and is never called, utilitiesKernels.cu:714-735):
 *   e = min_gamma ||gamma O - T||_2 / ||T||_2,  gamma = <O, T> / <O, O>,
 * over the pixels whose scan illumination sum_j sum_k |P_k|^2 (each frame at its ROI: the MLs
 * ROIs once ptyk_mls_begin has run, else those of ptyk_set_positions) exceeds frac times its
 * maximum (0.01 for the benchmark).  truth: rows x cols complex64, kept on the device.
 * out[4] = {e, Re gamma, Im gamma, pixels used}. */
int ptyk_set_true_object(ptyk_ctx* ctx, const float* truth);
int ptyk_object_error(ptyk_ctx* ctx, double frac, double* out);
/* Learned warm start (no reference counterpart): the object becomes the average of the given
 * patches (n x 2 x M x M: amplitude, phase; patch i for frame frames[i]) placed at the frames'
 * ROIs (the MLs ROIs after ptyk_mls_begin), weighted by sum_k |P_k|^2 of the current probe;
 * pixels whose weight sum is <= rel_threshold x the probe's peak intensity take fill. */
int ptyk_stitch_patches(ptyk_ctx* ctx, const float* patches, const int32_t* frames, int64_t n, float rel_threshold,
                        float fill_re, float fill_im);

/* ---------------------------------------------------------------- learned warm start
 * Convolutional encoder-decoder (no reference counterpart): four levels with c, 2c, 4c, 8c
 * channels (c = base_channels, 32 for the benchmark), 3x3 convolutions with ReLU, 2x2 max
 * pooling, nearest upsampling with skip concatenation, a final 1x1 convolution to two outputs
 * (object-patch amplitude and phase).  Convolutions are implicit GEMMs on the matrix cores
 * (BF16 operands, FP32 accumulation); master weights FP32 with Adam (0.9, 0.999, 1e-8).
 * Parameters: per layer W[cout][cin][k][k] then b[cout], layers in the order
 *   c<-1, c<-c, 2c<-c, 2c<-2c, 4c<-2c, 4c<-4c, 8c<-4c, 8c<-8c, 4c<-12c, 4c<-4c, 2c<-6c, 2c<-2c,
 *   c<-3c, c<-c (3x3), 2<-c (1x1).
 * Samples: inputs 1 x S x S, targets 2 x S x S (float32 on the host, stored BF16 on the device). */
typedef struct ptyk_net ptyk_net;
int ptyk_net_create(int32_t device, int32_t size, int32_t base_channels, int32_t batch, uint64_t seed, ptyk_net** out);
void ptyk_net_destroy(ptyk_net* net);
const char* ptyk_net_last_error(const ptyk_net* net);
int64_t ptyk_net_num_params(const ptyk_net* net);
int ptyk_net_get_params(ptyk_net* net, float* out);
int ptyk_net_set_params(ptyk_net* net, const float* in); /* also resets the Adam moments */
/* Loss = sum_p w(p) (y - t)^2 / (batch * 2 * sum_p w(p)); w is S x S, shared by both output
 * channels and every sample (NULL: uniform). */
int ptyk_net_set_loss_weight(ptyk_net* net, const float* weight);
int ptyk_net_set_data(ptyk_net* net, const float* inputs, const float* targets, int64_t n);
/* Streaming form of ptyk_net_set_data: room for n samples, then slices [offset, offset + count). */
int ptyk_net_reserve_data(ptyk_net* net, int64_t n);
int ptyk_net_put_data(ptyk_net* net, int64_t offset, const float* inputs, const float* targets, int64_t count);
/* Adam steps over stored samples order[0..count) (count a multiple of the batch). */
int ptyk_net_train(ptyk_net* net, const int64_t* order, int64_t count, float lr, double* mean_loss);
int ptyk_net_evaluate(ptyk_net* net, const int64_t* order, int64_t count, double* mean_loss);
int ptyk_net_predict(ptyk_net* net, const float* inputs, int64_t n, float* outputs);
/* Same, from raw diffraction amplitudes: the input log(1 + A) / max is formed on the device. */
int ptyk_net_predict_amplitudes(ptyk_net* net, const float* amplitudes, int64_t n, float* outputs);
/* Loss (mean squared error) and its parameter gradient for one batch. */
int ptyk_net_gradient(ptyk_net* net, const float* inputs, const float* targets, double* loss, float* grad);

/* Batched in-place 2D FFT of n frames of M x M complex64 (host buffers), forward
 * exp(-2 pi i kn/M), inverse unnormalised (CudaFFTPlan::getC2CPlan, CudaFFTPlan.cpp:63-101). */
int ptyk_fft2(ptyk_ctx* ctx, float* data, int64_t n, int inverse);

/* Simulator (Diffractions::simulate, Diffractions.cpp:80-116): for each of n positions,
 * |FFT2(bilinear(obj, pos) * probe)| (squared = 0) or its square (squared = 1).
 * obj is rows x cols, probe M x M, pos n x (row, col); out n x M x M float32. */
int ptyk_simulate(ptyk_ctx* ctx, const float* obj, int32_t rows, int32_t cols, const float* probe,
                  const float* pos_rc, int64_t n, int32_t squared, float* out);

/* Preprocessing of on-disk diffraction frames (DiffractionLoader::run, DiffractionLoader.cpp:55-131;
 * DiffractionsKernels.cu:100-118; utilitiesKernels.cu:841-851, 1945-1963) on the GPU: count raw
 * frames of raw_rows x raw_cols float32 (host) become count frames of M x M float32 (host, M = the
 * handle's frame size = symmetric_array_size): the S x S window centred at (center_row, center_col)
 * (zeros beyond the raw frame; a raw frame that is already M x M is taken whole), rotated by 90
 * degrees when rotate90 % 4 == 1 (the only rotation the reference implements), FFT-shifted when
 * fft_shift, and, when threshold_sqrt, the d_preprocessDiffractions step (non-finite -> 0, values
 * below threshold -> 0, else minus threshold, then the square root when square_root; the
 * reference's load path skips this step, Diffractions.cpp:226-229). */
typedef struct {
  int32_t raw_rows, raw_cols;     /* rawFileSize (-nx, -ny) */
  int32_t center_row, center_col; /* symmetric_array_center (-qx, -qy) */
  int32_t rotate90;               /* rotate_90_times */
  int32_t fft_shift;              /* FFT_SHIFT (-fftShiftData) */
  int32_t threshold_sqrt;
  int32_t square_root;            /* SQUARE_ROOT (-sqrtData) */
  float threshold;                /* threshold_raw_data (-threshold) */
  int32_t reserved[3];
} ptyk_ingest_params;
int ptyk_preprocess_frames(ptyk_ctx* ctx, const float* raw, int64_t count, const ptyk_ingest_params* params,
                           float* out);

/* Allocation-free simulator (no reference counterpart; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving or data-generation loops).  Same results as
 * ptyk_simulate, but obj (rows x cols complex64), probe (M x M complex64), pos (n x (row, col)
 * float32) and out (n x M x M float32) are DEVICE pointers, the call only enqueues on the handle's
 * stream (ptyk_synchronize waits), and the exit waves are staged in a fixed device array of
 * PTYK_NOALLOC_SCRATCH_BYTES per GPU instead of allocated memory: n x M x M x 8 bytes must fit in
 * it, else PTYK_EINVAL.  No hipMalloc/hipFree or allocating call is made.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams or threads) race.  Positions are
 * not validated (device memory); sample indices are clamped to the canvas for any value. */
#define PTYK_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int ptyk_simulate_noalloc(ptyk_ctx* ctx, const float* obj, int32_t rows, int32_t cols, const float* probe,
                          const float* pos, int64_t n, int32_t squared, float* out);

int ptyk_synchronize(ptyk_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* PTYK_H */
