// Internal host-side declarations shared by the engine's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace ptyk {

// Supported transform lengths (square frames M x M).
inline bool frame_size_supported(int m) {
  return m == 16 || m == 32 || m == 64 || m == 128 || m == 256 || m == 512;
}

// ---- batched 2D FFT of n frames of m x m complex64, in place (rows then columns).
hipError_t fft2_batch(float2* data, long n_frames, int m, bool inverse, const float2* tw, hipStream_t s);
// The column half of fft2_batch alone (m in {128, 256, 512}), for callers whose row half is fused
// into a kernel of their own.
hipError_t fft_cols_batch(float2* data, long n_frames, int m, bool inverse, const float2* tw, hipStream_t s);
// Forward 2D FFT of two frames a and b (any sizes fft2_batch takes) into dst[0..2 m^2), out of
// place: the row pass reads a and b, so they need not be copied next to each other first.
hipError_t fft2_pair(float2* dst, const float2* a, const float2* b, int m, const float2* tw, hipStream_t s);
// The row half of fft2_pair alone (m in {128, 256, 512}): dst holds the row spectra of a and b.
hipError_t fft_rows_pair(float2* dst, const float2* a, const float2* b, int m, const float2* tw, hipStream_t s);

// ---- DM (difference map) per-iteration launches; see ptyk_dm.hip for the equations.
struct DmArgs {
  float2* psi;            // N x M x M exit-wave state
  float2* obj;            // object canvas rows x cols
  float2* probe;          // mode 0, M x M
  const __half* amp;      // N x M x M amplitudes (FP16)
  const float* amp_scale; // N per-frame scales
  const int2* roi;        // N rounded ROI origins (row, col)
  const int* row_ptr;     // obj_rows + 1 : CSR of frames covering each object row
  const int* row_frames;  // N*M frame ids, ascending per object row
  double2* probe_partial; // probe_chunks x M x M
  double* obj_partial;    // obj_rows * obj_col_chunks
  double* probe_partial_sum;  // M (one per probe row block)
  float* factors;         // device: [0] = 1/objMax cached, [1] = 1/probeMax cached
  double* scal;           // device: [0] objSum, [1] probeSum
  const float2* tw;
  long n_frames;
  int m, obj_rows, obj_cols;
  int probe_chunks, frames_per_chunk;
  int phase_constraint, update_probe;
};
// HIP-event marks on the engine stream; the time from mark i to mark i+1 is charged to
// the category given at mark i (category -1 closes the sequence).
enum { DM_T_ROWS_FWD, DM_T_COLS, DM_T_ROWS_INV, DM_T_PROBE_GATHER, DM_T_OBJ_GATHER, DM_T_PROBE_APPLY,
       DM_T_OBJ_APPLY, DM_T_FINISH, DM_T_COUNT };
struct LaunchTimer {
  hipEvent_t* ev = nullptr;
  int* cat = nullptr;  // cap entries
  int cap = 0, n = 0;
  void mark(hipStream_t s, int category) {
    if (n < cap) {
      hipEventRecord(ev[n], s);
      cat[n++] = category;
    }
  }
  // Start/stop pairs recorded by a kernel's own dispatch (hipExtLaunchKernelGGL): the kernel's
  // own duration, without the stream's event processing that the marks' intervals carry.
  hipEvent_t* kev = nullptr;  // 2 kcap entries
  int* kcat = nullptr;        // kcap entries
  int kcap = 0, kn = 0;
  bool own(int category, hipEvent_t* start, hipEvent_t* stop) {
    if (kn >= kcap) return false;
    *start = kev[2 * kn];
    *stop = kev[2 * kn + 1];
    kcat[kn++] = category;
    return true;
  }
};
typedef LaunchTimer DmTimer;

hipError_t dm_iteration(const DmArgs& a, hipStream_t s, DmTimer* tm /* nullable */);
int dm_obj_col_chunks(int obj_cols);

// Wave-scheduled DM (ptyk_dm_wave.hip), M in {128, 256, 512}.
struct DmWaveArgs {
  float2* psi;
  float2* obj;
  float2* probe;
  const __half* amp;
  const float* amp_scale;
  const int2* roi;
  const int* wave_row_ptr;    // device: n_waves x (obj_rows + 1) offsets into wave_row_frames
  const int* wave_row_frames; // device: frame ids, ascending per (wave, object row)
  const int* wave_row_lo;     // host: first object row touched by each wave
  const int* wave_row_hi;     // host: one past the last
  double2* d_obj;             // FP64 object increment accumulator (canvas)
  double2* d_probe;           // FP64 probe increment accumulator (M x M)
  double* obj_partial;
  double* probe_partial_sum;
  float* factors;
  double* scal;
  const float2* tw;
  long n_frames;
  int m, obj_rows, obj_cols, wave, n_waves;
  int phase_constraint, update_probe;
};
bool dm_wave_supported(int m);
int dm_wave_probe_blocks(int m);
hipError_t dm_wave_iteration(const DmWaveArgs& a, hipStream_t s, DmTimer* tm /* nullable */);
// MLs forward FFT + multi-mode modulus step + inverse FFT of K mode-major planes of G frames in
// three passes (ptyk_dm_wave.hip). mls_modulus_blocks: error partials per frame, 0 when (m, K, cb)
// is not supported (m in {128, 256, 512}, K <= 8, cb in {8, 16, 32}: the column width of the
// middle pass's 64-row blocks).
// exit: when given, the first pass forms the exit waves itself (psi_k = P_k O_r, or O_r vp_j for
// mode 0 with the sub-pixel shift) and writes the object patches op (and op2); otherwise psi
// already holds them.
struct MlsExitArgs {
  const float2* obj;
  int cols;
  const float2* probe;  // K x M x M
  const float2* vp;     // G x M x M, mode 0's shifted probe (subpix)
  const int2* roi;
  const int* gf;
  int subpix;
  float2* op;
  float2* op2;  // nullable
  // nullable: mode 0's workgroups also write the column derivative of each object-patch row,
  // dX = IFFT_row(FFT_row(O_r) 2 pi i f_col) / M (the position step's first pass)
  float2* dx;
  // the exit-wave row pass covers planes [p0, p0 + np) of the K G (np = 0: all of them; a
  // partial range needs the wave-tile pass, m in {128, 256} with row_waves)
  int p0 = 0, np = 0;
  // mode 0's planes (wave-tile pass only): 0 its exit waves, patches and dX; 1 the patches and
  // dX only (no exit wave: its illumination probes may not be formed yet); 2 its exit waves only,
  // from the patches op already holds
  int mode0 = 0;
};
int mls_modulus_blocks(int m, int k, int cb);
// The exit-wave row pass alone over planes [p0, p0 + np) (wave tiles, m in {128, 256}): the
// probe modes 1..K-1 of a group can be formed before the group's mode 0, whose inputs (the
// illumination probes, the patches the position step still reads) become final later.
hipError_t mls_exit_rows(float2* psi, int G, int m, const MlsExitArgs* exit, int p0, int np, const float2* tw,
                         hipStream_t s, LaunchTimer* tm = nullptr);
// MLs launch categories of ptyk_mls_iteration_timed (main stream only; the second stream's
// launches overlap them and are not charged).
// MLS_T_OWN_A/B/C: the own kernel time of the exit-wave row pass, the modulus column pass and the
// inverse row pass (start/stop events of their dispatches), beside the marks' intervals.
enum { MLS_T_ITER, MLS_T_PROBES, MLS_T_PASS_A, MLS_T_PASS_B, MLS_T_PASS_C, MLS_T_PROBE_UPDATE, MLS_T_PRECOND,
       MLS_T_LSQ, MLS_T_STEP, MLS_T_MODES, MLS_T_JOIN, MLS_T_OWN_A, MLS_T_OWN_B, MLS_T_OWN_C, MLS_T_COUNT };
typedef LaunchTimer MlsTimer;
// pu_j = IFFT(exp(-2 pi i sign (sub f)) / M^2 FFT(chi_j conj(op_j))) for G planes in two passes,
// the separable ramp applied as row shifts then column shifts (m in {128, 256, 512}; ptyk_dm_wave.hip).
hipError_t mls_shift_passes(float2* pu, const float2* chi, const float2* op, int G, int m, const float2* sub,
                            const int* gf, float sign, const float2* tw, hipStream_t s);
// c_modes: mode planes the inverse row pass covers (K, or 1 when modes 2..K are left in their
// column-transformed state for mls_mode_updates).
// shift: when given with c_modes == 1, the inverse row pass also forms the probe update's row
// shifts of chi_0 conj(op) into pu (the first of mls_shift_passes' two passes; mls_shift_cols
// finishes them).
struct MlsShiftArgs {
  float2* pu;
  const float2* op;
  const float2* sub;
  float sign;
};
// row_waves: the exit-wave and inverse row passes run as wave tiles for m <= 256 (bitwise the
// same results as the workgroup-tile passes).
hipError_t mls_modulus_passes(float2* psi, int K, int G, int m, int cb, bool row_waves, const __half* amp,
                              const float* amp_scale, const int* gf, double* part, const float2* tw,
                              const MlsExitArgs* exit /* nullable */,
                              int c_modes, const MlsShiftArgs* shift /* nullable */, hipStream_t s,
                              MlsTimer* tm = nullptr);
// The column shifts of mls_shift_passes alone (after a fused row pass).
hipError_t mls_shift_cols(float2* pu, int G, int m, const float2* sub, const int* gf, float sign, const float2* tw,
                          hipStream_t s);
// Probe updates of modes 2..K (P_k += mean_j w_j chi_k,j conj(op_j) * beta) straight from the
// column-transformed planes: the inverse row pass and the frame sum in one kernel over chunks of
// fpc frames (partials in part, mls_pum_part_len elements), then the chunk sum in frame order
// (m in {128, 256, 512}, fpc >= MLS_PUM_FPC_MIN; ptyk_dm_wave.hip).  beta: the mean probe step
// (re, im).  row_waves (m <= 256): the chunk's frames spread over the waves of a workgroup, whose
// sums add in wave order.
constexpr int MLS_PUM_FPC_MIN = 2;
inline long mls_pum_part_len(int g, int k, long mm) {  // room for any fpc >= MLS_PUM_FPC_MIN
  return (long)((g + MLS_PUM_FPC_MIN - 1) / MLS_PUM_FPC_MIN) * (k > 1 ? k - 1 : 1) * mm;
}
hipError_t mls_mode_updates(float2* probe, const float2* psi, const float2* op, const float* illum, const int2* roi,
                            const int* gf, int K, int G, int m, int cols, float fw, const double* beta, float2* part,
                            int fpc, bool row_waves, const float2* tw, hipStream_t s);

// ---- ePIE: one persistent workgroup runs all positions of one iteration in order.
struct EpieArgs {
  float2* obj; float* obj_int;          // canvas and |O|^2 cache
  float2* probe; float* probe_int;      // K x M x M
  float2* psi_old; float2* psi;         // K x M x M scratch
  const __half* amp; const float* amp_scale;
  const float* amp_sum;                 // per frame sum of amplitudes (RMS normalisation)
  const float2* pos;                    // float positions (row, col)
  const int* order;                     // position order for this iteration
  double* scal;                         // [0] objMax, [1] probeMax, [2] error accumulator
  const float2* tw;
  int n_pos, m, k_use, k_total, obj_rows, obj_cols;
  int phase_constraint, update_probe, modal, rms, tex_quantize;
};
hipError_t epie_iteration(const EpieArgs& a, hipStream_t s);

// ---- MLs (maximum likelihood with least-squares steps), ptyk_mls.hip.
// A host thread that enqueues a group's side-stream work (event waits, launches, event records)
// while the calling thread goes on enqueueing the main stream's: the HIP calls of one iteration
// are shared between two host threads.  Tasks run in the order they were posted; drain() returns
// once every posted task has run, with the first error any of them returned (then cleared).
class SideEnqueuer {
 public:
  explicit SideEnqueuer(int device);
  ~SideEnqueuer();
  SideEnqueuer(const SideEnqueuer&) = delete;
  SideEnqueuer& operator=(const SideEnqueuer&) = delete;
  void post(std::function<hipError_t()> task);
  hipError_t drain();

 private:
  void run();
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::deque<std::function<hipError_t()>> q_;
  long posted_ = 0, done_ = 0;
  hipError_t err_ = hipSuccess;
  bool stop_ = false;
  int device_;
  std::thread th_;
};

// Slots of MlsDev::gscal (device FP64 scalars written and read by the MLs kernels).
enum { MLS_S_ONORM = 0, MLS_S_TMP = 1, MLS_S_BO = 2 /* re, im */, MLS_S_BP = 4 /* re, im */, MLS_S_PMAX = 6,
       MLS_S_COEF = 16 /* one per frame of the group */ };
struct MlsDev {
  // geometry
  int n = 0, m = 0, k = 0, rows = 0, cols = 0;
  int gmax = 0;  // capacity (frames) of the group scratch
  // MLs parameters (ReconstructionParams, Parameters.h:128-146)
  int pps = 20, probe_reconstruct = 1, object_reconstruct = 1, variable_probe = 1, variable_intensity = 1;
  int variable_probe_modes = 1, apply_subpix_shift = 1, apply_multimodal_update = 0, n_objects = 1, n_rec = 1;
  int box[4] = {0, 0, 0, 0};
  double delta_p = 0.1, beta_lsq = 0.9, beta_object = 1.0, beta_probe = 1.0;
  float max_illum = 0.f;  // Probe::MAX_ILLUM, fixed by initEvo
  // launch shapes (tuning knobs, same results): column width of the modulus pass's blocks, frames
  // per chunk of the probe-mode update
  int mod_cb = 16, pum_fpc = 4;
  bool row_waves = true;  // wave-tile row passes and probe-mode update (m <= 256)
  // context arrays (not owned)
  float2* obj = nullptr;
  float2* probe = nullptr;
  const __half* amp = nullptr;
  const float* amp_scale = nullptr;
  const float2* tw = nullptr;
  // MLs state (owned)
  float2* vmode = nullptr;      // variable (extra) probe mode, M x M
  double* evo = nullptr;        // 2 x N probe-evolution coefficients
  float* illum = nullptr;       // cached illumination canvas (Probe::p_object)
  float2* pos = nullptr;        // N scan positions, px, not min-shifted
  float2* pos_o = nullptr;      // N initial positions
  int2* roi = nullptr;          // N ROI origins (0-based)
  float2* sub = nullptr;        // N sub-pixel shifts (.x column, .y row)
  float2* step_obj = nullptr;   // N LSQ object steps
  float2* step_probe = nullptr; // N LSQ probe steps
  float* ferr = nullptr;        // N Fourier errors
  int* flag = nullptr;          // ROI outside the canvas
  // scratch
  float2 *op = nullptr, *vp = nullptr, *psi = nullptr, *pu = nullptr, *tmp = nullptr, *tmp2 = nullptr;
  float2 *spec = nullptr, *pum = nullptr, *acc = nullptr, *precond = nullptr;
  double* part = nullptr;
  long part_cap = 0;
  double* ftot = nullptr;       // per-frame totals of the group (gmax x 8)
  double* gscal = nullptr;
  int* gframes = nullptr;       // frames of all groups of one iteration
  // position step on a second stream (it reads only the group's exit waves, illumination
  // probes and object patches, and writes only these and the positions)
  float2* dxb = nullptr;        // column-derivative images, gmax frames
  // second object-patch and dX buffers: odd groups use them, so a group's patches can be formed
  // while the previous group's side work still reads its own
  float2* op_alt = nullptr;
  float2* dxb_alt = nullptr;
  double* pos_part = nullptr;   // partials, gmax x (m / 16) x 4
  double* pos_ftot = nullptr;   // per-frame totals, gmax x 4
  float2* pum_part = nullptr;   // chunk partials of the probe-mode updates (mls_pum_part_len)
  double* vp_part = nullptr;    // variable-probe update partials on the second stream
  double* vp_ftot = nullptr;    // its per-frame totals, gmax
  hipStream_t side = nullptr, side2 = nullptr;  // second-stream work of a group (streams A and B)
  hipEvent_t ev_side[4] = {nullptr, nullptr, nullptr, nullptr};  // fork points of the second stream's work
  hipEvent_t ev_vpdone = nullptr;                                 // end of the variable-mode update
  hipEvent_t ev_join[2] = {nullptr, nullptr};
  // Group graphs: the work of one group (both streams) captured once per group shape and flag set
  // and replayed for every group of that shape.  A replay takes its frame lists from gsel, which
  // the graph's first kernel fills from the iteration's table (gframes, goffs_dev) at the group
  // index gsel[0] and then advances.
  int* goffs_dev = nullptr;  // group offsets of the iteration (n + 1)
  int* gsel = nullptr;       // [0] next group index, [1, 1 + gmax) this group's frames, then the next group's
  std::map<uint64_t, hipGraphExec_t> graphs;
  uint64_t alloc_gen = 0;    // bumped when the group scratch moves (captured graphs hold its addresses)
  bool use_graphs = false;
  // side-stream work enqueued from a second host thread (not under graph capture)
  std::unique_ptr<SideEnqueuer> enq;
};
constexpr int MLS_SIDE_FORKS = 4;
int mls_chunks(int m);  // frame chunks of the per-frame reductions
// Probe::initEvo: normalisations and the initial illumination (host-synchronous).
hipError_t mls_init(MlsDev& d, const double* frame_power /* host, n */, hipStream_t s);
// One MLs iteration; group g covers d.gframes[goffs[g] .. goffs[g+1]) (goffs on the host).
hipError_t mls_iteration(MlsDev& d, int iter, const int* goffs, int n_groups, hipStream_t s,
                         MlsTimer* tm = nullptr);
// Partial-sum capacity (doubles) the MLs kernels need beyond the per-group reductions.
long mls_part_min(int k);
hipError_t mls_roi(MlsDev& d, hipStream_t s);
// Destroys the captured group graphs (before the scratch they address moves or is freed).
void mls_drop_graphs(MlsDev& d);

// ---- learned warm start: weighted average of (amplitude, phase) patches at the ROIs of the
// selected frames sel[0..n) into obj; pixels whose weight sum is <= rel_thr * max_q
// sum_k |P_k(q)|^2 take `fill` (ptyk_metrics.hip).
hipError_t stitch_patches(float2* obj, const float* patches, const float2* probe, int k, int m, const int2* roi,
                          const int* sel, int n, int rows, int cols, float* pint, float rel_thr, float2 fill,
                          hipStream_t s);
// ---- normalised object error (ptyk_metrics.hip); out[4] = {e, Re gamma, Im gamma, pixels}.
hipError_t object_error(const float2* obj, const float2* truth, const float2* probe, int k, int m, const int2* roi,
                        int n, int rows, int cols, double frac, float* cov, float* pint, double* part, int part_cap,
                        double* out, hipStream_t s);

// ---- frame ingest (ptyk_ingest.hip): raw count x raw_rows x raw_cols frames (device) -> count x
// size x size preprocessed frames (device).
struct IngestArgs {
  int raw_rows, raw_cols, size, center_row, center_col;
  int rotate, fft_shift, threshold_sqrt, square_root;
  float threshold;
};
hipError_t ingest_frames(const float* raw, float* out, long count, const IngestArgs& a, hipStream_t s);

// ---- allocation-free entry points (ptyk_noalloc.hip): the per-device address of the fixed
// scratch array, looked up by noalloc_setup (with the device current) at handle creation.
hipError_t noalloc_setup(int device);
void* noalloc_scratch(int device);
size_t noalloc_scratch_bytes();

// ---- misc kernels (ptyk_misc.hip)
hipError_t quantize_frames(const float* src_dev, __half* dst, float* scales, long n, int m, hipStream_t s);
hipError_t dequantize_frames(const __half* src, const float* scales, float* dst, long n, int m, hipStream_t s);
hipError_t frame_sums(const __half* amp, const float* scales, float* sums, long n, int m, hipStream_t s);
hipError_t mean_amplitude(const __half* amp, const float* scales, float* avg, long n, int m, hipStream_t s);
hipError_t probe_sinc_gauss(float2* probe, int m, float window, float factor, hipStream_t s);
hipError_t modulus_project_single(float2* f, const float* amp, int m, hipStream_t s);
hipError_t probe_normalize(float2* probe, float* probe_int, int k_total, int m, double* scal_out, hipStream_t s);
hipError_t sum_abs2(const float2* x, long count, double* out, hipStream_t s);
hipError_t max_abs2(const float2* x, long count, double* out, hipStream_t s);
hipError_t intensities(const float2* x, float* out, long count, hipStream_t s);
hipError_t simulate_exit_waves(const float2* obj, int obj_rows, int obj_cols, const float2* probe, const float2* pos,
                               float2* psi, long n, int m, int tex_quantize, hipStream_t s);
hipError_t abs_frames(const float2* psi, float* out, long count, int squared, hipStream_t s);
hipError_t modes_from_base(float2* probe, float* probe_int, int k_total, int m, float factor, hipStream_t s);
hipError_t orthogonalize_modes(float2* probe, float2* tmp, int k_total, int m, hipStream_t s);

}  // namespace ptyk
