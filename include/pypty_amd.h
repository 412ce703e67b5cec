/*
 * pypty_amd — C ABI of the MI355X (gfx950) iterative-ptychography hot path.
 *
 * Every pointer is a device pointer owned by the caller unless stated
 * otherwise; `stream` is a hipStream_t passed as void*.  No function
 * allocates device memory: scratch space is the caller's `workspace`, whose
 * size pty_workspace_bytes() reports.  Functions return PTY_OK (0) or a
 * negative status; pty_status_string() names it.  Complex data are complex64
 * (interleaved float re, im).  Windows are n x n with n in
 * {16, 32, 48, 64, 96, 128} for the LDS-resident kernels; the full pass
 * (pty_loss_and_direction) also takes n = 192 and 256, whose tiles live in
 * workspace memory.
 *
 * Data layouts
 *   object, planar     (S, K, Ny, Nx)   plane = s*K + k
 *   object, reference  (Ny, Nx, S, K)   the layout objective.py keeps
 *   probe              (M, n, n)        mode-major
 *   positions          (P, 2) int32     integer corner (y, x) of each window
 *   corrections        (P, 2) float32   sub-pixel (y, x) correction
 *   measured           (P, n, n) float32, DC at (n/2, n/2) (fftshifted)
 */
#ifndef PYPTY_AMD_H
#define PYPTY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTY_OK 0
#define PTY_ERR_ARG -1        /* bad argument / shape */
#define PTY_ERR_SIZE -2       /* window size not supported */
#define PTY_ERR_HIP -3        /* HIP runtime error (launch, event) */
#define PTY_ERR_BOUNDS -4     /* a window leaves the object */
#define PTY_ERR_WORKSPACE -5  /* workspace missing or too small */

/* most aberration coefficients per row (pty_problem.n_aberrations) */
#define PTY_MAX_ABERRATIONS 64

/* fft modes */
#define PTY_FFT_FWD 0        /* fft2                        fft.py:114 */
#define PTY_FFT_INV 1        /* ifft2 (1/n^2)               fft.py:136 */
#define PTY_FFT_FWD_SHIFT 2  /* fftshift(fft2(x))           fft.py:91  */
#define PTY_FFT_INV_ISHIFT 3 /* ifft2(ifftshift(x))         fft.py:158 */

/* loss families (objective.py:527-567) */
#define PTY_LOSS_LSQ_SQRT 0   /* (sqrt I - sqrt D)^2, the reference default */
#define PTY_LOSS_ML 1         /* Poisson: I - D log I */
#define PTY_LOSS_LSQ 2        /* (I - D)^2 */
#define PTY_LOSS_LSQ_SQRT_2 3 /* 0.5 (D - I)^2 / D for D > 0.5, I elsewhere */
/* compressed data (objective.py:560-585): measured rows hold the J virtual
 * detector coefficients D_j; the model's are c_j = sum_pix I * mask_j */
#define PTY_LOSS_LSQ_COMPRESSED 4     /* (c - D)^2 */
#define PTY_LOSS_GAUSS_COMPRESSED 5   /* (sqrt c - sqrt D)^2 */
#define PTY_LOSS_POISSON_COMPRESSED 6 /* c - D log c */

/* rows of `measured` */
#define PTY_MEASURED_BY_POSITION 0 /* row p holds position p's pattern */
#define PTY_MEASURED_BY_SEQUENCE 1 /* row i holds the pattern of sequence[i] (a streamed batch) */

/* object layouts */
#define PTY_LAYOUT_REFERENCE 0
#define PTY_LAYOUT_PLANAR 1

const char* pty_status_string(int status);
int pty_version(void);
/* 1 if the LDS-resident kernels support window size n, else 0 */
int pty_supported_window(int n);

/* Batched in-place 2D FFT of `batch` contiguous n x n complex64 tiles.
 * Replaces fft.fft2 / ifft2 / shift_fft2 / ifft2_ishift (pypty/fft.py:91-178)
 * for windows; the Python mirror moves the transformed axes innermost. */
int pty_fft2_c64(void* data, int64_t batch, int n, int mode, void* stream);

/* x <- ifft2(mask * fft2(x)) per tile, mask: n x n float32, natural
 * (unshifted) frequency order.  The window band-limit of
 * objective.py:439-441 and multislice_core.py:188-190. */
int pty_bandlimit_c64(void* data, int64_t batch, int n, const float* mask, void* stream);

/* full[p, y0+i, x0+j] += patches[b, p, i, j] for every window b, plane p
 * (overlapping windows accumulate).  full: planar (planes, ny, nx) or
 * reference (ny, nx, planes) per `layout`; patches: (batch, planes, n, n);
 * corners (batch, 2).
 * Replaces multislice_core.scatteradd (pypty/multislice_core.py:206-230). */
int pty_scatteradd_c64(void* full, int ny, int nx, int planes, int layout, const void* patches,
                       int batch, int n, const int32_t* corners, void* stream);

/* Geometry, physics and switches of one loss/gradient evaluation: the
 * arguments of objective.loss_and_direction (pypty/objective.py:47) that the
 * far-field multislice path reads. */
typedef struct pty_problem {
  int n;             /* window (probe) size */
  int n_slices;      /* S */
  int n_obj_modes;   /* K */
  int n_probe_modes; /* M */
  int obj_ny, obj_nx;
  int n_positions;   /* P: rows of positions / corrections / measured */
  int compute_batch; /* windows per launch */
  int algorithm;     /* PTY_LOSS_* */
  int obj_layout;    /* PTY_LAYOUT_* of obj and obj_grad */
  int step_obj, step_probe, step_pos; /* which gradients to produce */
  int shift_active;  /* sub-pixel probe shift (objective.py:352) */
  float loss_weight;
  float slice_distance;  /* dz, Angstrom (single distance) */
  float wavelength;      /* Angstrom */
  float pixel_size_x, pixel_size_y; /* Angstrom */
  float tan_x, tan_y;    /* in-specimen tilt of the propagator */
  float damping_cutoff;  /* band-mask cutoff (utils.py:153) */
  float edge_multiplier; /* first-slice propagator power (>1 active) */
  int n_masks;           /* J virtual detectors of the compressed losses */
  /* per-batch preprocessing of raw patterns (utils.py:587-653, applied by
   * objective.py:400-405 to every batch; not for compressed losses):
   * roll by (data_shift_y, data_shift_x) and zero the edge band the
   * reference zeroes, sum data_bin x data_bin blocks, scale by
   * data_multiplier, zero-pad data_pad on each side.  raw_n: side of the raw
   * patterns (0: n, no preprocessing unless other fields say so); the
   * result must be n x n. */
  int raw_n;
  int data_shift_y, data_shift_x;
  int data_bin;
  int data_pad;
  float data_multiplier; /* 0 is taken as 1 */
  /* Slice checkpointing (no counterpart in the reference, which keeps the
   * incident wave of every slice, multislice_core.py:66): with
   * 0 < checkpoint_slices < n_slices the forward keeps the wave entering
   * every checkpoint_slices-th slice (and every wave of the last segment),
   * and the object gradient recomputes each other segment of slices from its
   * checkpoint before running the adjoint through it.  The wave storage per
   * window and mode falls from n_slices slices to
   * ceil(n_slices / checkpoint_slices) + checkpoint_slices, for up to one
   * more forward pass of compute; the results are bit-identical to those
   * without checkpointing.  0 keeps every slice. */
  int checkpoint_slices;
  /* Batch lanes (no counterpart in the reference, which runs its batches one
   * after another, objective.py:383-725): with lanes = L in 2..4, batch i
   * runs in lane i % L, lane 0 on the caller's stream and lanes 1..L-1 on
   * library streams of the same device and default priority, each lane with
   * its own per-batch buffers, so that one batch's memory-bound kernels
   * overlap another's compute-bound kernels.  The kernels that add into the object, probe and
   * position gradients still run in batch order (events order them across
   * the lanes), so the results are bit-identical to lanes = 1.  The call
   * returns with all work joined to the caller's stream.  A call on a
   * stream that is being captured runs one lane.  0 or 1: one stream.  The
   * workspace holds one set of per-batch buffers per lane; a problem whose
   * per-lane buffers would take more than 128 GiB over all lanes runs fewer
   * lanes (pty_workspace_bytes reports the size for the lanes it runs). */
  int lanes;
  /* 1: also return the gradient of the static background
   * (objective.py:588-589, this_step_static_background); needs
   * pty_buffers.static_background and static_background_grad. */
  int step_static_background;
  /* 1: also return the gradient of the beam current (objective.py:688-698,
   * this_beam_current_step); needs pty_buffers.beam_current and
   * beam_current_grad. */
  int step_beam_current;
  /* Aberrations of the incident beam per position (objective.py:320-326,
   * 459-465): n_aberrations coefficients per row, in Krivanek order
   * (utils.py:1272-1308: C10, C12a, C12b, C21a, C21b, C23a, C23b, ...),
   * n_markers rows; 0 for none.  At most PTY_MAX_ABERRATIONS coefficients. */
  int n_aberrations;
  int n_markers;
  /* 1: also return the gradient of the aberrations (objective.py:673-681,
   * this_step_aberrations_array); needs pty_buffers.aberrations_grad. */
  int step_aberrations;
  /* Per-slice distances (multislice_core.py:69-71, 164-168): 0 or 1 uses
   * slice_distance for every propagation; more: pty_buffers.distances holds
   * n_distances >= n_slices - 1 distances and the propagation after slice s
   * covers distances[s]. */
  int n_distances;
  /* Beam tilts (objective.py:303-307, 328-350, 428-499): tilt_mode as the
   * reference's (iterative.py:135): 0 inside the specimen, 1 after, 2 before,
   * 3 inside and after, 4 all three; n_tilts rows (1, or one per position) of
   * pty_buffers.tilts; 0 for none (tan_x, tan_y then tilt the propagator).
   * With one row, or modes 1 and 2, every window's propagator takes row 0's
   * inside tilt; otherwise each window its own.  Tilts before the specimen
   * (modes 2, 4) are not on this path yet (PTY_ERR_ARG). */
  int tilt_mode;
  int n_tilts;
  /* 1: also return the tilt gradient (objective.py:163-178 of
   * multislice_core.py, objective.py:620-636); needs pty_buffers.tilts_grad. */
  int step_tilts;
} pty_problem;

typedef struct pty_buffers {
  const void* obj;           /* object, layout per obj_layout */
  const void* probe;         /* (M, n, n) */
  const int32_t* positions;  /* (P, 2) */
  const float* corrections;  /* (P, 2) */
  const float* measured;     /* (rows, n, n), (rows, raw_n, raw_n) or (rows, n_masks); rows per measured_order */
  const int32_t* sequence;   /* (n_seq) position indices, each at most once */
  int n_seq;
  double* loss_sse;          /* out [2]: loss (x loss_weight), sse */
  void* obj_grad;            /* out, like obj (written if step_obj) */
  void* probe_grad;          /* out (M, n, n) (written if step_probe) */
  float* pos_grad;           /* out (P, 2): rows of `sequence` assigned (if step_pos) */
  float* sim_patterns;       /* optional out (P, n, n) modelled intensity, or NULL */
  void* workspace;
  size_t workspace_bytes;
  const float* masks;        /* (n_masks, n, n) virtual detectors, zero frequency at (n/2, n/2) */
  int measured_order;        /* PTY_MEASURED_BY_* */
  /* 1: add this call's gradients to those of the previous call made with the
   * same problem and workspace (a pass split into batches whose patterns are
   * streamed in); loss_sse still holds this call's sums only */
  int accumulate;
  /* (n, n) static background b, zero frequency at (n/2, n/2), or NULL
   * (objective.py:359-361, 525-526): the modelled pattern is the mode-mean
   * intensity plus b^2 times the band mask.  Not with the compressed losses. */
  const float* static_background;
  /* (n, n) gradient with respect to b, summed over the pass
   * (objective.py:377, 588-589): 2 * b * loss_weight * the sum of dL/dI over
   * the positions, with dL/dI as the reference forms it (carrying the
   * 1/(M*K) of the mode mean when there is more than one mode). */
  float* static_background_grad;
  /* (n_positions) beam current or NULL (objective.py:446-450): the probe of
   * position i, before its sub-pixel shift, is |beam_current[i]| times the
   * probe. */
  const float* beam_current;
  /* (n_positions) out: rows of `sequence` assigned 2 sign(|bc|) times the
   * sum of Re(dL/dP* conj(P)) over the window and the probe modes, with P
   * the unscaled probe (objective.py:694-696; sign(|bc|) is 1 unless the
   * current is 0); the other rows zero unless accumulate. */
  float* beam_current_grad;
  /* (n_markers, n_aberrations) coefficients, Angstrom (objective.py:320-326)
   * and (n_positions) row of each position: the probe spectrum of position
   * i, after the beam current, is multiplied by
   * exp(-i sum_j A_j(q) aberrations[aberration_marker[i]][j]) times the band
   * mask, A_j = 2 pi / lambda (lambda q)^(n+1) / (n+1) times cos(m phi),
   * sin(m phi) or 1 (utils.py:1333-1367).  As in the reference, a sub-pixel
   * shift (shift_active) starts again from the spectrum before this phase
   * plate (objective.py:471-474), so with a shift the plates do not reach
   * the modelled patterns; the gradient is formed either way. */
  const float* aberrations;
  const int32_t* aberration_marker;
  /* (n_markers, n_aberrations) out: per row the sum over its positions and
   * the probe modes of 2/n^2 Re(G conj(F) conj(-i A_j)), with G the
   * spectrum of dL/dP* after the conjugate plate and F the probe spectrum
   * before it (objective.py:676-679); zero first unless accumulate. */
  float* aberrations_grad;
  /* (n_distances) per-slice distances, Angstrom, when n_distances > 1 */
  const float* distances;
  /* (n_tilts, 6) tilts in radians, columns (y, x) before, inside and after
   * the specimen (the reference's tilt array plus its correction) */
  const float* tilts;
  /* (n_tilts, 6) out: columns 2, 3 (inside) and 4, 5 (after) per the mode;
   * with one row the sum over the windows of the sequence, else the rows of
   * `sequence` (the others zero unless accumulate).  Inside: per window
   * -4 pi d / n^2 sum Im(conj(fft2 u_s) fft2(g_s) q) over the propagations
   * (into slice s from s-1, d its distance, times the edge multiplier for the
   * first), the pixels and the modes, u_s the wave entering slice s and g_s
   * dL/du_s*; after: 2 sum c Im(g conj(e)) with e the exit wave before the
   * tilt, g its gradient after it, c = 2 pi (j - n/2) p / lambda along y, x. */
  float* tilts_grad;
} pty_buffers;

int pty_workspace_bytes(const pty_problem* p, size_t* bytes);

/* One full forward + loss + adjoint pass over `sequence`:
 * replaces objective.loss_and_direction (pypty/objective.py:47-967) for
 * far-field multislice, the lsq_sqrt / ml / lsq / lsq_sqrt_2 losses and the
 * three compressed ones, object / probe / position gradients, one or per-slice
 * distances, tilts inside and after the specimen, with the reference's
 * per-batch data preprocessing. */
int pty_loss_and_direction(const pty_problem* p, const pty_buffers* b, void* stream);

/* Multislice forward on pre-cut windows (pypty/multislice_core.py:9-82).
 * probe_b (B, M, n, n); obj_b (B, S, K, n, n) band-limited windows;
 * propagator (n, n) natural frequency order, band mask included (the
 * reference's master_propagator_phase_space); raised to edge_multiplier for
 * the first slice when that is > 1.  Uses n, n_slices, n_obj_modes,
 * n_probe_modes, edge_multiplier of *p.
 * stash out (B, M, K, S, n, n) incident waves (not conjugated);
 * exit_wave out (B, M, K, n, n).  tables: scratch of pty_tables_bytes(n). */
int pty_tables_bytes(int n, size_t* bytes);
int pty_ms_forward(const pty_problem* p, int batch, const void* probe_b, const void* obj_b,
                   const void* propagator, void* stash, void* exit_wave, void* tables, void* stream);

/* Multislice adjoint (pypty/multislice_core.py:85-200).  g (B, M, K, n, n)
 * holds dL/d(exit)* on entry and dL/d(incident probe)* per (m, k) on exit;
 * stash (from pty_ms_forward) is overwritten with the per-mode window
 * gradients conj(u_s) g, as the reference overwrites waves_multislice;
 * mask (n, n) natural order (mask_clean); grad_b out (B, S, K, n, n):
 * band-limited sum over probe modes of the per-slice object gradient
 * (before the scatter-add). */
int pty_ms_adjoint(const pty_problem* p, int batch, void* g, void* stash, const void* obj_b,
                   const void* propagator, const float* mask, void* grad_b, void* tables,
                   void* stream);

/* Fourier band-limit of whole planes, in place: pypty/utils.py:102-150
 * (fourier_clean) applied over axes (0, 1) of every plane.  data planar
 * (planes, ny, nx); ny and nx products of primes <= 31 with 8*ny, 8*nx <= 64 KiB;
 * mask: normalised frequency radius <= cutoff/2 with a cosine taper of
 * width rolloff/2 when rolloff != 0. */
int pty_fourier_clean_c64(void* data, int64_t planes, int ny, int nx, double cutoff, double rolloff,
                          void* stream);

/* 2D FFT of whole planes (fft.py:91-178 for ny, nx up to 8192 whose prime
 * factors are <= 31), modes as pty_fft2_c64; scratch: device buffer of the
 * data's size. */
int pty_plane_fft2_c64(void* data, void* scratch, int64_t planes, int ny, int nx, int mode,
                       void* stream);

/* L-BFGS inner products (pypty/iterative.py:593-681, 683-690, 873-916):
 * out[p] = sum_i Re(a[p][i] * conj(b[p][i])) accumulated in FP64, for
 * npairs <= 16 pairs of n-element complex64 vectors, or of n-element float32
 * vectors when is_real (a, b: host arrays of device pointers).
 * partial: device scratch of npairs * 512 doubles; out: device, npairs doubles. */
int pty_multidot_c64(const void* const* a, const void* const* b, int npairs, int64_t n, int is_real,
                     double* out, double* partial, void* stream);

/* y <- alpha * y + beta * x over n complex64 (x NULL: y <- alpha * y):
 * the axpy steps of the two-loop recursion and of the trial point. */
int pty_axpby_c64(void* y, const void* x, int64_t n, float alpha, float beta, void* stream);

/* The same over n float32 (any n): the steps of the real blocks (positions,
 * static background, beam current, aberrations), iterative.py:573-681, 770. */
int pty_axpby_f32(float* y, const float* x, int64_t n, float alpha, float beta, void* stream);

/* L1 norms of the object's phase and log-amplitude (pypty/objective.py:1286-1320,
 * compute_full_l1_constraint, called at :811-830): *out = phase_weight *
 * sum |arg obj| + abs_weight * sum |log |obj|| (FP64 sums), and, when grad is
 * not NULL, grad += 0.5 (abs_weight sign(log |obj|) + i phase_weight
 * sign(arg obj)) / conj(obj).  A weight <= 0 contributes nothing.  n complex64
 * elements; partial: device scratch of 512 doubles; out: device, 1 double. */
int pty_l1_constraint_c64(const void* obj, int64_t n, float abs_weight, float phase_weight, void* grad,
                          double* out, double* partial, void* stream);

/* Centre of mass of each of `count` patterns (count, ny, nx) float32 about the
 * pattern centre, the first step of the DPC start (pypty/dpc.py:139-150,
 * fft_based_dpc): comx[i] = sum (x - (nx-1)/2) I_i, comy[i] likewise in y
 * (negated with flip_y, dpc.py:133-134), FP64 sums; comx, comy device arrays
 * of count doubles. */
int pty_center_of_mass(const float* patterns, int64_t count, int ny, int nx, int flip_y, double* comx, double* comy,
                       void* stream);

/* Charge flipping of the object, in place (pypty/objective.py:1014-1110,
 * charge_flip without fancy_sigma, applied by iterative.py:503-506): per plane
 * of obj (n complex64 in `planes` planes, layout PTY_LAYOUT_REFERENCE
 * (ny, nx, planes) or PTY_LAYOUT_PLANAR) the phase arg O and absorption
 * -log(|O| + 1e-10) below delta times their plane's largest value are
 * multiplied by beta, and O becomes exp(-absorption + i phase).
 * partial: device scratch of 130 * planes floats. */
int pty_charge_flip_c64(void* obj, int64_t n, int planes, int layout, float delta_phase, float delta_abs,
                        float beta_phase, float beta_abs, float* partial, void* stream);

/* Gram-Schmidt over the probe modes, in place (pypty/objective.py:1119-1140,
 * make_states_orthogonal, applied by iterative.py:498-499 when
 * keep_probe_states_orthogonal): probe (npix, modes) complex64 with the mode
 * innermost (the reference's (N, N, M)); for i = 1..M-1 and j < i,
 * p_i -= p_j sum(conj(p_j) p_i) / (1e-10 + sum |p_j|^2), FP64 sums. */
int pty_orthogonalize_modes_c64(void* probe, int64_t npix, int modes, void* stream);

/* A line-search step of a phase-only object (pypty/iterative.py:737-738):
 * out = |obj| exp(i (arg obj + step u)) over n complex64, u float32 */
int pty_phase_step_c64(const void* obj, const float* u, float step, void* out, int64_t n, void* stream);

/* out = 2 Re(-i conj(z) g): Wirtinger gradient -> phase gradient
 * (pypty/utils.py:260-276, complex_grad_to_phase_grad); out float32, n elements. */
int pty_phase_grad_c64(const void* g, const void* z, float* out, int64_t n, void* stream);

/* The same pass without a caller workspace and without any allocation (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  b->workspace is ignored: the workspace is a device
 * array of PTY_STATIC_WORKSPACE_BYTES inside the library, one per GPU, held
 * until the process ends.  Problems whose one-lane workspace
 * (pty_workspace_bytes with lanes = 1) exceeds it are rejected with
 * PTY_ERR_WORKSPACE.  The pass runs one batch lane and calls no hipMalloc,
 * hipFree or allocating routine.  Every call on one GPU shares that GPU's
 * array, which keeps the last call's contents: calls that overlap on one GPU
 * race (issue them on one stream), and b->accumulate adds to the gradients of
 * the previous call of the same problem on that GPU. */
#define PTY_STATIC_WORKSPACE_BYTES (64u << 20)
int pty_loss_and_direction_static(const pty_problem* p, const pty_buffers* b, void* stream);

/* Host-side check of the launch plan of pty_loss_and_direction (no
 * counterpart in the reference, whose objective.py:393-710 loop is serial):
 * `calls` passes of problem p over n_seq positions on up to `lanes` batch
 * lanes, issued back to back on one caller stream, are recorded instead of
 * launched (no HIP call is made).  Every buffer range each launch reads or
 * writes is checked to lie in its workspace buffer (per-batch buffers in the
 * launch's lane block), every pair of launches on different lanes that
 * touch one range, one of them writing, to be ordered by the pass's events,
 * and the batches to cover the sequence once.  Returns the number of
 * violations (0: none) or a negative status; report (report_bytes) receives
 * one line per violation, the plan itself with verbose.  omit plants faults
 * in the recorded plan only, to test the check (bit 0: no waits on the batch
 * ordering events, 1: on the fork, 2: on the joins; bit 3: a stash one window
 * smaller); 0 checks the plan the library issues. */
int pty_plan_check(const pty_problem* p, int n_seq, int lanes, int calls, int accumulate, int verbose, int omit,
                   char* report, size_t report_bytes);

/* Kernel timing (HIP events around every launch of the library). */
void pty_timing_enable(int on);
/* Writes up to `cap` entries: names (static strings), total ms, launches;
 * returns the number of kernel kinds; synchronises pending events. */
int pty_timing_read(const char** names, double* total_ms, int64_t* launches, int cap);
void pty_timing_reset(void);

/* ---- Learned warm start (no counterpart in the reference: a small CNN that
 * maps a diffraction pattern to the projected phase around the probe, whose
 * stitched predictions replace the flat starting object `obj` of
 * pypty/iterative.py:56-250).  BF16 operands are raw bit patterns
 * (uint16), rounded to nearest even. */

/* C (M x N, row stride ldc, FP32) = op(A) x op(B) with BF16 operands and FP32
 * accumulation on the matrix cores.  op(A)[m][k] = ta ? A[k*lda + m] :
 * A[m*lda + k]; op(B)[k][n] = tb ? B[n*ldb + k] : B[k*ldb + n].  A long
 * reduction over few output tiles is split into as many parts (up to 256) as
 * `partial` (partial_elems floats) holds M x N partial products of, summed in
 * a fixed order; without `partial` it runs unsplit. */
int pty_gemm_bf16(int ta, int tb, int M, int N, int K, const void* A, int lda, const void* B, int ldb, float* C,
                  int ldc, float* partial, size_t partial_elems, void* stream);
/* cols[p][(ky*ks + kx)*C + c] (BF16) = X[b][oy*stride - pad + ky][ox*stride - pad + kx][c]
 * (0 outside) for X (batch, H, W, C) FP32 and p = (b*Ho + oy)*Wo + ox; ks*ks*C must be a
 * multiple of 8, X and cols 16-byte aligned. */
int pty_cnn_im2col(const float* X, int batch, int H, int W, int C, int ks, int stride, int pad, void* cols,
                   void* stream);
/* the adjoint of im2col: dX (batch, H, W, C) = sum of the dcols entries that read each element
 * (C a multiple of 4, dcols and dX 16-byte aligned) */
int pty_cnn_col2im(const float* dcols, int batch, int H, int W, int C, int ks, int stride, int pad, float* dX,
                   void* stream);
/* A[p][c] = act(Y[p][c] + bias[c]), act = ReLU if relu else identity; A (FP32) and Abf (BF16) optional */
int pty_cnn_bias_act(const float* Y, int64_t P, int C, const float* bias, int relu, float* A, void* Abf,
                     void* stream);
/* dY = dA * [A > 0] (relu) or dA, as FP32 and BF16 */
int pty_cnn_act_grad(const float* dA, const float* A, int64_t n, int relu, float* dY, void* dYbf, void* stream);
/* db[c] = sum over p of dY[p][c] in a fixed order; partial: device scratch of
 * partial_elems floats (C per block of rows, up to 2048 blocks) */
int pty_cnn_bias_grad(const float* dY, int64_t P, int C, float* db, float* partial, size_t partial_elems,
                      void* stream);
int pty_cnn_to_bf16(const float* x, int64_t n, void* y, void* stream);
/* loss[0] = mean (Y - T)^2 (FP64, fixed order), dY = 2 (Y - T) / n; partial:
 * device scratch of partial_elems doubles (up to 1024 used) */
int pty_cnn_mse(const float* Y, const float* T, int64_t n, float* dY, double* loss, double* partial,
                size_t partial_elems, void* stream);
/* one bias-corrected Adam step on FP32 weights w (BF16 copy wbf optional) */
int pty_cnn_adam(float* w, void* wbf, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                 float beta2, float eps, int step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYPTY_AMD_H */
