/* C-ABI of libpty_hip.so: MI355X (gfx950) kernels for py4DSTEM's single-slice
 * gradient-descent ptychography.
 *
 * The reference has no native interface for this path: its seam is the
 * Python array module swap (py4DSTEM/process/phase/phase_base_class.py:46-84)
 * under SingleslicePtychography.reconstruct()
 * (py4DSTEM/process/phase/singleslice_ptychography.py:606-996).  Each entry
 * point below replaces the reference function(s) cited beside it; the Python
 * package py4dstem_amd binds them with ctypes.
 *
 * Conventions
 *  - Complex arrays are complex64: interleaved float32 (re, im), 8-byte aligned.
 *  - All buffers are device memory owned by the caller; the library keeps no
 *    pointer after a call returns and allocates no device memory (the one
 *    static array of pty_gd_run_noalloc aside).  pty_multi_run creates
 *    runtime objects (a stream, two events, a graph) and waits for its work
 *    before returning; a pty_gd_run with pty_gd_params.pipeline set creates
 *    a second stream and four events once per device (kept until the
 *    process ends) and stays asynchronous; every other entry is
 *    asynchronous and creates none.
 *  - Every call is asynchronous on `stream` (a hipStream_t, NULL = default
 *    stream) and returns 0 on success or a negative PTY_E* code; the message
 *    is available from pty_last_error() on the calling thread.
 *  - Frame edge S must be 32, 64, 128 or 256 (square frames).  256 px frames
 *    run through gradient descent (pty_gd_run and the single-batch entries);
 *    projection sets, probe modes, the batched runner pty_multi_run and the
 *    network entries take 32 to 128 (pty_multi_run and the network: 64 or
 *    128).  Objects need Px >= S and Py >= S.  "Kernel layout" amplitudes /
 *    probe spectra are stored in the FFT's permuted frequency order (see
 *    pty_permute_frames; natural order at 256 px).
 */
#ifndef PTY_HIP_H
#define PTY_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* pty_stream_t; /* hipStream_t */

enum {
  PTY_OK = 0,
  PTY_E_ARG = -1,      /* bad size, null pointer or misalignment */
  PTY_E_UNSUPPORTED = -2, /* frame size not compiled, option not on this path, an
                             object of 2^28 elements or more, or a staged batch
                             beyond 2^31 elements */
  PTY_E_LAUNCH = -3,   /* HIP launch / runtime error */
  PTY_E_WORKSPACE = -4 /* workspace too small */
};

int pty_version(void);
const char* pty_last_error(void);

/* Batched 2-D FFT of `batch` S x S frames, numpy.fft.fft2 / ifft2 semantics
 * (inverse scaled by 1/S^2), natural order in and out.  Replaces np.fft.fft2
 * / ifft2 at phase/utils.py:423, 526, 533 and ptychographic_methods.py:1600,
 * 1633. */
int pty_fft2_c64(const void* in, void* out, int batch, int S, int inverse, pty_stream_t stream);

/* Reorder `count` real S x S frames from natural to kernel layout
 * (direction 0) or back (direction 1). */
int pty_permute_frames(const float* in, float* out, int count, int S, int direction,
                       pty_stream_t stream);

/* probe_hat = fft2(probe) in kernel layout (the fft2 inside fft_shift,
 * phase/utils.py:526). */
int pty_probe_spectrum(const void* probe, void* probe_hat, int S, pty_stream_t stream);

/* Fused forward operator and exit-wave update for one mini-batch
 * (ptychographic_methods.py:1482-1527 _return_shifted_probes /
 * _overlap_projection, 1576-1635 _gradient_descent_fourier_projection,
 * 1736-1816 _forward).  For slot b < B with position idx = order[b]:
 *   ps[b]  = P_shift = ifft2(probe_hat * ramp(frac))    (natural order)
 *   chi[b] = ifft2(A e^{i angle F} - F), F = fft2(P_shift * O_patch)
 *   err[b] = sum (A - |F|)^2                             (float64)
 *   far[b] = F in the far-field staging layout           (skipped if NULL):
 *            with N1 = 8 (S=64) or 16 (S=128) and T = S*S/N1, element
 *            far[b][j*T + t] holds the kernel-layout entry at row t/(S/N1),
 *            column N1*(t % (S/N1)) + j (consecutive threads write
 *            consecutive words; only the position-correction pass reads it)
 * Patches are S x S, corner-centred. */
int pty_forward_adjoint_batch(const void* probe_hat, const void* object, int Px, int Py,
                              const float* positions, int N, const float* amplitudes,
                              const int* order, int B, int S, void* ps, void* chi,
                              double* err, void* far, pty_stream_t stream);

/* Deterministic scatter-add of staged patches into object-shaped sums, in
 * float64 and in batch order per pixel, cast to float32
 * (phase_base_class.py:1905-1968, _sum_overlapping_patches_bincounts, as
 * called at ptychographic_methods.py:1859-1886):
 *   grad = scatter(conj(ps) * chi), spsum = scatter(|ps|^2);
 *   tilemax[t] = max spsum over object tile t, pty_scatter_num_tiles(Px, Py)
 *   entries.  S may also be 32 (stitching of learned-initialiser patches). */
int pty_scatter_num_tiles(int Px, int Py);
int pty_scatter_patches(const void* ps, const void* chi, const float* positions, int N,
                        const int* order, int B, int S, int Px, int Py, void* grad, float* spsum,
                        float* tilemax, pty_stream_t stream);

/* object += step * grad / sqrt(1e-16 + ((1-n) spsum)^2 + (n max spsum)^2)
 * (ptychographic_methods.py:1864-1886). */
int pty_object_update(void* object, const void* grad, const float* spsum, const float* tilemax,
                      int num_tiles, int Px, int Py, float step_size, float normalization_min,
                      pty_stream_t stream);

/* Probe step sum_b conj(O_patch) * chi normalised by sum_b |O_patch|^2
 * (ptychographic_methods.py:1888-1906); reads the pre-update object.
 * `partials` is scratch of pty_probe_partials_bytes(S, B) bytes. */
size_t pty_probe_partials_bytes(int S, int B);
int pty_probe_update(void* probe, const void* chi, const void* object, int Px, int Py,
                     const float* positions, int N, const int* order, int B, int S,
                     float step_size, float normalization_min, void* partials,
                     pty_stream_t stream);

/* Object constraints for a complex object (ptychographic_constraints.py:
 * 29-97, 411-464): optional shrinkage by the mean phase over `mask` (uint8,
 * NULL = none) plus shrinkage_rad, then |O| <= 1 (or |O| = 1 when
 * pure_phase_object).  `partials` needs pty_constraint_partials_bytes(). */
size_t pty_constraint_partials_bytes(void);
int pty_object_constraints(void* object, int Px, int Py, const unsigned char* mask,
                           float shrinkage_rad, int pure_phase_object, void* partials,
                           pty_stream_t stream);

/* Gaussian smoothing of a complex object (ptychographic_constraints.py:
 * 119-151): scipy.ndimage.gaussian_filter along rows then columns with
 * reflect boundaries, float64 sums, float32 after each axis; the phase only
 * when pure_phase_object.  `weights`: device float64 [2 * radius + 1] (the
 * scipy kernel for sigma in pixels); `scratch`: Px * Py * 8 bytes. */
int pty_object_gaussian(void* object, int Px, int Py, const double* weights, int radius,
                        int pure_phase_object, void* scratch, pty_stream_t stream);

/* Butterworth filter of a complex object (ptychographic_constraints.py:
 * 153-201): mean removed, fft2 (any Px, Py: DFT products along rows and
 * columns), times the low/high-pass envelope on the fftfreq grid, ifft2,
 * mean restored.  `scratch`: Px * Py * 8 bytes; `partials`:
 * pty_constraint_partials_bytes(). */
int pty_object_butterworth(void* object, int Px, int Py, double q_lowpass, double q_highpass,
                           double butterworth_order, double sampling_x, double sampling_y,
                           void* scratch, void* partials, pty_stream_t stream);

/* Probe centre-of-mass recentring (ptychographic_constraints.py:894-918). */
int pty_probe_com_constraint(void* probe, int S, pty_stream_t stream);

/* ---- loss and gradients (backward of the forward model) ------------------ */

/* L = sum_b sum (|fft2(P_shift_b * O_patch_b)| - A_b)^2 over the num_slots
 * entries of `order` (indices into the N positions), processed `chunk` slots
 * at a time, and its gradients
 * as dL/dRe + i dL/dIm, each multiplied by `scale`:
 *   grad_object = -2 S^2 scatter(conj(P_shift) chi)
 *       (the un-normalised object term of _gradient_descent_adjoint,
 *        ptychographic_methods.py:1878-1884);
 *   grad_probe  = -2 S^2 ifft2(sum_b fft2(conj(O_patch_b) chi_b) conj(T_b))
 *       (chain rule through fft_shift, phase/utils.py:508-533; equals the
 *        reference's probe term, ptychographic_methods.py:1900-1906, only
 *        for integer positions).
 * `loss` is one float64 on the device; NULL grad pointers skip that part. */
size_t pty_loss_workspace_bytes(int S, int Px, int Py, int num_slots, int chunk);
int pty_loss_gradients(const void* object, const void* probe, int Px, int Py,
                       const float* positions, int N, const float* amplitudes,
                       const int* order, int num_slots, int chunk, int S, double scale,
                       double* loss,
                       void* grad_object, void* grad_probe, void* workspace,
                       size_t workspace_bytes, pty_stream_t stream);

/* Noise-free intensities |fft2(P_shift_b * O_patch_b)|^2 of positions 0..N-1
 * (float32 [N][S][S], natural corner-centred order) for synthetic data; the
 * forward model of _overlap_projection + fft2 (ptychographic_methods.py:
 * 1482-1527, 1597-1600). */
int pty_simulate_intensities(const void* probe_hat, const void* object, int Px, int Py,
                             const float* positions, int N, int S, float* out,
                             pty_stream_t stream);

/* ---- preprocessing ------------------------------------------------------- */

/* Per-pattern centre of mass (row, column) of float32 [N][S][S] intensities,
 * optionally multiplied by mask [S][S] (phase_base_class.py:613-767, the
 * vectorised branch); float64 accumulation. */
int pty_pattern_com(const float* intensities, int N, int S, const float* mask, float* com_x,
                    float* com_y, pty_stream_t stream);

/* Corner-centred amplitudes of the shifted patterns
 * (phase_base_class.py:1347-1460 with preprocess/utils.py:59-121): output
 * pattern k comes from input select[k] (NULL: k); it is shifted periodically
 * and bilinearly by (shift_x, shift_y)[input] (= -fitted CoM) with the
 * reference's float32 arithmetic, then sqrt(max(I, 0)) is written in the
 * kernel layout [N_out][S][S]; sums[k] = float64 sum of the shifted pattern
 * (the host accumulates the mean intensity from them in pattern order). */
int pty_shift_amplitudes(const float* intensities, int N_in, int S, const int* select, int N_out,
                         const float* shift_x, const float* shift_y, float* amplitudes,
                         double* sums, pty_stream_t stream);

/* Resampling and padding of the patterns (phase_base_class.py:223-420:
 * diffraction_intensities_shape with reshaping_method 'fourier', 'bilinear'
 * or 'bin', then padded_diffraction_intensities_shape).  Each is linear
 * along each axis, so pattern b (float32 [Qx][Qy]) becomes
 * out[b] = Re(Mx in[b] My^T) (float32 [Sx][Sy]); Mx complex64 [Sx][Qx] and
 * My complex64 [Sy][Qy] hold the per-axis maps (built on the host by applying
 * the reference's one-dimensional operation to unit vectors).  `workspace`:
 * pty_resample_workspace_bytes(batch, Qy, Sx); batch <= 65535 per call. */
size_t pty_resample_workspace_bytes(int batch, int Qy, int Sx);
int pty_resample_frames(const float* in, int batch, int Qx, int Qy, const void* Mx, const void* My,
                        int Sx, int Sy, float* out, void* workspace, pty_stream_t stream);

/* ---- whole reconstruction loop ------------------------------------------ */

/* pty_problem.object_type */
#define PTY_OBJECT_COMPLEX 0   /* object: complex64 [Px][Py] */
#define PTY_OBJECT_POTENTIAL 1 /* object: float32 potential V [Px][Py], patches exp(i V)
                                  (object_type='potential', ptychographic_methods.py:
                                  40-79, 1522-1523, 1868-1880); gradient descent only */

typedef struct pty_problem {
  int S, Px, Py, N;
  void* object;                          /* [Px][Py] (see object_type), updated in place */
  void* probe;                           /* complex64 [S][S], updated in place */
  const float* positions;                /* float32 [N][2] object pixels */
  const float* amplitudes;               /* float32 [N][S][S], kernel layout */
  const unsigned char* fov_mask_inverse; /* uint8 [Px][Py] or NULL */
  float mean_intensity;
  const float* fourier_mask;             /* float32 [S][S] kernel layout or NULL: the
                                            detector_fourier_mask of the GD projection
                                            (ptychographic_methods.py:1603-1616) */
  int object_type;                       /* PTY_OBJECT_COMPLEX or PTY_OBJECT_POTENTIAL */
} pty_problem;

typedef struct pty_gd_params {
  int num_iter;
  int max_batch;        /* positions per mini-batch, 1..N */
  int fix_probe;
  int fix_probe_com;    /* applied only when the probe is updated */
  int pure_phase_object;
  float step_size;
  float normalization_min;
  float shrinkage_rad;
  int stage;            /* positions per fused launch (0 = whole mini-batch); a
                           performance knob only: results are bitwise identical */
  /* probe constraints after the recentring, when the probe is updated
   * (ptychographic_constraints.py:1091-1147; with modes the aperture acts on
   * mode 0, the amplitude top-hat on every mode) */
  int fix_probe_aperture;         /* Fourier amplitude := probe_aperture (:1004-1033) */
  const float* probe_aperture;    /* float32 [S][S] kernel layout: |fft2(initial probe)| */
  int constrain_probe_amplitude;  /* real-space erf top-hat (:920-957) */
  double probe_amplitude_radius;  /* constrain_probe_amplitude_relative_radius */
  double probe_amplitude_width;   /* constrain_probe_amplitude_relative_width */
  /* object smoothing before the shrinkage and threshold
   * (ptychographic_constraints.py:119-151, 431-435): scipy.ndimage
   * gaussian_filter weights (float64, device, 2 * radius + 1 values), or NULL */
  const double* gaussian_weights;
  int gaussian_radius;
  /* Butterworth filter of the object after the smoothing
   * (ptychographic_constraints.py:153-201, 436-442): 0 = off.  A cut-off of
   * 0 leaves its factor out, as the reference's `if q_lowpass:` does. */
  int butterworth;
  double q_lowpass;          /* A^-1 */
  double q_highpass;         /* A^-1 */
  double butterworth_order;
  double sampling[2];        /* object pixel size (A) along axes 0 and 1 */
  /* potential objects: clip V at zero after the shrinkage
   * (ptychographic_constraints.py:99-118, 458-462) */
  int object_positivity;
  /* with stage < max_batch and no position correction: 1 runs the scatter
   * and probe sums of stage k on a second stream of the library's, beside
   * the fused pass of stage k + 1 (P_shift and chi staged double-buffered,
   * twice the staging workspace).  A performance knob only: results are
   * bitwise identical.  The library creates that stream and four events
   * once per device, and serialises pipelined calls' launch sequences on a
   * device with a lock. */
  int pipeline;
} pty_gd_params;

size_t pty_gd_workspace_bytes(const pty_problem* problem, const pty_gd_params* params);

/* num_iter iterations of phase/singleslice_ptychography.py:847-983 (GD):
 * `orders` holds num_iter permutations of 0..N-1 (int32, device), split into
 * batches by phase/utils.py:539-589; `errors` receives the normalised
 * per-iteration error (float64, device, num_iter entries). */
int pty_gd_run(const pty_problem* problem, const pty_gd_params* params, const int* orders,
               double* errors, void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* Bytes of the static per-GPU workspace of pty_gd_run_noalloc. */
#define PTY_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* pty_gd_run without a caller workspace and without any device allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The workspace is a static device array
 * of PTY_NOALLOC_SCRATCH_BYTES (64 MiB) on each GPU the library runs on, kept
 * until the process ends; its address on the current device is looked up
 * once, on the first call there.  Positions per fused launch
 * (params->stage) are lowered in multiples of 64 until the workspace fits,
 * which leaves the results bitwise equal to pty_gd_run's, and
 * params->pipeline is ignored (one stream, nothing created); a problem that
 * does not fit even at 64 positions per launch (the object's running sums
 * take 24 bytes per pixel) returns PTY_E_WORKSPACE.  Nothing in the array
 * needs to start at zero.  Every call on one GPU uses the same array, so
 * calls that overlap on one GPU (several streams or host threads) race; the
 * caller must order them. */
int pty_gd_run_noalloc(const pty_problem* problem, const pty_gd_params* params, const int* orders,
                       double* errors, pty_stream_t stream);

/* ---- position correction ---------------------------------------------------
 * reconstruct(fix_positions=False): after each mini-batch's object/probe
 * update, phase/ptychographic_methods.py:2071-2225 (_position_correction,
 * called at phase/singleslice_ptychography.py:913-926) moves every position
 * of the batch by the least-squares step of the measured-minus-estimated
 * intensities on the two one-pixel object-shift partials; at the end of each
 * iteration phase/ptychographic_constraints.py:1292-1311 optionally restores
 * the scan's centre of mass.  Limits are in object pixels (the reference's
 * Angstrom values divided by |sampling|). */
typedef struct pty_position_params {
  float* positions;               /* [N][2] = the problem's positions, updated in place */
  const float* positions_initial; /* [N][2] positions at the start of the run */
  float step_size;                /* positions_step_size */
  int clip_update;                /* max_position_update_distance given */
  double max_update;
  int clip_total;                 /* max_position_total_distance given */
  double max_total;
  int center_of_mass;             /* 1: positions -= mean(positions) - initial_com each iteration */
  float initial_com[2];           /* float32 mean of positions_initial */
  int* singular;                  /* device int or NULL: +1 per singular 2x2 system (the
                                     reference raises LinAlgError); the update is then 0 */
} pty_position_params;

size_t pty_gd_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params);

/* pty_gd_run with position correction (params->stage is ignored: every fused
 * launch covers the whole mini-batch). */
int pty_gd_run_positions(const pty_problem* problem, const pty_gd_params* params,
                         const pty_position_params* positions, const int* orders, double* errors,
                         void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* One batch of _position_correction (phase/ptychographic_methods.py:2071-2225):
 * ps = the batch's shifted probes and far = fft2(P_shift * O_patch) of the
 * forward pass (pty_forward_adjoint_batch's outputs, far in its staging
 * layout), object = the object
 * after the update; updates pp->positions[order[b]].  sums (optional, device,
 * [B][5] float64): sum p_x^2, p_x p_y, p_y^2, p_x d, p_y d per slot. */
int pty_position_correction_batch(const pty_position_params* pp, const void* ps, const void* far,
                                  const void* object, int Px, int Py, const float* amplitudes, int N,
                                  const int* order, int B, int S, double* sums, pty_stream_t stream);

/* positions -= float32 mean(positions) - initial_com
 * (phase/ptychographic_constraints.py:1292-1311); initial_com is a host pointer. */
int pty_positions_com(float* positions, int N, const float* initial_com, pty_stream_t stream);

/* ---- projection sets ---------------------------------------------------------
 * reconstruct(reconstruction_method='DM_AP' | 'RAAR' | 'RRR' | 'SUPERFLIP' |
 * 'generalized-projections'): phase/ptychographic_methods.py:1637-1734
 * (_projection_sets_fourier_projection) and :1910-1998
 * (_projection_sets_adjoint), with the (a, b, c) of
 * phase/phase_base_class.py:2005-2108.  Per position with overlap
 * psi = P_shift * O_patch and previous exit wave e (psi when !exit_valid):
 *   F = fft2(c psi + (1 - c) e),  err = sum (A - |F|)^2,
 *   e <- (1 - a - b) e + a psi + b ifft2(A e^{i angle F});
 * then O = scatter(conj(P_shift) e) / norm, P = sum conj(O_patch) e / norm
 * (replacements, no step size). */
typedef struct pty_projection_params {
  double a, b, c;    /* projection_a, projection_b, projection_c */
  void* exit_waves;  /* complex64 [N][S][S], natural order; slot k pairs with order[k];
                        persistent across calls (the reference's _exit_waves) */
  int exit_valid;    /* 0: the first iteration starts from e = psi */
} pty_projection_params;

size_t pty_projection_workspace_bytes(const pty_problem* problem, const pty_gd_params* params);

/* num_iter iterations with one batch of all N positions (params->max_batch
 * must be N; step_size unused; normalization_min, fix_probe, fix_probe_com,
 * pure_phase_object, shrinkage_rad and stage as for pty_gd_run).  `order` is
 * int32 [N] (the reference uses 0..N-1). */
int pty_projection_run(const pty_problem* problem, const pty_gd_params* params,
                       const pty_projection_params* projection, const int* order, double* errors,
                       void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* pty_projection_run with position correction (fix_positions=False;
 * singleslice_ptychography.py:913-926 after the projection-set adjoint,
 * ptychographic_methods.py:2071-2225): the far field of P_shift * O from the
 * iteration's forward object and probe, the updated object, the same
 * position parameters and centre-of-mass constraint as
 * pty_gd_run_positions.  Workspace: pty_projection_positions_workspace_bytes
 * (whole-batch P_shift, chi and far-field staging). */
size_t pty_projection_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params);
int pty_projection_run_positions(const pty_problem* problem, const pty_gd_params* params,
                                 const pty_projection_params* projection, const pty_position_params* positions,
                                 const int* order, double* errors, void* workspace, size_t workspace_bytes,
                                 pty_stream_t stream);

/* The exit-wave step of one batch (slots 0..B-1 of order; exit_waves[k] is
 * slot k): ps[b] = P_shift (natural), err[b], exit waves advanced in place. */
int pty_projection_batch(const void* probe_hat, const void* object, int Px, int Py,
                         const float* positions, int N, const float* amplitudes, const int* order,
                         int B, int S, const pty_projection_params* projection, void* ps,
                         double* err, pty_stream_t stream);

/* ---- mixed state (several incoherent probe modes) ---------------------------
 * MixedstatePtychography (phase/mixedstate_ptychography.py:45-1027) on the GD
 * path: probe [K][S][S]; per position and mode P_k = shifted probe,
 * F_k = fft2(P_k O_patch); far = sqrt(sum_k |F_k|^2); chi_k = ifft2(A/far F_k - F_k)
 * (ptychographic_methods.py:2696-2805); object += (sum_k step scatter(conj(P_k)
 * chi_k)) norm(sum_k scatter(|P_k|^2)), probe_k += step sum_b conj(O_b) chi_bk
 * norm(sum_b |O_b|^2) (:2906-3000); per iteration the object constraints and
 * the recentring of every mode by mode 0's centre of mass
 * (ptychographic_constraints.py:1156-1178).  The mode orthogonalisation
 * (:1180-1221) is left to the host between calls. */
size_t pty_mixed_workspace_bytes(const pty_problem* problem, const pty_gd_params* params, int num_probes);
int pty_mixed_run(const pty_problem* problem, const pty_gd_params* params, int num_probes, const int* orders,
                  double* errors, void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* pty_mixed_run with position correction after every mini-batch's update
 * (mixedstate_ptychography.py:942-956 with ptychographic_methods.py:
 * 2071-2225): per mode the two perturbed transforms add p_x, p_y and |F_k|^2
 * into per-pixel accumulators in mode order; the least squares then uses the
 * mode sums and the intensity sqrt(sum_k |F_k|^2)^2 (:2736-2740).  The
 * position parameters and constraints are those of pty_gd_run_positions.
 * Workspace: pty_mixed_positions_workspace_bytes (adds [B][3][S][S] float32). */
size_t pty_mixed_positions_workspace_bytes(const pty_problem* problem, const pty_gd_params* params,
                                           int num_probes);
int pty_mixed_run_positions(const pty_problem* problem, const pty_gd_params* params, int num_probes,
                            const pty_position_params* positions, const int* orders, double* errors,
                            void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* One batch of the mixed-state forward/adjoint terms: probe_hat [K][S][S]
 * (kernel layout); ps, chi [K][B][S][S] natural (mode-major); far [B][K][S*S]
 * far fields in the staging layout of pty_forward_adjoint_batch; err[b]. */
int pty_mixed_forward_adjoint_batch(const void* probe_hat, int num_probes, const void* object, int Px,
                                    int Py, const float* positions, int N, const float* amplitudes,
                                    const int* order, int B, int S, void* ps, void* chi, double* err,
                                    void* far, pty_stream_t stream);

/* Every mode shifted by the centre of mass of mode 0's intensity
 * (ptychographic_constraints.py:1156-1178); `shift` is 2 floats of device
 * scratch. */
int pty_probe_com_modes(void* probe, int num_probes, int S, float* shift, pty_stream_t stream);

/* ---- many reconstructions together --------------------------------------- */

/* One entry per independent reconstruction; params.num_iter is ignored. */
typedef struct pty_multi_problem {
  pty_problem problem;
  pty_gd_params params;
  const int* orders;   /* int32 [num_iter][N], device */
  double* errors;      /* float64 [num_iter], device */
} pty_multi_problem;

/* Advances `count` reconstructions (frames 64 and/or 128, any sizes) by
 * num_iter GD iterations.  Every stage of an iteration is one launch for all
 * problems; with use_graph the iteration's launch sequence is captured once
 * into a hipGraph and replayed num_iter times (returns after completion).
 * Each problem's result is bitwise identical to its own pty_gd_run.  No
 * reference counterpart: the reference runs reconstructions one at a time. */
size_t pty_multi_workspace_bytes(const pty_multi_problem* problems, int count);
int pty_multi_run(const pty_multi_problem* problems, int count, int num_iter, int use_graph,
                  void* workspace, size_t workspace_bytes, pty_stream_t stream);

/* ---- learned object initialiser ------------------------------------------ */

/* No reference counterpart: the reference starts from a flat object or a
 * caller-supplied guess (ptychographic_methods.py:40-79, the object the
 * initialiser replaces).  These entry points run a 3x3 convolutional network
 * (BF16 operands on MFMA, FP32 accumulation) that maps 3x3 scan
 * neighbourhoods of diffraction patterns to 32 x 32 object patches.
 *
 * Activations are NHWC bfloat16 (raw 16-bit words) with C = 16 or 64; 16-byte
 * aligned.  A 3x3 convolution has zero padding equal to its dilation:
 *   in_y = out_y * stride + (ky - 1) * dil, Ho = (Hi - 1) / stride + 1.
 * Master weights are float32 [Co][9][Ci] (tap = 3 ky + kx); the kernels read
 * two bfloat16 copies: forward wf[Co][Kf] (k = tap Ci + ci) and backward
 * wb[Ci][Kb] (k = tap Co + co), K rounded up to 32, padding zero. */
typedef struct pty_conv_desc {
  int n;               /* samples */
  int Hi, Wi, Ci;      /* input */
  int Ho, Wo, Co;      /* output */
  int stride, dil;     /* stride 1 or 2, dilation 1..16 */
} pty_conv_desc;

/* 16-bit words of the forward (backward_copy = 0) or backward weight copy */
size_t pty_conv_weight_words(const pty_conv_desc* d, int backward_copy);
/* y = act(conv(x) + bias), act = ReLU when relu != 0; outputs bf16 and/or f32 */
int pty_conv_forward(const pty_conv_desc* d, const unsigned short* x, const unsigned short* wf,
                     const float* bias, int relu, unsigned short* y_bf16, float* y_f32,
                     pty_stream_t stream);
/* dx = conv^T(dy), times (x_act > 0) when x_act is given (ReLU of the input) */
int pty_conv_backward_data(const pty_conv_desc* d, const unsigned short* dy,
                           const unsigned short* wb, const unsigned short* x_act,
                           unsigned short* dx_bf16, float* dx_f32, pty_stream_t stream);
/* dw [Co][9][Ci] and dbias [Co] (optional), deterministic reduction */
size_t pty_conv_backward_weight_workspace_bytes(const pty_conv_desc* d);
int pty_conv_backward_weight(const pty_conv_desc* d, const unsigned short* x,
                             const unsigned short* dy, float* dw, float* dbias, void* workspace,
                             size_t workspace_bytes, pty_stream_t stream);
/* one Adam step (bias-corrected, step >= 1) on weights and biases; rewrites wf, wb */
int pty_conv_adam(const pty_conv_desc* d, float lr, float beta1, float beta2, float eps, int step,
                  float* w, float* m, float* v, const float* grad, float* bias, float* bias_m,
                  float* bias_v, const float* bias_grad, unsigned short* wf, unsigned short* wb,
                  pty_stream_t stream);
/* wf, wb from master weights */
int pty_conv_pack_weights(const pty_conv_desc* d, const float* w, unsigned short* wf,
                          unsigned short* wb, pty_stream_t stream);
/* Network input [count][64][64][16]: channel j < 9 is frame neighbours[s][j]
 * binned S/64 x S/64 from the centred pattern, log1p(counts) minus the
 * sample mean; channels 9..15 zero.  frames: counts, natural corner-centred
 * order (kernel_layout = 0) or amplitudes in the kernel layout (counts = A^2,
 * kernel_layout = 1).  S = 64 or 128. */
size_t pty_net_inputs_workspace_bytes(int count);
int pty_net_inputs(const float* frames, int N, int S, int kernel_layout, const int* neighbours,
                   int count, unsigned short* x, void* workspace, size_t workspace_bytes,
                   pty_stream_t stream);
/* loss = mean over [count][32][32][2] of (y - target)^2, y = channels 0 (amplitude)
 * and 1 (phase) of the last layer's f32 output [count][32][32][16];
 * dy (bf16, 16 channels) = d loss / d y */
size_t pty_net_loss_workspace_bytes(int count);
int pty_net_loss(const float* y, const float* target, int count, unsigned short* dy,
                 double* loss, void* workspace, size_t workspace_bytes, pty_stream_t stream);
/* stitching operands for pty_scatter_patches with S = 32: ps = sqrt(w),
 * chi = sqrt(w) a e^{i phi}, corner-centred, w = sin^2 taper */
int pty_net_patches(const float* y, int count, void* ps, void* chi, pty_stream_t stream);
/* object = weighted_sum / weight_sum where weight_sum > floor_weight, else 1 */
int pty_net_finish(const void* weighted_sum, const float* weight_sum, int Px, int Py,
                   float floor_weight, void* object, pty_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* PTY_HIP_H */
