/*
 * rbc_abi.h -- C-ABI of the MI355X-native Rayleigh-Benard time stepper (libsdc_rbc.so).
 *
 * Drop-in boundary for gpu_flow's hot path (jceyy/sdc, gpu_flow/code): the Boussinesq
 * RBC IMEX time stepper with its nonlinear-term evaluation, inverse-B, implicit linear
 * step, and the Nusselt diagnostics.  Plain C types only; all calls return an int status
 * (0 = OK, < 0 = error; rbc_last_error() gives the text).  Nothing calls exit().
 *
 * Field layout at the boundary is the reference's (matrix/matrix_device.h:13-63):
 *   theta, f, g : (M/2+1) x N x L complex, kx column fastest, then row (ky), then mode m;
 *   F, G        : L complex.
 * Complex values are interleaved (re, im) doubles at the boundary for every precision.
 *
 * Threading: one HIP stream per context; calls on one context are not thread-safe;
 * distinct contexts may be driven from distinct host threads.  All device memory is
 * allocated in rbc_create; rbc_step and the operator calls allocate nothing.
 */
#ifndef RBC_ABI_H
#define RBC_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBC_ABI_VERSION 1

/* Status codes */
#define RBC_OK 0
#define RBC_ERR_ARG (-1)
#define RBC_ERR_HIP (-2)
#define RBC_ERR_NOMEM (-3)
#define RBC_ERR_NUMERIC (-4)
#define RBC_ERR_STATE (-5)
#define RBC_ERR_IO (-6)

/* Arithmetic of the GPU path.
 *   RBC_F64    : every GPU stage in FP64 (parity form; PRECISION_DOUBLE, cuda_defines.h:9-27)
 *   RBC_F32    : FFTs, products, vertical transforms and implicit apply in FP32
 *   RBC_BF16X3 : as RBC_F32, but the vertical transforms run as BF16 hi/lo split MFMA
 *                products accumulated in FP32 (L <= 32, and one level combination of
 *                2 nq P floats below 2^31 bytes; rbc_create refuses larger grids with
 *                RBC_ERR_ARG, where RBC_F32 falls back to VALU vertical transforms)
 * Per-wavenumber factors are always built in FP64 on the host. */
typedef enum { RBC_F64 = 0, RBC_F32 = 1, RBC_BF16X3 = 2 } rbc_precision;

typedef struct {
    int M, N, L;     /* physical grid M x N, L vertical modes (main.cu: -nmodes M N L) */
    double Lx, Ly;   /* horizontal box lengths, height 1 (main.cu: -clength) */
    double Ra, Pr;   /* Rayleigh and Prandtl numbers (-ra, -pr) */
    double dt;       /* time step, baked into the implicit factors (-dt) */
    int precision;   /* rbc_precision */
    int nq;          /* vertical collocation levels; 0 = smallest exact rule for L */
    int device;      /* HIP device ordinal (-device) */
} rbc_params;

typedef struct rbc_ctx rbc_ctx;
typedef struct rbc_tracer rbc_tracer;  /* passive tracers (below); destroy before its context */

/* ≙ Coeff<double>(L) + linear_implicit_operator::init_operator + inverse_B_matrix::
 *   init_operator + nonlinear_operator_rayleigh_noslip::init_operator +
 *   implicit_explicit_step::init_timestepping (main.cu:431-454). */
int rbc_create(const rbc_params* p, rbc_ctx** out);
void rbc_destroy(rbc_ctx* c);
const char* rbc_last_error(const rbc_ctx* c); /* c may be NULL: last create error */
int rbc_get_info(const rbc_ctx* c, int* nq, int* kx_kept, int* ky_kept, int* n_unique_k);

/* Host state exchange in the reference layout (see above).  NULL skips a field.
 * The stepper dealiases on entry (implicit_explicit_stepping.cu:40), so set_state keeps
 * only the modes the dealias mask keeps; get_state returns zeros in the removed band. */
int rbc_set_state(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G);
int rbc_get_state(rbc_ctx* c, double* theta, double* f, double* g, double* F, double* G);

/* ≙ implicit_explicit_step::step_time + make_real_signal, nsteps times (main.cu:612-616).
 * The context owns the AB2 history; the first step after create/reset is Euler
 * (implicit_explicit_stepping.cu:43-52).
 * Non-finite state (≙ find_NaN, util/util.cu:83-105, which the reference defines but never
 * calls): the end-of-step kernel records the first step that ends with a non-finite value;
 * rbc_synchronize, rbc_get_state and rbc_nusselt then return RBC_ERR_NUMERIC (the text names
 * the step, counted from rbc_create; get_state still copies the state out), and rbc_step
 * refuses with RBC_ERR_NUMERIC until rbc_set_state or rbc_read_state_files replaces the state
 * (reset the AB2 history too: it holds the non-finite step's terms).  rbc_step enqueues and
 * does not wait, so the report comes from the next call that waits. */
int rbc_step(rbc_ctx* c, int nsteps);
int rbc_reset_history(rbc_ctx* c);

/* Operator-level calls on host buffers in the reference layout (inputs are dealiased):
 *   rbc_nonlinear  ≙ nonlinear_operator_rayleigh_noslip::calculate_operator
 *                    (nonlinear_operator_rayleigh_noslip.h:56)
 *   rbc_inverse_B  ≙ inverse_B_matrix::calculate_inverse (inverse_B_matrix.h:18)
 *   rbc_implicit   ≙ linear_implicit_operator::calculate_operator
 *                    (linear_implicit_operator.h:65)
 * Outputs are written for the kept modes and zero elsewhere. */
int rbc_nonlinear(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                  double* o_F, double* o_G);
int rbc_inverse_B(rbc_ctx* c, const double* theta, const double* f, const double* g,
                  const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                  double* o_F, double* o_G);
int rbc_implicit(rbc_ctx* c, const double* theta, const double* f, const double* g,
                 const double* F, const double* G, double* o_theta, double* o_f, double* o_g,
                 double* o_F, double* o_G);

/* ---- Device-resident exchange --------------------------------------------------------------
 * The same calls on device memory of the context's GPU, in the reference layout above with
 * the element type of the context's precision: complex128 (interleaved doubles) for RBC_F64,
 * complex64 (interleaved floats) for RBC_F32 and RBC_BF16X3 -- the matrix_folder data the
 * reference's operators take and return (nonlinear_operator_rayleigh_noslip.h:56,
 * inverse_B_matrix.h:18, linear_implicit_operator.h:65; implicit_explicit_stepping.cu:32-97).
 * Work is enqueued on the context's stream (rbc_stream) and the call returns without waiting:
 * inputs must be ready before the call (synchronise the producing stream first) and outputs
 * are ready once rbc_synchronize returns.  No call allocates.  Outputs hold zeros in the
 * removed band; set_state NULL fields are left unchanged, get_state NULL fields are skipped.
 * rbc_set_state_device ≙ rbc_set_state (clears a recorded non-finite step). */
int rbc_set_state_device(rbc_ctx* c, const void* theta, const void* f, const void* g, const void* F,
                         const void* G);
int rbc_get_state_device(rbc_ctx* c, void* theta, void* f, void* g, void* F, void* G);
/* ≙ calculate_operator (nonlinear_operator_rayleigh_noslip.h:56) */
int rbc_nonlinear_device(rbc_ctx* c, const void* theta, const void* f, const void* g, const void* F,
                         const void* G, void* o_theta, void* o_f, void* o_g, void* o_F, void* o_G);
/* ≙ calculate_inverse (inverse_B_matrix.h:18) */
int rbc_inverse_B_device(rbc_ctx* c, const void* theta, const void* f, const void* g, const void* F,
                         const void* G, void* o_theta, void* o_f, void* o_g, void* o_F, void* o_G);
/* ≙ linear_implicit_operator::calculate_operator (linear_implicit_operator.h:65) */
int rbc_implicit_device(rbc_ctx* c, const void* theta, const void* f, const void* g, const void* F,
                        const void* G, void* o_theta, void* o_f, void* o_g, void* o_F, void* o_G);
/* The context's HIP stream (a hipStream_t), for callers that order their own work after it. */
void* rbc_stream(rbc_ctx* c);

/* ---- Initial states and mean-flow forcing ---------------------------------------------------
 * rbc_init_state ≙ the initial data main.cu:290-352 creates when no -files are given, followed
 * by make_real_signal (main.cu:396); clears the AB2 history (the first step is Euler).
 *   RBC_INIT_ZERO   every field zero
 *   RBC_INIT_RANDOM matrix_device_init::init_random (init.cu:9-20, 348-365): every theta
 *                   coefficient uniform in [-amplitude, amplitude] (real and imaginary part;
 *                   the reference's default amplitude is 1e-6), f, g, F, G zero.  The numbers
 *                   come from SplitMix64 counters on (seed, reference-layout index): cuRAND's
 *                   XORWOW stream cannot be reproduced.
 *   RBC_INIT_TEST   matrix_device_init::init_test (init.cu:263-310): theta and f hold
 *                   1e-6 sin(kx) at ky = 0, kx != 0; F_m = 1e-6 cos(m); g = G = 0 (amplitude and
 *                   seed unused).  The one initial state that needs no random numbers.
 * Only the kept modes are stored (the stepper dealiases on entry). */
#define RBC_INIT_ZERO 0
#define RBC_INIT_RANDOM 1
#define RBC_INIT_TEST 2
int rbc_init_state(rbc_ctx* c, int kind, unsigned long long seed, double amplitude);
/* ≙ the pattern starts of main.cu:326-352 (-init_rectangle, -init_dislocation, -init_isr, -init_sv,
 * -init_cr, -init_giantspiral; util/init.cu:22-261): theta starts random (RBC_INIT_RANDOM,
 * amplitude 1e-6, all levels) and its vertical mode 0 is replaced by the unnormalised 2-D R2C of a
 * real pattern of amplitude 1e-6 on the M x N grid, then make_real.  init_mode ≙ -init_mode (the
 * reference's default is Lx / 2), truncated to int as the init_* parameters are; the rectangle
 * takes its x mode from Lx and M as main.cu:330 does.  The patterns are evaluated in FP64; the
 * rectangle's noise columns and the isr phase come from the project RNG (the reference uses C
 * rand()).  A setup call (allocates one M x N plane for its duration; waits for the stream). */
#define RBC_PATTERN_RECTANGLE 1
#define RBC_PATTERN_DISLOCATION 2
#define RBC_PATTERN_ISR 3
#define RBC_PATTERN_SV 4
#define RBC_PATTERN_CR 5
#define RBC_PATTERN_GIANTSPIRAL 6
int rbc_init_pattern(rbc_ctx* c, int kind, double init_mode, unsigned long long seed);
/* ---- Penalization and plane operators ------------------------------------------------------
 * rbc_set_penalization ≙ the PENALIZED builds (nonlinear_operator_rayleigh_noslip.cu:
 * 1097-1343): Brinkman terms of a mask chi(x, y) added to the nonlinear term with penalization
 * parameter eta (main.cu -eta; eta = 0 turns them off).  Masks (:1652-1740): RBC_MASK_RECTANGLE
 * (3-point border; the default PENALIZED build), RBC_MASK_CIRCLE (-DCIRCLE), RBC_MASK_DISKARRAY
 * (3 x 2 disks, -DDISKARRAY), RBC_MASK_EVERYWHERE, or RBC_MASK_CUSTOM with `mask` holding
 * M x N values (x fastest).  A setup call (allocates; waits for the stream). */
#define RBC_MASK_RECTANGLE 0
#define RBC_MASK_CIRCLE 1
#define RBC_MASK_DISKARRAY 2
#define RBC_MASK_EVERYWHERE 3
#define RBC_MASK_CUSTOM 4
int rbc_set_penalization(rbc_ctx* c, double eta, int shape, const double* mask);
/* Operators on the state's mode planes, applied between steps as main.cu:569-590 does:
 *   rbc_freeze        ≙ freeze_operator::freeze_circle (radius) / freeze_rectangle (width):
 *                       captures the physical planes of theta, f, g and the mask
 *   rbc_apply_freeze  ≙ freeze_operator::calculate_operator: outside the mask the planes
 *                       return to the captured ones (nothing before rbc_freeze)
 *   rbc_temp_mask     ≙ masktemp::temp_circle / temp_rectangle (the heat-source region)
 *   rbc_apply_temp_mask ≙ masktemp::calculate_operator: theta + 1.5/100 in the region
 *   rbc_mask_theta    ≙ masking::set_mask_parameters + masker: theta times `factor` inside a
 *                       circle at (cx, cy) (fractions of the box) of radius r (fraction of min/2)
 * Shapes: RBC_PLANE_CIRCLE (param = radius as a fraction of min(M, N) / 2), RBC_PLANE_RECTANGLE
 * (param = border width as a fraction of the side).  Enqueued on the context's stream. */
#define RBC_PLANE_CIRCLE 0
#define RBC_PLANE_RECTANGLE 1
int rbc_freeze(rbc_ctx* c, int shape, double param);
int rbc_apply_freeze(rbc_ctx* c);
int rbc_temp_mask(rbc_ctx* c, int shape, double param);
int rbc_apply_temp_mask(rbc_ctx* c);
int rbc_mask_theta(rbc_ctx* c, double cx, double cy, double radius, double factor);

/* ≙ calculate_temperature_operator::calculate_operator_at (calculate_temperature_operator.cu:
 * 27-78): theta(x, y, z) = sum_i S_i(z) theta_i(x, y) on the M x N grid at nz heights, FP32,
 * [l][y][x]. */
int rbc_temperature_grid(rbc_ctx* c, int nz, const double* z, float* out);
/* Packets of the visual stream (output/server_writer.cu:79-252, output/datastruct.h:4-72): the
 * frame server_writer::Prepare builds for one request, written to `out` as NumPackets packets
 * of 2048 bytes (packet 0: the simulation header; then the image sampled at Nx x Ny points,
 * res_req <= 0 for the full grid): RBC_IMG_TEMPERATURE (theta at z = 0), RBC_IMG_PARTICLES
 * (particle counts on a res_req^2 grid, 128^2 by default; tracer t), RBC_IMG_TEMP_PARTICLES
 * (temperature with the particles' points set to 0).  Nt is the step index, eta and t_final go
 * into the header.  *nbytes receives the size (out = NULL: query).  The socket that sends
 * them (RBCServer) is the caller's. */
#define RBC_IMG_TEMPERATURE 1
#define RBC_IMG_PARTICLES 2
#define RBC_IMG_TEMP_PARTICLES 3
int rbc_stream_packets(rbc_ctx* c, rbc_tracer* t, int img_type, int res_req, int Nt, double eta,
                       double t_final, void* out, size_t cap, size_t* nbytes);

/* ≙ force_mean (util/util.cu:74-81), which main.cu:567-568 applies before every step under
 * -mean_F: F_0 = mean_F + 0i.  Enqueued on the context's stream. */
int rbc_force_mean(rbc_ctx* c, double mean_F);

/* Diagnostics (FP64 reductions on the device; the reference computes none):
 *   nu[0] plate Nusselt number at z=-1/2, nu[1] at z=+1/2, nu[2] = 1 + <w theta>. */
int rbc_nusselt(rbc_ctx* c, double* nu3);

/* Per-stage timing with HIP events on the context's stream: runs nsteps steps and writes
 * the mean milliseconds per call of each stage into ms[RBC_NSTAGES] (stage names from
 * rbc_stage_name). */
#define RBC_NSTAGES 7
int rbc_profile_steps(rbc_ctx* c, int nsteps, double* ms);
const char* rbc_stage_name(int i);

/* Blocks until all work queued on the context's stream has finished; RBC_ERR_NUMERIC once a
 * step has ended non-finite (see rbc_step). */
int rbc_synchronize(rbc_ctx* c);

/* ---- State files in the reference's binary format ---------------------------------------
 * ≙ matrix_folder_writer::read_binary_file / write_binary_file (output/matrix_folder_writer.cu:
 * 3-103) for the five files theta, f, g, F, G (main.cu:359-381 read, :700-713 write).  Format:
 * int32 1, int32 3, int32 dims {M/2+1, N, L} (F, G: {1, 1, L}), then complex values
 * (complex64 from an FP32 build, complex128 from a PRECISION_DOUBLE build; the reader infers
 * which from the file size).
 *   which = 0: the state; reading it applies make_real as main.cu:396 does.
 *   which = 1: the AB2 history (the previous step's B^-1 N, implicit_explicit_stepping.h
 *              res_1), which the reference never writes; restoring it makes a restart continue
 *              with AB2 instead of an Euler step.
 * elem_bytes: 8 (complex64), 16 (complex128) or 0 (the context's precision: 16 for RBC_F64). */
int rbc_read_state_files(rbc_ctx* c, const char* const paths[5], int which);
int rbc_write_state_files(rbc_ctx* c, const char* const paths[5], int which, int elem_bytes);
/* AB2 history in the reference layout (see rbc_set_state).  get fails with RBC_ERR_STATE
 * before the first step; set makes the next step an AB2 step. */
int rbc_get_history(rbc_ctx* c, double* theta, double* f, double* g, double* F, double* G);
int rbc_set_history(rbc_ctx* c, const double* theta, const double* f, const double* g,
                    const double* F, const double* G);

/* ---- Velocity operator ------------------------------------------------------------------
 * ≙ calculate_velocity_operator::calculate_operator_at (calculate_velocity_operator.cu:
 * 106-158): u_x, u_y, u_z of the current state on the M x N grid at nz heights z (clamped to
 * [-1/2, 1/2]), FP32, layout [l][y][x] per component.  The scratch (3 nz M N floats and the
 * level spectra) belongs to the context and grows only when nz exceeds every earlier call's. */
int rbc_velocity_grid(rbc_ctx* c, int nz, const double* z, float* ux, float* uy, float* uz);
/* Characteristic velocity U = <|u|> of the current state (main_extract_vel.cu:182-218): the mean
 * of |u| over the M x N grid at nz heights linspace(-1/2, 1/2, nz) (the reference tool uses
 * nz = 10), components in FP32 as the reference's velocity operator gives them, squared, summed
 * and averaged in FP64, points with a NaN skipped.  Uses the rbc_velocity_grid scratch. */
int rbc_characteristic_velocity(rbc_ctx* c, int nz, double* U);
/* The velocity of the current state at n arbitrary points, summed exactly over the kept
 * modes (no counterpart in the reference; the spectral tracer's evaluation).  L <= 32.
 * Host arrays; the context's point scratch grows only when n exceeds every earlier call's. */
int rbc_velocity_at(rbc_ctx* c, int n, const float* x, const float* y, const float* z,
                    float* u, float* v, float* w);
/* rbc_velocity_at on device arrays of the context's GPU, enqueued on its stream without a
 * wait (results ready after rbc_synchronize); the partial sums live in the context's point
 * scratch, grown (with a wait) when n exceeds every earlier call's. */
int rbc_velocity_at_device(rbc_ctx* c, int n, const float* x, const float* y, const float* z,
                           float* u, float* v, float* w);
/* rbc_velocity_at_device without any allocation: no counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).  The
 * partial sums live in a file-scope device array of 16 MiB, one per GPU, held until the process
 * ends; n above 1,398,101 points is rejected with RBC_ERR_ARG.  All contexts on one GPU share
 * that array: calls that overlap in time on one GPU (other streams, other threads) race, so
 * order them.  Results equal rbc_velocity_at_device's. */
int rbc_velocity_at_device_noalloc(rbc_ctx* c, int n, const float* x, const float* y, const float* z,
                                   float* u, float* v, float* w);

/* ---- Passive tracers (≙ particle_tracer, particle/particle_tracer.h:21-89) ----------------
 * Positions are FP32, stored as a ring of buffer_size time slots (the reference's layout
 * [slot][particle]).  diff_coeff ≙ the constructor's pDiffCoeff (the Lewis-number argument,
 * main.cu:497); random numbers come from SplitMix64 counters on `seed` (the reference's
 * cuRAND streams cannot be reproduced).
 *   RBC_TRACER_RK4_TRILINEAR : the reference: RK4 in the frozen velocity on a 128-level grid
 *                              with trilinear, periodic interpolation (particle_tracer.cu:
 *                              141-212, 508-687) and reflection at z = +-(1/2 - 1e-3).
 *   RBC_TRACER_SPECTRAL_RK2  : midpoint RK2 with the velocity summed exactly at each
 *                              particle (rbc_velocity_at); L <= 32. */
#define RBC_TRACER_RK4_TRILINEAR 0
#define RBC_TRACER_SPECTRAL_RK2 1
/* init types in the order of particle_tracer::initType (particle_tracer.h:50) */
#define RBC_PART_RANDOM 0
#define RBC_PART_CENTER 1
#define RBC_PART_CENTER_SIGMA 2
#define RBC_PART_DIAGONAL 3
#define RBC_PART_PAIRS 4
#define RBC_PART_COORD_SIGMA 5
/* ≙ particle_tracer(M, N, L, Np, buffer, cube_length, Le) (particle_tracer.cu:7-44): all
 * particles start at the centre.  Device memory is allocated here, once: the position ring
 * (3 buffer_size num_particles floats) and the mode's work -- RK4: the 128-level velocity grid;
 * spectral RK2: velocities and midpoints (6 num_particles floats), the partial sums of the mode
 * slices and, for FP32 / BF16X3 contexts, the basis values at the particle heights
 * (16 L num_particles bytes). */
int rbc_tracer_create(rbc_ctx* c, int num_particles, int buffer_size, int mode, double diff_coeff,
                      unsigned long long seed, rbc_tracer** out);
void rbc_tracer_destroy(rbc_tracer* t);
const char* rbc_tracer_last_error(const rbc_tracer* t);
/* ≙ init_positions (particle_tracer.cu:60-135); (cx, cy, cz, sigma) for the coord/sigma types */
int rbc_tracer_init(rbc_tracer* t, int type, double cx, double cy, double cz, double sigma);
/* overwrite the current slot */
int rbc_tracer_set_positions(rbc_tracer* t, const float* x, const float* y, const float* z);
/* ≙ step_particles (particle_tracer.cu:141-212) in the context's current velocity; fails with
 * RBC_ERR_STATE when the ring is full (is_buffer_full). */
int rbc_tracer_step(rbc_tracer* t, double dt);
/* ≙ get_particle_positions (particle_tracer.cu:272-311): copies num_timesteps slots
 * ([slot][particle], at most buffer_size - 1) and, with clear, moves the current slot to 0. */
int rbc_tracer_get_positions(rbc_tracer* t, int clear, float* x, float* y, float* z, int* num_timesteps);
int rbc_tracer_current(rbc_tracer* t, float* x, float* y, float* z);
int rbc_tracer_info(const rbc_tracer* t, int* iteration_index, int* particle_time);
/* FP64 moments of the current slot: {<x>, <y>, calc_secondmoment, calc_pair_secondmoment}
 * (particle_tracer.cu:217-258, 771-866, PERIODIC minimum image for pairs). */
int rbc_tracer_moments(rbc_tracer* t, double* out4);

/* ---- Kinetic-energy spectra (FP64) -------------------------------------------------------
 * rbc_energy_spectrum: E_h(k) = shell sums over |k| in bins of width min(2 pi/Lx, 2 pi/Ly)
 * (bin s holds |k| in [(s - 1/2) dk, (s + 1/2) dk)) of (1/2)|u_h,k|^2 integrated over z by the
 * context's Gauss-Legendre rule; totals = {E_h, E_z} (volume integrals of (u^2+v^2)/2 and w^2/2
 * over unit area).  No counterpart in the reference.
 * rbc_energy_spectrum_ref: the reference's binning (energy_spectrum_2d.cu:53-129) of |c_k|^2
 * for each velocity component c at height z: E3[c * nbins + b].
 * Both return the bin count in *nbins; pass NULL output to query it. */
int rbc_energy_spectrum(rbc_ctx* c, int cap, double* E, int* nbins, double* totals);
int rbc_energy_spectrum_ref(rbc_ctx* c, double z, int cap, double* E3, int* nbins);

/* ---- Ensembles of independent runs --------------------------------------------------------
 * No counterpart in the reference, which runs one simulation per process
 * (run/script_onset_sdc.sh:8).  Members may differ in M, N, L, Ra, Pr, dt and precision but
 * share one device; each member is a full context (rbc_ensemble_member gives it for
 * rbc_set_state, rbc_get_state, rbc_nusselt, ...; it stays owned by the ensemble).
 * rbc_ensemble_step advances every member by nsteps: when a member has no AB2 history, the
 * first of the nsteps runs eagerly for every member (an Euler step for members without
 * history, an AB2 step for the others), then steps_per_launch steps of all members are
 * captured once into one
 * HIP graph and replayed (steps_per_launch = 0: eager launches).  Results equal stepping each
 * member alone.  rbc_ensemble_tune times launch shapes of the generic FFT kernels per size
 * class (M, N, L, precision) on the first member of the class (its state is restored) and
 * applies the fastest to the class; report (optional, 3 doubles per member) receives
 * {y-FFT columns per block, row-kernel threads, ms per step}.
 * In the graph, the members of one size class (same M, N, L, nq, precision and tuned shapes)
 * on the FP32 / BF16X3 path of 64..256-point grids with L <= 32 advance together, one launch
 * per stage for the class (same per-member arithmetic as a run of its own); others are graph
 * branches of their own.  rbc_ensemble_batching(e, 0) turns this off (1: on, the default);
 * rbc_ensemble_batched_members gives how many members the current graph steps in batches. */
typedef struct rbc_ensemble rbc_ensemble;
int rbc_ensemble_create(int n, const rbc_params* params, rbc_ensemble** out);
void rbc_ensemble_destroy(rbc_ensemble* e);
const char* rbc_ensemble_last_error(const rbc_ensemble* e);
int rbc_ensemble_size(const rbc_ensemble* e);
rbc_ctx* rbc_ensemble_member(rbc_ensemble* e, int i);
int rbc_ensemble_tune(rbc_ensemble* e, double* report);
int rbc_ensemble_step(rbc_ensemble* e, int nsteps, int steps_per_launch);
int rbc_ensemble_batching(rbc_ensemble* e, int on);
int rbc_ensemble_batched_members(const rbc_ensemble* e);

/* ---- Learned closure for coarse runs -------------------------------------------------------
 * No counterpart in the reference.  A pointwise MLP (inputs u, v, w, theta, d_x theta,
 * d_y theta, d_z theta, z; two tanh layers of 64; outputs r_theta, r_f, r_g; 4931 parameters,
 * BF16 weights, FP32 arithmetic) attached to a coarse context.  The points are the coarse
 * M x N grid at nz Gauss-Legendre levels (0: 2 L; at most 512).
 *   rbc_closure_sample        appends to the dataset (capacity_points) the inputs and the
 *                             targets r of R = P_c N_fine(X_fine) - N_coarse(P_c X_fine) at the
 *                             points, from a finer context's state (P_c: truncation to the
 *                             coarse modes; the coarse context's state is not touched)
 *   rbc_closure_restrict_state  coarse state <- P_c (fine state); AB2 history cleared
 *   rbc_closure_fit_normalization  input mean / 1/std and target rms from the first n points
 *   rbc_closure_train         Adam steps on random minibatches of dataset points
 *                             [first, first + n) with the GPU forward/backward kernels;
 *                             loss[steps] receives the batch losses (normalised MSE)
 *   rbc_closure_gradient      gradient (and loss) of the normalised MSE over n given points
 *   rbc_closure_predict       r at n given input points
 *   rbc_closure_eval          normalised MSE of the model and of predicting 0 on dataset points
 *   rbc_closure_enable        attach (on = 1) / detach the correction to the coarse context's
 *                             steps (rbc_step then runs the corrected method)
 *   rbc_closure_interval      evaluate the network on every `every`-th corrected step and add
 *                             the last evaluated correction on the steps between (default 1)
 *   rbc_closure_damping       add gain times the correction on the modes with |k| <= kfrac k_max
 *                             (k_max: the largest kept |k|) and none above (default 1, 1)
 *   rbc_closure_step          corrected steps in chunks of check_every; a chunk whose last
 *                             correction has |dN| / |N| above max_ratio (or not finite) is redone
 *                             with the plain method, which then continues (fallback)
 *   rbc_closure_status        attached?, fallback taken?, last |dN| / |N| */
typedef struct rbc_closure rbc_closure;  /* destroy before its coarse context */
int rbc_closure_create(rbc_ctx* coarse, int nz, int capacity_points, rbc_closure** out);
void rbc_closure_destroy(rbc_closure* cl);
const char* rbc_closure_last_error(const rbc_closure* cl);
int rbc_closure_info(const rbc_closure* cl, int* npts, int* nweights, int* count, int* nz);
int rbc_closure_restrict_state(rbc_closure* cl, rbc_ctx* fine);
int rbc_closure_sample(rbc_closure* cl, rbc_ctx* fine, int* count);
int rbc_closure_points(rbc_closure* cl, rbc_ctx* fine, float* X, float* Y);
int rbc_closure_dataset(rbc_closure* cl, int first, int n, float* X, float* Y);
int rbc_closure_load_dataset(rbc_closure* cl, int n, const float* X, const float* Y); /* replaces it */
int rbc_closure_clear_dataset(rbc_closure* cl);
int rbc_closure_fit_normalization(rbc_closure* cl, int n);
int rbc_closure_get_normalization(const rbc_closure* cl, float* in_mean, float* in_inv, float* out_scale);
int rbc_closure_set_normalization(rbc_closure* cl, const float* in_mean, const float* in_inv,
                                  const float* out_scale);
int rbc_closure_get_weights(rbc_closure* cl, float* w);
int rbc_closure_set_weights(rbc_closure* cl, const float* w);
int rbc_closure_gradient(rbc_closure* cl, int n, const float* X, const float* Y, float* grad, double* loss);
int rbc_closure_train(rbc_closure* cl, int steps, int batch, double lr, unsigned long long seed, int first,
                      int n, double* loss);
int rbc_closure_predict(rbc_closure* cl, int n, const float* X, float* Y);
int rbc_closure_eval(rbc_closure* cl, int first, int n, double* mse_model, double* mse_zero);
int rbc_closure_enable(rbc_closure* cl, int on, double max_ratio);
int rbc_closure_interval(rbc_closure* cl, int every);
int rbc_closure_damping(rbc_closure* cl, double gain, double kfrac);
int rbc_closure_step(rbc_closure* cl, int nsteps, int check_every);
int rbc_closure_status(const rbc_closure* cl, int* active, int* fallback, double* last_ratio);

/* ---- Simulation driver ---------------------------------------------------------------------
 * ≙ main() of gpu_flow/code/main.cu: the reference's argument set (main.cu:61-229; -delta_t and
 * -final_t, the names README.txt:38-39 lists, are accepted for -dt and -t), iCeil(t/dt) steps
 * (main.cu:279), -mean_F forcing before each step (main.cu:567-568), tracers per -le value,
 * state files theta_%08d.bin .. G_%08d.bin every steps/nfiles steps (main.cu:697-728), tracks
 * (.trkss) and second-moment files, simulation.log; plus -precision f32|f64|bf16x3,
 * -tracer_mode rk4|spectral, -write_history / -history_files, -nq.  Returns the process exit
 * status (0, or 1 with the message on stderr).  argv[0] is the program name. */
int rbc_run_main(int argc, const char* const* argv);
/* ≙ main() of gpu_flow/code/main_extract_vel.cu for its characteristic-velocity output:
 * -clength, -files (five paths), -write_charvel_nz n > 0 with -outfile_charvel writes <|u|> over
 * 10 heights (the reference tool's fixed count) as a data-point file; -description. */
int rbc_extract_vel_main(int argc, const char* const* argv);

#ifdef __cplusplus
}
#endif
#endif /* RBC_ABI_H */
