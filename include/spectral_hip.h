/*
 * spectral_hip.h — C ABI of the MI355X (gfx950) pseudo-spectral Navier-Stokes
 * RK3 hot path. Plain pointers, sizes and status codes only; no C++ or torch
 * types cross this boundary.
 *
 * Every function returns 0 on success and a nonzero status on failure; the
 * failure text ("what: detail", the reference's std::runtime_error messages)
 * is then available from sph_last_error() on the calling thread.
 *
 * Layouts
 *   physical field : real N^3 array, id = ix + nx*(iy + ny*iz) (x fastest),
 *                    as in src/ava_taylor_green_impl.hpp:340-412 and
 *                    src/fftw_fft.hpp:30-40.
 *   spectral field : half spectrum (R2C), (nx/2+1)*ny*nz complex values,
 *                    id = ky + ny*(kz + nz*kx) (ky fastest, kx slowest);
 *                    mode (mx,my,mz) = (kx, signed(ky), signed(kz)) with the
 *                    Nyquist index mapped to +n/2 (impl.hpp:123-125).
 *                    Interleaved (re, im). Query the mode of every storage
 *                    slot with sph_fft_spectral_modes().
 *   precision      : SPH_FP32 or SPH_FP64 selects float or double storage
 *                    and arithmetic (the reference's SPECTRAL_PRECISION,
 *                    impl.hpp:48-52, becomes a run-time choice).
 */
#ifndef SPECTRAL_HIP_H
#define SPECTRAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SPH_FP32 = 0, SPH_FP64 = 1 };

typedef struct sph_fft_s* sph_fft_t;
typedef struct sph_solver_s* sph_solver_t;

/* Mirrors spectral::ava_solver::Diagnostics (src/ava_taylor_green.hpp:8-19). */
typedef struct sph_diagnostics {
    double t;
    double kinetic_energy;
    double dissipation;
    double enstrophy;
    double max_spectral_divergence;
    double max_velocity;
    double dt;
    double cfl;
    double main_dt;
    double main_cfl;
} sph_diagnostics;

/* Mirrors the per-section timers printed by print_profile
 * (src/ava_taylor_green_impl.hpp:1181-1286 and 563-594). Seconds. */
typedef struct sph_profile {
    double step;
    double diagnostics;
    double max_velocity;
    long long step_calls;
    long long rhs_calls;
    long long diagnostics_calls;
    long long max_velocity_calls;
    long long forward_calls;  /* 3D forward transforms, counted as the reference counts them */
    long long backward_calls; /* 3D backward transforms */
} sph_profile;

/* Thread-local text of the last failure on this thread ("" if none). */
const char* sph_last_error(void);

/* Number of visible GPUs (0 when no HIP device is present). */
int sph_device_count(int* count);

/* 1 if n is a transform length the library was compiled for, else 0. */
int sph_supported_length(int n);

/* ---------------------------------------------------------------------------
 * FFT backend — replaces spectral::FftwFft3D / FlupsFft3D
 * (src/fftw_fft.hpp:13-81, src/fftw_fft.cpp:8-87, src/flups_fft.hpp:28-133)
 * and the internal seam detail::FftPlan (src/ava_taylor_green_impl.hpp:255-652).
 * ------------------------------------------------------------------------- */

/* FftwFft3D(shape, lengths, comm) (src/fftw_fft.cpp:8-36): shape is
 * {nx, ny, nz}; every n must be > 1 and a supported length, every length > 0. */
int sph_fft_create(const int shape[3], const double lengths[3], int precision, sph_fft_t* out);
int sph_fft_destroy(sph_fft_t fft);

/* physical_size() / spectral storage in complex values (fftw_fft.hpp:25-28). */
int sph_fft_sizes(sph_fft_t fft, int64_t* physical_size, int64_t* spectral_modes);

/* for_each_spectral(...) (fftw_fft.hpp:45-66) as a table: mode indices of
 * every spectral storage slot, each array spectral_modes long. */
int sph_fft_spectral_modes(sph_fft_t fft, int* mx, int* my, int* mz);

/* Device-pointer transforms (FlupsFft3D::forward_device/backward_device,
 * flups_fft.hpp:60-61; FftPlan::forward/backward, impl.hpp:414-561).
 * Asynchronous on `stream` (a hipStream_t; NULL = the handle's own stream).
 * forward: real physical -> half spectrum, unnormalised, sign e^{-i}.
 * backward: half spectrum -> real physical, sign e^{+i}, multiplied by
 * 1/(nx*ny*nz); the kx = 0 and kx = nx/2 planes contribute their real part,
 * which is what Re(C2C backward) does (impl.hpp:555-561, fftw_fft.cpp:70-85).
 * Buffers are caller-owned device memory in the handle's precision. */
int sph_fft_forward(sph_fft_t fft, const void* physical, void* spectral, void* stream);
int sph_fft_backward(sph_fft_t fft, const void* spectral, void* physical, void* stream);

/* Backward transform that allocates nothing. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The same
 * kernels and results as sph_fft_backward of a handle of this shape and
 * precision, without a handle: the intermediate half-spectrum field and the
 * twiddle tables live in a static device array of the library, which every
 * GPU it runs on holds (80 MiB + 64 KiB) until the process ends. The
 * intermediate needs (shape[0]/2+1)*shape[1]*shape[2] complex values of the
 * precision; above 80 MiB the call fails ("... bytes of scratch ...") before
 * launching anything. Every call on one GPU shares that array and leaves its
 * contents behind (the twiddles are reloaded only when the shape or precision
 * changes), so calls that overlap on one GPU race: issue them on one stream,
 * or order them. stream: a hipStream_t (NULL = the default stream). */
int sph_fft_backward_noalloc(const int shape[3], int precision, const void* spectral, void* physical,
                             void* stream);

/* Host-vector transforms with the reference's std::vector<double> contract
 * (fftw_fft.cpp:53-85): physical is physical_size doubles; spectral is
 * 2*spectral_modes interleaved doubles. Synchronous. */
int sph_fft_forward_host(sph_fft_t fft, const double* physical, double* spectral);
int sph_fft_backward_host(sph_fft_t fft, const double* spectral, double* physical);

/* ---------------------------------------------------------------------------
 * Step API — replaces spectral::ava_solver::TaylorGreenSolver
 * (src/ava_taylor_green.hpp:36-53, src/ava_taylor_green_impl.hpp:1289-1308).
 * ------------------------------------------------------------------------- */

/* TaylorGreenSolver(n, reynolds) (impl.hpp:703-716): allocates the state,
 * sets the Taylor-Green field at cell centres, transforms and projects it.
 * Fails with "grid size must be positive" / "Reynolds number must be positive". */
int sph_solver_create(int n, double reynolds, int precision, sph_solver_t* out);
int sph_solver_destroy(sph_solver_t solver);

/* step_rk3(dt) (impl.hpp:979-1045): one Williamson low-storage RK3 step. */
int sph_solver_step_rk3(sph_solver_t solver, double dt);

/* diagnostics(t) (impl.hpp:1122-1179); dt/cfl fields are left 0. */
int sph_solver_diagnostics(sph_solver_t solver, double t, sph_diagnostics* out);

/* max_velocity() (impl.hpp:1105-1120). */
int sph_solver_max_velocity(sph_solver_t solver, double* out);

/* Sizes of the state: n, spectral modes per component (half spectrum). */
int sph_solver_sizes(sph_solver_t solver, int* n, int64_t* spectral_modes);

/* Copy component c (0..2) of u_hat to / from host memory as interleaved
 * doubles in the spectral layout above (parity fixtures, checkpoints). */
int sph_solver_get_uhat(sph_solver_t solver, int component, double* host);
int sph_solver_set_uhat(sph_solver_t solver, int component, const double* host);

/* Counters and timers behind print_profile (impl.hpp:1181-1286). */
int sph_solver_profile(sph_solver_t solver, sph_profile* out);

/* GPU time in ms of each of the most recent RK3 steps (HIP events recorded
 * on the solver's stream before and after each step, graph replay
 * included), oldest first: up to `max` values into ms, their number in
 * *count (at most the last 4096 steps are kept). No reference counterpart
 * (the reference accumulates wall time only, impl.hpp:980,1043); bench.py
 * reports the median. */
int sph_solver_step_times(sph_solver_t solver, double* ms, int max, int* count);

/* print_profile(rank, init, run) (impl.hpp:1181-1286): prints the
 * profile_solver / profile_breakdown / profile_fft lines to stdout. */
int sph_solver_print_profile(sph_solver_t solver, int rank, double init_time, double run_time);

/* Capture the RK3 step into a hipGraph (1) or launch kernels directly (0).
 * Default 1. Results are identical either way. */
int sph_solver_set_graph(sph_solver_t solver, int enabled);

/* Wait for all work queued by the solver. */
int sph_solver_synchronize(sph_solver_t solver);

/* Measure the pass kernels of an RK3 stage with HIP events on the solver's
 * stream: each of the five pass kernels is launched `reps` times back to
 * back on the current state (the RK-update pass on a scratch copy of the
 * state) and its mean duration in milliseconds is written to ms[5] in the
 * order inverse-z, inverse-y, x (inverse x + products + forward x),
 * forward-y, forward-z + update + the next stage's inverse z. A stage runs
 * the last four; the standalone inverse z runs only before a step whose
 * state was set from outside (creation, set_uhat). Also returns the
 * algorithmic HBM bytes per launch of each pass in bytes[5]. The state is
 * unchanged. */
int sph_solver_time_passes(sph_solver_t solver, int reps, double* ms, double* bytes);

/* The pass kernels inside real steps: `steps` RK3 steps of dt with the
 * kernels launched directly (not through the graph) and a HIP event between
 * consecutive launches on the solver's stream. ms[6]: [0] the standalone
 * inverse z if the velocity pairs had to be formed first (else 0), [1..4] the
 * mean duration of inverse-y, x, forward-y and forward-z + update + next
 * inverse z over the steps' 3*steps stages, [5] the mean step time;
 * bytes[5]: algorithmic HBM bytes per launch as sph_solver_time_passes.
 * The state advances by `steps` steps (as sph_solver_step_rk3 would). */
int sph_solver_time_step_passes(sph_solver_t solver, double dt, int steps, double* ms, double* bytes);

/* One adaptive RK3 step as run_taylor_green's CFL loop takes it
 * (impl.hpp:1363-1378, without the clipping to output times):
 * dt = min(dt_max, cfl * (2 pi / n) / max|u|) (dt_max when max|u| <= 1e-14),
 * max|u| of the state before the step. Where the x pass's fused reduction
 * runs (*fused = 1; may be NULL), max|u| comes from the step's own stage-0 x
 * pass, otherwise from max_velocity(). Returns max|u| and the dt taken. */
int sph_solver_step_adaptive(sph_solver_t solver, double dt_max, double cfl, double* max_velocity, double* dt,
                             int* fused);

/* The choices of the solver's autotuner (the reference plans its FFTs with
 * FFTW_ESTIMATE / cuFFT, impl.hpp:257-412; here the y and x passes come in
 * several tile widths, kernel forms and radix orders, timed at creation on
 * the solver's own buffers). For the tuned passes inverse-y, x, forward-y:
 * variant[3] = the variant kept, lines[3] = its lines per tile, ms[12] = mean
 * launch time of each variant, {pass0 v0, .. pass0 v3, pass1 v0, ...} (0
 * where a pass has fewer variants, or tuning was switched off with
 * SPH_AUTOTUNE=0). Variants: y passes 0 = default width, 1 = half width,
 * 2 = a quarter; x pass 0 = register-direct default width, 1 =
 * register-direct half width, 2 = LDS-staged default width, 3 =
 * register-direct default width in the alternative radix order (lengths that
 * have one, sph_radix_plan). */
int sph_solver_tuning(sph_solver_t solver, int* variant, int* lines, double* ms);

/* One tuned pass of the autotuner, any of the four: pass 0 inverse y, 1 x,
 * 2 forward y (as sph_solver_tuning), 3 forward z + update + next inverse z
 * (variants: 0 = default z-tile width, 1 = half of it, 2 = default width in
 * the alternative radix order; the standalone inverse z takes the same width
 * and order). variant, lines per tile, ms[4] = mean launch time of each
 * variant (0 where the pass has fewer variants). */
int sph_solver_tuned_pass(sph_solver_t solver, int pass, int* variant, int* lines, double* ms);

/* The radix stages of the transforms in the variant the autotuner kept for
 * tuned pass `pass` (0..3 as sph_solver_tuned_pass), first stage first:
 * radices[0 .. *count-1], radices holds 16 values. */
int sph_solver_radix_plan(sph_solver_t solver, int pass, int* radices, int* count);

/* The radix stages of the z-type passes (z transforms, register-direct x
 * pass) for transform length n: order 0 = the default plan, 1 = the
 * alternative order the autotuner times beside it (the reference leaves plan
 * choice to FFTW / cuFFT, src/fftw_fft.cpp:31-32, impl.hpp:288,300-301).
 * radices holds 16 values; *count = 0 when n is not a compiled length. No GPU
 * needed. */
int sph_radix_plan(int n, int order, int* radices, int* count);

/* The work layout the autotuner kept (*layout):
 *   0  line layout: the pair fields between the y passes and the x pass are
 *      [kx][z][y] (an x tile reads n/2+1 rows n*n pairs apart), the state
 *      and the spectral pair fields [kx][kz][ky];
 *   1  transposed: [z][y][kx] between the y passes and the x pass (an x
 *      tile is one contiguous block; the y passes transpose on the way);
 *   2  ky-blocked: the state and the spectral pair fields in blocks of kb ky
 *      ([kx][ky/kb][kz][ky%kb], kb = the z tile's ky count, so a z tile is
 *      one contiguous block), the x pass in the line layout (the y passes
 *      move between the two).
 * 1 and 2 exist at lengths >= 256 where a second work buffer of 12
 * half-spectrum fields fits. ms (4 doubles): summed launch time of passes
 * 1-3 (inverse y, x, forward y) with the fastest tile variants of layout 0
 * and of layout 1, of passes 1-4 in layout 2, and of passes 1-4 with the
 * faster of layouts 0 and 1 (0 where a layout was not timed). SPH_LAYOUT=0 /
 * 1 / 2 in the environment forces a layout. The layout is internal: every
 * state access of this API uses the plain mode order. */
int sph_solver_layout(sph_solver_t solver, int* layout, double* ms);

/* Device copy bandwidth: a streaming 16-byte-per-lane copy of `bytes` bytes,
 * timed with HIP events over `reps` launches; *gbs = 2*bytes/time (read +
 * write) in GB/s. The measured ceiling against which the pass kernels are
 * compared, and the calibration kernel for hardware byte counters. */
int sph_measure_copy_bandwidth(int64_t bytes, int reps, double* gbs);

/* run_taylor_green(config) (impl.hpp:1310-1410): fixed-dt or adaptive-CFL
 * driver writing <output_dir>/taylor_green_diagnostics.csv. The fields mirror
 * RunConfig (src/ava_taylor_green.hpp:21-34); precision selects FP32/FP64.
 * final_row (may be NULL) receives the last diagnostics row. */
typedef struct sph_run_config {
    int n;
    int steps;
    int max_steps;
    int output_every;
    int print_every;
    double dt;
    double cfl;
    double t_end;
    double output_dt;
    double reynolds;
    int adaptive_cfl;
    int precision;
    const char* output_dir;
} sph_run_config;

int sph_run_taylor_green(const sph_run_config* config, sph_diagnostics* final_row);

/* ------------------------------------------------------------------------
 * Batched ensemble of independent periodic solves.
 *
 * The reference runs one TaylorGreenSolver per process (impl.hpp:658-1308);
 * this is that solver batched: `count` members, each with its own grid size
 * n[j] (any compiled length), viscosity nu[j] and seeded random-phase
 * solenoidal initial field with E(k) ~ k^4 exp(-2 (k/4)^2), scaled to kinetic
 * energy `energy`. Members of equal size share every pass launch, sizes up
 * to 128 can share one launch per pass (sph_ensemble_launch_mode), and a
 * step is one hipGraph (launch structure 2, cooperative, runs uncaptured).
 * Each member's state follows the same RK3 stage as TaylorGreenSolver::step.
 * ------------------------------------------------------------------------ */
typedef struct sph_ensemble_s* sph_ensemble_t;

int sph_ensemble_create(int count, const int* n, const double* nu, const uint64_t* seeds, double energy,
                        int precision, sph_ensemble_t* out);
int sph_ensemble_destroy(sph_ensemble_t ens);

/* One RK3 step of every member; dt[j] is member j's time step. */
int sph_ensemble_step_rk3(sph_ensemble_t ens, const double* dt);

/* Same meaning as the solver calls, for member `member`. */
int sph_ensemble_diagnostics(sph_ensemble_t ens, int member, double t, sph_diagnostics* out);
int sph_ensemble_max_velocity(sph_ensemble_t ens, int member, double* out);
int sph_ensemble_get_uhat(sph_ensemble_t ens, int member, int component, double* host);
int sph_ensemble_set_uhat(sph_ensemble_t ens, int member, int component, const double* host);
int sph_ensemble_set_graph(sph_ensemble_t ens, int enabled);
int sph_ensemble_synchronize(sph_ensemble_t ens);

/* GPU seconds spent in sph_ensemble_step_rk3 and the number of steps. */
int sph_ensemble_step_seconds(sph_ensemble_t ens, double* seconds, int64_t* steps);

/* Pass timings as sph_solver_time_passes, summed over the distinct sizes. */
int sph_ensemble_time_passes(sph_ensemble_t ens, int reps, double* ms, double* bytes);

/* How the ensemble launches a step (*mode):
 *   0  one chain of launches per size (four per RK stage) on parallel
 *      branches of the step's graph;
 *   1  one launch per pass for all members of the sizes 16, 32, 48, 64, 96
 *      and 128 (four launches per RK stage, one-wavefront tiles of every size
 *      in each launch); other sizes on their own chains beside it;
 *   2  one launch per RK stage for those sizes: a cooperative launch of
 *      resident work-groups that walk the tiles of the four passes with a
 *      grid-wide barrier between passes (not captured into the graph).
 * Chosen at creation by timing each (SPH_ENSEMBLE_FUSED=0 / 1 / 2 forces
 * one). step_ms (3 doubles) = the timed step of each (0 where not timed). */
int sph_ensemble_launch_mode(sph_ensemble_t ens, int* mode, double* step_ms);

/* The ensemble's initial field for one member, on the host: three
 * components of n*n*(n/2+1) interleaved (re, im) doubles, layout as
 * sph_solver_get_uhat. */
int sph_random_solenoidal_field(int n, uint64_t seed, double energy, double* u0, double* u1, double* u2);

#ifdef __cplusplus
}
#endif

#endif /* SPECTRAL_HIP_H */
