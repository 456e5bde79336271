/*
 * qgw_hip.h -- the device C-ABI of the MI355X timestep hot path
 * (libqgw_hip.so).  Plain C types only.
 *
 * The reference exposes its hot path as global C functions called from main
 * (src/qg.c:153-200) over process-global state.  Here the same calls take an
 * explicit context so that several runs (an ensemble) can live in one
 * process; each entry point names the reference function it replaces.
 *
 * Host arrays always use the reference storage layout
 *   idx(i,j,k) = k*(n+2)*(m+2) + j*(n+2) + i        (src/domain.h:62)
 * with a one-point ring (i = 0, n+1 and j = 0, m+1).  Device arrays are owned
 * by the context and are never exposed.
 *
 * All calls return QGW_OK (0) or a negative status (qgw_host.h); none calls
 * exit().  Calls on one context are not thread-safe; calls on different
 * contexts are independent (each owns a HIP stream).
 */
#ifndef QGW_HIP_H
#define QGW_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "qgw_params.h"
#include "qgw_host.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Largest layer count per run the device library accepts: the reference's
 * nl_max (src/qg.c:45).  Up to 32 layers the layer<->mode projections hold a
 * point's layers in registers (or feed 32 x 32 matrix-core tiles in FP32 from
 * 16 layers); above 32 they run over 32 x 32 blocks of the matrix, each block
 * of 32 modes stored by its first block of layers and accumulated by the
 * others (ceil(nl/32)^2 passes over the spectral array). */
#define QGW_DEV_NL_MAX 1000

/* Arithmetic forms of the step. */
#define QGW_FP64 0        /* every field and pass in FP64 */
#define QGW_FP32 1        /* FP32 fields and passes, AB3 increment with Kahan compensation */
#define QGW_FP32_PLAIN 2  /* FP32 fields and passes, uncompensated AB3 */

/* Fields for qgw_upload / qgw_download (full storage layout, nl layers
 * unless noted). */
#define QGW_FIELD_Q 1      /* potential vorticity q */
#define QGW_FIELD_PSI 2    /* streamfunction psi */
#define QGW_FIELD_TOPO 3   /* bottom topography, one layer */
#define QGW_FIELD_F0 4     /* newest AB3 tendency (f0_ab, src/timestep.h:10) */
#define QGW_FIELD_F1 5     /* previous tendency (f1_ab) */
#define QGW_FIELD_F2 6     /* tendency two steps back (f2_ab) */

typedef struct qgw_ctx qgw_ctx;

/* Create a run on the current HIP device: the grid and arrays of
 * init_domain/init_vars (src/domain.h:64,178), the time-step state of
 * init_timestep/check_timestep (src/timestep.h:56,18) and the transform set-up
 * of init_elliptic (src/elliptic.h:29).  `precision` is QGW_FP64, QGW_FP32 or
 * QGW_FP32_PLAIN.  q, psi and the tendency history start at zero, topography
 * from the bathtub formula.  The vertical modes must be supplied with
 * qgw_set_modes before the first inversion.  Returns NULL and sets *status
 * on failure (1 <= nl <= QGW_DEV_NL_MAX; bounded (bc_fac >= 0): any grid,
 * with the FFT sine transforms when 2(n+1) and 2(m+1) are in {2^a: 8..8192},
 * {3*2^a: 24..6144} or {200, 400, 1000, 2000, 4000, 448} (NX = 100, 200,
 * 500, 1000, 2000, 224), else dense sine transforms, matrix products with
 * S_ij = 2 sin(pi i j/(N+1)) on the matrix cores through rocBLAS (O(N)
 * flops per point instead of O(log N), for any NX, NY the reference's FFTW
 * plans accept); periodic (bc_fac = -1): NX = NY even, in the FFT length
 * sets; n_hyper even and <= 18 (above 8 the update runs in its streaming form
 * only); with
 * stochastic forcing (k_f != 0), NX = NY; with file forcing, dt_forc_period a
 * positive multiple of dt_forc > 0).  When the environment variable
 * QGW_MAX_DEVICE_BYTES holds a positive byte count, a context (or batch group)
 * that would allocate more device memory fails with QGW_ERR_NOMEM and frees
 * what it had allocated.
 *
 * Failure detection: every step's update kernel raises a device flag when a
 * new q value (or, in qgw_rhs, a tendency) is NaN or Inf.  The flag is
 * checked, and QGW_ERR_NUMERIC returned, at the end of qgw_run, qgw_time_run,
 * qgw_synchronize and qgw_diagnostics (and of their batch forms); on the
 * adaptive step the CFL reduction also fails on a non-finite psi.  The flag
 * stays raised until q is uploaded again (a single run). */
qgw_ctx *qgw_create(const qgw_params *p, int precision, int *status);

/* Free the run (clean_fft, clean_timestep, clean_eigmode: src/elliptic.h:157,
 * src/timestep.h:169, src/eigmode.h:28). */
int qgw_destroy(qgw_ctx *ctx);

/* Install the vertical modes computed on the host (outputs of compute_eigmode,
 * src/eigmode.h:87-271; see qgwh_eigmode). cl2m[k*nl+l], cm2l[l*nl+k], iRd2[k]. */
int qgw_set_modes(qgw_ctx *ctx, const double *cl2m, const double *cm2l, const double *iRd2);

/* Copy a host array (reference layout, FP64) to the device, or back.  In the
 * FP32 forms the upload rounds to FP32 and, for q, keeps the rounding residual
 * in the compensation array, so a download returns q to ~2^-48 relative. */
int qgw_upload(qgw_ctx *ctx, int field, const double *host);
int qgw_download(qgw_ctx *ctx, int field, double *host);

/* psi from q: layer->mode projection, 2-D sine transform, Helmholtz division,
 * inverse transform, mode->layer projection, then the wall ring of psi and q
 * (invert_pv, src/elliptic.h:76-155). */
int qgw_invert_pv(qgw_ctx *ctx);

/* Right-hand side into the newest history slot F0 without advancing time:
 * invert_pv, then -J(psi,q) - beta psi_x + nu lap q + nu_hyper lap^(n_hyper+2) psi,
 * wind on layer 0, Ekman drag and topography on layer nl-1
 * (rhs + calc_diff, src/dynamics.h:47-73, 13-44). */
int qgw_rhs(qgw_ctx *ctx);

/* CFL-limited time step from the current psi with the reference's doubling,
 * DT_max cap and snap to the next output time (adjust_timestep,
 * src/timestep.h:69-111).  Updates the context's dt and returns it in *dt_out.
 * With dt_fixed > 0 it returns dt_fixed and changes nothing. */
int qgw_adjust_timestep(qgw_ctx *ctx, double *dt_out);

/* One step: rotate the tendency history, rhs, new dt, variable-step AB3 update
 * of the interior of q, t += dt (timestep, src/timestep.h:114-167). */
int qgw_timestep(qgw_ctx *ctx);

/* nsteps consecutive qgw_timestep calls.  With a fixed dt and no 4-D forcing
 * record change the whole sequence is enqueued without host synchronisation
 * (the tendency buffers rotate by pointer); the call then waits for the
 * stream once, to check the non-finite flag (QGW_ERR_NUMERIC).  The
 * graph-replayed step is the batch form (qgw_batch_step). */
int qgw_run(qgw_ctx *ctx, int nsteps);

/* Energy and enstrophy of the current (psi, q) pair, accumulated in FP64:
 * per layer E_layer[l] = -sum psi*q*Delta^2/Lx^2 and Z_layer[l] = sum q^2*Delta^2/Lx^2
 * over all stored points (test/test_energy_injection.py:24-25), and the
 * thickness-weighted totals E = sum_l dh_l E_l/(2H), Z = sum_l dh_l Z_l/H.
 * Any output pointer may be NULL.  Call after qgw_invert_pv for the state at
 * the current time. */
int qgw_diagnostics(qgw_ctx *ctx, double *E_layer, double *Z_layer, double *E_tot, double *Z_tot);

/* Allocation-free form of qgw_diagnostics.  No counterpart in the reference:
 * it is for callers that cannot allocate or wait (code under HIP graph
 * capture, tight serving loops).  Enqueues on the context's stream, without
 * waiting, the per-layer FP64 sums of the current (psi, q) pair into the
 * caller's DEVICE array sums[2*nl]: sums[2l] = sum psi*q and
 * sums[2l+1] = sum q^2 over all stored points of layer l (the compensated q
 * in the QGW_FP32 form), so that E_layer[l] = -sums[2l]*Delta^2/Lx^2 and
 * Z_layer[l] = sums[2l+1]*Delta^2/Lx^2 are exactly what qgw_diagnostics
 * reports.  The partial sums live in a fixed scratch array of the library
 * (256 KiB per GPU, for at most 64 layers; more layers return QGW_ERR_ARG);
 * it calls no allocator.  Calls that overlap on one GPU (several contexts'
 * streams at once) race on that scratch: order them on one stream or with
 * events. */
int qgw_diagnostics_device(qgw_ctx *ctx, double *sums);

/* Time-step state (the reference globals t, dt, dt0, dt1, it; src/qg.c:63-70,
 * src/timestep.h:14). */
int qgw_get_time(qgw_ctx *ctx, double *t, double *dt, double *dt0, double *dt1, int *it);
int qgw_set_time(qgw_ctx *ctx, double t, double dt, double dt0, double dt1, int it);

/* Time of the next output: the variable step is shortened so that an integer
 * number of steps lands on it (t_out, src/timestep.h:98-109, advanced by the
 * driver at each output, src/qg.c:185-188). */
int qgw_set_next_output(qgw_ctx *ctx, double t_out);

/* File of the 4-D forcing (init_4d_forcing/calc_4d_forcing,
 * src/forcing.h:226-289), used when dt_forc_period != 0: variable
 * "q_forcing_3d" (t, level, y, x) with dt_forc_period/dt_forc records, read
 * by qgwh_nc_read_var.  Default "input_vars.nc" in the working directory, as
 * in the reference.  Each step adds (w1 F[r] + w2 F[r+1 mod nrec]) dt to the
 * interior of q after the AB3 update, r = floor(t/dt_forc) mod nrec and
 * w2 = t/dt_forc - floor(t/dt_forc) at the time t before the step; the two
 * records are read when r changes.  A record that cannot be read fails the
 * step with QGW_ERR_IO before the state changes (the reference prints the
 * error and carries on with the previous records). */
int qgw_set_forcing_file(qgw_ctx *ctx, const char *path);

/* Stochastic ring forcing (init_stoch_forc/calc_forc, src/forcing.h:67-208),
 * on when k_f != 0: each step draws sigma_f N(0,1) e^{i theta} times the
 * envelope for every forced point of qgwh_stoch_ring (qgwh_stoch_coeffs, on
 * the host), the device forms the real field of that half spectrum (the
 * unnormalised c2r of the reference's N x N plan, N = NX) and adds it times
 * sqrt(dt) to the interior of layer 0 after the AB3 update, before the 4-D
 * forcing (src/timestep.h:143-150).  The stream starts as the C library's
 * rand() stream does (seed 1), so a run draws the reference's sequence; it
 * is private to the context.  This call restarts it from another seed. */
int qgw_set_stoch_seed(qgw_ctx *ctx, unsigned seed);

/* Wait for all work queued on the context's stream; QGW_ERR_NUMERIC if a
 * step has produced a non-finite q. */
int qgw_synchronize(qgw_ctx *ctx);

/* Time nsteps steps of qgw_run between two HIP events recorded on the
 * context's stream; *ms is the elapsed time. */
int qgw_time_run(qgw_ctx *ctx, int nsteps, double *ms);

/* Per-pass timing: while enabled, every kernel launch of the step is
 * bracketed by HIP events on the context's stream.  qgw_profile_read returns,
 * per pass id (QGW_PASS_*), the accumulated milliseconds and launch counts,
 * and resets them. */
#define QGW_PASS_P1 0      /* projection + forward row transform */
#define QGW_PASS_T1 1      /* transpose to column-major spectral layout */
#define QGW_PASS_P2 2      /* column transform, Helmholtz division, inverse column transform */
#define QGW_PASS_T2 3      /* transpose back */
#define QGW_PASS_P3 4      /* inverse row transform + back-projection + wall ring */
#define QGW_PASS_P4 5      /* tendency + AB3 update */
#define QGW_PASS_CFL 6     /* CFL reduction (adaptive dt only) */
#define QGW_PASS_DIAG 7    /* diagnostics reduction */
#define QGW_PASS_FORC 8    /* forcing after the AB3 update (stochastic synthesis, 4-D records) */
#define QGW_NPASS 9
int qgw_profile_enable(qgw_ctx *ctx, int on);
int qgw_profile_read(qgw_ctx *ctx, double *ms, int *count, int npass);

/* Tuning knobs.  QGW_TUNE_P4_TILE selects the tile of the tendency/AB3
 * kernel: 0: 64x8, 1: 64x16, 2: 64x32, 3: 128x16 (columns x rows) LDS
 * tiles, 4: register-streaming wavefront strips (default), 5: the same
 * strips in blocks of 4 over 224 output columns whose boundaries fall on
 * 128-byte lines (faster at 4096^2). */
#define QGW_TUNE_P4_TILE 1
/* QGW_TUNE_T_KERNEL selects the layer<->mode projection: 0 automatic (matrix
 * cores for FP32 at 16 or more layers, else VALU), 1 VALU, 2 matrix cores on
 * 32-column tiles, 3 matrix cores on 16-column tiles (2 and 3: FP32 forms
 * only). */
#define QGW_TUNE_T_KERNEL 2
/* QGW_TUNE_P4_WAVES sets the row chunk of the streaming P4 so that about
 * 2, 4, 8, 16, 32, 40 (default), 48, 64, 96, 128, 192, 256, 384, 512 or 768
 * wavefronts per SIMD are launched (chunks of at least 16 rows).  Shorter
 * chunks are faster down to about 16 rows at 2048^2 x 10 (40-48) and 26 rows
 * at 4096^2 x 32 (384); the autotuner tries 16 to 384. */
#define QGW_TUNE_P4_WAVES 3
/* QGW_TUNE_DST selects the sine transforms of P1/P2/P3: 1 through complex
 * FFTs of length n+1 (m+1) with a pre/post pass (both half lengths n+1 and
 * m+1 in 64..4096 or 96..3072), 2 (default where 1 is provided) as 1 with
 * the row passes P1/P3 in persistent blocks that load the next row pair
 * while transforming the current one, 0 through complex FFTs of the odd
 * extension, length 2(n+1) (2(m+1)); 3 (reported, and the only choice) for
 * the dense sine transforms of grids the FFT lengths do not cover. */
#define QGW_TUNE_DST 4
/* QGW_TUNE_FUSE: 1 runs the inversion in four passes over HBM -- the
 * layer<->mode projections inside the row transforms, the column transforms
 * on column tiles of the mode-major spectral array, no transposed copy
 * (where the half-length transforms apply on a bounded grid); 0, the
 * default, in six: the projections fused into two transposes.  Fewer bytes,
 * but the column-tile pass is slower than the transposed column pass: at
 * 2048^2 x 10 the two forms take about the same time with plain psi stores,
 * the four-pass form is the faster (by 2%) with the non-temporal psi stores
 * of QGW_TUNE_PSI_NT; the autotuner tries both. */
#define QGW_TUNE_FUSE 7
/* QGW_TUNE_P2M_PAIRS: column pairs per block of the four-pass column pass,
 * 1, 2, 4 or 8 (capped by the 1024-thread block and the LDS; 0: the most
 * that fit, 64-byte row segments).  Default: 4 in the FP32 forms, 2 in FP64,
 * the fastest measured at 2048^2 x 10; the autotuner tries 2, 4 and 8. */
#define QGW_TUNE_P2M_PAIRS 8
/* QGW_TUNE_PSI_NT: 1 (default) stores psi from the half-length inverse row
 * passes (six- and four-pass forms) with non-temporal stores, 0 with plain
 * ones.  The tendency pass that reads psi next runs faster after the
 * non-temporal stores at 2048^2 x 10 (by more than the row pass loses), not at
 * 4096^2 x 32; the autotuner tries both.  Results are the same either way. */
#define QGW_TUNE_PSI_NT 9
/* QGW_TUNE_P4_REV: 1 (default) runs the interior walks of the odd row chunks
 * of the streaming tendency pass (QGW_TUNE_P4_TILE 4 and 5) up their chunk,
 * 0 every walk down.  Two neighbouring chunks both load the psi and q rows
 * around their common boundary; walking in opposite directions they load
 * them at about the same time, so the second load meets L2.  Every stencil
 * is evaluated in the same order either way, so results are the same bits. */
#define QGW_TUNE_P4_REV 10
/* QGW_TUNE_P4_ORDER: the order in which the walks of the streaming tendency
 * pass (QGW_TUNE_P4_TILE 4 and 5) are dealt, strips fastest: 1 the layers of
 * a row chunk before the next chunk, 0 the chunks of a layer before the next
 * layer.  Default 1 up to 16 layers (4% faster at 2048^2 x 10), else 0 (3-4%
 * faster at 4096^2 x 32); the autotuner tries both.  Results are the same
 * either way. */
#define QGW_TUNE_P4_ORDER 11
/* QGW_TUNE_FFT_E: radix schedule of the half-length FFTs of power-of-two
 * lengths from 1024: 16 (radix-16 stages; the default) or 8 (radix-8 stages,
 * twice the threads per row), for the row passes P1 and P3. */
#define QGW_TUNE_FFT_E 5
/* QGW_TUNE_FFT_E_COL: the same choice for the column pass P2 (forward
 * transform, Helmholtz division, inverse transform): 16 is the default of the
 * FP32 forms, 8 of QGW_FP64. */
#define QGW_TUNE_FFT_E_COL 6
int qgw_tune_set(qgw_ctx *ctx, int knob, int value);
int qgw_tune_get(qgw_ctx *ctx, int knob, int *value);

/* Autotuner: times `reps` steps with each P4 variant (and, for the streaming
 * one, each QGW_TUNE_P4_WAVES value) and keeps the fastest (returned in
 * *p4_tile, its ms per step in *best_ms); it also chooses QGW_TUNE_DST and
 * QGW_TUNE_FFT_E and QGW_TUNE_FFT_E_COL, and QGW_TUNE_T_KERNEL for the FP32 forms.  The state (fields,
 * buffer rotation and time) is copied before the timing runs and restored
 * after them, so the run continues as if the call had not been made; the copy
 * takes one more state's worth of device memory while it runs. */
int qgw_autotune(qgw_ctx *ctx, int reps, int *p4_tile, double *best_ms);

/* Bytes of device memory held by the context. */
int qgw_device_bytes(qgw_ctx *ctx, size_t *bytes);

/* Human-readable text for a status code; the last HIP error text of the
 * context is appended by qgw_last_error. */
const char *qgw_strerror(int status);
const char *qgw_last_error(qgw_ctx *ctx);

/* ---- Ensembles ------------------------------------------------------------
 * No counterpart in the reference, which runs one model per process (its
 * state is process-global, src/qg.c:48-52).  A batch advances independent
 * runs ("members") together on one stream.  Members whose grid and step
 * scalars agree (NX, NY, nl, Lx, n_hyper, nu_hyper, nu, beta, f0, hEkb,
 * dh[nl-1], bc_fac, dt_fixed) form a group that shares every launch of a
 * step: its members are stacked along the layer axis.  Each member keeps its
 * own modes, wind (tau0, forc_mode), topography and state.  Every member needs
 * a fixed step (dt_fixed > 0).  qgw_batch_step replays a hipGraph of six
 * steps of every group (the period of the buffer rotation), captured on first
 * use; remainders run as ordinary launches.  Members with stochastic
 * (k_f != 0) or 4-D file forcing (dt_forc_period != 0) form groups of their
 * own and are stepped launch by launch on the batch stream beside the
 * replayed graph (their steps take host work: the random draws and the
 * record reads); each starts the random stream of the C library, as a run of
 * the reference does (qgw_batch_set_stoch_seed gives it another), and reads
 * input_vars.nc unless qgw_batch_set_forcing_file names another file.
 * Members are numbered in the order of `params`. */
typedef struct qgw_batch qgw_batch;

qgw_batch *qgw_batch_create(int nmembers, const qgw_params *params, int precision, int *status);
int qgw_batch_destroy(qgw_batch *b);
/* Number of groups, and each member's group (group_of_member[nmembers], may be NULL). */
int qgw_batch_groups(qgw_batch *b, int *ngroups, int *group_of_member);
int qgw_batch_set_modes(qgw_batch *b, int member, const double *cl2m, const double *cm2l, const double *iRd2);
int qgw_batch_upload(qgw_batch *b, int member, int field, const double *host);
int qgw_batch_download(qgw_batch *b, int member, int field, double *host);
/* invert_pv (src/elliptic.h:76) of every member. */
int qgw_batch_invert_pv(qgw_batch *b);
/* nsteps of timestep (src/timestep.h:114) for every member; waits for the
 * stream at the end and returns QGW_ERR_NUMERIC if any group's q has become
 * non-finite (the flag of a group of several members stays raised). */
int qgw_batch_step(qgw_batch *b, int nsteps);
/* As qgw_batch_step, timed with HIP events on the batch stream.  The graph of
 * the current state is captured and instantiated before the start event, so
 * the timed region holds only launches. */
int qgw_batch_time_step(qgw_batch *b, int nsteps, double *ms);
/* Capture and instantiate the graph for the current state now, instead of at
 * the first qgw_batch_step that replays it (no step is run; nothing is done
 * when the batch cannot replay a graph, e.g. while profiling). */
int qgw_batch_prepare(qgw_batch *b);
int qgw_batch_synchronize(qgw_batch *b);
int qgw_batch_diagnostics(qgw_batch *b, int member, double *E_layer, double *Z_layer, double *E_tot,
                          double *Z_tot);
int qgw_batch_get_time(qgw_batch *b, double *t, int *it);
/* The forcing of one member (qgw_set_stoch_seed, qgw_set_forcing_file). */
int qgw_batch_set_stoch_seed(qgw_batch *b, int member, unsigned seed);
int qgw_batch_set_forcing_file(qgw_batch *b, int member, const char *path);
/* Graph replay on (default) or off (every step launched one kernel at a time). */
int qgw_batch_use_graph(qgw_batch *b, int on);
/* Number of graph replays so far. */
int qgw_batch_graph_launches(qgw_batch *b, int *count);
/* Autotune every group (as qgw_autotune; state preserved); the choice of
 * group k is returned in p4_tile_per_group[k] (may be NULL). */
int qgw_batch_autotune(qgw_batch *b, int reps, int *p4_tile_per_group);
int qgw_batch_tune_set(qgw_batch *b, int group, int knob, int value);
int qgw_batch_tune_get(qgw_batch *b, int group, int knob, int *value);
/* Per-pass timing summed over groups (steps then run without the graph). */
int qgw_batch_profile_enable(qgw_batch *b, int on);
int qgw_batch_profile_read(qgw_batch *b, double *ms, int *count, int npass);
int qgw_batch_device_bytes(qgw_batch *b, size_t *bytes);
const char *qgw_batch_last_error(qgw_batch *b);

#ifdef __cplusplus
}
#endif

#endif /* QGW_HIP_H */
