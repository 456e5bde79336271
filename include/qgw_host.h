/*
 * qgw_host.h -- host-side C API (libqgw_host.so): parameter file reader,
 * grid set-up, vertical-mode decomposition and the seeded initial condition.
 *
 * Everything here runs once per run, in FP64, on the host.  The per-step hot
 * path lives behind qgw_hip.h.
 */
#ifndef QGW_HOST_H
#define QGW_HOST_H

#include <stddef.h>
#include <stdint.h>
#include "qgw_params.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes shared with qgw_hip.h. */
#define QGW_OK 0
#define QGW_ERR_ARG -1        /* invalid argument or unsupported configuration */
#define QGW_ERR_IO -2         /* file not found / unreadable */
#define QGW_ERR_NOMEM -3      /* host or device allocation failed */
#define QGW_ERR_HIP -4        /* a HIP runtime call failed */
#define QGW_ERR_NUMERIC -5    /* eigen-solver failure, NaN/Inf in the state */
#define QGW_ERR_STATE -6      /* call out of order (e.g. modes not set) */

/* Fill *p with the reference's defaults (src/qg.c:58-83, src/forcing.h:13-14,40-41)
 * and the extension defaults.  Replaces the static initialisers of the
 * reference's globals. */
void qgwh_default_params(qgw_params *p);

/* Read a "key = value" parameter file into *p (only the keys present are
 * changed).  Same line syntax as the reference's read_params
 * (src/extra.h:98-153): spaces are removed, a line splits at '=', arrays are
 * written [a,b,c], unknown keys and comment lines are ignored.  Unlike the
 * reference (which calls exit(0) when the file is missing, src/extra.h:149-151)
 * this returns QGW_ERR_IO. */
int qgwh_read_params(const char *path, qgw_params *p);

/* Grid geometry (src/domain.h:64-176).  n, m are the interior point counts:
 * NX-1, NY-1 for bounded runs (bc_fac in [0, 1]), NX, NY for periodic runs
 * (bc_fac = -1); storage per layer is (n+2)*(m+2) with a one-point ring. */
typedef struct qgw_grid {
  int n, m;           /* interior points in x and y */
  int nxp2, nyp2;     /* storage extents (n+2, m+2) */
  double Delta;       /* grid spacing Lx/NX */
  double Lx, Ly;      /* domain lengths (Ly = NY*Delta) */
} qgw_grid;

/* Compute the grid geometry; X (nxp2), Y (nyp2), Kw (nyp2) and Lw (nxp2) may
 * be NULL.  Bounded: X[i] = i*Delta, Kw[j] = pi*j/Lx for j=1..m and
 * Lw[i] = pi*i/Lx for i=1..n, exactly as src/domain.h:169-174 (including the
 * use of Lx for K).  Periodic (src/domain.h:163-167): X[i] = (i-1)*Delta,
 * Kw[i] = 2 pi i/Lx for i = 0..NX/2 (the half-spectrum x wavenumbers) and
 * Lw[j] = 2 pi/Lx * (((j + NX/2) mod NX) - NX/2) for j = 0..NX-1 (signed, y);
 * a periodic grid must be square with NX even (the serial reference's r2c
 * layout, src/elliptic.h:32-47).  Returns QGW_ERR_ARG for inconsistent
 * inputs. */
int qgwh_grid(const qgw_params *p, qgw_grid *g, double *X, double *Y, double *Kw, double *Lw);

/* Bathtub topography on the full storage grid, (m+2)*(n+2) values, row-major
 * (src/domain.h:200-207). */
int qgwh_topo(const qgw_params *p, double *topo);

/* Vertical modes of the stretching matrix (src/eigmode.h:87-271):
 *   cl2m[k*nl+l]  layer->mode projection (row k = left eigenvector of mode k)
 *   cm2l[l*nl+k]  mode->layer projection (column k = right eigenvector)
 *   iRd2[k]       eigenvalue of mode k (1/Rd^2), iRd2[0] = 0 (or 1/Ld^2)
 * Modes are sorted by ascending eigenvalue, right vectors carry the Flierl
 * (1978) normalisation sum_l dh_l v_l^2 = H with a positive surface value, and
 * the left vectors are scaled so that cl2m * cm2l = I.
 * The stretching matrix is symmetric in the dh-weighted inner product, so it
 * is solved as a symmetric tridiagonal problem by implicit QL (in place of
 * the reference's LAPACKE_dgeev call, src/eigmode.h:145). */
int qgwh_eigmode(const qgw_params *p, double *cl2m, double *cm2l, double *iRd2);

/* The stretching matrix Gamma itself, row-major nl*nl (src/eigmode.h:99-122). */
int qgwh_gamma(const qgw_params *p, double *gamma);

/* Wind forcing of the top layer per grid row j (0..m+1): the value of the
 * reference's forcing_q macro (src/domain.h:33). */
int qgwh_wind_row(const qgw_params *p, double *forc_row);

/* Seeded initial condition (no counterpart in the reference, which reads q
 * from restart.nc, src/qg.c:160-164).  Per layer, a sine series (bounded) or
 * a Fourier series (periodic, zero mean, periodic rings) with Gaussian
 * coefficients of variance ~ E(kappa)/kappa^3,
 * E(kappa) = kappa^3/(kappa0^6 + kappa^6), kappa0 = 2 pi/ic_lambda0, is
 * transformed to psi, rescaled so that the RMS of the forward-difference
 * velocity (src/timestep.h:77-78) equals ic_urms, and q is formed as the exact
 * spectral inverse of the PV inversion (so invert_pv(q) returns psi).
 * psi and q are full storage arrays nl*(m+2)*(n+2) (ring included; zero ring
 * for psi, q ring = 2*bc_fac/Delta^2 * adjacent psi).  RNG: SplitMix64 with
 * Box-Muller, seeded with `seed`. */
int qgwh_make_ic(const qgw_params *p, const double *cl2m, const double *cm2l,
                 const double *iRd2, uint64_t seed, double *psi, double *q);

/* Unnormalised 2-D sine transform (type I on both axes, FFTW RODFT00
 * convention: Y_k = 2 sum_j X_j sin(pi (j+1)(k+1)/(n+1))) of an m-by-n
 * row-major array (x fastest), in place.  Used by the initial condition. */
int qgwh_dst2d(double *a, int n, int m);

/* Domain-integrated diagnostics of a full storage state, FP64:
 *   E_layer[l] = -sum psi*q * Delta^2/Lx^2,  Z_layer[l] = sum q^2 * Delta^2/Lx^2
 * over all stored points (the energy and enstrophy formulas of
 * test/test_energy_injection.py:24-25), and thickness-weighted totals
 *   E = sum_l dh_l E_layer[l] / (2H),  Z = sum_l dh_l Z_layer[l] / H. */
int qgwh_diagnostics(const qgw_params *p, const double *psi, const double *q,
                     double *E_layer, double *Z_layer, double *E_tot, double *Z_tot);

/* Initial dt and DT_max with the reference's exact expression quirks
 * (check_timestep, src/timestep.h:18-54, with the unparenthesised min/max
 * macros of src/qg.c:40-41).  nu_hyper must already carry the model's sign
 * (qgwh_nu_hyper_signed). */
void qgwh_check_timestep(const qgw_params *p, double nu_hyper_signed, double *dt, double *DT_max);

/* nu_hyper with the sign rule of src/domain.h:210. */
double qgwh_nu_hyper_signed(const qgw_params *p);

/* ---- stochastic ring forcing (src/forcing.h:67-208) ---------------------- */

/* Forced wavenumbers of init_stoch_forc (src/forcing.h:131-149): with
 * N = NX (the reference's N_F: NX periodic, NXp1 = NX bounded), k = k_f/(2 pi)
 * and dk = 1/Lx, the half-spectrum points (i, j), i in [0, N/2], j in [0, N),
 * whose l = ((j + N/2) mod N - N/2) dk and k_i = i dk lie within dk/2 of the
 * circle |(k_i, l)| = k along x or along y, in the reference's order (j outer,
 * i inner).  *n_forced gets their number N_P; up to cap of them go to ii, jj
 * (may be NULL).  *envelope (may be NULL) gets sqrt(8/(2(N_P-1))) pi k, the
 * amplitude of calc_forc (src/forcing.h:180), or 0 when N_P <= 1 (then the
 * reference adds nothing). */
int qgwh_stoch_ring(const qgw_params *p, int cap, int *ii, int *jj, int *n_forced, double *envelope);

/* Random stream of the forcing.  Seed 1 is the stream of the C library's
 * rand() without srand, which the reference draws from (src/forcing.h:62-63);
 * each stream is private to its handle. */
typedef struct qgwh_rng qgwh_rng;
qgwh_rng *qgwh_rng_create(unsigned seed);
void qgwh_rng_destroy(qgwh_rng *rng);
/* The next value, as rand() would return it. */
int qgwh_rng_next(qgwh_rng *rng);
/* A copy of a stream's state (the copy continues where the source is). */
qgwh_rng *qgwh_rng_clone(const qgwh_rng *rng);
int qgwh_rng_assign(qgwh_rng *dst, const qgwh_rng *src);

/* One step of calc_forc's coefficients (src/forcing.h:174-190) for count forced
 * points: per point, in order, magnitude = sigma_f * normal_noise() and
 * phase = noise() * 2 pi, coef[2p] = envelope * magnitude * cos(phase) and
 * coef[2p+1] = envelope * magnitude * sin(phase).  normal_noise() draws two
 * values (the log argument first: the order GCC evaluates the reference's
 * macro in), noise() one. */
int qgwh_stoch_coeffs(qgwh_rng *rng, double sigma_f, double envelope, int count, double *coef);

/* Human-readable text for a status code. */
const char *qgwh_strerror(int status);

/* sizeof(qgw_params), for bindings that mirror the struct. */
size_t qgwh_params_size(void);

#ifdef __cplusplus
}
#endif

#endif /* QGW_HOST_H */
