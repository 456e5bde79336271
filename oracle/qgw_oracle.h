/*
 * qgw_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle is a CPU FP64 restatement of the reference's timestep hot path
 * (bounded branch), written from the reference sources in their loop order,
 * with its own FFT-based sine transform and its own (Jacobi) eigen-solver in
 * place of FFTW and LAPACKE, which are absent here.  Only the tests, the
 * smoke check and the benchmark's CPU-baseline leg may load it; the product
 * (libqgw_hip.so, libqgw_host.so) never does.
 *
 * Function -> reference map (file:line):
 *   orc_create ............. init_domain src/domain.h:64-176, init_elliptic
 *                            src/elliptic.h:29-74, init_vars src/domain.h:178-212,
 *                            init_timestep/check_timestep src/timestep.h:56-67,18-54
 *   orc_modes .............. compute_eigmode src/eigmode.h:87-271
 *   orc_invert_pv .......... invert_pv src/elliptic.h:76-155 (bounded branch)
 *   orc_adjust_bc .......... adjust_bc src/domain.h:221-349 (bounded branch)
 *   orc_calc_diff .......... calc_diff src/dynamics.h:13-44
 *   orc_rhs ................ rhs src/dynamics.h:47-73 with the stencil macros
 *                            src/domain.h:12-36
 *   orc_adjust_timestep .... adjust_timestep src/timestep.h:69-111
 *   orc_timestep ........... timestep src/timestep.h:114-167 (no stochastic
 *                            forcing), AB3 coefficients doc/timestep.org:95-97
 *   orc_set_forcing4d ...... the records of init_4d_forcing/calc_4d_forcing
 *                            src/forcing.h:226-289, applied in orc_timestep
 *   orc_set_stochastic ..... init_stoch_forc src/forcing.h:67-166; calc_forc
 *                            src/forcing.h:168-208 in orc_timestep
 *   orc_c2r2d .............. the c2r plan of src/forcing.h:119,196
 *   orc_dst2d / orc_dst1d .. FFTW RODFT00 definition (src/elliptic.h:63-66)
 *   orc_diagnostics ........ test/test_energy_injection.py:24-25
 *
 * Pins (each checked in tests/): DST against scipy.fft.dst(type=1) fixtures,
 * the O(n^2) direct sum and DST o DST = 2(n+1) I; eigenmodes against
 * scipy.linalg.eigh fixtures, the nl=2 closed form and the invariants;
 * invert_pv against single-sine-mode closed forms; the Jacobian against its
 * antisymmetry and exactness identities and a brute-force 12-term sum; AB3
 * coefficients against an independent sympy integration (the construction of
 * doc/timestep.org); check_timestep against hand-expanded cases of the macro
 * quirks; adjust_timestep against hand-computed CFL/doubling/snap cases;
 * diagnostics against a sine-mode closed form; the 4-D forcing against the
 * hand-computed triangle wave of a three-record ramp on a sine mode, where
 * the dynamics vanish exactly; the stochastic ring against a hand-counted
 * ring and the integer geometry of the selection rule, its c2r against
 * numpy.fft.irfft2, its energy injection against the sigma^2 rate of
 * doc/forcing.org.  Each term of orc_rhs alone against a closed form on a
 * sine mode (beta as the centred x-difference of f = f0 + beta*y, nu lap q
 * through the 5-point eigenvalue, the wind curl of doc/usage.org at worked
 * points, the drag as rate r = hEkb*f0/(2 dh[nl-1]) on the bottom vorticity,
 * J(psi, topo) of a linear ramp as a topographic beta), their superposition,
 * the bc_fac > 0 wall ring against Thom's no-slip wall vorticity 2 psi_1/Delta^2
 * and the W/E corner ownership, and the bathtub topography at worked points;
 * each of these pins is shown to fail under the planted mistakes ORC_MUT_*.
 * No function is unpinned.
 */
#ifndef QGW_ORACLE_H
#define QGW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NL_MAX 1000

typedef struct orc_params {
  int NX, NY, nl, n_hyper;
  double Lx;
  double dh[ORC_NL_MAX];
  double N2[ORC_NL_MAX];
  double tau0, forc_mode, f0, beta, nu, nu_hyper, hEkb, Ld, h_topo, w_topo, bc_fac;
  double cfl, sigma_f, tend, dt_out;
  double dt_fixed; /* > 0: fixed step extension */
} orc_params;

typedef struct orc orc;

/* fields for orc_field */
#define ORC_Q 1
#define ORC_PSI 2
#define ORC_TOPO 3
#define ORC_F0 4
#define ORC_F1 5
#define ORC_F2 6
#define ORC_LAPN 7 /* lap_n_diff */
#define ORC_FORC 8 /* stochastic forcing field of the last step (layer 0) */

/* Planted mistakes for orc_mutate (tests only): each names one plausible
 * misreading of the reference that a pin in tests/ must catch. */
#define ORC_MUT_NONE 0
#define ORC_MUT_BETA_SIGN 1    /* +beta*psi_x instead of -beta*psi_x */
#define ORC_MUT_BETA_AXIS 2    /* beta term differenced along y */
#define ORC_MUT_NU_PSI 3       /* nu * lap(psi) instead of nu * lap(q) */
#define ORC_MUT_WIND_LAYER 4   /* wind forcing on the bottom layer */
#define ORC_MUT_WIND_DH 5      /* wind forcing divided by dh[nl-1] */
#define ORC_MUT_WIND_LX 6      /* wind forcing with Lx in place of Ly */
#define ORC_MUT_DRAG_DH 7      /* drag divided by dh[0] */
#define ORC_MUT_DRAG_HALF 8    /* drag without the factor 1/2 */
#define ORC_MUT_DRAG_Q 9       /* drag on lap(q) instead of lap(psi) */
#define ORC_MUT_TOPO_SWAP 10   /* J(topo, psi) instead of J(psi, topo) */
#define ORC_MUT_TOPO_LAYER 11  /* J(psi[0], topo) instead of J(psi[nl-1], topo) */
#define ORC_MUT_BC_FACTOR 12   /* wall ring bc_fac/Delta^2 instead of 2 bc_fac/Delta^2 */
#define ORC_MUT_BC_ORDER 13    /* S/N written after W/E (S/N own the corners) */
#define ORC_MUT_TOPO_LX 14     /* bathtub shelf at Lx - y instead of Ly - y */
#define ORC_MUT_SHELF_WIDTH 15 /* shelf exp(-x^2/d^2) instead of exp(-x^2/(2 d^2)) */
void orc_mutate(int id);  /* set before orc_create for the topography ones */

orc *orc_create(const orc_params *p, int *status);
void orc_destroy(orc *s);
int orc_modes(orc *s, double *cl2m, double *cm2l, double *iRd2);
int orc_set_modes(orc *s, const double *cl2m, const double *cm2l, const double *iRd2);
double *orc_field(orc *s, int which);
void orc_sizes(orc *s, int *n, int *m, int *nl);
void orc_invert_pv(orc *s);
void orc_adjust_bc(orc *s, double *field, int type);
void orc_calc_diff(orc *s);
void orc_rhs(orc *s, int which_out);
double orc_adjust_timestep(orc *s);
void orc_timestep(orc *s);
void orc_run(orc *s, int nsteps);
/* 4-D forcing records: n_rec x (nl storage fields), spaced dt_forc (copied);
 * n_rec = 0 switches it off */
int orc_set_forcing4d(orc *s, int n_rec, double dt_forc, const double *records);
/* stochastic ring forcing (k_f = 0 switches it off); returns the number of
 * forced points N_P.  Resets the C library's rand() stream (srand(1)). */
int orc_set_stochastic(orc *s, double k_f, double sigma_f);
int orc_stoch_points(orc *s, int *ii, int *jj);
int orc_stoch_draw(orc *s, double *coef);   /* one step's coefficients, 2 per point */
int orc_c2r2d(const double *spec, int N, double *out); /* N x (N/2+1) complex -> N x N */
void orc_get_time(orc *s, double *t, double *dt, double *dt0, double *dt1, int *it, double *DT_max);
void orc_set_time(orc *s, double t, double dt, double dt0, double dt1, int it, double t_out);
void orc_diagnostics(orc *s, double *E_layer, double *Z_layer, double *E_tot, double *Z_tot);
void orc_ab3_coeffs(double dt, double dt0, double dt1, double *c);
double orc_nu_hyper(orc *s);

/* stand-alone transforms and solvers */
int orc_dst1d(double *x, int n);           /* unnormalised RODFT00, in place */
int orc_dst2d(double *a, int n, int m);    /* m rows of n (x fastest), in place */
int orc_fft(double *re, double *im, int L);/* forward complex DFT, in place */
int orc_eig_sym(int n, const double *A, double *w, double *V); /* Jacobi, ascending */
void orc_check_timestep(const orc_params *p, double *dt, double *DT_max);
int orc_threads(int n);                    /* set OpenMP threads (0: default); returns count */
void orc_jacobian_field(orc *s, const double *p, const double *q, double *out);
void orc_laplacian_field(orc *s, const double *p, double *out);

#ifdef __cplusplus
}
#endif

#endif
