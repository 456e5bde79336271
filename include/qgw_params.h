/*
 * qgw_params.h -- the run parameters of the multi-layer QG model, shared by
 * the host library (libqgw_host.so) and the device library (libqgw_hip.so).
 *
 * The registered keys, their types and their defaults follow the reference's
 * globals and its parameter registry:
 *   defaults  src/qg.c:58-83, src/forcing.h:13-14, src/forcing.h:40-41
 *   registry  src/qg.c:113-137
 * Keys marked "extension" are not read by the reference; the reference's own
 * reader ignores unknown keys (src/extra.h:129-145), so files written for the
 * reference read identically here, and the extension keys are ignored there.
 */
#ifndef QGW_PARAMS_H
#define QGW_PARAMS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Length of the array parameters dh[] and N2[] (src/qg.c:45). */
#define QGW_NL_MAX 1000

typedef struct qgw_params {
  /* ---- reference keys (src/qg.c:113-137) ---- */
  int NX;                 /* grid points in x (cell count; bounded grids use NX-1 interior points) */
  int NY;                 /* grid points in y */
  int nl;                 /* number of layers */
  int n_hyper;            /* hyperviscosity order (even, >= 0) */
  double Lx;              /* domain length in x [m] */
  double dh[QGW_NL_MAX];  /* layer thicknesses [m], top first */
  double tau0;            /* wind stress / rho0 [m^2 s^-2] */
  double forc_mode;       /* 1: single gyre, 2: double gyre */
  double f0;              /* Coriolis parameter [s^-1] */
  double beta;            /* beta [m^-1 s^-1] */
  double nu;              /* harmonic viscosity on q [m^2/s] */
  double nu_hyper;        /* hyperviscosity [m^n_hyper/s]; sign is set by the model */
  double hEkb;            /* bottom Ekman layer thickness [m] */
  double N2[QGW_NL_MAX];  /* Brunt-Vaisala frequency^2 at the nl-1 interfaces [s^-2] */
  double Ld;              /* deformation radius for a 1.5-layer run [m] */
  double h_topo;          /* bathtub topography height [m] */
  double w_topo;          /* bathtub topography width [m] */
  double bc_fac;          /* 0 free slip, 1 no slip, (0,1) partial slip, -1 periodic */
  double tend;            /* run length [s] */
  double dt_out;          /* output interval [s] */
  double dt_forc;         /* 4-D forcing record spacing [s] */
  double dt_forc_period;  /* 4-D forcing period [s] (0: off) */
  double sigma_f;         /* stochastic forcing amplitude */
  double k_f;             /* stochastic forcing wavenumber (0: off) */
  double cfl;             /* CFL number */

  /* ---- extensions ---- */
  double dt_fixed;        /* > 0: fixed time step [s]; 0: the reference's adaptive step */
  int nsteps;             /* > 0: number of steps to run (driver); 0: run to tend */
  int ic_kind;            /* 0: q and topo from restart.nc when present, else rest; 1: seeded k^-3 field */
  uint64_t seed;          /* RNG seed of the seeded initial condition */
  double ic_urms;         /* target RMS velocity of the initial condition [m/s] */
  double ic_lambda0;      /* peak wavelength of the initial spectrum [m] */
  int precision;          /* 0: FP64, 1: FP32 fields + compensated AB3, 2: FP32 plain */
  int verbose;            /* print initialisation info like the reference's `print` flag */
} qgw_params;

#ifdef __cplusplus
}
#endif

#endif /* QGW_PARAMS_H */
