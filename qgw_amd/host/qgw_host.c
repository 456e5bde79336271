/*
 * qgw_host.c -- host-side set-up of a qgw run, in C (FP64, once per run).
 *
 *   parameter file reader ........ behaviour of src/extra.h:98-153
 *   grid / wavenumbers / topo .... src/domain.h:64-176, 200-207
 *   stretching matrix and modes .. src/eigmode.h:87-271 (own symmetric QL solver)
 *   wind row ..................... src/domain.h:33
 *   initial-dt quirks ............ src/timestep.h:18-54 with src/qg.c:40-41
 *   seeded initial condition ..... no reference counterpart (the reference reads restart.nc)
 *   diagnostics .................. test/test_energy_injection.py:24-25
 */
#define _GNU_SOURCE /* sincos */
#include <ctype.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#include "qgw_host.h"

/* The reference's pi literal (src/qg.c:44); equal to M_PI in binary64. */
static const double QGW_PI = 3.141592653589793;

/* ------------------------------------------------------------------------ */
/* defaults                                                                   */
/* ------------------------------------------------------------------------ */

void qgwh_default_params(qgw_params *p) {
  memset(p, 0, sizeof(*p));
  /* src/qg.c:58-83 */
  p->nl = 1;
  for (int i = 0; i < QGW_NL_MAX; i++) { p->dh[i] = 0.0; p->N2[i] = 0.0; }
  p->dh[0] = 1.0;
  p->N2[0] = 1.0;
  p->tend = 1.0;
  p->dt_out = 0.1;
  p->cfl = 0.2;
  p->beta = 0.0;
  p->nu = 0.0;
  p->n_hyper = 0;
  p->nu_hyper = 0.0;
  p->hEkb = 0.0;
  p->tau0 = 0.0;
  p->forc_mode = 1.0;
  p->f0 = 1.e-5;
  p->bc_fac = 0.0;
  p->Ld = 0.0;
  p->h_topo = 0.0;
  p->w_topo = 1.0;
  /* src/forcing.h:13-14, 40-41 */
  p->sigma_f = 1.0;
  p->k_f = 0.0;
  p->dt_forc = 0.0;
  p->dt_forc_period = 0.0;
  /* extensions */
  p->dt_fixed = 0.0;
  p->nsteps = 0;
  p->ic_kind = 0;
  p->seed = 20261203ULL;
  p->ic_urms = 0.1;
  p->ic_lambda0 = 200.0e3;
  p->precision = 0;
  p->verbose = 0;
}

/* ------------------------------------------------------------------------ */
/* parameter file reader                                                      */
/* ------------------------------------------------------------------------ */

enum { PK_INT, PK_DOUBLE, PK_ARRAY, PK_U64 };

typedef struct {
  const char *name;
  int kind;
  size_t offset;
} param_key;

#define PKEY(nm, kd) {#nm, kd, offsetof(qgw_params, nm)}

/* Registered keys: the reference's 25 (src/qg.c:113-137) and the extensions. */
static const param_key k_keys[] = {
    PKEY(NX, PK_INT),          PKEY(NY, PK_INT),         PKEY(nl, PK_INT),
    PKEY(n_hyper, PK_INT),     PKEY(Lx, PK_DOUBLE),      PKEY(dh, PK_ARRAY),
    PKEY(tau0, PK_DOUBLE),     PKEY(forc_mode, PK_DOUBLE), PKEY(f0, PK_DOUBLE),
    PKEY(beta, PK_DOUBLE),     PKEY(nu, PK_DOUBLE),      PKEY(nu_hyper, PK_DOUBLE),
    PKEY(hEkb, PK_DOUBLE),     PKEY(N2, PK_ARRAY),       PKEY(Ld, PK_DOUBLE),
    PKEY(h_topo, PK_DOUBLE),   PKEY(w_topo, PK_DOUBLE),  PKEY(bc_fac, PK_DOUBLE),
    PKEY(tend, PK_DOUBLE),     PKEY(dt_out, PK_DOUBLE),  PKEY(dt_forc, PK_DOUBLE),
    PKEY(dt_forc_period, PK_DOUBLE), PKEY(sigma_f, PK_DOUBLE), PKEY(k_f, PK_DOUBLE),
    PKEY(cfl, PK_DOUBLE),
    /* extensions */
    PKEY(dt_fixed, PK_DOUBLE), PKEY(nsteps, PK_INT),     PKEY(ic_kind, PK_INT),
    PKEY(seed, PK_U64),        PKEY(ic_urms, PK_DOUBLE), PKEY(ic_lambda0, PK_DOUBLE),
    PKEY(precision, PK_INT),   PKEY(verbose, PK_INT),
};

/* Array values: the text after '=' with the newline cut; the first field is
 * delimited by any of "[,]" and the following ones by "," only, and each
 * field is converted with atof (so "2900.]" reads as 2900).  Same field rule
 * as src/extra.h:98-115. */
static void parse_array(char *text, double *arr) {
  char *nl = strchr(text, '\n');
  if (nl) *nl = '\0';
  int n = 0;
  char *save = NULL;
  char *tok = strtok_r(text, "[,]", &save);
  while (tok != NULL && n < QGW_NL_MAX) {
    arr[n++] = atof(tok);
    tok = strtok_r(NULL, ",", &save);
  }
}

int qgwh_read_params(const char *path, qgw_params *p) {
  FILE *fp = fopen(path, "rt");
  if (!fp) return QGW_ERR_IO;
  char line[300];
  while (fgets(line, sizeof line, fp)) {
    /* drop every space (not tabs), as src/extra.h:26-33 */
    char *w = line;
    for (char *r = line; *r; r++)
      if (*r != ' ') *w++ = *r;
    *w = '\0';
    char *save = NULL;
    char *key = strtok_r(line, "=", &save);
    char *val = strtok_r(NULL, "=", &save);
    if (!key) continue;
    for (size_t i = 0; i < sizeof k_keys / sizeof k_keys[0]; i++) {
      if (strcmp(k_keys[i].name, key) != 0) continue;
      char *dst = (char *)p + k_keys[i].offset;
      const char *v = val ? val : "";
      switch (k_keys[i].kind) {
        case PK_INT: *(int *)dst = atoi(v); break;
        case PK_DOUBLE: *(double *)dst = atof(v); break;
        case PK_U64: *(uint64_t *)dst = strtoull(v, NULL, 10); break;
        case PK_ARRAY:
          if (val) parse_array(val, (double *)dst);
          break;
      }
      if (p->verbose) fprintf(stdout, "scan param %s\n", k_keys[i].name);
    }
  }
  fclose(fp);
  return QGW_OK;
}

/* ------------------------------------------------------------------------ */
/* grid                                                                       */
/* ------------------------------------------------------------------------ */

int qgwh_grid(const qgw_params *p, qgw_grid *g, double *X, double *Y, double *Kw, double *Lw) {
  if (!p || !g) return QGW_ERR_ARG;
  if (p->NX < 3 || p->NY < 3 || !(p->Lx > 0.0)) return QGW_ERR_ARG;
  if (p->bc_fac == -1.0) {
    /* periodic (src/domain.h:71-75,144-147,163-167): NX x NY points, the zero
     * of X and Y at the second grid point, Fourier wavenumbers.  The serial
     * reference transforms an in-place r2c array whose row length is NX + 2
     * with dimensions (NX, NY) (src/elliptic.h:32-47), which is consistent
     * only for a square grid, and its half-spectrum wavenumbers need an even
     * NX. */
    if (p->NX != p->NY || (p->NX % 2) != 0) return QGW_ERR_ARG;
    const int N = p->NX;
    g->Delta = p->Lx / N;
    g->Lx = p->Lx;
    g->Ly = p->NY * g->Delta;
    g->n = N;
    g->m = p->NY;
    g->nxp2 = g->n + 2;
    g->nyp2 = g->m + 2;
    if (X)
      for (int i = 0; i < g->nxp2; i++) X[i] = (i - 1) * g->Delta;
    if (Y)
      for (int j = 0; j < g->nyp2; j++) Y[j] = (j - 1) * g->Delta;
    if (Kw) { /* K[i], i < N/2 + 1: the half-spectrum (x) wavenumbers */
      for (int j = 0; j < g->nyp2; j++) Kw[j] = 0.0;
      for (int i = 0; i < N / 2 + 1; i++) Kw[i] = 2.0 * QGW_PI * i / p->Lx;
    }
    if (Lw) { /* L[j], j < N: the full (y) dimension, signed */
      for (int i = 0; i < g->nxp2; i++) Lw[i] = 0.0;
      for (int j = 0; j < N; j++) Lw[j] = 2.0 * QGW_PI / p->Lx * (((j + N / 2) % N) - N / 2);
    }
    return QGW_OK;
  }
  if (p->bc_fac < 0.0) return QGW_ERR_ARG; /* bc_fac is -1 (periodic) or in [0, 1] */
  g->Delta = p->Lx / p->NX;
  g->Lx = p->Lx;
  g->Ly = p->NY * g->Delta;
  g->n = p->NX - 1;
  g->m = p->NY - 1;
  g->nxp2 = g->n + 2;
  g->nyp2 = g->m + 2;
  if (X)
    for (int i = 0; i < g->nxp2; i++) X[i] = i * g->Delta;
  if (Y)
    for (int j = 0; j < g->nyp2; j++) Y[j] = j * g->Delta;
  if (Kw) {
    for (int j = 0; j < g->nyp2; j++) Kw[j] = 0.0;
    for (int j = 1; j < g->m + 1; j++) Kw[j] = QGW_PI * j / p->Lx;
  }
  if (Lw) {
    for (int i = 0; i < g->nxp2; i++) Lw[i] = 0.0;
    for (int i = 1; i < g->n + 1; i++) Lw[i] = QGW_PI * i / p->Lx;
  }
  return QGW_OK;
}

static double shelf(double x, double d) { return 1.0 - exp(-(x * x) / (2.0 * (d * d))); }

int qgwh_topo(const qgw_params *p, double *topo) {
  qgw_grid g;
  int st = qgwh_grid(p, &g, NULL, NULL, NULL, NULL);
  if (st) return st;
  const double off = (p->bc_fac == -1.0) ? -1.0 : 0.0; /* X[i] = (i + I0) Delta */
  for (int j = 0; j < g.nyp2; j++) {
    double y = (j + off) * g.Delta;
    for (int i = 0; i < g.nxp2; i++) {
      double x = (i + off) * g.Delta;
      double s = shelf(x, p->w_topo) * shelf(p->Lx - x, p->w_topo) * shelf(y, p->w_topo) *
                 shelf(g.Ly - y, p->w_topo);
      topo[(size_t)j * g.nxp2 + i] = p->h_topo * (1.0 - s);
    }
  }
  return QGW_OK;
}

int qgwh_wind_row(const qgw_params *p, double *forc_row) {
  qgw_grid g;
  int st = qgwh_grid(p, &g, NULL, NULL, NULL, NULL);
  if (st) return st;
  const double off = (p->bc_fac == -1.0) ? -1.0 : 0.0; /* Y[j] = (j + J0) Delta */
  for (int j = 0; j < g.nyp2; j++) {
    double y = (j + off) * g.Delta;
    forc_row[j] = -p->tau0 / p->dh[0] * p->forc_mode * QGW_PI / g.Ly *
                  sin(p->forc_mode * QGW_PI * y / g.Ly);
  }
  return QGW_OK;
}

double qgwh_nu_hyper_signed(const qgw_params *p) {
  return pow(-1.0, p->n_hyper / 2 - 1) * fabs(p->nu_hyper);
}

void qgwh_check_timestep(const qgw_params *p, double nu_hyper_signed, double *dt_out,
                         double *DT_max_out) {
  qgw_grid g;
  if (qgwh_grid(p, &g, NULL, NULL, NULL, NULL)) {
    *dt_out = 0.0;
    *DT_max_out = 0.0;
    return;
  }
  /* Lx/Nx uses the interior count (Nx = NX-1 bounded, NX periodic, after init_domain). */
  double h = p->Lx / g.n;
  double a = nu_hyper_signed * 20. / pow(h, p->n_hyper);
  double b = p->nu * 20. / pow(h, 2);
  double c = 2. * p->beta * p->Lx;
  /* the unparenthesised max(max(a,b),c) of src/timestep.h:27 evaluates as */
  double mx = (a < b) ? b : ((a < c) ? c : a);
  double DT_max = pow(mx, -1.);
  double dt = 0.0;
  double Delta = g.Delta, Ly = g.Ly;
  if (p->tau0 != 0 && p->sigma_f != 0) {
    double s1 = sqrt(p->cfl * Delta / (p->tau0 / p->dh[0] * p->forc_mode * QGW_PI / Ly * Ly));
    double s2 = pow((Delta * p->cfl) * (Delta * p->cfl) / p->sigma_f, 1. / 3.);
    /* 1e-3*min(s1,s2) expands to (1e-3*s1 > s2) ? s2 : s1 */
    dt = (1e-3 * s1 > s2) ? s2 : s1;
  } else if (p->tau0 != 0 && p->sigma_f == 0) {
    dt = 1e-3 * sqrt(p->cfl * Delta / (p->tau0 / p->dh[0] * p->forc_mode * QGW_PI / Ly * Ly));
  } else if (p->tau0 == 0 && p->sigma_f != 0) {
    dt = 1e-3 * pow((Delta * p->cfl) * (Delta * p->cfl) / p->sigma_f, 1. / 3.);
  }
  if (DT_max != 0 && dt > DT_max) dt = DT_max;
  *dt_out = dt;
  *DT_max_out = DT_max;
}

/* ------------------------------------------------------------------------ */
/* vertical modes                                                             */
/* ------------------------------------------------------------------------ */

int qgwh_gamma(const qgw_params *p, double *G) {
  int nl = p->nl;
  if (nl < 1 || nl > QGW_NL_MAX) return QGW_ERR_ARG;
  double dhc[QGW_NL_MAX];
  for (int l = 0; l < nl - 1; l++) dhc[l] = 0.5 * (p->dh[l] + p->dh[l + 1]);
  for (int i = 0; i < nl * nl; i++) G[i] = 0.0;
  if (nl == 1) {
    G[0] = 1.0;
    return QGW_OK;
  }
  double f02 = p->f0 * p->f0;
  for (int l = 0; l < nl; l++) {
    double lo = 0.0, up = 0.0;
    if (l > 0) {
      lo = -f02 / p->N2[l - 1] / (dhc[l - 1] * p->dh[l]);
      G[nl * l + l - 1] = lo;
    }
    if (l < nl - 1) {
      up = -f02 / p->N2[l] / (dhc[l] * p->dh[l]);
      G[nl * l + l + 1] = up;
    }
    G[nl * l + l] = -lo - up;
  }
  /* the first and last rows carry a single off-diagonal term */
  G[0] = -G[1];
  G[nl * (nl - 1) + nl - 1] = -G[nl * (nl - 1) + nl - 2];
  return QGW_OK;
}

/* Symmetric tridiagonal eigen-solver: the implicit symmetric QR algorithm
 * with Wilkinson shifts (Golub & Van Loan, Matrix Computations, 4th ed.,
 * Algorithms 8.3.2 and 8.3.3), written from the mathematics below.
 *
 * T has diagonal d[0..n-1] and sub-diagonal e[0..n-2] (e[k] = T[k+1][k]).
 * Each sweep works on the lowest unreduced block p..q: the shift mu is the
 * eigenvalue of the block's trailing 2x2 closer to its last entry; a plane
 * rotation R_k = [[c, s], [-s, c]] in rows/columns (k, k+1) is applied as
 * T <- R_k T R_k^T, the first one chosen to zero the second entry of
 * (d[p] - mu, e[p]), each later one to zero the bulge T[k+1][k-1] left by
 * the one before.  The rotations accumulate into Z <- Z R_k^T, so on return
 * d holds the eigenvalues and column k of Z (row-major, Z[i*n+k]) the unit
 * eigenvector of d[k].  An off-diagonal entry below eps*(|d_k| + |d_k+1|)
 * is set to zero (deflation). */
static void rot_cols(double *Z, int n, int k, double c, double s) {
  for (int r = 0; r < n; r++) {
    double a = Z[r * n + k], b = Z[r * n + k + 1];
    Z[r * n + k] = c * a + s * b;
    Z[r * n + k + 1] = -s * a + c * b;
  }
}

static int tridiag_eig(int n, double *d, const double *e_in, double *Z) {
  const double eps = 2.2204460492503131e-16;
  double *e = (double *)calloc((size_t)n + 1, sizeof(double));
  if (!e) return QGW_ERR_NOMEM;
  for (int k = 0; k < n - 1; k++) e[k] = e_in[k];
  for (int i = 0; i < n * n; i++) Z[i] = 0.0;
  for (int i = 0; i < n; i++) Z[i * n + i] = 1.0;
  int q = n - 1, sweeps = 0;
  while (q > 0) {
    /* deflate at the bottom */
    if (fabs(e[q - 1]) <= eps * (fabs(d[q - 1]) + fabs(d[q])) || e[q - 1] == 0.0) {
      e[q - 1] = 0.0;
      q--;
      continue;
    }
    /* top of the unreduced block ending at q */
    int p = q - 1;
    while (p > 0 && fabs(e[p - 1]) > eps * (fabs(d[p - 1]) + fabs(d[p]))) p--;
    if (p > 0) e[p - 1] = 0.0;
    if (++sweeps > 60 * n) {
      free(e);
      return QGW_ERR_NUMERIC;
    }
    /* Wilkinson shift from [[d[q-1], e[q-1]], [e[q-1], d[q]]] */
    double half = 0.5 * (d[q - 1] - d[q]);
    double b2 = e[q - 1] * e[q - 1];
    double mu = d[q] - b2 / (half + (half >= 0 ? 1.0 : -1.0) * hypot(half, e[q - 1]));
    double x = d[p] - mu, z = e[p], bulge = 0.0;
    for (int k = p; k < q; k++) {
      if (k > p) {
        x = e[k - 1];
        z = bulge;
      }
      double r = hypot(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = z / r;
      }
      if (k > p) e[k - 1] = r; /* the bulge is annihilated */
      double a = d[k], b = e[k], cc = d[k + 1];
      d[k] = c * c * a + 2.0 * c * s * b + s * s * cc;
      d[k + 1] = s * s * a - 2.0 * c * s * b + c * c * cc;
      e[k] = c * s * (cc - a) + (c * c - s * s) * b;
      if (k + 1 < q) {
        bulge = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      rot_cols(Z, n, k, c, s);
    }
  }
  free(e);
  return QGW_OK;
}

int qgwh_eigmode(const qgw_params *p, double *cl2m, double *cm2l, double *iRd2) {
  int nl = p->nl;
  if (nl < 1 || nl > QGW_NL_MAX) return QGW_ERR_ARG;
  double H = 0.0;
  for (int k = 0; k < nl; k++) H += p->dh[k];
  if (nl == 1) {
    /* Gamma = [1]: both vectors are 1 after the normalisations */
    cl2m[0] = 1.0;
    cm2l[0] = 1.0;
    iRd2[0] = (p->Ld != 0) ? 1.0 / (p->Ld * p->Ld) : 0.0;
    return QGW_OK;
  }
  for (int l = 0; l < nl; l++) {
    if (!(p->dh[l] > 0.0)) return QGW_ERR_ARG;
    if (l < nl - 1 && !(p->N2[l] != 0.0)) return QGW_ERR_ARG;
  }
  size_t nn = (size_t)nl * nl;
  double *G = (double *)malloc(nn * sizeof(double));
  double *Z = (double *)malloc(nn * sizeof(double));
  double *vr = (double *)malloc(nn * sizeof(double));
  double *vl = (double *)malloc(nn * sizeof(double));
  double *d = (double *)malloc(nl * sizeof(double));
  double *e = (double *)malloc(nl * sizeof(double));
  int *ord = (int *)malloc(nl * sizeof(int));
  int st = QGW_OK;
  if (!G || !Z || !vr || !vl || !d || !e || !ord) {
    st = QGW_ERR_NOMEM;
    goto done;
  }
  qgwh_gamma(p, G);
  /* D^(1/2) Gamma D^(-1/2) is symmetric (dh_l G_{l,l+1} = dh_{l+1} G_{l+1,l});
   * its off-diagonal is -sqrt(G_{l,l+1} G_{l+1,l}) (both factors are negative
   * for positive N2) */
  for (int l = 0; l < nl; l++) d[l] = G[nl * l + l];
  for (int l = 0; l < nl - 1; l++) {
    double prod = G[nl * l + l + 1] * G[nl * (l + 1) + l];
    double s = (G[nl * l + l + 1] < 0) ? -1.0 : 1.0;
    e[l] = s * sqrt(fabs(prod));
  }
  st = tridiag_eig(nl, d, e, Z);
  if (st) goto done;
  /* ascending order (src/eigmode.h:154-185); a stable insertion sort */
  for (int i = 0; i < nl; i++) ord[i] = i;
  for (int i = 1; i < nl; i++) {
    int v = ord[i], j = i - 1;
    while (j >= 0 && d[ord[j]] > d[v]) {
      ord[j + 1] = ord[j];
      j--;
    }
    ord[j + 1] = v;
  }
  /* right vectors D^(-1/2) u, left vectors D^(1/2) u, columns sorted */
  for (int m = 0; m < nl; m++) {
    int src = ord[m];
    for (int k = 0; k < nl; k++) {
      double sq = sqrt(p->dh[k]);
      vr[k * nl + m] = Z[k * nl + src] / sq;
      vl[k * nl + m] = Z[k * nl + src] * sq;
    }
  }
  /* Flierl (1978) normalisation, surface-positive (src/eigmode.h:192-214) */
  for (int m = 0; m < nl; m++) {
    double dot = 0.0;
    for (int k = 0; k < nl; k++) dot += p->dh[k] * vr[k * nl + m] * vr[k * nl + m];
    double fl = ((vr[m] > 0) ? 1.0 : -1.0) * sqrt(H / dot);
    for (int k = 0; k < nl; k++) vr[k * nl + m] *= fl;
  }
  /* left vectors scaled to vl^T vr = 1 (src/eigmode.h:215-223) */
  for (int m = 0; m < nl; m++) {
    double dot = 0.0;
    for (int k = 0; k < nl; k++) dot += vr[k * nl + m] * vl[k * nl + m];
    for (int k = 0; k < nl; k++) vl[k * nl + m] /= dot;
  }
  for (int m = 0; m < nl; m++)
    for (int k = 0; k < nl; k++) {
      cl2m[k * nl + m] = vl[m * nl + k];
      cm2l[k * nl + m] = vr[k * nl + m];
    }
  for (int k = 1; k < nl; k++) iRd2[k] = d[ord[k]];
  iRd2[0] = 0.0;
  if (p->verbose) {
    fprintf(stdout, "Deformation radii:\n");
    for (int k = 0; k < nl; k++)
      fprintf(stdout, "iRd2: %g , def radius: %g\n", iRd2[k], iRd2[k] > 0 ? sqrt(1 / iRd2[k]) : 0);
  }
done:
  free(G);
  free(Z);
  free(vr);
  free(vl);
  free(d);
  free(e);
  free(ord);
  return st;
}

/* ------------------------------------------------------------------------ */
/* host FFT and 2-D sine transform (initial condition only)                  */
/* ------------------------------------------------------------------------ */

typedef double _Complex cplx;

typedef struct {
  int L;
  int nr;
  int radix[32];
  cplx *tw; /* tw[m] = exp(-2 pi i m / L) */
} hplan;

static int hplan_make(hplan *pl, int L) {
  pl->L = L;
  pl->nr = 0;
  int r = L;
  while (r % 4 == 0) { pl->radix[pl->nr++] = 4; r /= 4; }
  while (r % 2 == 0) { pl->radix[pl->nr++] = 2; r /= 2; }
  while (r % 3 == 0) { pl->radix[pl->nr++] = 3; r /= 3; }
  while (r % 5 == 0) { pl->radix[pl->nr++] = 5; r /= 5; }
  while (r % 7 == 0) { pl->radix[pl->nr++] = 7; r /= 7; }
  if (r != 1) return QGW_ERR_ARG;
  pl->tw = (cplx *)malloc(sizeof(cplx) * L);
  if (!pl->tw) return QGW_ERR_NOMEM;
  for (int m = 0; m < L; m++) {
    double a = -2.0 * QGW_PI * (double)m / (double)L;
    pl->tw[m] = cos(a) + I * sin(a);
  }
  return QGW_OK;
}

/* Stockham autosort FFT (forward, e^{-2 pi i jk/L}); result is returned in
 * whichever of x/y holds it. */
static cplx *hfft(const hplan *pl, cplx *x, cplx *y) {
  int L = pl->L, Ns = 1;
  cplx *src = x, *dst = y;
  for (int s = 0; s < pl->nr; s++) {
    int R = pl->radix[s];
    int stride = L / R;
    int twstep = L / (Ns * R);
    for (int j = 0; j < stride; j++) {
      int k = j % Ns;
      cplx v[7], o[7];
      for (int r = 0; r < R; r++) v[r] = src[j + r * stride] * pl->tw[(r * k * twstep) % L];
      if (R == 4) {
        cplx a0 = v[0] + v[2], a1 = v[0] - v[2], a2 = v[1] + v[3], a3 = (v[1] - v[3]) * (-I);
        o[0] = a0 + a2;
        o[2] = a0 - a2;
        o[1] = a1 + a3;
        o[3] = a1 - a3;
      } else if (R == 2) {
        o[0] = v[0] + v[1];
        o[1] = v[0] - v[1];
      } else {
        for (int q = 0; q < R; q++) {
          cplx acc = 0;
          for (int r = 0; r < R; r++) acc += v[r] * pl->tw[((r * q) % R) * (L / R)];
          o[q] = acc;
        }
      }
      int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[base + r * Ns] = o[r];
    }
    Ns *= R;
    cplx *t = src;
    src = dst;
    dst = t;
  }
  return src;
}

/* Unnormalised DST-I of two real sequences a, b of length n (strided) at once:
 * the complex sequence a + i b, odd-extended to length 2(n+1), is transformed
 * and DST(a) = -Im, DST(b) = Re of the result. */
static void dst_pair(const hplan *pl, cplx *w1, cplx *w2, double *a, double *b, int n,
                     ptrdiff_t stride) {
  int L = pl->L;
  w1[0] = 0;
  w1[n + 1] = 0;
  for (int j = 1; j <= n; j++) {
    cplx z = a[(ptrdiff_t)(j - 1) * stride] + I * (b ? b[(ptrdiff_t)(j - 1) * stride] : 0.0);
    w1[j] = z;
    w1[L - j] = -z;
  }
  cplx *r = hfft(pl, w1, w2);
  for (int k = 1; k <= n; k++) {
    a[(ptrdiff_t)(k - 1) * stride] = -cimag(r[k]);
    if (b) b[(ptrdiff_t)(k - 1) * stride] = creal(r[k]);
  }
}

int qgwh_dst2d(double *a, int n, int m) {
  if (n < 1 || m < 1) return QGW_ERR_ARG;
  hplan px, py;
  int st = hplan_make(&px, 2 * (n + 1));
  if (st) return st;
  st = hplan_make(&py, 2 * (m + 1));
  if (st) {
    free(px.tw);
    return st;
  }
  int err = 0;
#pragma omp parallel
  {
    cplx *w1 = (cplx *)malloc(sizeof(cplx) * (size_t)(2 * (n > m ? n : m) + 4));
    cplx *w2 = (cplx *)malloc(sizeof(cplx) * (size_t)(2 * (n > m ? n : m) + 4));
    if (!w1 || !w2) {
#pragma omp atomic write
      err = 1;
    } else {
      /* rows (x direction, contiguous), two at a time */
#pragma omp for schedule(static)
      for (int j = 0; j < m; j += 2) {
        double *r0 = a + (size_t)j * n;
        double *r1 = (j + 1 < m) ? a + (size_t)(j + 1) * n : NULL;
        dst_pair(&px, w1, w2, r0, r1, n, 1);
      }
      /* columns (y direction, stride n), two at a time */
#pragma omp for schedule(static)
      for (int i = 0; i < n; i += 2) {
        double *c0 = a + i;
        double *c1 = (i + 1 < n) ? a + i + 1 : NULL;
        dst_pair(&py, w1, w2, c0, c1, m, n);
      }
    }
    free(w1);
    free(w2);
  }
  free(px.tw);
  free(py.tw);
  return err ? QGW_ERR_NOMEM : QGW_OK;
}

/* ------------------------------------------------------------------------ */
/* seeded initial condition                                                   */
/* ------------------------------------------------------------------------ */

/* SplitMix64 (Steele, Lea & Flood 2014). */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* uniform in (0, 1] */
static double unif(uint64_t *s) { return ((double)(splitmix64(s) >> 11) + 1.0) * 0x1.0p-53; }

/* draw k (0-based) of the stream seeded with s0: SplitMix64 advances its
 * state by a constant, so any draw is reached directly (parallel fills give
 * the sequence of the serial stream) */
static double unif_at(uint64_t s0, uint64_t k) {
  uint64_t s = s0 + k * 0x9e3779b97f4a7c15ULL;
  return unif(&s);
}

/* In-place 2-D complex FFT of an m x n array (rows of n), forward or inverse
 * (unnormalised), rows and columns in parallel. */
static int fft2_complex(cplx *a, int n, int m, int inverse) {
  hplan px, py;
  int st = hplan_make(&px, n);
  if (st) return st;
  st = hplan_make(&py, m);
  if (st) {
    free(px.tw);
    return st;
  }
  int err = 0;
#pragma omp parallel
  {
    int L = n > m ? n : m;
    cplx *w1 = (cplx *)malloc(sizeof(cplx) * (size_t)L);
    cplx *w2 = (cplx *)malloc(sizeof(cplx) * (size_t)L);
    if (!w1 || !w2) {
#pragma omp atomic write
      err = 1;
    } else {
#pragma omp for schedule(static)
      for (int j = 0; j < m; j++) {
        cplx *r = a + (size_t)j * n;
        for (int i = 0; i < n; i++) w1[i] = inverse ? conj(r[i]) : r[i];
        cplx *o = hfft(&px, w1, w2);
        for (int i = 0; i < n; i++) r[i] = inverse ? conj(o[i]) : o[i];
      }
#pragma omp for schedule(static)
      for (int i = 0; i < n; i++) {
        for (int j = 0; j < m; j++) w1[j] = inverse ? conj(a[(size_t)j * n + i]) : a[(size_t)j * n + i];
        cplx *o = hfft(&py, w1, w2);
        for (int j = 0; j < m; j++) a[(size_t)j * n + i] = inverse ? conj(o[j]) : o[j];
      }
    }
    free(w1);
    free(w2);
  }
  free(px.tw);
  free(py.tw);
  return err ? QGW_ERR_NOMEM : QGW_OK;
}

/* Periodic form of the seeded initial condition: the same k^-3 spectrum on
 * the Fourier wavenumbers (2 pi k / L), psi the real part of the inverse 2-D
 * transform of seeded complex Gaussian coefficients (zero mean), q the modal
 * Helmholtz operator applied in Fourier space, so that the periodic inversion
 * (src/elliptic.h:103-118) reproduces psi.  The rings are periodic copies
 * (src/domain.h:225-256). */
static int make_ic_periodic(const qgw_params *p, const qgw_grid *g, const double *cl2m, const double *cm2l,
                            const double *iRd2, uint64_t seed, double *psi, double *q) {
  const int n = g->n, m = g->m, nl = p->nl;
  const size_t NI = (size_t)n * m, S = (size_t)g->nxp2 * g->nyp2;
  cplx *C = (cplx *)malloc(sizeof(cplx) * NI * nl);
  cplx *Q = (cplx *)malloc(sizeof(cplx) * NI * nl);
  double *Xm = (double *)malloc(sizeof(double) * 2 * nl);
  if (!C || !Q || !Xm) {
    free(C);
    free(Q);
    free(Xm);
    return QGW_ERR_NOMEM;
  }
  const double kappa0 = 2.0 * QGW_PI / p->ic_lambda0, k06 = pow(kappa0, 6);
  uint64_t s = seed;
  for (int l = 0; l < nl; l++)
    for (int b = 0; b < m; b++)
      for (int a = 0; a < n; a++) {
        double u1 = unif(&s), u2 = unif(&s);
        double r = sqrt(-2.0 * log(u1));
        double ka = 2.0 * QGW_PI / p->Lx * (((a + n / 2) % n) - n / 2);
        double kb = 2.0 * QGW_PI / p->Lx * (((b + m / 2) % m) - m / 2);
        double kap = sqrt(ka * ka + kb * kb);
        double amp = (a == 0 && b == 0) ? 0.0 : 1.0 / sqrt(k06 + pow(kap, 6));
        C[(size_t)l * NI + (size_t)b * n + a] = amp * r * (cos(2.0 * QGW_PI * u2) + I * sin(2.0 * QGW_PI * u2));
      }
  int st = QGW_OK;
  for (int l = 0; l < nl && !st; l++) st = fft2_complex(C + (size_t)l * NI, n, m, 1);
  if (st) goto out;
  for (size_t t = 0; t < NI * nl; t++) C[t] = creal(C[t]); /* the real field psi */
  /* RMS of the forward-difference velocity, periodic */
  double sum = 0.0;
  for (int l = 0; l < nl; l++)
    for (int j = 0; j < m; j++)
      for (int i = 0; i < n; i++) {
        const cplx *f = C + (size_t)l * NI;
        double c = creal(f[(size_t)j * n + i]);
        double e = creal(f[(size_t)j * n + (i + 1) % n]);
        double nn = creal(f[(size_t)((j + 1) % m) * n + i]);
        double u = -(nn - c) / g->Delta, v = (e - c) / g->Delta;
        sum += u * u + v * v;
      }
  double rms = sqrt(sum / ((double)NI * nl));
  double scale = (rms > 0.0 && p->ic_urms > 0.0) ? p->ic_urms / rms : 1.0;
  for (size_t t = 0; t < NI * nl; t++) {
    C[t] *= scale;
    Q[t] = C[t];
  }
  for (int l = 0; l < nl && !st; l++) st = fft2_complex(Q + (size_t)l * NI, n, m, 0);
  if (st) goto out;
  for (int b = 0; b < m; b++)
    for (int a = 0; a < n; a++) {
      size_t o = (size_t)b * n + a;
      double ka = 2.0 * QGW_PI / p->Lx * (((a + n / 2) % n) - n / 2);
      double kb = 2.0 * QGW_PI / p->Lx * (((b + m / 2) % m) - m / 2);
      for (int k = 0; k < nl; k++) {
        cplx acc = 0.0;
        for (int l = 0; l < nl; l++) acc += cl2m[k * nl + l] * Q[(size_t)l * NI + o];
        acc *= -(ka * ka + kb * kb + iRd2[k]);
        Xm[2 * k] = creal(acc);
        Xm[2 * k + 1] = cimag(acc);
      }
      for (int l = 0; l < nl; l++) {
        cplx acc = 0.0;
        for (int k = 0; k < nl; k++) acc += cm2l[l * nl + k] * (Xm[2 * k] + I * Xm[2 * k + 1]);
        Q[(size_t)l * NI + o] = acc;
      }
    }
  for (int l = 0; l < nl && !st; l++) st = fft2_complex(Q + (size_t)l * NI, n, m, 1);
  if (st) goto out;
  for (int l = 0; l < nl; l++)
    for (int j = 0; j < g->nyp2; j++)
      for (int i = 0; i < g->nxp2; i++) {
        const int jj = (j == 0) ? m - 1 : (j == m + 1) ? 0 : j - 1; /* periodic copies on the ring */
        const int ii = (i == 0) ? n - 1 : (i == n + 1) ? 0 : i - 1;
        const size_t o = (size_t)l * NI + (size_t)jj * n + ii, d = (size_t)l * S + (size_t)j * g->nxp2 + i;
        psi[d] = creal(C[o]);
        q[d] = creal(Q[o]) / (double)NI;
      }
out:
  free(C);
  free(Q);
  free(Xm);
  return st;
}

int qgwh_make_ic(const qgw_params *p, const double *cl2m, const double *cm2l, const double *iRd2,
                 uint64_t seed, double *psi, double *q) {
  qgw_grid g;
  int st = qgwh_grid(p, &g, NULL, NULL, NULL, NULL);
  if (st) return st;
  if (p->bc_fac == -1.0) return make_ic_periodic(p, &g, cl2m, cm2l, iRd2, seed, psi, q);
  int n = g.n, m = g.m, nl = p->nl;
  size_t NI = (size_t)n * m;
  size_t S = (size_t)g.nxp2 * g.nyp2;
  double *X = (double *)malloc(sizeof(double) * NI * nl);
  double *W = (double *)malloc(sizeof(double) * NI * nl);
  if (!X || !W) {
    free(X);
    free(W);
    return QGW_ERR_NOMEM;
  }
  double kappa0 = 2.0 * QGW_PI / p->ic_lambda0;
  double k06 = pow(kappa0, 6);
  /* coefficient e (layer-major, then b, then a) takes the cosine (e even) or
   * the sine (e odd) of the Box-Muller pair made of draws e-e%2 and e-e%2+1 */
  const long long NT = (long long)NI * nl;
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < NT; e++) {
    const uint64_t d = (uint64_t)(e - (e & 1));
    const double u1 = unif_at(seed, d), u2 = unif_at(seed, d + 1);
    const double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincos(2.0 * QGW_PI * u2, &sn, &cs);
    const double gauss = (e & 1) ? r * sn : r * cs;
    const size_t o = (size_t)e % NI;
    const int a = (int)(o % (size_t)n) + 1, b = (int)(o / (size_t)n) + 1;
    double ka = QGW_PI * a / p->Lx, kb = QGW_PI * b / p->Lx;
    double kap = sqrt(ka * ka + kb * kb);
    double amp = 1.0 / sqrt(k06 + pow(kap, 6));
    X[e] = gauss * amp;
  }
  double norm = 4.0 * (n + 1) * (double)(m + 1);
  /* psi per layer */
  memcpy(W, X, sizeof(double) * NI * nl);
  for (int l = 0; l < nl && !st; l++) st = qgwh_dst2d(W + (size_t)l * NI, n, m);
  if (st) goto out;
#pragma omp parallel for schedule(static)
  for (long long t = 0; t < NT; t++) W[t] /= norm;
  /* RMS of the forward-difference velocity over the interior (zero ring) */
  double sum = 0.0;
  for (int l = 0; l < nl; l++)
    for (int j = 1; j <= m; j++)
      for (int i = 1; i <= n; i++) {
        double c = W[(size_t)l * NI + (size_t)(j - 1) * n + (i - 1)];
        double e = (i < n) ? W[(size_t)l * NI + (size_t)(j - 1) * n + i] : 0.0;
        double nn = (j < m) ? W[(size_t)l * NI + (size_t)j * n + (i - 1)] : 0.0;
        double u = -(e - c) / g.Delta, v = (nn - c) / g.Delta;
        sum += u * u + v * v;
      }
  double rms = sqrt(sum / ((double)NI * nl));
  double scale = (rms > 0.0 && p->ic_urms > 0.0) ? p->ic_urms / rms : 1.0;
#pragma omp parallel for schedule(static)
  for (long long t = 0; t < NT; t++) {
    X[t] *= scale;
    W[t] *= scale;
  }
  memset(psi, 0, sizeof(double) * S * nl);
#pragma omp parallel for collapse(2) schedule(static)
  for (int l = 0; l < nl; l++)
    for (int j = 1; j <= m; j++)
      for (int i = 1; i <= n; i++)
        psi[(size_t)l * S + (size_t)j * g.nxp2 + i] = W[(size_t)l * NI + (size_t)(j - 1) * n + (i - 1)];
  /* q: spectral modal Helmholtz operator applied to the coefficients */
  int nomem = 0;
#pragma omp parallel
  {
    double *Xm = (double *)malloc(sizeof(double) * nl);
    double *Qm = (double *)malloc(sizeof(double) * nl);
    if (!Xm || !Qm) {
#pragma omp atomic write
      nomem = 1;
    }
#pragma omp barrier
    if (!nomem) {
#pragma omp for schedule(static)
    for (int b = 1; b <= m; b++)
      for (int a = 1; a <= n; a++) {
        size_t o = (size_t)(b - 1) * n + (a - 1);
        double La = QGW_PI * a / p->Lx, Kb = QGW_PI * b / p->Lx;
        for (int k = 0; k < nl; k++) {
          double acc = 0.0;
          for (int l = 0; l < nl; l++) acc += cl2m[k * nl + l] * X[(size_t)l * NI + o];
          Xm[k] = acc;
          Qm[k] = -(La * La + Kb * Kb + iRd2[k]) * acc;
        }
        for (int l = 0; l < nl; l++) {
          double acc = 0.0;
          for (int k = 0; k < nl; k++) acc += cm2l[l * nl + k] * Qm[k];
          W[(size_t)l * NI + o] = acc;
        }
      }
    }
    free(Xm);
    free(Qm);
  }
  if (nomem) {
    st = QGW_ERR_NOMEM;
    goto out;
  }
  for (int l = 0; l < nl && !st; l++) st = qgwh_dst2d(W + (size_t)l * NI, n, m);
  if (st) goto out;
  memset(q, 0, sizeof(double) * S * nl);
#pragma omp parallel for collapse(2) schedule(static)
  for (int l = 0; l < nl; l++)
    for (int j = 1; j <= m; j++)
      for (int i = 1; i <= n; i++)
        q[(size_t)l * S + (size_t)j * g.nxp2 + i] = W[(size_t)l * NI + (size_t)(j - 1) * n + (i - 1)] / norm;
  /* wall ring of q (src/domain.h:308-348 with a zero psi ring) */
  {
    double c = 2.0 * p->bc_fac / (g.Delta * g.Delta);
    for (int l = 0; l < nl; l++) {
      double *qq = q + (size_t)l * S, *pp = psi + (size_t)l * S;
      for (int i = 0; i < g.nxp2; i++) {
        qq[i] = c * pp[g.nxp2 + i];
        qq[(size_t)(m + 1) * g.nxp2 + i] = c * pp[(size_t)m * g.nxp2 + i];
      }
      for (int j = 0; j < g.nyp2; j++) {
        qq[(size_t)j * g.nxp2] = c * pp[(size_t)j * g.nxp2 + 1];
        qq[(size_t)j * g.nxp2 + n + 1] = c * pp[(size_t)j * g.nxp2 + n];
      }
    }
  }
out:
  free(X);
  free(W);
  return st;
}

/* ------------------------------------------------------------------------ */
/* diagnostics                                                                */
/* ------------------------------------------------------------------------ */

int qgwh_diagnostics(const qgw_params *p, const double *psi, const double *q, double *E_layer,
                     double *Z_layer, double *E_tot, double *Z_tot) {
  qgw_grid g;
  int st = qgwh_grid(p, &g, NULL, NULL, NULL, NULL);
  if (st) return st;
  size_t S = (size_t)g.nxp2 * g.nyp2;
  double w = g.Delta * g.Delta / (p->Lx * p->Lx);
  double H = 0.0, Et = 0.0, Zt = 0.0;
  for (int l = 0; l < p->nl; l++) H += p->dh[l];
  for (int l = 0; l < p->nl; l++) {
    double e = 0.0, z = 0.0;
    for (size_t t = 0; t < S; t++) {
      e += psi[(size_t)l * S + t] * q[(size_t)l * S + t];
      z += q[(size_t)l * S + t] * q[(size_t)l * S + t];
    }
    e = -e * w;
    z = z * w;
    if (E_layer) E_layer[l] = e;
    if (Z_layer) Z_layer[l] = z;
    Et += p->dh[l] * e;
    Zt += p->dh[l] * z;
  }
  if (E_tot) *E_tot = Et / (2.0 * H);
  if (Z_tot) *Z_tot = Zt / H;
  return QGW_OK;
}

const char *qgwh_strerror(int s) {
  switch (s) {
    case QGW_OK: return "ok";
    case QGW_ERR_ARG: return "invalid argument or unsupported configuration";
    case QGW_ERR_IO: return "file not found or unreadable";
    case QGW_ERR_NOMEM: return "allocation failed";
    case QGW_ERR_HIP: return "HIP runtime error";
    case QGW_ERR_NUMERIC: return "numerical failure (eigen-solver or non-finite state)";
    case QGW_ERR_STATE: return "call out of order";
    default: return "unknown status";
  }
}

/* Size of the parameter struct, so that bindings can check their layout. */
size_t qgwh_params_size(void) { return sizeof(qgw_params); }
