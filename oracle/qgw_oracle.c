/*
 * qgw_oracle.c -- TEST INFRASTRUCTURE ONLY: CPU FP64 restatement of the
 * reference's timestep hot path.  See qgw_oracle.h for the function map and
 * for what pins each function.  Loaded by tests, the smoke check and the
 * CPU-baseline leg of the benchmark; never by the product.
 *
 * Compile with -ffp-contract=off so that a*b+c is rounded twice, as the
 * reference's plain `gcc -O3` x86-64 build does.
 */
#include "qgw_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

static const double PI_REF = 3.141592653589793; /* src/qg.c:44 */
static const double HUGE_REF = 1e30;            /* src/qg.c:43 */
static const double TEPS = 1e-9;                /* src/timestep.h:16 */

/* A planted mistake (ORC_MUT_*), for the tests that show which mistake each
 * pin of the right-hand side, the wall ring and the topography catches.
 * 0, the default, is the faithful restatement. */
static int g_mut = ORC_MUT_NONE;
void orc_mutate(int id) { g_mut = id; }

int orc_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time (Cooley & Tukey 1965)      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int L;
  cplx *W; /* W[m] = exp(-2 pi i m / L) */
} fftplan;

static int plan_init(fftplan *p, int L) {
  p->L = L;
  p->W = (cplx *)malloc(sizeof(cplx) * (size_t)L);
  if (!p->W) return -1;
  for (int m = 0; m < L; m++) {
    double a = -2.0 * PI_REF * (double)m / (double)L;
    p->W[m] = cos(a) + I * sin(a);
  }
  return 0;
}

static int smallest_factor(int N) {
  for (int f = 2; f * f <= N; f++)
    if (N % f == 0) return f;
  return N;
}

/* out[0..N-1] = DFT of in[0], in[stride], ..., in[(N-1) stride];
 * wstep = L/N indexes the size-L root table. */
static void fft_rec(const fftplan *pl, cplx *out, const cplx *in, int N, int stride, int wstep) {
  if (N == 1) {
    out[0] = in[0];
    return;
  }
  int p = smallest_factor(N);
  int M = N / p;
  for (int r = 0; r < p; r++) fft_rec(pl, out + r * M, in + r * stride, M, stride * p, wstep * p);
  cplx v[p], res[p]; /* radix p may be a large prime */
  for (int k = 0; k < M; k++) {
    for (int r = 0; r < p; r++) v[r] = out[r * M + k] * pl->W[((long)r * k * wstep) % pl->L];
    for (int t = 0; t < p; t++) {
      cplx acc = 0;
      for (int r = 0; r < p; r++) acc += v[r] * pl->W[((long)r * t * M * wstep) % pl->L];
      res[t] = acc;
    }
    for (int t = 0; t < p; t++) out[k + M * t] = res[t];
  }
}

int orc_fft(double *re, double *im, int L) {
  if (L < 1) return -1;
  fftplan pl;
  if (plan_init(&pl, L)) return -1;
  cplx *in = (cplx *)malloc(sizeof(cplx) * L), *out = (cplx *)malloc(sizeof(cplx) * L);
  for (int j = 0; j < L; j++) in[j] = re[j] + I * im[j];
  fft_rec(&pl, out, in, L, 1, 1);
  for (int j = 0; j < L; j++) {
    re[j] = creal(out[j]);
    im[j] = cimag(out[j]);
  }
  free(in);
  free(out);
  free(pl.W);
  return 0;
}

/* Two RODFT00 transforms at once: x and y (length n, strided) become the real
 * and imaginary parts of a sequence odd-extended to length 2(n+1); the DFT of
 * an odd real sequence is -i times its sine transform, hence
 * DST(x) = -Im, DST(y) = Re. */
static void dst_two(const fftplan *pl, cplx *bin, cplx *bout, double *x, double *y, int n, long stride) {
  int L = pl->L;
  bin[0] = 0;
  bin[n + 1] = 0;
  for (int j = 1; j <= n; j++) {
    cplx z = x[(long)(j - 1) * stride] + I * (y ? y[(long)(j - 1) * stride] : 0.0);
    bin[j] = z;
    bin[L - j] = -z;
  }
  fft_rec(pl, bout, bin, L, 1, 1);
  for (int k = 1; k <= n; k++) {
    x[(long)(k - 1) * stride] = -cimag(bout[k]);
    if (y) y[(long)(k - 1) * stride] = creal(bout[k]);
  }
}

int orc_dst1d(double *x, int n) {
  fftplan pl;
  if (plan_init(&pl, 2 * (n + 1))) return -1;
  cplx *a = (cplx *)malloc(sizeof(cplx) * 2 * (n + 1)), *b = (cplx *)malloc(sizeof(cplx) * 2 * (n + 1));
  dst_two(&pl, a, b, x, NULL, n, 1);
  free(a);
  free(b);
  free(pl.W);
  return 0;
}

/* 2-D RODFT00 on an m x n row-major array: fftw_plan_r2r_2d(Ny, Nx, ...) of
 * src/elliptic.h:63-66 (n0 = Ny slow, n1 = Nx fast). */
static int dst2d_plans(double *a, int n, int m, const fftplan *px, const fftplan *py) {
  int Lmax = 2 * ((n > m ? n : m) + 1);
#pragma omp parallel
  {
    cplx *b1 = (cplx *)malloc(sizeof(cplx) * Lmax), *b2 = (cplx *)malloc(sizeof(cplx) * Lmax);
#pragma omp for schedule(static)
    for (int j = 0; j < m; j += 2)
      dst_two(px, b1, b2, a + (long)j * n, (j + 1 < m) ? a + (long)(j + 1) * n : NULL, n, 1);
#pragma omp for schedule(static)
    for (int i = 0; i < n; i += 2)
      dst_two(py, b1, b2, a + i, (i + 1 < n) ? a + i + 1 : NULL, m, n);
    free(b1);
    free(b2);
  }
  return 0;
}

int orc_dst2d(double *a, int n, int m) {
  fftplan px, py;
  if (plan_init(&px, 2 * (n + 1))) return -1;
  if (plan_init(&py, 2 * (m + 1))) return -1;
  dst2d_plans(a, n, m, &px, &py);
  free(px.W);
  free(py.W);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Symmetric eigen-solver: cyclic Jacobi (Golub & Van Loan sec. 8.5)        */
/* ------------------------------------------------------------------------ */

int orc_eig_sym(int n, const double *A0, double *w, double *V) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  memcpy(A, A0, sizeof(double) * n * n);
  for (int i = 0; i < n * n; i++) V[i] = 0.0;
  for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        tot += A[i * n + j] * A[i * n + j];
        if (i != j) off += A[i * n + j] * A[i * n + j];
      }
    if (off <= 1e-34 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) { /* A <- A J */
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) { /* A <- J^T A */
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; i++) w[i] = A[i * n + i];
  /* ascending selection sort of (w, columns of V) */
  for (int i = 0; i < n; i++) {
    int mi = i;
    for (int j = i + 1; j < n; j++)
      if (w[j] < w[mi]) mi = j;
    if (mi != i) {
      double t = w[i];
      w[i] = w[mi];
      w[mi] = t;
      for (int k = 0; k < n; k++) {
        double tv = V[k * n + i];
        V[k * n + i] = V[k * n + mi];
        V[k * n + mi] = tv;
      }
    }
  }
  free(A);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* state                                                                      */
/* ------------------------------------------------------------------------ */

struct orc {
  orc_params p;
  int Nx, Ny, Nxp1, Nyp1, Nxp2, Nyp2;
  double Delta, Lx, Ly, norm;
  double *X, *Y, *K, *L;
  double *psi, *q, *diff, *lapn, *topo, *f0, *f1, *f2, *wrk1;
  int periodic; /* bc_fac == -1 */
  cplx *cw;     /* periodic: the full 2-D spectrum, Ny x Nx */
  double *cl2m, *cm2l, *iRd2;
  double nu_hyper;
  double t, dt, dt0, dt1, DT_max, t_out;
  int it;
  fftplan px, py;
  /* 4-D forcing: n_forc records of nl storage fields, spaced dt_forc */
  int n_forc;
  double dt_forc;
  double *forc;
  /* stochastic ring forcing: n_P forced half-spectrum points (st_i, st_j) of
   * an NF x NF grid, amplitude st_env, the spectrum and the field forc_s */
  int stoch, NF, n_P;
  int *st_i, *st_j;
  double st_kf, st_sigma;
  cplx *st_spec;
  double *forc_p, *forc_s;
};

#define IDX(s, i, j, k) ((long)(k) * (s)->Nxp2 * (s)->Nyp2 + (long)(j) * (s)->Nxp2 + (i))

void orc_check_timestep(const orc_params *p, double *dt_out, double *DT_max_out) {
  /* init_domain: Delta from the cell count, Ly from NY; Nx = NX-1 (bounded)
   * or NX (periodic), src/domain.h:71-79 */
  double Delta = p->Lx / p->NX;
  double Ly = p->NY * Delta;
  int Nx = (p->bc_fac == -1) ? p->NX : p->NX - 1;
  double nu_hyper = pow(-1, p->n_hyper / 2 - 1) * fabs(p->nu_hyper); /* src/domain.h:210 */
  double inv_t_nu_hyper = nu_hyper * 20. / pow(p->Lx / Nx, p->n_hyper);
  double inv_t_nu = p->nu * 20. / pow(p->Lx / Nx, 2);
  double inv_t_beta = 2. * p->beta * p->Lx;
  /* max(max(a,b),c) with max(p,q) = p < q ? q : p unparenthesised */
  double mx = inv_t_nu_hyper < inv_t_nu ? inv_t_nu
              : inv_t_nu_hyper < inv_t_beta ? inv_t_beta
              : inv_t_nu_hyper < inv_t_nu ? inv_t_nu : inv_t_nu_hyper;
  double DT_max = pow(mx, -1.);
  double dt = 0.0;
  double cfl = p->cfl, sigma_f = p->sigma_f;
  if (p->tau0 != 0 && sigma_f != 0) {
    double A = sqrt(cfl * Delta / (p->tau0 / p->dh[0] * p->forc_mode * PI_REF / Ly * Ly));
    double B = pow((Delta * cfl) * (Delta * cfl) / sigma_f, 1. / 3.);
    /* 1e-3*min(A,B) with min(p,q) = p > q ? q : p */
    dt = 1e-3 * A > B ? B : A;
  } else if (p->tau0 != 0 && sigma_f == 0) {
    dt = 1e-3 * sqrt(cfl * Delta / (p->tau0 / p->dh[0] * p->forc_mode * PI_REF / Ly * Ly));
  } else if (p->tau0 == 0 && sigma_f != 0) {
    dt = 1e-3 * pow((Delta * cfl) * (Delta * cfl) / sigma_f, 1. / 3.);
  }
  if (DT_max != 0 && dt > DT_max) dt = DT_max;
  *dt_out = dt;
  *DT_max_out = DT_max;
}

static int compute_modes(orc *s) {
  const orc_params *p = &s->p;
  int nl = p->nl;
  double htotal = 0;
  for (int k = 0; k < nl; k++) htotal += p->dh[k];
  double *G = (double *)calloc((size_t)nl * nl, sizeof(double));
  double dhc[ORC_NL_MAX];
  for (int l = 0; l < nl - 1; l++) dhc[l] = 0.5 * (p->dh[l] + p->dh[l + 1]);
  if (nl > 1) {
    int l = 0;
    G[nl * l + l + 1] = -(p->f0 * p->f0) / p->N2[l] / (dhc[l] * p->dh[l]);
    G[nl * l + l] = -G[nl * l + l + 1];
    for (l = 1; l < nl - 1; l++) {
      G[nl * l + l - 1] = -(p->f0 * p->f0) / p->N2[l - 1] / (dhc[l - 1] * p->dh[l]);
      G[nl * l + l + 1] = -(p->f0 * p->f0) / p->N2[l] / (dhc[l] * p->dh[l]);
      G[nl * l + l] = -G[nl * l + l - 1] - G[nl * l + l + 1];
    }
    l = nl - 1;
    G[nl * l + l - 1] = -(p->f0 * p->f0) / p->N2[l - 1] / (dhc[l - 1] * p->dh[l]);
    G[nl * l + l] = -G[nl * l + l - 1];
  } else {
    G[0] = 1.;
  }
  /* Gamma = D^-1 S with S symmetric: eigen-decompose T = D^1/2 Gamma D^-1/2 */
  double *T = (double *)malloc(sizeof(double) * nl * nl), *V = (double *)malloc(sizeof(double) * nl * nl);
  double *w = (double *)malloc(sizeof(double) * nl);
  double *vr = (double *)malloc(sizeof(double) * nl * nl), *vl = (double *)malloc(sizeof(double) * nl * nl);
  for (int i = 0; i < nl; i++)
    for (int j = 0; j < nl; j++) T[i * nl + j] = 0.0;
  for (int i = 0; i < nl; i++) T[i * nl + i] = G[i * nl + i];
  for (int i = 0; i + 1 < nl; i++) {
    double a = G[i * nl + i + 1], b = G[(i + 1) * nl + i];
    double v = (a < 0 ? -1.0 : 1.0) * sqrt(fabs(a * b));
    T[i * nl + i + 1] = v;
    T[(i + 1) * nl + i] = v;
  }
  orc_eig_sym(nl, T, w, V); /* ascending, as the qsort of src/eigmode.h:154-160 */
  for (int k = 0; k < nl; k++)
    for (int m = 0; m < nl; m++) {
      vr[k * nl + m] = V[k * nl + m] / sqrt(p->dh[k]);
      vl[k * nl + m] = V[k * nl + m] * sqrt(p->dh[k]);
    }
  /* Flierl normalisation, sign(x) = x > 0 ? 1 : -1 (src/qg.c:42) */
  for (int m = 0; m < nl; m++) {
    double dotp = 0.;
    for (int k = 0; k < nl; k++) dotp += p->dh[k] * vr[k * nl + m] * vr[k * nl + m];
    double flfac = (vr[m] > 0 ? 1 : -1) * sqrt(htotal / dotp);
    for (int k = 0; k < nl; k++) vr[k * nl + m] = flfac * vr[k * nl + m];
  }
  for (int m = 0; m < nl; m++) {
    double dotp = 0.;
    for (int k = 0; k < nl; k++) dotp += vr[k * nl + m] * vl[k * nl + m];
    for (int k = 0; k < nl; k++) vl[k * nl + m] = vl[k * nl + m] / dotp;
  }
  for (int m = 0; m < nl; m++)
    for (int k = 0; k < nl; k++) {
      s->cl2m[k * nl + m] = vl[m * nl + k];
      s->cm2l[k * nl + m] = vr[k * nl + m];
    }
  for (int k = 1; k < nl; k++) s->iRd2[k] = w[k];
  s->iRd2[0] = 0.;
  if (nl == 1 && p->Ld != 0) s->iRd2[0] = 1 / (p->Ld * p->Ld);
  free(G);
  free(T);
  free(V);
  free(w);
  free(vr);
  free(vl);
  return 0;
}

orc *orc_create(const orc_params *p, int *status) {
  *status = 0;
  const int per = (p->bc_fac == -1);
  if ((!per && p->bc_fac < 0) || p->NX < 3 || p->NY < 3 || p->nl < 1 || p->nl > ORC_NL_MAX ||
      (per && (p->NX != p->NY || p->NX % 2))) {
    *status = -1;
    return NULL;
  }
  orc *s = (orc *)calloc(1, sizeof(orc));
  s->p = *p;
  s->periodic = per;
  /* init_domain (src/domain.h:64-176): bounded runs drop the wall point */
  s->Delta = p->Lx / p->NX;
  s->Lx = p->Lx;
  s->Ly = p->NY * s->Delta;
  s->Nx = per ? p->NX : p->NX - 1;
  s->Ny = per ? p->NY : p->NY - 1;
  s->Nxp1 = s->Nx + 1;
  s->Nyp1 = s->Ny + 1;
  s->Nxp2 = s->Nx + 2;
  s->Nyp2 = s->Ny + 2;
  s->X = (double *)calloc(s->Nxp2, sizeof(double));
  s->Y = (double *)calloc(s->Nyp2, sizeof(double));
  s->L = (double *)calloc(s->Nxp2, sizeof(double));
  s->K = (double *)calloc(s->Nyp2, sizeof(double));
  const int I0 = per ? -1 : 0; /* periodic: the zero of X and Y at the second point */
  for (int i = 0; i < s->Nxp2; i++) s->X[i] = (i + I0) * s->Delta;
  for (int j = 0; j < s->Nyp2; j++) s->Y[j] = (j + I0) * s->Delta;
  if (per) {
    for (int j = 0; j < s->Nx; j++) s->L[j] = 2 * PI_REF / p->Lx * (((j + s->Nx / 2) % s->Nx) - s->Nx / 2);
    for (int i = 0; i < s->Nx / 2 + 1; i++) s->K[i] = 2 * PI_REF * i / p->Lx;
  } else {
    for (int j = 1; j < s->Nyp1; j++) s->K[j] = PI_REF * j / p->Lx;
    for (int i = 1; i < s->Nxp1; i++) s->L[i] = PI_REF * i / p->Lx;
  }
  /* init_elliptic (src/elliptic.h:29-71) */
  if (per) {
    s->norm = (double)p->NX * p->NY;
    s->cw = (cplx *)calloc((size_t)s->Nx * s->Ny, sizeof(cplx));
    plan_init(&s->px, s->Nx);
    plan_init(&s->py, s->Ny);
  } else {
    s->norm = 4.0 * s->Nxp1 * s->Nyp1;
    plan_init(&s->px, 2 * (s->Nx + 1));
    plan_init(&s->py, 2 * (s->Ny + 1));
  }
  s->wrk1 = (double *)calloc((size_t)s->Nx * s->Ny, sizeof(double));
  s->cl2m = (double *)calloc((size_t)p->nl * p->nl, sizeof(double));
  s->cm2l = (double *)calloc((size_t)p->nl * p->nl, sizeof(double));
  s->iRd2 = (double *)calloc(p->nl, sizeof(double));
  compute_modes(s);
  /* init_vars */
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * p->nl;
  s->psi = (double *)calloc(S, sizeof(double));
  s->q = (double *)calloc(S, sizeof(double));
  s->diff = (double *)calloc(S, sizeof(double));
  s->lapn = (double *)calloc(S, sizeof(double));
  s->topo = (double *)calloc((size_t)s->Nxp2 * s->Nyp2, sizeof(double));
  for (int j = 0; j < s->Nyp2; j++)
    for (int i = 0; i < s->Nxp2; i++) {
#define SHELF(x, d) (1 - exp(-((x) * (x)) / ((g_mut == ORC_MUT_SHELF_WIDTH ? 1 : 2) * ((d) * (d)))))
      double Ly_top = g_mut == ORC_MUT_TOPO_LX ? p->Lx : s->Ly;
      s->topo[IDX(s, i, j, 0)] =
          p->h_topo * (1 - SHELF(s->X[i], p->w_topo) * SHELF(p->Lx - s->X[i], p->w_topo) *
                               SHELF(s->Y[j], p->w_topo) * SHELF(Ly_top - s->Y[j], p->w_topo));
#undef SHELF
    }
  s->nu_hyper = pow(-1, p->n_hyper / 2 - 1) * fabs(p->nu_hyper);
  /* init_timestep */
  s->f0 = (double *)calloc(S, sizeof(double));
  s->f1 = (double *)calloc(S, sizeof(double));
  s->f2 = (double *)calloc(S, sizeof(double));
  orc_check_timestep(p, &s->dt, &s->DT_max);
  if (p->dt_fixed > 0) s->dt = p->dt_fixed;
  s->dt0 = s->dt;
  s->dt1 = s->dt;
  s->t = 0;
  s->t_out = 0;
  s->it = 0;
  return s;
}

void orc_destroy(orc *s) {
  if (!s) return;
  free(s->X); free(s->Y); free(s->K); free(s->L);
  free(s->psi); free(s->q); free(s->diff); free(s->lapn); free(s->topo);
  free(s->f0); free(s->f1); free(s->f2); free(s->wrk1); free(s->cw);
  free(s->cl2m); free(s->cm2l); free(s->iRd2);
  free(s->px.W); free(s->py.W); free(s->forc);
  free(s->st_i); free(s->st_j); free(s->st_spec); free(s->forc_p); free(s->forc_s);
  free(s);
}

int orc_modes(orc *s, double *cl2m, double *cm2l, double *iRd2) {
  int nl = s->p.nl;
  memcpy(cl2m, s->cl2m, sizeof(double) * nl * nl);
  memcpy(cm2l, s->cm2l, sizeof(double) * nl * nl);
  memcpy(iRd2, s->iRd2, sizeof(double) * nl);
  return 0;
}

int orc_set_modes(orc *s, const double *cl2m, const double *cm2l, const double *iRd2) {
  int nl = s->p.nl;
  memcpy(s->cl2m, cl2m, sizeof(double) * nl * nl);
  memcpy(s->cm2l, cm2l, sizeof(double) * nl * nl);
  memcpy(s->iRd2, iRd2, sizeof(double) * nl);
  return 0;
}

double *orc_field(orc *s, int which) {
  switch (which) {
    case ORC_Q: return s->q;
    case ORC_PSI: return s->psi;
    case ORC_TOPO: return s->topo;
    case ORC_FORC: return s->forc_s;
    case ORC_F0: return s->f0;
    case ORC_F1: return s->f1;
    case ORC_F2: return s->f2;
    case ORC_LAPN: return s->lapn;
  }
  return NULL;
}

void orc_sizes(orc *s, int *n, int *m, int *nl) {
  *n = s->Nx;
  *m = s->Ny;
  *nl = s->p.nl;
}

double orc_nu_hyper(orc *s) { return s->nu_hyper; }

/* adjust_bc (src/domain.h:221-349): periodic copies, or the bounded wall
 * values; S, N, W, E in that order, so W/E own the corners. */
void orc_adjust_bc(orc *s, double *field, int field_type) {
  if (s->periodic) {
    for (int k = 0; k < s->p.nl; k++) {
      for (int i = 0; i < s->Nxp2; i++) field[IDX(s, i, 0, k)] = field[IDX(s, i, s->Ny, k)];
      for (int i = 0; i < s->Nxp2; i++) field[IDX(s, i, s->Nyp1, k)] = field[IDX(s, i, 1, k)];
      for (int j = 0; j < s->Nyp2; j++) field[IDX(s, 0, j, k)] = field[IDX(s, s->Nx, j, k)];
      for (int j = 0; j < s->Nyp2; j++) field[IDX(s, s->Nxp1, j, k)] = field[IDX(s, 1, j, k)];
    }
    return;
  }
  double psi_bc = 0.;
  double *psi = s->psi;
  double c = (g_mut == ORC_MUT_BC_FACTOR ? 1 : 2) * s->p.bc_fac / (s->Delta * s->Delta);
  for (int k = 0; k < s->p.nl; k++) {
    for (int pass = 0; pass < 2; pass++) {
      /* S and N first, then W and E, which therefore own the corners */
      int sn = (pass == 0) != (g_mut == ORC_MUT_BC_ORDER);
      if (sn) {
        for (int i = 0; i < s->Nxp2; i++) {
          int j = 0;
          field[IDX(s, i, j, k)] = (field_type ? c * (psi[IDX(s, i, j + 1, k)] - psi_bc) : psi_bc);
        }
        for (int i = 0; i < s->Nxp2; i++) {
          int j = s->Nyp1;
          field[IDX(s, i, j, k)] = (field_type ? c * (psi[IDX(s, i, j - 1, k)] - psi_bc) : psi_bc);
        }
      } else {
        for (int j = 0; j < s->Nyp2; j++) {
          int i = 0;
          field[IDX(s, i, j, k)] = (field_type ? c * (psi[IDX(s, i + 1, j, k)] - psi_bc) : psi_bc);
        }
        for (int j = 0; j < s->Nyp2; j++) {
          int i = s->Nxp1;
          field[IDX(s, i, j, k)] = (field_type ? c * (psi[IDX(s, i - 1, j, k)] - psi_bc) : psi_bc);
        }
      }
    }
  }
}

/* 2-D complex DFT of an Ny x Nx array (rows of Nx), forward or inverse
 * (unnormalised; inverse by conjugation). */
static void fft2_plans(cplx *a, int Nx, int Ny, const fftplan *px, const fftplan *py, int inverse) {
#pragma omp parallel
  {
    int Lm = Nx > Ny ? Nx : Ny;
    cplx *b1 = (cplx *)malloc(sizeof(cplx) * Lm), *b2 = (cplx *)malloc(sizeof(cplx) * Lm);
#pragma omp for schedule(static)
    for (int j = 0; j < Ny; j++) {
      for (int i = 0; i < Nx; i++) b1[i] = inverse ? conj(a[(long)j * Nx + i]) : a[(long)j * Nx + i];
      fft_rec(px, b2, b1, Nx, 1, 1);
      for (int i = 0; i < Nx; i++) a[(long)j * Nx + i] = inverse ? conj(b2[i]) : b2[i];
    }
#pragma omp for schedule(static)
    for (int i = 0; i < Nx; i++) {
      for (int j = 0; j < Ny; j++) b1[j] = inverse ? conj(a[(long)j * Nx + i]) : a[(long)j * Nx + i];
      fft_rec(py, b2, b1, Ny, 1, 1);
      for (int j = 0; j < Ny; j++) a[(long)j * Nx + i] = inverse ? conj(b2[j]) : b2[j];
    }
    free(b1);
    free(b2);
  }
}

/* invert_pv, periodic branch (src/elliptic.h:76-155 with bc_fac = -1): the
 * reference's r2c/c2r pair is restated on the full complex spectrum, which is
 * Hermitian (the divisor is real and even in both wavenumbers), so its real
 * part is the c2r result.  Division where the divisor is non-zero; the (0,0)
 * coefficient is zeroed for every mode; normalisation by NX*NY. */
static void invert_pv_periodic(orc *s) {
  int nl = s->p.nl, Nx = s->Nx, Ny = s->Ny;
  double *psi = s->psi, *q = s->q;
  cplx *cw = s->cw;
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * nl;
  for (size_t t = 0; t < S; t++) psi[t] = 0.;
  for (int k = 0; k < nl; k++) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        double acc = 0;
        for (int l = 0; l < nl; l++) acc += s->cl2m[k * nl + l] * q[IDX(s, i, j, l)];
        cw[(long)(j - 1) * Nx + (i - 1)] = acc;
      }
    fft2_plans(cw, Nx, Ny, &s->px, &s->py, 0);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < Ny; j++)
      for (int i = 0; i < Nx; i++) {
        int ih = i <= Nx / 2 ? i : Nx - i; /* K of the Hermitian partner */
        double fact = -(s->L[j] * s->L[j] + s->K[ih] * s->K[ih] + s->iRd2[k]);
        if (fact != 0) cw[(long)j * Nx + i] = cw[(long)j * Nx + i] / fact;
      }
    cw[0] = 0;
    fft2_plans(cw, Nx, Ny, &s->px, &s->py, 1);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        double w = creal(cw[(long)(j - 1) * Nx + (i - 1)]) / s->norm;
        for (int l = 0; l < nl; l++) psi[IDX(s, i, j, l)] += s->cm2l[l * nl + k] * w;
      }
  }
  orc_adjust_bc(s, psi, 0);
  orc_adjust_bc(s, q, 1);
}

void orc_invert_pv(orc *s) {
  if (s->periodic) {
    invert_pv_periodic(s);
    return;
  }
  int nl = s->p.nl, Nx = s->Nx, Ny = s->Ny;
  double *psi = s->psi, *q = s->q, *wrk1 = s->wrk1;
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * nl;
  for (size_t t = 0; t < S; t++) psi[t] = 0.;
  for (int k = 0; k < nl; k++) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        long f = (long)(j - 1) * Nx + (i - 1);
        wrk1[f] = 0;
        for (int l = 0; l < nl; l++) wrk1[f] += s->cl2m[k * nl + l] * q[IDX(s, i, j, l)];
      }
    dst2d_plans(wrk1, Nx, Ny, &s->px, &s->py);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        double fact = -(s->L[i] * s->L[i] + s->K[j] * s->K[j] + s->iRd2[k]);
        long f = (long)(j - 1) * Nx + (i - 1);
        wrk1[f] = wrk1[f] / fact;
      }
    dst2d_plans(wrk1, Nx, Ny, &s->px, &s->py);
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        long f = (long)(j - 1) * Nx + (i - 1);
        wrk1[f] = wrk1[f] / s->norm;
      }
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        long f = (long)(j - 1) * Nx + (i - 1);
        for (int l = 0; l < nl; l++) psi[IDX(s, i, j, l)] += s->cm2l[l * nl + k] * wrk1[f];
      }
  }
  orc_adjust_bc(s, psi, 0);
  orc_adjust_bc(s, q, 1);
}

/* laplacian macro of src/domain.h:12 */
static inline double lap(const orc *s, const double *p, int i, int j, int k) {
  return (p[IDX(s, i + 1, j, k)] + p[IDX(s, i - 1, j, k)] + p[IDX(s, i, j + 1, k)] +
          p[IDX(s, i, j - 1, k)] - 4 * p[IDX(s, i, j, k)]) /
         (s->Delta * s->Delta);
}

/* jacobian_lev(p, q, k1, k2) of src/domain.h:17-27 (Arakawa 1966) */
static inline double jac(const orc *s, const double *p, const double *q, int i, int j, int k1, int k2) {
#define P(a, b) p[IDX(s, a, b, k1)]
#define Q(a, b) q[IDX(s, a, b, k2)]
  return (((P(i + 1, j) - P(i - 1, j)) * (Q(i, j + 1) - Q(i, j - 1)) +
           (P(i, j - 1) - P(i, j + 1)) * (Q(i + 1, j) - Q(i - 1, j)) +
           P(i + 1, j) * (Q(i + 1, j + 1) - Q(i + 1, j - 1)) -
           P(i - 1, j) * (Q(i - 1, j + 1) - Q(i - 1, j - 1)) -
           P(i, j + 1) * (Q(i + 1, j + 1) - Q(i - 1, j + 1)) +
           P(i, j - 1) * (Q(i + 1, j - 1) - Q(i - 1, j - 1)) +
           Q(i, j + 1) * (P(i + 1, j + 1) - P(i - 1, j + 1)) -
           Q(i, j - 1) * (P(i + 1, j - 1) - P(i - 1, j - 1)) -
           Q(i + 1, j) * (P(i + 1, j + 1) - P(i + 1, j - 1)) +
           Q(i - 1, j) * (P(i - 1, j + 1) - P(i - 1, j - 1))) /
          (12. * s->Delta * s->Delta));
#undef P
#undef Q
}

void orc_calc_diff(orc *s) {
  int nl = s->p.nl;
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * nl;
  memcpy(s->diff, s->psi, S * sizeof(double));
  for (int n = 0; n <= s->p.n_hyper; n += 2) {
    for (int k = 0; k < nl; k++) {
#pragma omp parallel for schedule(static)
      for (int j = 1; j < s->Nyp1; j++)
        for (int i = 1; i < s->Nxp1; i++) s->lapn[IDX(s, i, j, k)] = lap(s, s->diff, i, j, k);
    }
    orc_adjust_bc(s, s->lapn, 1);
    memcpy(s->diff, s->lapn, S * sizeof(double));
  }
}

static double *field_ptr(orc *s, int which) { return orc_field(s, which); }

void orc_rhs(orc *s, int which_out) {
  double *f1 = field_ptr(s, which_out);
  const orc_params *p = &s->p;
  int nl = p->nl;
  double *psi = s->psi, *q = s->q;
  orc_invert_pv(s);
  orc_calc_diff(s);
  double Delta = s->Delta;
  /* planted mistakes (g_mut): the faithful terms are the ones with g_mut = 0 */
  const double bsign = g_mut == ORC_MUT_BETA_SIGN ? -1 : 1;
  const int bdi = g_mut == ORC_MUT_BETA_AXIS ? 0 : 1, bdj = 1 - bdi;
  const double *nuf = g_mut == ORC_MUT_NU_PSI ? psi : q;
  for (int k = 0; k < nl; k++) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        f1[IDX(s, i, j, k)] = -jac(s, psi, q, i, j, k, k) -
                              bsign * (p->beta * (psi[IDX(s, i + bdi, j + bdj, k)] - psi[IDX(s, i - bdi, j - bdj, k)]) /
                                       (2 * Delta)) +
                              p->nu * (nuf[IDX(s, i + 1, j, k)] + nuf[IDX(s, i - 1, j, k)] + nuf[IDX(s, i, j + 1, k)] +
                                       nuf[IDX(s, i, j - 1, k)] - 4 * nuf[IDX(s, i, j, k)]) /
                                  (Delta * Delta) +
                              s->nu_hyper * s->lapn[IDX(s, i, j, k)];
      }
  }
  double dhb = p->dh[nl - 1];
  const int wl = g_mut == ORC_MUT_WIND_LAYER ? nl - 1 : 0;
  const double wdh = g_mut == ORC_MUT_WIND_DH ? p->dh[nl - 1] : p->dh[0];
  const double wLy = g_mut == ORC_MUT_WIND_LX ? p->Lx : s->Ly;
  const double ddh = g_mut == ORC_MUT_DRAG_DH ? p->dh[0] : dhb;
  const double dhalf = g_mut == ORC_MUT_DRAG_HALF ? 1 : 2;
  const double *dfield = g_mut == ORC_MUT_DRAG_Q ? q : psi;
  const int tk = g_mut == ORC_MUT_TOPO_LAYER ? 0 : nl - 1;
#pragma omp parallel for schedule(static)
  for (int j = 1; j < s->Nyp1; j++)
    for (int i = 1; i < s->Nxp1; i++) {
      /* forcing_q (src/domain.h:33) */
      f1[IDX(s, i, j, wl)] += (-p->tau0 / wdh * p->forc_mode * PI_REF / wLy *
                               sin(p->forc_mode * PI_REF * (s->Y[j]) / wLy));
      int k = nl - 1;
      double jt = g_mut == ORC_MUT_TOPO_SWAP ? jac(s, s->topo, psi, i, j, 0, tk) : jac(s, psi, s->topo, i, j, tk, 0);
      f1[IDX(s, i, j, nl - 1)] +=
          -p->hEkb * p->f0 / (dhalf * ddh) *
              (dfield[IDX(s, i + 1, j, k)] + dfield[IDX(s, i - 1, j, k)] + dfield[IDX(s, i, j + 1, k)] +
               dfield[IDX(s, i, j - 1, k)] - 4 * dfield[IDX(s, i, j, k)]) /
              (Delta * Delta) -
          jt * p->f0 / dhb;
    }
}

double orc_adjust_timestep(orc *s) {
  double dt_max = HUGE_REF;
  double *psi = s->psi;
  double Delta = s->Delta, cfl = s->p.cfl;
  for (int k = 0; k < s->p.nl; k++)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        double u = -(psi[IDX(s, i + 1, j, k)] - psi[IDX(s, i, j, k)]) / Delta;
        double v = (psi[IDX(s, i, j + 1, k)] - psi[IDX(s, i, j, k)]) / Delta;
        u = fabs(u) < fabs(v) ? fabs(v) : fabs(u);
        if (u != 0.) {
          double dt_loc = cfl * Delta / u;
          if (dt_loc < dt_max) dt_max = dt_loc;
        }
      }
  double dt = s->dt;
  if (dt < dt_max) dt = 2. * dt;
  if (dt > s->DT_max) dt = s->DT_max;
  if (dt > dt_max) dt = dt_max;
  double t = s->t, t_out = s->t_out;
  if (t_out > t + TEPS) {
    unsigned int n_step = (t_out - t) / dt;
    if (n_step == 0)
      dt = t_out - t;
    else {
      double dt1 = (t_out - t) / n_step;
      if (dt1 > dt * (1. + TEPS))
        dt = (t_out - t) / (n_step + 1);
      else if (dt1 < dt)
        dt = dt1;
    }
  }
  return dt;
}

void orc_ab3_coeffs(double dt, double dt0, double dt1, double *c) {
  c[0] = dt * (2 * (dt * dt) + 3 * dt * (2 * dt0 + dt1) + 6 * dt0 * (dt0 + dt1)) / (6 * dt0 * (dt0 + dt1));
  c[1] = (dt * dt) * (-dt / 3 - dt0 / 2 - dt1 / 2) / (dt0 * dt1);
  c[2] = (dt * dt) * (dt / 3 + dt0 / 2) / (dt1 * (dt0 + dt1));
}

/* c2r of the reference's FFTW plan (unnormalised backward transform of an
 * N x N real field from its N x (N/2+1) half spectrum, src/forcing.h:119,196)
 * restated for a half spectrum that need not be Hermitian: inverse complex
 * DFTs along j for every column i, then along i the real inverse of each
 * half-spectrum row, which uses only the real parts of its i = 0 and (N even)
 * i = N/2 entries. */
int orc_c2r2d(const double *spec, int N, double *out) {
  if (N < 1) return -1;
  const int NK = N / 2 + 1;
  fftplan pl;
  if (plan_init(&pl, N)) return -1;
  cplx *G = (cplx *)malloc(sizeof(cplx) * (size_t)N * NK);
  cplx *b1 = (cplx *)malloc(sizeof(cplx) * N), *b2 = (cplx *)malloc(sizeof(cplx) * N);
  for (int i = 0; i < NK; i++) {
    for (int j = 0; j < N; j++) b1[j] = conj(spec[2 * ((long)j * NK + i)] + I * spec[2 * ((long)j * NK + i) + 1]);
    fft_rec(&pl, b2, b1, N, 1, 1);
    for (int j = 0; j < N; j++) G[(long)j * NK + i] = conj(b2[j]);
  }
  for (int j = 0; j < N; j++) {
    const cplx *g = G + (long)j * NK;
    b1[0] = creal(g[0]);
    for (int i = 1; i < NK; i++) {
      if (2 * i == N) {
        b1[i] = creal(g[i]);
      } else {
        b1[i] = g[i];
        b1[N - i] = conj(g[i]);
      }
    }
    for (int i = 0; i < N; i++) b1[i] = conj(b1[i]);
    fft_rec(&pl, b2, b1, N, 1, 1);
    for (int i = 0; i < N; i++) out[(long)j * N + i] = creal(conj(b2[i]));
  }
  free(G); free(b1); free(b2); free(pl.W);
  return 0;
}

/* init_stoch_forc (src/forcing.h:67-166): the forced points of the half
 * spectrum, j outer and i inner.  The C library's rand() stream is reset
 * (srand(1) is the stream a program starts with), so one stochastic oracle
 * at a time draws the reference's sequence. */
int orc_set_stochastic(orc *s, double k_f, double sigma_f) {
  free(s->st_i); free(s->st_j); free(s->st_spec); free(s->forc_p); free(s->forc_s);
  s->st_i = s->st_j = NULL;
  s->st_spec = NULL;
  s->forc_p = s->forc_s = NULL;
  s->stoch = 0;
  if (k_f == 0) return 0;
  const int N = s->p.NX; /* N_F: NX periodic, NXp1 = NX bounded */
  if (N < 2 || s->Nx + 1 < N || s->Ny + 1 < N) return -1; /* the N x N field must fit */
  const int NK = N / 2 + 1;
  s->NF = N;
  s->st_kf = k_f / (2 * PI_REF);
  s->st_sigma = sigma_f;
  const double kf = s->st_kf, dk = 1. / s->p.Lx;
  s->st_i = (int *)malloc(sizeof(int) * (size_t)N * NK);
  s->st_j = (int *)malloc(sizeof(int) * (size_t)N * NK);
  int np = 0;
  for (int j = 0; j < N; j++)
    for (int i = 0; i < NK; i++) {
      double l = (((j + N / 2) % N) - N / 2) * dk;
      double k = i * dk;
      double dk_ = fabs(k - sqrt(kf * kf - l * l));
      double m1 = fabs(l - sqrt(kf * kf - k * k)), m2 = fabs(l + sqrt(kf * kf - k * k));
      double dl_ = m1 > m2 ? m2 : m1;
      if (dk_ < dk / 2 || dl_ < dk / 2) {
        s->st_i[np] = i;
        s->st_j[np] = j;
        np++;
      }
    }
  s->n_P = np;
  s->st_spec = (cplx *)calloc((size_t)N * NK, sizeof(cplx));
  s->forc_p = (double *)calloc((size_t)N * N, sizeof(double));
  s->forc_s = (double *)calloc((size_t)s->Nxp2 * s->Nyp2, sizeof(double));
  s->stoch = 1;
  srand(1);
  return np;
}

int orc_stoch_points(orc *s, int *ii, int *jj) {
  for (int p = 0; s->stoch && p < s->n_P; p++) {
    if (ii) ii[p] = s->st_i[p];
    if (jj) jj[p] = s->st_j[p];
  }
  return s->stoch ? s->n_P : 0;
}

/* One step of calc_forc's random coefficients (src/forcing.h:178-190), from
 * the C library's rand(); the Box-Muller normal is written as one expression
 * of two rand() calls, as the reference's macro is, so the compiler orders
 * the two draws as it does there. */
static void stoch_draw(orc *s, double *coef) {
  const double pi = PI_REF, k_f = s->st_kf;
  for (int p = 0; p < s->n_P; p++) {
    double envelope = sqrt((double)8 / ((s->n_P - 1) * 2)) * pi * k_f;
    double magnitude = s->st_sigma * (sqrt(-2. * log(((double)(rand()) + 1.) / ((double)(RAND_MAX) + 2.))) *
                                      cos(2 * pi * rand() / (double)RAND_MAX));
    double phase = (((double)rand() / RAND_MAX)) * 2 * pi;
    coef[2 * p] = envelope * magnitude * cos(phase);
    coef[2 * p + 1] = envelope * magnitude * sin(phase);
  }
}

int orc_stoch_draw(orc *s, double *coef) {
  if (!s->stoch) return -1;
  stoch_draw(s, coef);
  return s->n_P;
}

/* calc_forc (src/forcing.h:168-208): the spectrum of this step's draws, its
 * c2r, copied to forc_s[i+1][j+1] of layer 0. */
static void calc_forc(orc *s) {
  if (s->n_P <= 1) return;
  const int N = s->NF, NK = N / 2 + 1;
  double *coef = (double *)malloc(sizeof(double) * 2 * s->n_P);
  double *spec = (double *)calloc((size_t)2 * N * NK, sizeof(double));
  stoch_draw(s, coef);
  for (int p = 0; p < s->n_P; p++) {
    long o = (long)s->st_j[p] * NK + s->st_i[p];
    spec[2 * o] = coef[2 * p];
    spec[2 * o + 1] = coef[2 * p + 1];
  }
  orc_c2r2d(spec, N, s->forc_p);
  for (int j = 0; j < N; j++)
    for (int i = 0; i < N; i++) s->forc_s[IDX(s, i + 1, j + 1, 0)] = s->forc_p[(long)j * N + i];
  free(coef);
  free(spec);
}

void orc_timestep(orc *s) {
  double *swap = s->f2;
  s->f2 = s->f1;
  s->f1 = s->f0;
  s->f0 = swap;
  orc_rhs(s, ORC_F0);
  s->dt1 = s->dt0;
  s->dt0 = s->dt;
  if (s->p.dt_fixed > 0)
    s->dt = s->p.dt_fixed;
  else
    s->dt = orc_adjust_timestep(s);
  double c[3];
  orc_ab3_coeffs(s->dt, s->dt0, s->dt1, c);
  double *q = s->q, *f0 = s->f0, *f1 = s->f1, *f2 = s->f2;
  for (int k = 0; k < s->p.nl; k++) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) {
        long o = IDX(s, i, j, k);
        q[o] = q[o] + c[0] * f0[o] + c[1] * f1[o] + c[2] * f2[o];
      }
  }
  if (s->stoch) {
    /* stochastic forcing on layer 0, times sqrt(dt) (src/timestep.h:143-150) */
    calc_forc(s);
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) q[IDX(s, i, j, 0)] += s->forc_s[IDX(s, i, j, 0)] * sqrt(s->dt);
  }
  if (s->n_forc > 0) {
    /* calc_4d_forcing (src/forcing.h:246-289) at the time before the step,
     * added with the new dt (src/timestep.h:152-163) */
    const double t = s->t;
    const int irec1 = ((int)floor(t / s->dt_forc)) % s->n_forc;
    const int irec2 = (irec1 + 1) % s->n_forc;
    const double w2 = t / s->dt_forc - (int)floor(t / s->dt_forc);
    const double w1 = 1 - w2;
    const long S = (long)s->Nxp2 * s->Nyp2 * s->p.nl;
    const double *F1 = s->forc + irec1 * S, *F2 = s->forc + irec2 * S;
    for (int k = 0; k < s->p.nl; k++)
      for (int j = 1; j < s->Nyp1; j++)
        for (int i = 1; i < s->Nxp1; i++) {
          const long o = IDX(s, i, j, k);
          const double qf = w1 * F1[o] + w2 * F2[o];
          q[o] += qf * s->dt;
        }
  }
  s->t = s->t + s->dt;
  s->it++;
}

int orc_set_forcing4d(orc *s, int n_rec, double dt_forc, const double *records) {
  free(s->forc);
  s->forc = NULL;
  s->n_forc = 0;
  if (n_rec <= 0) return 0;
  if (!(dt_forc > 0) || !records) return -1;
  const size_t S = (size_t)s->Nxp2 * s->Nyp2 * s->p.nl * n_rec;
  s->forc = (double *)malloc(S * sizeof(double));
  if (!s->forc) return -1;
  memcpy(s->forc, records, S * sizeof(double));
  s->n_forc = n_rec;
  s->dt_forc = dt_forc;
  return 0;
}

void orc_run(orc *s, int nsteps) {
  for (int i = 0; i < nsteps; i++) orc_timestep(s);
}

void orc_get_time(orc *s, double *t, double *dt, double *dt0, double *dt1, int *it, double *DT_max) {
  if (t) *t = s->t;
  if (dt) *dt = s->dt;
  if (dt0) *dt0 = s->dt0;
  if (dt1) *dt1 = s->dt1;
  if (it) *it = s->it;
  if (DT_max) *DT_max = s->DT_max;
}

void orc_set_time(orc *s, double t, double dt, double dt0, double dt1, int it, double t_out) {
  s->t = t;
  s->dt = dt;
  s->dt0 = dt0;
  s->dt1 = dt1;
  s->it = it;
  s->t_out = t_out;
}

void orc_diagnostics(orc *s, double *E_layer, double *Z_layer, double *E_tot, double *Z_tot) {
  int nl = s->p.nl;
  size_t S = (size_t)s->Nxp2 * s->Nyp2;
  double w = s->Delta * s->Delta / (s->Lx * s->Lx);
  double H = 0, Et = 0, Zt = 0;
  for (int l = 0; l < nl; l++) H += s->p.dh[l];
  for (int l = 0; l < nl; l++) {
    double e = 0, z = 0;
    for (size_t t = 0; t < S; t++) {
      e += s->psi[l * S + t] * s->q[l * S + t];
      z += s->q[l * S + t] * s->q[l * S + t];
    }
    e = -e * w;
    z = z * w;
    if (E_layer) E_layer[l] = e;
    if (Z_layer) Z_layer[l] = z;
    Et += s->p.dh[l] * e;
    Zt += s->p.dh[l] * z;
  }
  if (E_tot) *E_tot = Et / (2 * H);
  if (Z_tot) *Z_tot = Zt / H;
}

/* The stencil macros on caller-supplied fields (storage layout of this
 * state, nl layers), for the identity checks of the tests: out = J(p,q) or
 * laplacian(p) at interior points, zero elsewhere. */
void orc_jacobian_field(orc *s, const double *p, const double *q, double *out) {
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * s->p.nl;
  for (size_t t = 0; t < S; t++) out[t] = 0.0;
  for (int k = 0; k < s->p.nl; k++)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) out[IDX(s, i, j, k)] = jac(s, p, q, i, j, k, k);
}

void orc_laplacian_field(orc *s, const double *p, double *out) {
  size_t S = (size_t)s->Nxp2 * s->Nyp2 * s->p.nl;
  for (size_t t = 0; t < S; t++) out[t] = 0.0;
  for (int k = 0; k < s->p.nl; k++)
    for (int j = 1; j < s->Nyp1; j++)
      for (int i = 1; i < s->Nxp1; i++) out[IDX(s, i, j, k)] = lap(s, p, i, j, k);
}
