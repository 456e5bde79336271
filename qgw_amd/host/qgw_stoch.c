/* qgw_stoch.c -- host part of the stochastic ring forcing (see
 * include/qgw_host.h): the forced wavenumbers and the per-step random
 * coefficients.  The device turns the coefficients into the forcing field. */
#define _GNU_SOURCE
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "qgw_host.h"

static const double pi_ref = 3.141592653589793; /* src/qg.c:44 */

int qgwh_stoch_ring(const qgw_params *p, int cap, int *ii, int *jj, int *n_forced, double *envelope) {
  if (!p || !n_forced || cap < 0 || p->NX < 2 || !(p->Lx > 0)) return QGW_ERR_ARG;
  const int N = p->NX;             /* N_F (src/forcing.h:84-88) */
  const int NK = N / 2 + 1;        /* NK_f */
  const double k_f = p->k_f / (2 * pi_ref);
  const double dk = 1. / p->Lx;
  int np = 0;
  for (int j = 0; j < N; j++)
    for (int i = 0; i < NK; i++) {
      const double l = (((j + N / 2) % N) - N / 2) * dk;
      const double k = i * dk;
      const double dist_k = fabs(k - sqrt(k_f * k_f - l * l));
      const double a = fabs(l - sqrt(k_f * k_f - k * k)), b = fabs(l + sqrt(k_f * k_f - k * k));
      const double dist_l = a > b ? b : a; /* min(p,q), src/qg.c:40 */
      if (dist_k < dk / 2 || dist_l < dk / 2) {
        if (np < cap) {
          if (ii) ii[np] = i;
          if (jj) jj[np] = j;
        }
        np++;
      }
    }
  *n_forced = np;
  if (envelope) *envelope = (np > 1) ? sqrt((double)8 / ((np - 1) * 2)) * pi_ref * k_f : 0.0;
  return QGW_OK;
}

struct qgwh_rng {
  struct random_data rd;
  char state[128]; /* 128 bytes: the additive-feedback generator rand() uses */
};

qgwh_rng *qgwh_rng_create(unsigned seed) {
  qgwh_rng *r = (qgwh_rng *)calloc(1, sizeof(qgwh_rng));
  if (!r) return NULL;
  if (initstate_r(seed, r->state, sizeof r->state, &r->rd) != 0) {
    free(r);
    return NULL;
  }
  return r;
}

void qgwh_rng_destroy(qgwh_rng *r) { free(r); }

/* The generator's pointers (struct random_data, <stdlib.h>) point into the
 * handle's own state array, so a copy moves them by the handles' distance. */
int qgwh_rng_assign(qgwh_rng *dst, const qgwh_rng *src) {
  if (!dst || !src) return QGW_ERR_ARG;
  if (dst == src) return QGW_OK;
  memcpy(dst, src, sizeof *dst);
  const ptrdiff_t d = (const char *)dst - (const char *)src;
  dst->rd.fptr = (int32_t *)((char *)src->rd.fptr + d);
  dst->rd.rptr = (int32_t *)((char *)src->rd.rptr + d);
  dst->rd.state = (int32_t *)((char *)src->rd.state + d);
  dst->rd.end_ptr = (int32_t *)((char *)src->rd.end_ptr + d);
  return QGW_OK;
}

qgwh_rng *qgwh_rng_clone(const qgwh_rng *src) {
  if (!src) return NULL;
  qgwh_rng *r = (qgwh_rng *)malloc(sizeof(qgwh_rng));
  if (r) qgwh_rng_assign(r, src);
  return r;
}

int qgwh_rng_next(qgwh_rng *r) {
  int32_t v = 0;
  random_r(&r->rd, &v);
  return (int)v;
}

int qgwh_stoch_coeffs(qgwh_rng *r, double sigma_f, double envelope, int count, double *coef) {
  if (!r || count < 0 || (count && !coef)) return QGW_ERR_ARG;
  for (int p = 0; p < count; p++) {
    /* normal_noise() and noise() of src/forcing.h:62-63 */
    const int u1 = qgwh_rng_next(r);
    const int u2 = qgwh_rng_next(r);
    const double nn = sqrt(-2. * log(((double)(u1) + 1.) / ((double)(RAND_MAX) + 2.))) *
                      cos(2 * pi_ref * u2 / (double)RAND_MAX);
    const double magnitude = sigma_f * nn;
    const double phase = ((double)qgwh_rng_next(r) / RAND_MAX) * 2 * pi_ref;
    coef[2 * p] = envelope * magnitude * cos(phase);
    coef[2 * p + 1] = envelope * magnitude * sin(phase);
  }
  return QGW_OK;
}
