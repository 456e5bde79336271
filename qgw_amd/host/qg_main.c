/*
 * qg_main.c -- the C driver: the reference's main program (src/qg.c:103-226)
 * restated over the device C-ABI.
 *
 *   ./qg_hip [params.in]
 *
 * Sequence (src/qg.c:140-200): read parameters, create outdir_NNNN/ and
 * copy the parameter file there (src/extra.h:163-193), set up the run
 * (init_domain/init_elliptic/init_vars/init_timestep -> qgw_create +
 * qgwh_eigmode + qgw_set_modes), read q and topo from restart.nc when it
 * exists, first inversion, create outdir_NNNN/vars.nc, then
 * `while (t < tend)`: at output times invert, append psi and q to vars.nc
 * and report, then one timestep.
 *
 * Differences from the reference driver, all through keys the reference
 * ignores:
 *   ic_kind = 1      seeded k^-3 initial condition instead of restart.nc
 *   dt_fixed > 0     fixed step (the reference adapts dt every step)
 *   nsteps > 0       stop after nsteps steps
 *   precision        0 FP64, 1 FP32 + compensated AB3, 2 FP32
 * Each output also prints energy and enstrophy; errors end the run with a
 * status (the reference returns from the failing routine and carries on).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <sys/stat.h>
#include <sys/types.h>

#include "qgw_hip.h"
#include "qgw_host.h"
#include "qgw_nc.h"

static const double TEPS = 1e-9; /* src/timestep.h:16 */

static int fail(const char *what, int st, qgw_ctx *ctx) {
  fprintf(stderr, "qg_hip: %s: %s %s\n", what, qgw_strerror(st), ctx ? qgw_last_error(ctx) : "");
  return 1;
}

/* create_outdir + backup_config (src/extra.h:163-193) */
static int make_outdir(const char *param_file, char *dir, size_t len) {
  for (int i = 1; i < 10000; i++) {
    snprintf(dir, len, "outdir_%04d/", i);
    if (mkdir(dir, 0777) == 0) {
      fprintf(stdout, "Writing output in %s\n", dir);
      char name[600];
      snprintf(name, sizeof name, "%sparams.in", dir);
      FILE *src = fopen(param_file, "rb"), *dst = fopen(name, "wb");
      int ok = src && dst;
      char buf[4096];
      size_t n;
      while (ok && (n = fread(buf, 1, sizeof buf, src)) > 0) ok = fwrite(buf, 1, n, dst) == n;
      if (src) fclose(src);
      if (dst && fclose(dst)) ok = 0;
      return ok ? QGW_OK : QGW_ERR_IO;
    }
  }
  return QGW_ERR_IO;
}

int main(int argc, char **argv) {
  qgw_params p;
  qgwh_default_params(&p);
  const char *file = (argc >= 2) ? argv[1] : "params.in";
  int st = qgwh_read_params(file, &p);
  if (st) return fail(file, st, NULL);
  char dir[512];
  if ((st = make_outdir(file, dir, sizeof dir))) return fail("output directory", st, NULL);

  qgw_grid g;
  if ((st = qgwh_grid(&p, &g, NULL, NULL, NULL, NULL))) return fail("grid", st, NULL);
  const int nl = p.nl;
  double *cl2m = malloc(sizeof(double) * nl * nl), *cm2l = malloc(sizeof(double) * nl * nl);
  double *iRd2 = malloc(sizeof(double) * nl);
  if ((st = qgwh_eigmode(&p, cl2m, cm2l, iRd2))) return fail("eigmode", st, NULL);
  if (p.verbose)
    for (int k = 0; k < nl; k++)
      fprintf(stdout, "iRd2: %g , def radius: %g\n", iRd2[k], iRd2[k] > 0 ? sqrt(1 / iRd2[k]) : 0);

  qgw_ctx *ctx = qgw_create(&p, p.precision, &st);
  if (!ctx) return fail("create", st, NULL);
  if ((st = qgw_set_modes(ctx, cl2m, cm2l, iRd2))) return fail("set_modes", st, ctx);

  size_t S = (size_t)g.nxp2 * g.nyp2 * nl;
  double *q = calloc(S, sizeof(double)), *psi = calloc(S, sizeof(double));
  if (p.ic_kind == 1) {
    if ((st = qgwh_make_ic(&p, cl2m, cm2l, iRd2, p.seed, psi, q))) return fail("initial condition", st, ctx);
    if ((st = qgw_upload(ctx, QGW_FIELD_Q, q))) return fail("upload", st, ctx);
  } else {
    /* restart (src/qg.c:160-164): q and topo from record 0 of restart.nc;
     * without the file the run starts from rest */
    double *topo = malloc(sizeof(double) * g.nxp2 * g.nyp2);
    int found = 0;
    qgwh_topo(&p, topo);
    struct stat sb;
    if (stat("restart.nc", &sb) == 0) {
      if ((st = qgwh_nc_read("restart.nc", &p, 0, q, topo, &found))) return fail("restart.nc", st, ctx);
      fprintf(stdout, "restart.nc:%s%s\n", (found & 1) ? " q" : "", (found & 2) ? " topo" : "");
    }
    if ((found & 1) && (st = qgw_upload(ctx, QGW_FIELD_Q, q))) return fail("upload", st, ctx);
    if ((found & 2) && (st = qgw_upload(ctx, QGW_FIELD_TOPO, topo))) return fail("upload", st, ctx);
    free(topo);
  }
  if ((st = qgw_invert_pv(ctx))) return fail("invert_pv", st, ctx); /* first inversion */
  char vars[600];
  snprintf(vars, sizeof vars, "%svars.nc", dir);
  qgw_nc *nc = qgwh_nc_create(vars, &p, &st);
  if (!nc) return fail(vars, st, ctx);

  double t = 0, dt = 0, dt0, dt1, t_out = 0;
  int it = 0;
  qgw_get_time(ctx, &t, &dt, &dt0, &dt1, &it);
  const double tend = (p.nsteps > 0 && p.dt_fixed > 0) ? p.nsteps * p.dt_fixed : p.tend;
  while (t < tend) {
    if (p.verbose) fprintf(stdout, "i = %d, t = %e dt = %e \n", it, t, dt);
    if (fabs(t - t_out) < TEPS * (dt > 0 ? dt : 1.0) || (it == 0 && t_out == 0)) {
      t_out += p.dt_out;
      qgw_set_next_output(ctx, t_out);
      if ((st = qgw_invert_pv(ctx))) return fail("invert_pv", st, ctx);
      if ((st = qgw_download(ctx, QGW_FIELD_PSI, psi)) || (st = qgw_download(ctx, QGW_FIELD_Q, q)))
        return fail("download", st, ctx);
      if ((st = qgwh_nc_write(nc, t, psi, q))) return fail("write vars.nc", st, ctx);
      double E, Z;
      if ((st = qgw_diagnostics(ctx, NULL, NULL, &E, &Z))) return fail("diagnostics", st, ctx);
      fprintf(stdout, "output t = %e  it = %d  E = %.12e  Z = %.12e\n", t, it, E, Z);
      if (!isfinite(E) || !isfinite(Z)) return fail("non-finite state", QGW_ERR_NUMERIC, ctx);
    }
    if ((st = qgw_timestep(ctx))) return fail("timestep", st, ctx);
    qgw_get_time(ctx, &t, &dt, &dt0, &dt1, &it);
    if (p.nsteps > 0 && it >= p.nsteps) break;
  }
  if ((st = qgw_invert_pv(ctx))) return fail("invert_pv", st, ctx);
  double E, Z;
  qgw_diagnostics(ctx, NULL, NULL, &E, &Z);
  fprintf(stdout, "final t = %e  it = %d  E = %.12e  Z = %.12e\n", t, it, E, Z);
  if ((st = qgwh_nc_close(nc))) return fail("close vars.nc", st, ctx);
  qgw_destroy(ctx);
  free(q);
  free(psi);
  free(cl2m);
  free(cm2l);
  free(iRd2);
  return 0;
}
