/* qgw_nc.c -- netCDF classic-format writer and reader (see include/qgw_nc.h).
 *
 * Layout (the published netCDF classic format specification):
 *   header = 'C' 'D' 'F' version, numrecs, dim_list, gatt_list, var_list
 *   dim  = name, length (0 for the record dimension)
 *   var  = name, ndims, dimids, vatt_list, nc_type, vsize, begin
 *   name = length, characters padded with zeros to 4 bytes
 * All integers and values are big-endian.  Fixed-size variables are stored
 * whole at their begin offsets; records follow, each holding one slab (vsize
 * bytes) of every record variable in definition order.  vsize is rounded up
 * to a multiple of 4, except when a file has a single record variable.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "qgw_host.h"
#include "qgw_nc.h"

enum { NC_DIMENSION = 10, NC_VARIABLE = 11, NC_ATTRIBUTE = 12 };
enum { NC_BYTE = 1, NC_CHAR = 2, NC_SHORT = 3, NC_INT = 4, NC_FLOAT = 5, NC_DOUBLE = 6 };

static const int type_size[7] = {0, 1, 1, 2, 4, 4, 8};

/* ---- big-endian encoding ------------------------------------------------ */

typedef struct {
  unsigned char *p;
  size_t n, cap;
  int err;
} wbuf;

static void wb_bytes(wbuf *b, const void *src, size_t n) {
  if (b->err) return;
  if (b->n + n > b->cap) {
    size_t c = b->cap ? b->cap * 2 : 1024;
    while (c < b->n + n) c *= 2;
    unsigned char *q = (unsigned char *)realloc(b->p, c);
    if (!q) {
      b->err = 1;
      return;
    }
    b->p = q;
    b->cap = c;
  }
  memcpy(b->p + b->n, src, n);
  b->n += n;
}

static void wb_u32(wbuf *b, uint32_t v) {
  unsigned char c[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                        (unsigned char)v};
  wb_bytes(b, c, 4);
}

static void wb_name(wbuf *b, const char *s) {
  uint32_t n = (uint32_t)strlen(s);
  wb_u32(b, n);
  wb_bytes(b, s, n);
  static const unsigned char z[4] = {0, 0, 0, 0};
  wb_bytes(b, z, (4 - n % 4) % 4);
}

static void put_be32(unsigned char *c, uint32_t v) {
  c[0] = (unsigned char)(v >> 24);
  c[1] = (unsigned char)(v >> 16);
  c[2] = (unsigned char)(v >> 8);
  c[3] = (unsigned char)v;
}

static uint32_t get_be32(const unsigned char *c) {
  return ((uint32_t)c[0] << 24) | ((uint32_t)c[1] << 16) | ((uint32_t)c[2] << 8) | (uint32_t)c[3];
}

static void put_f32(unsigned char *c, float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  put_be32(c, u);
}

/* ---- output geometry ------------------------------------------------------ */

typedef struct {
  int nxp2, nyp2, nl;
  int ibc;           /* 1: periodic (the ring is not written) */
  int nxo, nyo;      /* x and y extents in the file */
  double Delta;
} geom;

static int geom_of(const qgw_params *p, geom *g) {
  qgw_grid gr;
  int st = qgwh_grid(p, &gr, NULL, NULL, NULL, NULL);
  if (st) return st;
  g->nxp2 = gr.nxp2;
  g->nyp2 = gr.nyp2;
  g->nl = p->nl;
  g->ibc = (p->bc_fac == -1.0) ? 1 : 0; /* src/netcdf_io.h:42-47 */
  g->nxo = gr.nxp2 - 2 * g->ibc;
  g->nyo = gr.nyp2 - 2 * g->ibc;
  g->Delta = gr.Delta;
  return QGW_OK;
}

/* ---- writer -------------------------------------------------------------- */

struct qgw_nc {
  FILE *f;
  geom g;
  int nrec;
  long rec_start, recsize, slab;
  unsigned char *tmp; /* one field slab, encoded */
};

qgw_nc *qgwh_nc_create(const char *path, const qgw_params *p, int *status) {
  int dummy;
  if (!status) status = &dummy;
  *status = QGW_ERR_ARG;
  if (!path || !p) return NULL;
  geom g;
  if (geom_of(p, &g)) return NULL;
  const long slab = 4L * g.nl * g.nyo * g.nxo; /* vsize of psi and q */
  const long fx = 4L * g.nxo, fy = 4L * g.nyo, fz = 4L * g.nl;
  /* header with zero begins, then patched */
  wbuf h = {0};
  wb_bytes(&h, "CDF\001", 4);
  wb_u32(&h, 0); /* numrecs */
  wb_u32(&h, NC_DIMENSION);
  wb_u32(&h, 4);
  wb_name(&h, "t");
  wb_u32(&h, 0);
  wb_name(&h, "level");
  wb_u32(&h, (uint32_t)g.nl);
  wb_name(&h, "y");
  wb_u32(&h, (uint32_t)g.nyo);
  wb_name(&h, "x");
  wb_u32(&h, (uint32_t)g.nxo);
  wb_u32(&h, 0); /* no global attributes */
  wb_u32(&h, 0);
  wb_u32(&h, NC_VARIABLE);
  wb_u32(&h, 6);
  struct {
    const char *name;
    int nd;
    int dims[4];
    long vsize;
  } vars[6] = {{"t", 1, {0}, 4},          {"y", 1, {2}, fy},          {"x", 1, {3}, fx},
               {"level", 1, {1}, fz},     {"psi", 4, {0, 1, 2, 3}, slab}, {"q", 4, {0, 1, 2, 3}, slab}};
  size_t begin_at[6];
  for (int v = 0; v < 6; v++) {
    wb_name(&h, vars[v].name);
    wb_u32(&h, (uint32_t)vars[v].nd);
    for (int d = 0; d < vars[v].nd; d++) wb_u32(&h, (uint32_t)vars[v].dims[d]);
    wb_u32(&h, 0); /* no variable attributes */
    wb_u32(&h, 0);
    wb_u32(&h, NC_FLOAT);
    wb_u32(&h, (uint32_t)vars[v].vsize);
    begin_at[v] = h.n;
    wb_u32(&h, 0);
  }
  if (h.err) {
    free(h.p);
    *status = QGW_ERR_NOMEM;
    return NULL;
  }
  /* data offsets: fixed variables y, x, level after the header, then records */
  const long off_y = (long)h.n, off_x = off_y + fy, off_z = off_x + fx, rec_start = off_z + fz;
  const long recsize = 4 + 2 * slab;
  /* CDF-1 stores begin offsets in 32 bits; records are located by 64-bit
   * arithmetic (begin + rec * recsize), so the file itself may grow past 2 GiB */
  if (rec_start + 4 + slab > 0x7fffffffL) {
    free(h.p);
    return NULL;
  }
  const long begins[6] = {rec_start, off_y, off_x, off_z, rec_start + 4, rec_start + 4 + slab};
  for (int v = 0; v < 6; v++) put_be32(h.p + begin_at[v], (uint32_t)begins[v]);
  qgw_nc *nc = (qgw_nc *)calloc(1, sizeof(qgw_nc));
  unsigned char *coords = (unsigned char *)malloc((size_t)(fx + fy + fz));
  if (!nc || !coords) {
    free(h.p);
    free(nc);
    free(coords);
    *status = QGW_ERR_NOMEM;
    return NULL;
  }
  nc->tmp = (unsigned char *)malloc((size_t)slab);
  nc->f = fopen(path, "wb+");
  if (!nc->tmp || !nc->f) {
    *status = nc->f ? QGW_ERR_NOMEM : QGW_ERR_IO;
    if (nc->f) fclose(nc->f);
    free(nc->tmp);
    free(nc);
    free(h.p);
    free(coords);
    return NULL;
  }
  /* coordinates (src/netcdf_io.h:134-154): y, x = i*Delta, level = i */
  unsigned char *c = coords;
  for (int j = 0; j < g.nyo; j++, c += 4) put_f32(c, (float)(j * g.Delta));
  for (int i = 0; i < g.nxo; i++, c += 4) put_f32(c, (float)(i * g.Delta));
  for (int k = 0; k < g.nl; k++, c += 4) put_f32(c, (float)k);
  int ok = fwrite(h.p, 1, h.n, nc->f) == h.n && fwrite(coords, 1, (size_t)(fx + fy + fz), nc->f) == (size_t)(fx + fy + fz) &&
           fflush(nc->f) == 0;
  free(h.p);
  free(coords);
  if (!ok) {
    fclose(nc->f);
    free(nc->tmp);
    free(nc);
    *status = QGW_ERR_IO;
    return NULL;
  }
  nc->g = g;
  nc->nrec = 0;
  nc->rec_start = rec_start;
  nc->recsize = recsize;
  nc->slab = slab;
  *status = QGW_OK;
  return nc;
}

/* One storage field as the file's float slab (src/netcdf_io.h:231-241). */
static void encode_field(const qgw_nc *nc, const double *a) {
  const geom *g = &nc->g;
  unsigned char *c = nc->tmp;
  for (int k = 0; k < g->nl; k++)
    for (int j = g->ibc; j < g->nyp2 - g->ibc; j++)
      for (int i = g->ibc; i < g->nxp2 - g->ibc; i++, c += 4)
        put_f32(c, (float)a[(size_t)k * g->nxp2 * g->nyp2 + (size_t)j * g->nxp2 + i]);
}

int qgwh_nc_write(qgw_nc *nc, double t, const double *psi, const double *q) {
  if (!nc || !psi || !q) return QGW_ERR_ARG;
  const long off = nc->rec_start + (long)nc->nrec * nc->recsize;
  unsigned char tb[4];
  put_f32(tb, (float)t);
  if (fseek(nc->f, off, SEEK_SET) || fwrite(tb, 1, 4, nc->f) != 4) return QGW_ERR_IO;
  encode_field(nc, psi);
  if (fwrite(nc->tmp, 1, (size_t)nc->slab, nc->f) != (size_t)nc->slab) return QGW_ERR_IO;
  encode_field(nc, q);
  if (fwrite(nc->tmp, 1, (size_t)nc->slab, nc->f) != (size_t)nc->slab) return QGW_ERR_IO;
  nc->nrec++;
  unsigned char nr[4];
  put_be32(nr, (uint32_t)nc->nrec);
  if (fseek(nc->f, 4, SEEK_SET) || fwrite(nr, 1, 4, nc->f) != 4 || fflush(nc->f)) return QGW_ERR_IO;
  return QGW_OK;
}

int qgwh_nc_records(qgw_nc *nc) { return nc ? nc->nrec : QGW_ERR_ARG; }

int qgwh_nc_close(qgw_nc *nc) {
  if (!nc) return QGW_ERR_ARG;
  int st = fclose(nc->f) ? QGW_ERR_IO : QGW_OK;
  free(nc->tmp);
  free(nc);
  return st;
}

/* ---- reader -------------------------------------------------------------- */

typedef struct {
  const unsigned char *p;
  size_t n, at;
  int err;
} rbuf;

static uint32_t rb_u32(rbuf *b) {
  if (b->at + 4 > b->n) {
    b->err = 1;
    return 0;
  }
  uint32_t v = get_be32(b->p + b->at);
  b->at += 4;
  return v;
}

static uint64_t rb_off(rbuf *b, int v64) {
  if (!v64) return rb_u32(b);
  uint64_t hi = rb_u32(b);
  return (hi << 32) | rb_u32(b);
}

/* name: returns a pointer into the buffer and its length */
static const char *rb_name(rbuf *b, uint32_t *len) {
  *len = rb_u32(b);
  if (b->err || b->at + *len > b->n) {
    b->err = 1;
    return NULL;
  }
  const char *s = (const char *)b->p + b->at;
  b->at += *len + (4 - *len % 4) % 4;
  return s;
}

static void rb_skip_atts(rbuf *b) {
  uint32_t tag = rb_u32(b), n = rb_u32(b);
  if (b->err || tag == 0) return;
  if (tag != NC_ATTRIBUTE) {
    b->err = 1;
    return;
  }
  for (uint32_t a = 0; a < n && !b->err; a++) {
    uint32_t len;
    rb_name(b, &len);
    uint32_t ty = rb_u32(b), ne = rb_u32(b);
    if (ty < 1 || ty > 6) {
      b->err = 1;
      return;
    }
    size_t bytes = (size_t)ne * type_size[ty];
    b->at += bytes + (4 - bytes % 4) % 4;
    if (b->at > b->n) b->err = 1;
  }
}

typedef struct {
  char name[64];
  int nd;
  uint32_t dims[8];
  int type;
  uint64_t vsize, begin;
  int is_rec;
} ncvar;

/* A file opened for reading: its header parsed. */
typedef struct {
  FILE *f;
  uint32_t numrecs;
  uint32_t dimlen[64];
  ncvar *vars;
  uint32_t nvar;
  uint64_t recsize;
} ncfile;

static void nc_release(ncfile *F) {
  if (F->f) fclose(F->f);
  free(F->vars);
  memset(F, 0, sizeof *F);
}

/* Parses the header (read whole; up to 1 MiB) of a CDF-1 or CDF-2 file. */
static int nc_open_read(const char *path, ncfile *F) {
  memset(F, 0, sizeof *F);
  F->f = fopen(path, "rb");
  if (!F->f) return QGW_ERR_IO;
  const size_t hcap = 1 << 20;
  unsigned char *hb = (unsigned char *)malloc(hcap);
  if (!hb) {
    nc_release(F);
    return QGW_ERR_NOMEM;
  }
  const size_t hn = fread(hb, 1, hcap, F->f);
  int st = QGW_OK;
  if (hn < 8 || memcmp(hb, "CDF", 3) || (hb[3] != 1 && hb[3] != 2)) {
    st = QGW_ERR_IO;
    goto out;
  }
  {
    const int v64 = (hb[3] == 2);
    rbuf b = {hb, hn, 4, 0};
    F->numrecs = rb_u32(&b);
    int recdim = -1;
    uint32_t tag = rb_u32(&b), ndim = rb_u32(&b);
    if (tag != 0 && (tag != NC_DIMENSION || ndim > 64)) b.err = 1;
    for (uint32_t d = 0; d < ndim && !b.err && tag; d++) {
      uint32_t len;
      rb_name(&b, &len);
      F->dimlen[d] = rb_u32(&b);
      if (F->dimlen[d] == 0) recdim = (int)d;
    }
    rb_skip_atts(&b);
    tag = rb_u32(&b);
    const uint32_t nvar = rb_u32(&b);
    if ((tag != 0 && tag != NC_VARIABLE) || nvar > 4096) b.err = 1;
    if (b.err) {
      st = QGW_ERR_IO;
      goto out;
    }
    F->vars = (ncvar *)calloc(nvar ? nvar : 1, sizeof(ncvar));
    if (!F->vars) {
      st = QGW_ERR_NOMEM;
      goto out;
    }
    F->nvar = nvar;
    int nrecvars = 0;
    for (uint32_t v = 0; v < nvar && !b.err; v++) {
      ncvar *V = &F->vars[v];
      uint32_t len;
      const char *nm = rb_name(&b, &len);
      if (b.err) break;
      const size_t cl = len < 63 ? len : 63;
      memcpy(V->name, nm, cl);
      V->name[cl] = 0;
      const uint32_t nd = rb_u32(&b);
      if (nd > 8) {
        b.err = 1;
        break;
      }
      V->nd = (int)nd;
      for (uint32_t d = 0; d < nd; d++) {
        V->dims[d] = rb_u32(&b);
        if (V->dims[d] >= ndim) b.err = 1;
      }
      rb_skip_atts(&b);
      V->type = (int)rb_u32(&b);
      V->vsize = rb_u32(&b);
      V->begin = rb_off(&b, v64);
      if (V->type < 1 || V->type > 6) b.err = 1;
      V->is_rec = (nd > 0 && (int)V->dims[0] == recdim);
      if (V->is_rec) {
        F->recsize += V->vsize;
        nrecvars++;
      }
    }
    if (b.err) {
      st = QGW_ERR_IO;
      goto out;
    }
    /* one record variable: its slab is not padded */
    if (nrecvars == 1)
      for (uint32_t v = 0; v < nvar; v++)
        if (F->vars[v].is_rec) {
          uint64_t e = 1;
          for (int d = 1; d < F->vars[v].nd; d++) e *= F->dimlen[F->vars[v].dims[d]];
          F->recsize = e * (uint64_t)type_size[F->vars[v].type];
        }
  }
out:
  free(hb);
  if (st) nc_release(F);
  return st;
}

/* Variable `name` (layers levels of the output grid) at record rec, or its
 * single slab, into storage arrays (src/netcdf_io.h:330-366).  Returns
 * QGW_OK with *got = 0 when the file has no such variable. */
static int nc_get(ncfile *F, const geom *g, const char *name, int layers, int rec, double *dst, int *got) {
  *got = 0;
  for (uint32_t v = 0; v < F->nvar; v++) {
    const ncvar *V = &F->vars[v];
    if (strcmp(V->name, name)) continue;
    if (V->type != NC_FLOAT && V->type != NC_DOUBLE) return QGW_ERR_ARG;
    const size_t per = (size_t)g->nxo * g->nyo;
    uint64_t cnt = 1;
    for (int d = V->is_rec ? 1 : 0; d < V->nd; d++) cnt *= F->dimlen[V->dims[d]];
    /* the count the reference asks for: layers x the output extents */
    if (cnt != (uint64_t)layers * per || (V->is_rec && (uint32_t)rec >= F->numrecs)) return QGW_ERR_ARG;
    const uint64_t off = V->begin + (V->is_rec ? (uint64_t)rec * F->recsize : 0);
    const int ts = type_size[V->type];
    unsigned char *raw = (unsigned char *)malloc((size_t)cnt * ts);
    if (!raw) return QGW_ERR_NOMEM;
    if (fseek(F->f, (long)off, SEEK_SET) || fread(raw, (size_t)ts, (size_t)cnt, F->f) != cnt) {
      free(raw);
      return QGW_ERR_IO;
    }
    for (int k = 0; k < layers; k++)
      for (int j = g->ibc; j < g->nyp2 - g->ibc; j++)
        for (int i = g->ibc; i < g->nxp2 - g->ibc; i++) {
          const size_t e = (size_t)k * per + (size_t)(j - g->ibc) * g->nxo + (i - g->ibc);
          const unsigned char *c = raw + e * ts;
          double x;
          if (ts == 4) {
            const uint32_t u = get_be32(c);
            float xf;
            memcpy(&xf, &u, 4);
            x = xf;
          } else {
            const uint64_t u = ((uint64_t)get_be32(c) << 32) | get_be32(c + 4);
            double xd;
            memcpy(&xd, &u, 8);
            x = (float)xd; /* nc_get_vara_float converts to float */
          }
          dst[(size_t)k * g->nxp2 * g->nyp2 + (size_t)j * g->nxp2 + i] = x;
        }
    free(raw);
    *got = 1;
    return QGW_OK;
  }
  return QGW_OK;
}

int qgwh_nc_read(const char *path, const qgw_params *p, int rec, double *q, double *topo, int *found) {
  if (found) *found = 0;
  if (!path || !p || rec < 0) return QGW_ERR_ARG;
  geom g;
  if (geom_of(p, &g)) return QGW_ERR_ARG;
  ncfile F;
  int st = nc_open_read(path, &F);
  if (st) return st;
  int got = 0, bits = 0;
  /* read_nc's list of src/qg.c:161-163: q (3d_field), topo (2d_field) */
  if (q && !(st = nc_get(&F, &g, "q", g.nl, rec, q, &got)) && got) bits |= 1;
  if (!st && topo && !(st = nc_get(&F, &g, "topo", 1, rec, topo, &got)) && got) bits |= 2;
  nc_release(&F);
  if (found) *found = bits;
  return st;
}

int qgwh_nc_read_var(const char *path, const qgw_params *p, const char *name, int layers, int rec, double *out,
                     int *found) {
  if (found) *found = 0;
  if (!path || !p || !name || !out || layers < 1 || rec < 0) return QGW_ERR_ARG;
  geom g;
  if (geom_of(p, &g)) return QGW_ERR_ARG;
  ncfile F;
  int st = nc_open_read(path, &F);
  if (st) return st;
  int got = 0;
  st = nc_get(&F, &g, name, layers, rec, out, &got);
  nc_release(&F);
  if (found) *found = got;
  return st;
}
