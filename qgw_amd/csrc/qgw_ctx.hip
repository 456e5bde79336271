// qgw_ctx.hip -- the C-ABI of the MI355X timestep hot path (libqgw_hip.so).
//
// Host side of every entry point declared in include/qgw_hip.h: context set-up
// (init_domain/init_vars/init_timestep/init_elliptic, src/domain.h:64,178,
// src/timestep.h:56, src/elliptic.h:29), upload/download in the reference idx
// layout, the per-call hot-path boundaries (invert_pv, rhs, adjust_timestep,
// timestep: src/elliptic.h:76, src/dynamics.h:47, src/timestep.h:69,114) and
// the host logic of the variable step (src/timestep.h:69-111, 126-133).
// The kernels are in kernels.hpp / fft.hpp and are instantiated per precision
// in qgw_launch_f32.hip and qgw_launch_f64.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <mutex>
#include <utility>
#include <vector>

#include <rocblas/rocblas.h>

#include "qgw_hip.h"
#include "qgw_host.h"
#include "qgw_nc.h"
#include "launch.hpp"

namespace {

constexpr double kPi = 3.141592653589793;  // src/qg.c:44
constexpr double kHuge = 1e30;             // src/qg.c:43
constexpr double kTeps = 1e-9;             // src/timestep.h:16

bool supported_length(int L) {
  int r = L;
  while (r % 2 == 0) r /= 2;
  while (r % 3 == 0) r /= 3;
  while (r % 5 == 0) r /= 5;
  while (r % 7 == 0) r /= 7;
  if (r != 1) return false;
  return qgw::launch_has_length(L);
}

}  // namespace

// One group of ensemble members of equal shape and step scalars (a plain run
// is a group of one).  Members are stacked along the layer axis: member b's
// layer l is layer b*nl + l of every field, and each member has its own
// modes, wind row and topography.
struct qgw_ctx {
  qgw_params p;                 // member 0; the scalars every member shares
  int nb;                       // members
  std::vector<qgw_params> mp;   // per-member parameters
  std::vector<char> mset;       // per-member: modes set
  bool own_stream;
  int prec;
  size_t esz;
  qgw::Geo g;
  qgw_grid grid;
  int Lx, Ly;
  hipStream_t stream;
  void *q[2];
  int qcur;
  void *psi, *fa[3], *comp, *topo, *wind, *what, *whatT, *cl2m, *cm2l, *twx, *twy, *twxh, *twyh;
  double *iRd2_d, *La2, *Kb2, *diag_part, *diag_out, *stage;
  unsigned long long *cfl_max;
  unsigned *nonfinite;  // raised by P4 when a new q (or a tendency) is NaN or Inf; sticky until q is uploaded
  size_t byte_limit;    // device bytes the context may allocate (QGW_MAX_DEVICE_BYTES; 0: no limit)
  int nblk_diag;
  int p4_tile;
  int t_kernel;  // projection kernel choice (QGW_TUNE_T_KERNEL)
  int p4_waves;  // streaming P4 wavefronts per SIMD (QGW_TUNE_P4_WAVES)
  int dst_half;  // sine transforms (QGW_TUNE_DST: 0, 1 half-length, 2 pipelined rows); -1: half lengths not provided
  int cus;       // compute units of the device
  double dk;     // y wavenumber step of the bounded transforms, pi / Lx (src/domain.h:170-173)
  int fft_e;     // elements per thread of the half-length FFTs of the row passes (QGW_TUNE_FFT_E)
  int fft_e_col; // and of the column pass (QGW_TUNE_FFT_E_COL)
  int fuse;      // QGW_TUNE_FUSE: 1 four-pass inversion, 0 six passes; -1: four passes not available
  int p2m_g;     // QGW_TUNE_P2M_PAIRS: column pairs per block of the four-pass column pass (0: most)
  int psi_nt;    // QGW_TUNE_PSI_NT: non-temporal psi stores of the half-length inverse row passes
  int p4_rev;    // QGW_TUNE_P4_REV: streaming P4 walks of odd row chunks run up their chunk
  int p4_order;  // QGW_TUNE_P4_ORDER: streaming P4 walks dealt layers (1) or row chunks (0) first
  int per;       // periodic domain (bc_fac = -1)
  // dense sine transforms (bounded grids whose transform lengths the FFT
  // kernels do not cover): S matrices n x n and m x m, a whatT-shaped temporary
  int dst_gemm;
  rocblas_handle blas;
  void *Sx, *Sy, *wtmp;
  bool modes_set;
  double t, dt, dt0, dt1, DT_max, t_out;
  int it;
  double nu_hyper;
  // 4-D forcing (init_4d_forcing/calc_4d_forcing, src/forcing.h:226-289):
  // nrec4 records spaced dt_forc in file forc_path; slots forc4[0..1] hold
  // records irec4[0..1] (-1: empty)
  int nrec4;
  void *forc4[2];
  int irec4[2];
  std::string forc_path;
  // stochastic ring forcing (init_stoch_forc/calc_forc, src/forcing.h:67-208):
  // nst forced points (coefficients drawn on the host from rng each step,
  // synthesised and added on the device); nothing is added when nst <= 1
  int nst;
  double st_env;
  qgwh_rng *rng;
  qgw::StochArgs sa;
  double *st_host[2];  // pinned staging of the coefficients, alternating
  hipEvent_t st_ev[2];
  int st_slot;
  size_t bytes;
  // profiling
  bool prof;
  struct Rec {
    int pass;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> evpool;
  std::vector<void *> allocs;  // every device allocation (field pointers may be offset or rotated)
  double prof_ms[QGW_NPASS];
  int prof_cnt[QGW_NPASS];
  double *scratch;  // this device's copy of the file-scope `scratch` array (qgw_diagnostics_device)
  char err[256];
};

// ---- scratch of the allocation-free entry point --------------------------
// Partial sums of qgw_diagnostics_device: 2 doubles per block and layer,
// kNoAllocBlocks blocks per layer, at most kNoAllocLayers layers: 256 KiB,
// held by every GPU the library runs on until the process ends.  Host code
// never uses the array's name as an address: each device's copy is resolved
// once, with that device current, by hipGetSymbolAddress (scratch_on_device).
constexpr int kNoAllocBlocks = 256;
constexpr int kNoAllocLayers = 64;
static __device__ __align__(256) double scratch[2 * kNoAllocBlocks * kNoAllocLayers];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on device `dev` (the current device),
// resolved on the first call for that device and kept.
hipError_t scratch_on_device(int dev, double **out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (double *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int fail_hip(qgw_ctx *c, hipError_t e, const char *what) {
  if (c) snprintf(c->err, sizeof c->err, "%s: %s", what, hipGetErrorString(e));
  return QGW_ERR_HIP;
}

#define HIPCHK(c, call)                                  \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return fail_hip((c), _e, #call); \
  } while (0)

int dalloc(qgw_ctx *c, void **ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) bytes = 16;
  if (c->byte_limit && c->bytes + bytes > c->byte_limit) {
    snprintf(c->err, sizeof c->err, "device memory limit QGW_MAX_DEVICE_BYTES=%zu reached (%zu + %zu bytes)",
             c->byte_limit, c->bytes, bytes);
    return QGW_ERR_NOMEM;
  }
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    fail_hip(c, e, "hipMalloc");
    return QGW_ERR_NOMEM;
  }
  c->allocs.push_back(*ptr);
  e = hipMemset(*ptr, 0, bytes);
  if (e != hipSuccess) return fail_hip(c, e, "hipMemset");
  c->bytes += bytes;
  return QGW_OK;
}

// A storage field (rows with the one-point wall ring): the pointer is offset
// so that interior element i = 1 of every row starts a 256-byte line (the
// pitch is a multiple of 256 bytes), which keeps the row passes and the
// stencil tiles on whole cache lines.
int dalloc_field(qgw_ctx *c, void **ptr, size_t bytes) {
  void *raw = nullptr;
  int st = dalloc(c, &raw, bytes + 256);
  if (st) return st;
  *ptr = (char *)raw + 256 - c->esz;
  return QGW_OK;
}

hipEvent_t ev_get(qgw_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int prof_flush(qgw_ctx *c) {
  if (c->recs.empty()) return QGW_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &r : c->recs) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      c->prof_ms[r.pass] += ms;
      c->prof_cnt[r.pass] += 1;
    }
    c->evpool.push_back(r.a);
    c->evpool.push_back(r.b);
  }
  c->recs.clear();
  return QGW_OK;
}

// Bracket one launch with events when profiling.
struct PassScope {
  qgw_ctx *c;
  int pass;
  hipEvent_t a = nullptr;
  PassScope(qgw_ctx *cc, int ps) : c(cc), pass(ps) {
    if (c->prof) {
      if (c->recs.size() > 4000) prof_flush(c);
      a = ev_get(c);
      if (a) hipEventRecord(a, c->stream);
    }
  }
  ~PassScope() {
    if (c->prof && a) {
      hipEvent_t b = ev_get(c);
      if (b) {
        hipEventRecord(b, c->stream);
        c->recs.push_back({pass, a, b});
      }
    }
  }
};

int check_launch(qgw_ctx *c, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_hip(c, e, what);
  return QGW_OK;
}

template <typename T>
int make_twiddles(qgw_ctx *c, int L, void **out) {
  std::vector<T> h(2 * (size_t)L);
  for (int k = 0; k < L; k++) {
    double a = -2.0 * kPi * (double)k / (double)L;
    h[2 * k] = (T)cos(a);
    h[2 * k + 1] = (T)sin(a);
  }
  int st = dalloc(c, out, sizeof(T) * 2 * L);
  if (st) return st;
  HIPCHK(c, hipMemcpy(*out, h.data(), sizeof(T) * 2 * L, hipMemcpyHostToDevice));
  return QGW_OK;
}

template <typename T>
int upload_vec(qgw_ctx *c, void *dst, const double *src, size_t cnt) {
  std::vector<T> h(cnt);
  for (size_t i = 0; i < cnt; i++) h[i] = (T)src[i];
  HIPCHK(c, hipMemcpy(dst, h.data(), sizeof(T) * cnt, hipMemcpyHostToDevice));
  return QGW_OK;
}

int upload_typed(qgw_ctx *c, void *dst, const double *src, size_t cnt) {
  return c->prec == QGW_FP64 ? upload_vec<double>(c, dst, src, cnt) : upload_vec<float>(c, dst, src, cnt);
}

// Storage pitch: n+2 rounded up to 64 elements, and 64 more when that is a
// multiple of 1024 elements, so that the rows a tile touches at once do not
// all start at the same offset modulo a large power of two (HBM channel
// interleave).
int pitch_of(int w) {
  const int p = (w + 63) / 64 * 64;
  return (p % 1024 == 0) ? p + 64 : p;
}

void release(qgw_ctx *c) {
  if (!c) return;
  if (c->blas) rocblas_destroy_handle(c->blas);
  if (c->rng) qgwh_rng_destroy(c->rng);
  for (int i = 0; i < 2; i++) {
    if (c->st_host[i]) hipHostFree(c->st_host[i]);
    if (c->st_ev[i]) hipEventDestroy(c->st_ev[i]);
  }
  for (void *p : c->allocs)
    if (p) hipFree(p);
  for (auto &r : c->recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  for (auto e : c->evpool) hipEventDestroy(e);
  if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
  delete c;
}

// ---- hot-path sequences ---------------------------------------------------

// C (column-major M x N, ldc) = S (M x K, ld M) * B (K x N, ldb), batched over
// `batch` matrices B, C with strides sB, sC (elements): a row-major array of
// rows of K values (pitch ldb) times a symmetric S, i.e. a DST-I of every row.
int dense_gemm(qgw_ctx *c, const void *S, int M, int N, int K, const void *B, int ldb, long sB, void *C, int ldc,
               long sC, int batch) {
  rocblas_status r = rocblas_set_stream(c->blas, c->stream);
  if (r == rocblas_status_success) {
    if (c->prec == QGW_FP64) {
      const double one = 1.0, zero = 0.0;
      r = rocblas_dgemm_strided_batched(c->blas, rocblas_operation_none, rocblas_operation_none, M, N, K, &one,
                                        (const double *)S, M, 0, (const double *)B, ldb, sB, &zero, (double *)C,
                                        ldc, sC, batch);
    } else {
      const float one = 1.f, zero = 0.f;
      r = rocblas_sgemm_strided_batched(c->blas, rocblas_operation_none, rocblas_operation_none, M, N, K, &one,
                                        (const float *)S, M, 0, (const float *)B, ldb, sB, &zero, (float *)C, ldc,
                                        sC, batch);
    }
  }
  if (r != rocblas_status_success) {
    snprintf(c->err, sizeof c->err, "dense sine transform: %s", rocblas_status_to_string(r));
    return QGW_ERR_HIP;
  }
  return QGW_OK;
}

// invert_pv with dense sine transforms: P1 rows (what[l] = q[l] S_x), T1,
// P2 columns (S_y on every whatT row, division, S_y again), T2, P3 rows
// (psi[l] = what[l] S_x) and the wall rings.
int run_invert_dense(qgw_ctx *c, qgw::LaunchArgs &a) {
  const qgw::Geo &g = c->g;
  const size_t e = c->esz;
  char *q = (char *)c->q[c->qcur], *psi = (char *)c->psi;
  int st;
  {
    PassScope ps(c, QGW_PASS_P1);
    st = dense_gemm(c, c->Sx, g.n, g.m, g.n, q + e * (g.P + 1), g.P, g.LS, c->what, g.Pn, g.WS, g.nl);
  }
  if (st) return st;
  {
    PassScope ps(c, QGW_PASS_T1);
    st = qgw::launch_t1(a);
  }
  if (st || (st = check_launch(c, "t1"))) return st;
  {
    PassScope ps(c, QGW_PASS_P2);
    st = dense_gemm(c, c->Sy, g.m, g.n, g.m, c->whatT, g.Pm, g.WTS, c->wtmp, g.Pm, g.WTS, g.nl);
    if (!st) {
      st = qgw::launch_helm_div(a, c->wtmp);
      if (!st) st = check_launch(c, "helm_div");
    }
    if (!st) st = dense_gemm(c, c->Sy, g.m, g.n, g.m, c->wtmp, g.Pm, g.WTS, c->whatT, g.Pm, g.WTS, g.nl);
  }
  if (st) return st;
  {
    PassScope ps(c, QGW_PASS_T2);
    st = qgw::launch_t2(a);
  }
  if (st || (st = check_launch(c, "t2"))) return st;
  {
    PassScope ps(c, QGW_PASS_P3);
    st = dense_gemm(c, c->Sx, g.n, g.m, g.n, c->what, g.Pn, g.WS, psi + e * (g.P + 1), g.P, g.LS, g.nl);
    if (!st) {
      st = qgw::launch_rings(a);
      if (!st) st = check_launch(c, "rings");
    }
  }
  return st;
}

int run_invert(qgw_ctx *c) {
  if (!c->modes_set) {
    snprintf(c->err, sizeof c->err, "vertical modes not set (qgw_set_modes)");
    return QGW_ERR_STATE;
  }
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.Lx = c->Lx;
  a.Ly = c->Ly;
  a.stream = c->stream;
  a.q = c->q[c->qcur];
  a.psi = c->psi;
  a.what = c->what;
  a.whatT = c->whatT;
  a.cl2m = c->cl2m;
  a.cm2l = c->cm2l;
  a.twx = c->twx;
  a.twy = c->twy;
  a.La2 = c->La2;
  a.Kb2 = c->Kb2;
  a.iRd2 = c->iRd2_d;
  a.t_kernel = c->t_kernel;
  a.twxh = c->twxh;
  a.twyh = c->twyh;
  a.dst_half = c->dst_half > 0 ? c->dst_half : 0;
  a.fft_e = c->fft_e;
  a.fft_e_col = c->fft_e_col;
  a.p2m_g = c->p2m_g;
  a.psi_nt = c->psi_nt;
  a.p4_rev = c->p4_rev;
  a.p4_order = c->p4_order;
  a.cus = c->cus;
  a.dk = c->dk;
  a.per = c->per;
  // normalisation of the transform pair (src/elliptic.h:36,51)
  a.norm = c->per ? (double)c->g.n * (double)c->g.m : 4.0 * (double)(c->g.n + 1) * (double)(c->g.m + 1);
  a.ring_c = 2.0 * c->p.bc_fac / (c->grid.Delta * c->grid.Delta);
  if (c->dst_gemm) return run_invert_dense(c, a);
  int st;
  if (c->fuse == 1 && a.dst_half > 0) {
    {
      PassScope ps(c, QGW_PASS_P1);
      st = qgw::launch_p1m(a);
    }
    if (st || (st = check_launch(c, "p1m"))) return st;
    {
      PassScope ps(c, QGW_PASS_P2);
      st = qgw::launch_p2m(a);
    }
    if (st || (st = check_launch(c, "p2m"))) return st;
    {
      PassScope ps(c, QGW_PASS_P3);
      st = qgw::launch_p3m(a);
    }
    if (st || (st = check_launch(c, "p3m"))) return st;
    return QGW_OK;
  }
  {
    PassScope ps(c, QGW_PASS_P1);
    st = qgw::launch_p1(a);
  }
  if (st || (st = check_launch(c, "p1"))) return st;
  {
    PassScope ps(c, QGW_PASS_T1);
    st = qgw::launch_t1(a);
  }
  if (st || (st = check_launch(c, "t1"))) return st;
  {
    PassScope ps(c, QGW_PASS_P2);
    st = qgw::launch_p2(a);
  }
  if (st || (st = check_launch(c, "p2"))) return st;
  {
    PassScope ps(c, QGW_PASS_T2);
    st = qgw::launch_t2(a);
  }
  if (st || (st = check_launch(c, "t2"))) return st;
  {
    PassScope ps(c, QGW_PASS_P3);
    st = qgw::launch_p3(a);
  }
  if (st || (st = check_launch(c, "p3"))) return st;
  return QGW_OK;
}

int run_p4(qgw_ctx *c, const double *cab3, int update) {
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  a.q = c->q[c->qcur];
  a.qn = c->q[c->qcur ^ 1];
  a.psi = c->psi;
  a.fnew = c->fa[0];
  a.f1 = c->fa[1];
  a.f2 = c->fa[2];
  a.comp = (c->prec == QGW_FP32) ? c->comp : nullptr;
  a.topo = c->topo;
  a.wind = c->wind;
  const double D = c->grid.Delta;
  qgw::StepCoef &cf = a.cf;
  cf.c0 = cab3[0];
  cf.c1 = cab3[1];
  cf.c2 = cab3[2];
  cf.beta = c->p.beta;
  cf.nu = c->p.nu;
  cf.nu_hyper = c->nu_hyper;
  cf.inv_d2 = 1.0 / (D * D);
  cf.inv_12d2 = 1.0 / (12. * D * D);
  cf.inv_2d = 1.0 / (2 * D);
  const double dhb = c->p.dh[c->p.nl - 1];
  cf.drag = -c->p.hEkb * c->p.f0 / (2 * dhb);
  cf.topo_fac = c->p.f0 / dhb;
  cf.ring_c = 2.0 * c->p.bc_fac / (D * D);
  cf.update = update;
  cf.nonfinite = c->nonfinite;
  a.nlap = (c->nu_hyper != 0.0) ? (c->p.n_hyper / 2 + 1) : 0;
  a.p4_tile = c->p4_tile;
  a.p4_waves = c->p4_waves;
  a.per = c->per;
  int st;
  {
    PassScope ps(c, QGW_PASS_P4);
    st = qgw::launch_p4(a);
  }
  if (st || (st = check_launch(c, "p4"))) return st;
  if (update) c->qcur ^= 1;
  return QGW_OK;
}

// The P4 non-finite flag: QGW_ERR_NUMERIC once a step has produced a NaN or
// Inf in q (or in the tendency of qgw_rhs).  Waits for the stream.
int check_finite(qgw_ctx *c) {
  unsigned h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->nonfinite, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h) {
    snprintf(c->err, sizeof c->err, "non-finite q (NaN or Inf) produced by a step at or before step %d", c->it);
    return QGW_ERR_NUMERIC;
  }
  return QGW_OK;
}

void ab3_coeffs(double dt, double dt0, double dt1, double *c) {
  // variable-step AB3 (src/timestep.h:131-133, doc/timestep.org:95-97)
  c[0] = dt * (2 * (dt * dt) + 3 * dt * (2 * dt0 + dt1) + 6 * dt0 * (dt0 + dt1)) / (6 * dt0 * (dt0 + dt1));
  c[1] = (dt * dt) * (-dt / 3 - dt0 / 2 - dt1 / 2) / (dt0 * dt1);
  c[2] = (dt * dt) * (dt / 3 + dt0 / 2) / (dt1 * (dt0 + dt1));
}

// adjust_timestep (src/timestep.h:69-111): CFL reduction on the device, the
// doubling / cap / snap-to-output logic on the host.
int adaptive_dt(qgw_ctx *c, double *dt_new) {
  HIPCHK(c, hipMemsetAsync(c->cfl_max, 0, sizeof(unsigned long long), c->stream));
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  a.psi = c->psi;
  a.cfl_max = c->cfl_max;
  int st;
  {
    PassScope ps(c, QGW_PASS_CFL);
    st = qgw::launch_cfl(a);
  }
  if (st || (st = check_launch(c, "cfl"))) return st;
  unsigned long long bits = 0;
  HIPCHK(c, hipMemcpyAsync(&bits, c->cfl_max, sizeof bits, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double dmax;
  memcpy(&dmax, &bits, sizeof dmax);
  const double Delta = c->grid.Delta, cfl = c->p.cfl;
  // min over points of cfl*Delta/u equals cfl*Delta/max(u): division is monotone
  double dt_max = kHuge;
  const double umax = dmax / Delta;
  if (umax != 0.) dt_max = cfl * Delta / umax;
  if (!isfinite(dmax)) {
    snprintf(c->err, sizeof c->err, "non-finite psi in the CFL reduction");
    return QGW_ERR_NUMERIC;
  }
  double dt = c->dt;
  if (dt < dt_max) dt = 2. * dt;
  if (dt > c->DT_max) dt = c->DT_max;
  if (dt > dt_max) dt = dt_max;
  const double t = c->t, t_out = c->t_out;
  if (t_out > t + kTeps) {
    unsigned int n_step = (t_out - t) / dt;
    if (n_step == 0)
      dt = t_out - t;
    else {
      double dt1 = (t_out - t) / n_step;
      if (dt1 > dt * (1. + kTeps))
        dt = (t_out - t) / (n_step + 1);
      else if (dt1 < dt)
        dt = dt1;
    }
  }
  *dt_new = dt;
  return QGW_OK;
}

int ensure_stage(qgw_ctx *c);

// Record rec of the 4-D forcing file into slot `slot` (read_nc of
// "q_forcing_3d", src/forcing.h:262-263).
int load_forcing_record(qgw_ctx *c, int slot, int rec) {
  const size_t cnt = (size_t)(c->g.n + 2) * (c->g.m + 2) * c->p.nl;
  std::vector<double> h(cnt, 0.0);
  int found = 0;
  int st = qgwh_nc_read_var(c->forc_path.c_str(), &c->p, "q_forcing_3d", c->p.nl, rec, h.data(), &found);
  if (st || !found) {
    snprintf(c->err, sizeof c->err, "4-D forcing: %s record %d of %s", st ? "cannot read" : "no q_forcing_3d in",
             rec, c->forc_path.c_str());
    return st ? st : QGW_ERR_IO;
  }
  if ((st = ensure_stage(c))) return st;
  HIPCHK(c, hipMemcpyAsync(c->stage, h.data(), cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  st = qgw::launch_from_f64(a, c->p.nl, c->stage, c->forc4[slot], nullptr);
  if (st || (st = check_launch(c, "from_f64"))) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));  // h is released on return
  c->irec4[slot] = rec;
  return QGW_OK;
}

// The two records around time t (calc_4d_forcing, src/forcing.h:253-270):
// irec1 = floor(t/dt_forc) mod nrec and its successor.  A slot already
// holding a wanted record is reused instead of read again.
int forcing_records_at(qgw_ctx *c, double t) {
  const long long fl = (long long)floor(t / c->p.dt_forc);
  const int r1 = (int)(((fl % c->nrec4) + c->nrec4) % c->nrec4), r2 = (r1 + 1) % c->nrec4;
  if (c->irec4[0] == r1 && c->irec4[1] == r2) return QGW_OK;
  if (c->irec4[1] == r1) {
    std::swap(c->forc4[0], c->forc4[1]);
    std::swap(c->irec4[0], c->irec4[1]);
  }
  int st = QGW_OK;
  if (c->irec4[0] != r1) st = load_forcing_record(c, 0, r1);
  if (!st && c->irec4[1] != r2) st = load_forcing_record(c, 1, r2);
  return st;
}

// The forcing after the AB3 update (src/timestep.h:143-163), in the
// reference's order: the stochastic field of this step's draws times sqrt(dt)
// on layer 0, then (w1 F1 + w2 F2) dt on every layer with the weights of the
// time before the step (src/forcing.h:273-286); dt is the new step.
int run_forcing(qgw_ctx *c, double t_old) {
  const bool stoch = c->nst > 1, f4 = c->nrec4 > 0;
  if (!stoch && !f4) return QGW_OK;
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  a.q = c->q[c->qcur];
  a.comp = (c->prec == QGW_FP32) ? c->comp : nullptr;
  double w1 = 0, w2 = 0;
  if (f4) {
    const double x = t_old / c->p.dt_forc;
    w2 = x - floor(x);
    w1 = 1 - w2;
  }
  int st = QGW_OK;
  {
    PassScope ps(c, QGW_PASS_FORC);
    if (stoch) {
      // this step's coefficients (calc_forc's draws) through an alternating
      // pinned buffer: the copy that last read it must be done
      const int slot = c->st_slot;
      c->st_slot ^= 1;
      HIPCHK(c, hipEventSynchronize(c->st_ev[slot]));
      st = qgwh_stoch_coeffs(c->rng, c->p.sigma_f, c->st_env, c->nst, c->st_host[slot]);
      if (st) return st;
      HIPCHK(c, hipMemcpyAsync((void *)c->sa.coef, c->st_host[slot], sizeof(double) * 2 * c->nst,
                               hipMemcpyHostToDevice, c->stream));
      HIPCHK(c, hipEventRecord(c->st_ev[slot], c->stream));
      st = qgw::launch_stoch(a, c->sa, sqrt(c->dt));
    }
    if (!st && f4) st = qgw::launch_forcing(a, c->forc4[0], c->forc4[1], w1, w2, c->dt);
  }
  if (st || (st = check_launch(c, "forcing"))) return st;
  return QGW_OK;
}

int one_step(qgw_ctx *c) {
  // records of the 4-D forcing first, so that a failed read leaves the state as it was
  const double t_old = c->t;
  if (c->nrec4 > 0) {
    int st = forcing_records_at(c, t_old);
    if (st) return st;
  }
  // rotate the tendency history (src/timestep.h:117-120)
  void *swap = c->fa[2];
  c->fa[2] = c->fa[1];
  c->fa[1] = c->fa[0];
  c->fa[0] = swap;
  int st = run_invert(c);
  if (st) return st;
  c->dt1 = c->dt0;
  c->dt0 = c->dt;
  if (c->p.dt_fixed > 0) {
    c->dt = c->p.dt_fixed;
  } else {
    double nd;
    st = adaptive_dt(c, &nd);
    if (st) return st;
    c->dt = nd;
  }
  double cab[3];
  ab3_coeffs(c->dt, c->dt0, c->dt1, cab);
  st = run_p4(c, cab, 1);
  if (st) return st;
  if ((st = run_forcing(c, t_old))) return st;
  c->t = c->t + c->dt;
  c->it++;
  return QGW_OK;
}

// Members that may share one group: equal grid and the scalars that enter
// the step kernels, and a fixed step (an adaptive step is per run).
bool same_group(const qgw_params &a, const qgw_params &b) {
  if (a.dt_fixed <= 0 || b.dt_fixed <= 0) return false;
  // forced members keep a context of their own (host-fed draws or records)
  if (a.dt_forc_period != 0 || a.k_f != 0 || b.dt_forc_period != 0 || b.k_f != 0) return false;
  return a.NX == b.NX && a.NY == b.NY && a.nl == b.nl && a.Lx == b.Lx && a.n_hyper == b.n_hyper &&
         a.nu_hyper == b.nu_hyper && a.nu == b.nu && a.beta == b.beta && a.f0 == b.f0 && a.hEkb == b.hEkb &&
         a.dh[a.nl - 1] == b.dh[b.nl - 1] && a.bc_fac == b.bc_fac && a.dt_fixed == b.dt_fixed;
}

// Records of the 4-D forcing: dt_forc_period must be a positive multiple of
// dt_forc > 0 (init_4d_forcing, src/forcing.h:238-251); 0 when it is not.
int forcing_records(const qgw_params *p) {
  if (!(p->dt_forc > 0) || !(p->dt_forc_period > 0)) return 0;
  const double r = floor(p->dt_forc_period / p->dt_forc);
  if (r < 1 || r > 1e9 || r * p->dt_forc != p->dt_forc_period) return 0;
  return (int)r;
}

int check_params(const qgw_params *p, qgw_grid *grid, int *Lx, int *Ly, int *dense) {
  if (p->nl < 1 || p->nl > qgw::kMaxLayers) return QGW_ERR_ARG;
  // n_hyper/2 + 1 nested Laplacians, at most 10 (the streaming update kernel's register windows)
  if (p->n_hyper < 0 || (p->n_hyper % 2) != 0 || p->n_hyper > 18) return QGW_ERR_ARG;
  // stochastic forcing: the reference's N x N field (N = NX) must be the grid's
  if (p->k_f != 0 && (p->NX != p->NY || !(p->Lx > 0))) return QGW_ERR_ARG;
  if (p->dt_forc_period != 0 && forcing_records(p) < 1) return QGW_ERR_ARG;
  if (qgwh_grid(p, grid, nullptr, nullptr, nullptr, nullptr) != QGW_OK) return QGW_ERR_ARG;
  // transform lengths: the odd extension 2(n+1) (bounded) or the period N (periodic)
  const bool per = (p->bc_fac == -1.0);
  *Lx = per ? grid->n : 2 * (grid->n + 1);
  *Ly = per ? grid->m : 2 * (grid->m + 1);
  *dense = 0;
  if (!supported_length(*Lx) || !supported_length(*Ly)) {
    // bounded grids of any size: dense sine transforms; the periodic branch needs the FFT lengths
    if (per || grid->n < 1 || grid->m < 1 || grid->n > 65535) return QGW_ERR_ARG;
    *dense = 1;
  }
  return QGW_OK;
}

void *member_field(qgw_ctx *c, int member, int field, int *layers, void **resid) {
  *resid = nullptr;
  *layers = c->p.nl;
  const size_t lay = c->esz * (size_t)c->g.LS;  // bytes per layer
  const size_t off = lay * (size_t)c->p.nl * member;
  switch (field) {
    case QGW_FIELD_Q:
      // the compensation has 16-bit elements at the same element offsets
      if (c->prec == QGW_FP32) *resid = (char *)c->comp + off / c->esz * sizeof(short);
      return (char *)c->q[c->qcur] + off;
    case QGW_FIELD_PSI: return (char *)c->psi + off;
    case QGW_FIELD_TOPO: *layers = 1; return (char *)c->topo + lay * member;
    case QGW_FIELD_F0: return (char *)c->fa[0] + off;
    case QGW_FIELD_F1: return (char *)c->fa[1] + off;
    case QGW_FIELD_F2: return (char *)c->fa[2] + off;
  }
  return nullptr;
}

int ensure_stage(qgw_ctx *c) {
  if (c->stage) return QGW_OK;
  size_t b = sizeof(double) * (size_t)(c->g.n + 2) * (c->g.m + 2) * c->p.nl;
  hipError_t e = hipMalloc((void **)&c->stage, b);
  if (e != hipSuccess) return fail_hip(c, e, "hipMalloc(stage)");
  c->allocs.push_back(c->stage);
  c->bytes += b;
  return QGW_OK;
}

int member_upload(qgw_ctx *c, int member, int field, const double *host) {
  if (!c || !host || member < 0 || member >= c->nb) return QGW_ERR_ARG;
  int layers;
  void *resid;
  void *dst = member_field(c, member, field, &layers, &resid);
  if (!dst) return QGW_ERR_ARG;
  int st = ensure_stage(c);
  if (st) return st;
  size_t cnt = (size_t)(c->g.n + 2) * (c->g.m + 2) * layers;
  HIPCHK(c, hipMemcpyAsync(c->stage, host, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  st = qgw::launch_from_f64(a, layers, c->stage, dst, resid);
  if (st || (st = check_launch(c, "from_f64"))) return st;
  // a new state: the non-finite flag of the previous one no longer applies
  if (field == QGW_FIELD_Q && c->nb == 1) HIPCHK(c, hipMemsetAsync(c->nonfinite, 0, sizeof(unsigned), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QGW_OK;
}

int member_download(qgw_ctx *c, int member, int field, double *host) {
  if (!c || !host || member < 0 || member >= c->nb) return QGW_ERR_ARG;
  int layers;
  void *resid;
  void *src = member_field(c, member, field, &layers, &resid);
  if (!src) return QGW_ERR_ARG;
  int st = ensure_stage(c);
  if (st) return st;
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  st = qgw::launch_to_f64(a, layers, src, resid, c->stage);
  if (st || (st = check_launch(c, "to_f64"))) return st;
  size_t cnt = (size_t)(c->g.n + 2) * (c->g.m + 2) * layers;
  HIPCHK(c, hipMemcpyAsync(host, c->stage, cnt * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QGW_OK;
}

int member_set_modes(qgw_ctx *c, int member, const double *cl2m, const double *cm2l, const double *iRd2) {
  if (!c || !cl2m || !cm2l || !iRd2 || member < 0 || member >= c->nb) return QGW_ERR_ARG;
  const int nl = c->p.nl;
  for (int i = 0; i < nl * nl; i++)
    if (!isfinite(cl2m[i]) || !isfinite(cm2l[i])) return QGW_ERR_ARG;
  for (int i = 0; i < nl; i++)
    if (!isfinite(iRd2[i])) return QGW_ERR_ARG;
  const size_t mo = c->esz * (size_t)nl * nl * member;
  int st = upload_typed(c, (char *)c->cl2m + mo, cl2m, (size_t)nl * nl);
  if (st) return st;
  st = upload_typed(c, (char *)c->cm2l + mo, cm2l, (size_t)nl * nl);
  if (st) return st;
  HIPCHK(c, hipMemcpy(c->iRd2_d + (size_t)nl * member, iRd2, sizeof(double) * nl, hipMemcpyHostToDevice));
  c->mset[member] = 1;
  bool all = true;
  for (int b = 0; b < c->nb; b++) all = all && c->mset[b];
  c->modes_set = all;
  return QGW_OK;
}

// Energy and enstrophy of every member of the group (one reduction launch).
int group_diag(qgw_ctx *c, std::vector<double> &h) {
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  a.psi = c->psi;
  a.q = c->q[c->qcur];
  a.comp = (c->prec == QGW_FP32) ? c->comp : nullptr;
  int st;
  {
    PassScope ps(c, QGW_PASS_DIAG);
    st = qgw::launch_diag(a, c->diag_part, c->nblk_diag, c->diag_out);
  }
  if (st || (st = check_launch(c, "diag"))) return st;
  h.assign(2 * (size_t)c->g.nl, 0.0);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->diag_out, sizeof(double) * 2 * c->g.nl, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return QGW_OK;
}

void member_energy(const qgw_ctx *c, int member, const double *h, double *E_layer, double *Z_layer, double *E_tot,
                   double *Z_tot) {
  const qgw_params &p = c->mp[member];
  const int nl = p.nl;
  const double *hm = h + 2 * (size_t)nl * member;
  const double w = c->grid.Delta * c->grid.Delta / (p.Lx * p.Lx);
  double H = 0, Et = 0, Zt = 0;
  for (int l = 0; l < nl; l++) H += p.dh[l];
  for (int l = 0; l < nl; l++) {
    const double e = -hm[2 * l] * w, z = hm[2 * l + 1] * w;
    if (E_layer) E_layer[l] = e;
    if (Z_layer) Z_layer[l] = z;
    Et += p.dh[l] * e;
    Zt += p.dh[l] * z;
  }
  if (E_tot) *E_tot = Et / (2 * H);
  if (Z_tot) *Z_tot = Zt / H;
}

// Stochastic ring forcing: the forced points grouped by column for the
// synthesis kernels, the twiddles e^{+2 pi i t/N}, the random stream (seed 1:
// the C library's rand() stream the reference draws from) and the buffers.
int setup_stochastic(qgw_ctx *c, const qgw_params *p) {
  const int N = p->NX;
  int np = 0;
  double env = 0;
  if (qgwh_stoch_ring(p, 0, nullptr, nullptr, &np, &env) != QGW_OK) return QGW_ERR_ARG;
  c->nst = np;
  c->st_env = env;
  if (np <= 1) return QGW_OK;  // nothing is forced (src/forcing.h:170)
  std::vector<int> ii(np), jj(np);
  qgwh_stoch_ring(p, np, ii.data(), jj.data(), &np, &env);
  // columns in increasing i, their points in draw order
  std::vector<int> cols;
  for (int i = 0; i <= N / 2; i++)
    for (int k = 0; k < np; k++)
      if (ii[k] == i) {
        cols.push_back(i);
        break;
      }
  const int nc = (int)cols.size();
  std::vector<int> idx;  // col_i[nc], col_start[nc+1], mode_j[np], mode_k[np]
  idx.insert(idx.end(), cols.begin(), cols.end());
  std::vector<int> mj, mk;
  idx.push_back(0);
  for (int ci = 0; ci < nc; ci++) {
    for (int k = 0; k < np; k++)
      if (ii[k] == cols[ci]) {
        mj.push_back(jj[k]);
        mk.push_back(k);
      }
    idx.push_back((int)mj.size());
  }
  idx.insert(idx.end(), mj.begin(), mj.end());
  idx.insert(idx.end(), mk.begin(), mk.end());
  std::vector<double> tw(2 * (size_t)N);
  for (int t = 0; t < N; t++) {
    const double a = 2.0 * M_PI * (double)t / (double)N;
    tw[2 * t] = cos(a);
    tw[2 * t + 1] = sin(a);
  }
  // E[c][i'] = w_c e^{+2 pi i col_i[c] i'/N} over the n interior columns
  const int n = c->g.n;
  std::vector<double> E(2 * (size_t)nc * n);
  for (int ci = 0; ci < nc; ci++) {
    const double w = (cols[ci] == 0 || 2 * cols[ci] == N) ? 1.0 : 2.0;
    for (int i = 0; i < n; i++) {
      const size_t t = ((size_t)cols[ci] * i) % N, o = 2 * ((size_t)ci * n + i);
      E[o] = w * tw[2 * t];
      E[o + 1] = w * tw[2 * t + 1];
    }
  }
  void *d_idx = nullptr, *d_tw = nullptr, *d_coef = nullptr, *d_G = nullptr, *d_E = nullptr;
  int st = dalloc(c, &d_idx, sizeof(int) * idx.size());
  st = st ? st : dalloc(c, &d_tw, sizeof(double) * tw.size());
  st = st ? st : dalloc(c, &d_E, sizeof(double) * E.size());
  st = st ? st : dalloc(c, &d_coef, sizeof(double) * 2 * np);
  st = st ? st : dalloc(c, &d_G, sizeof(double) * 2 * (size_t)N * nc);
  if (st) return st;
  if (hipMemcpy(d_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d_E, E.data(), sizeof(double) * E.size(), hipMemcpyHostToDevice) != hipSuccess)
    return QGW_ERR_HIP;
  const int *di = (const int *)d_idx;
  c->sa.N = N;
  c->sa.nc = nc;
  c->sa.col_start = di + nc;
  c->sa.mode_j = di + 2 * nc + 1;
  c->sa.mode_k = di + 2 * nc + 1 + np;
  c->sa.coef = (const double *)d_coef;
  c->sa.tw = (const double *)d_tw;
  c->sa.E = (const double *)d_E;
  c->sa.G = (double *)d_G;
  for (int i = 0; i < 2; i++) {
    if (hipHostMalloc((void **)&c->st_host[i], sizeof(double) * 2 * np, hipHostMallocDefault) != hipSuccess)
      return QGW_ERR_NOMEM;
    if (hipEventCreateWithFlags(&c->st_ev[i], hipEventDisableTiming) != hipSuccess) return QGW_ERR_HIP;
  }
  c->rng = qgwh_rng_create(1);
  return c->rng ? QGW_OK : QGW_ERR_NOMEM;
}

// A group of nb members (ps[0..nb-1], all same_group, or nb = 1) on stream
// `ext` (owned by the caller) or on a stream of its own (ext = null).
// The dense sine transforms: a BLAS handle, S_x (n x n) and S_y (m x m) with
// S_ij = 2 sin(pi i j / (N+1)) (i, j = 1..N; the unnormalised RODFT00 of
// src/elliptic.h:63-66, formed in FP64), and the temporary of the column pass.
int setup_dense(qgw_ctx *c) {
  if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
    c->blas = nullptr;
    snprintf(c->err, sizeof c->err, "rocblas_create_handle failed");
    return QGW_ERR_HIP;
  }
  const int dims[2] = {c->g.n, c->g.m};
  void **dst[2] = {&c->Sx, &c->Sy};
  for (int d = 0; d < 2; d++) {
    const int N = dims[d];
    std::vector<double> S((size_t)N * N);
    for (int i = 1; i <= N; i++)
      for (int j = 1; j <= N; j++) {
        // sin(pi i j/(N+1)) with the argument reduced exactly: i j mod 2(N+1)
        const long r = ((long)i * j) % (2L * (N + 1));
        S[(size_t)(i - 1) * N + (j - 1)] = 2.0 * sin(kPi * (double)r / (double)(N + 1));
      }
    int st = dalloc(c, dst[d], c->esz * (size_t)N * N);
    if (!st) st = upload_typed(c, *dst[d], S.data(), S.size());
    if (st) return st;
  }
  return dalloc(c, &c->wtmp, c->esz * (size_t)c->g.WTS * c->g.nl);
}

qgw_ctx *create_group(const qgw_params *ps, int nb, int precision, hipStream_t ext, int *status) {
  *status = QGW_ERR_ARG;
  if (!ps || nb < 1) return nullptr;
  if (precision != QGW_FP64 && precision != QGW_FP32 && precision != QGW_FP32_PLAIN) return nullptr;
  const qgw_params *p = &ps[0];
  qgw_grid grid;
  int Lx, Ly, dense;
  if (check_params(p, &grid, &Lx, &Ly, &dense) != QGW_OK) return nullptr;
  for (int b = 1; b < nb; b++)
    if (!same_group(ps[0], ps[b])) return nullptr;
  const int n = grid.n, m = grid.m, nl = p->nl;
  if ((long)nb * nl > (1L << 20)) return nullptr;

  qgw_ctx *c = new qgw_ctx();
  memset(c->err, 0, sizeof c->err);
  {
    // an optional ceiling on the device memory of one context (callers that share the GPU)
    const char *lim = getenv("QGW_MAX_DEVICE_BYTES");
    c->byte_limit = lim ? strtoull(lim, nullptr, 10) : 0;
  }
  c->p = *p;
  c->nb = nb;
  c->mp.assign(ps, ps + nb);
  c->mset.assign(nb, 0);
  c->prec = precision;
  c->esz = (precision == QGW_FP64) ? 8 : 4;
  c->grid = grid;
  c->Lx = Lx;
  c->Ly = Ly;
  const int NL = nb * nl;  // layers of all members
  qgw::Geo &g = c->g;
  g.n = n;
  g.m = m;
  g.nl = NL;
  g.nlm = nl;
  const bool per = (p->bc_fac == -1.0);
  // spectral rows: n values (bounded), or N/2 + 1 complex = n + 2 reals (periodic)
  const int wc = per ? n + 2 : n;
  g.P = pitch_of(n + 2);
  g.LS = (long)(m + 2) * g.P;
  g.Pn = pitch_of(wc);
  g.WS = (long)m * g.Pn;
  g.Pm = pitch_of(m);
  g.WTS = (long)wc * g.Pm;
  *status = QGW_ERR_HIP;
  if (ext) {
    c->stream = ext;
    c->own_stream = false;
  } else {
    c->own_stream = true;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      c->stream = nullptr;
      release(c);
      return nullptr;
    }
  }
  {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) {
      release(c);
      return nullptr;
    }
    c->cus = cus;
    // this device's copy of the allocation-free entry point's scratch
    if (scratch_on_device(dev, &c->scratch) != hipSuccess) {
      release(c);
      return nullptr;
    }
  }
  c->dk = 3.141592653589793 / p->Lx;
  const size_t field = c->esz * (size_t)g.LS * NL;
  int st = QGW_OK;
  st = st ? st : dalloc_field(c, &c->q[0], field);
  st = st ? st : dalloc_field(c, &c->q[1], field);
  st = st ? st : dalloc_field(c, &c->psi, field);
  for (int i = 0; i < 3; i++) st = st ? st : dalloc_field(c, &c->fa[i], field);
  // (16-bit elements; a full field is allocated so that state copies treat it like the others)
  if (precision == QGW_FP32) st = st ? st : dalloc_field(c, &c->comp, field);
  st = st ? st : dalloc_field(c, &c->topo, c->esz * (size_t)g.LS * nb);
  st = st ? st : dalloc(c, &c->wind, c->esz * (size_t)(m + 2) * nb);
  st = st ? st : dalloc(c, &c->what, c->esz * (size_t)g.WS * NL);
  st = st ? st : dalloc(c, &c->whatT, c->esz * (size_t)g.WTS * NL);
  st = st ? st : dalloc(c, &c->cl2m, c->esz * (size_t)nl * nl * nb);
  st = st ? st : dalloc(c, &c->cm2l, c->esz * (size_t)nl * nl * nb);
  st = st ? st : dalloc(c, (void **)&c->iRd2_d, sizeof(double) * NL);
  st = st ? st : dalloc(c, (void **)&c->La2, sizeof(double) * (n + 2));
  st = st ? st : dalloc(c, (void **)&c->Kb2, sizeof(double) * (m + 2));
  c->nblk_diag = 256;
  st = st ? st : dalloc(c, (void **)&c->diag_part, sizeof(double) * 2 * c->nblk_diag * NL);
  st = st ? st : dalloc(c, (void **)&c->diag_out, sizeof(double) * 2 * NL);
  st = st ? st : dalloc(c, (void **)&c->cfl_max, sizeof(unsigned long long));
  st = st ? st : dalloc(c, (void **)&c->nonfinite, sizeof(unsigned));
  c->nrec4 = (p->dt_forc_period != 0) ? forcing_records(p) : 0;
  c->irec4[0] = c->irec4[1] = -1;
  c->forc4[0] = c->forc4[1] = nullptr;
  c->forc_path = "input_vars.nc";  // src/forcing.h:262
  if (c->nrec4 > 0 && nb != 1) st = st ? st : QGW_ERR_ARG;  // one run per context
  for (int i = 0; i < 2 && c->nrec4 > 0; i++) st = st ? st : dalloc_field(c, &c->forc4[i], field);
  if (!st && p->k_f != 0) {
    if (nb != 1) st = QGW_ERR_ARG;
    if (!st) st = setup_stochastic(c, p);
  }
  c->dst_gemm = dense;
  if (!st && dense) st = setup_dense(c);
  if (!st) {
    if (precision == QGW_FP64) {
      st = make_twiddles<double>(c, Lx, &c->twx);
      if (!st) st = make_twiddles<double>(c, Ly, &c->twy);
      if (!st) st = make_twiddles<double>(c, Lx / 2, &c->twxh);
      if (!st) st = make_twiddles<double>(c, Ly / 2, &c->twyh);
    } else {
      st = make_twiddles<float>(c, Lx, &c->twx);
      if (!st) st = make_twiddles<float>(c, Ly, &c->twy);
      if (!st) st = make_twiddles<float>(c, Lx / 2, &c->twxh);
      if (!st) st = make_twiddles<float>(c, Ly / 2, &c->twyh);
    }
  }
  if (st) {
    *status = st;
    release(c);
    return nullptr;
  }
  // wavenumbers squared, La2 along x and Kb2 along y: bounded K[j] and L[i]
  // (src/domain.h:169-174; K uses Lx as the reference does), periodic K[i]
  // (x, half spectrum) and L[j] (y, signed) (src/domain.h:163-167)
  {
    std::vector<double> Lw(n + 2), Kw(m + 2), La2(n + 2, 0.0), Kb2(m + 2, 0.0);
    qgwh_grid(p, &grid, nullptr, nullptr, Kw.data(), Lw.data());
    if (per) {
      for (int i = 0; i < n / 2 + 1; i++) La2[i] = Kw[i] * Kw[i];
      for (int j = 0; j < m; j++) Kb2[j] = Lw[j] * Lw[j];
    } else {
      for (int i = 0; i < n + 2; i++) La2[i] = Lw[i] * Lw[i];
      for (int j = 0; j < m + 2; j++) Kb2[j] = Kw[j] * Kw[j];
    }
    if (hipMemcpy(c->La2, La2.data(), sizeof(double) * (n + 2), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->Kb2, Kb2.data(), sizeof(double) * (m + 2), hipMemcpyHostToDevice) != hipSuccess) {
      release(c);
      return nullptr;
    }
  }
  // per-member wind row and topography (src/domain.h:33,200-207)
  {
    std::vector<double> wind((size_t)(m + 2) * nb);
    for (int b = 0; b < nb; b++) qgwh_wind_row(&ps[b], wind.data() + (size_t)(m + 2) * b);
    if (upload_typed(c, c->wind, wind.data(), wind.size())) {
      release(c);
      return nullptr;
    }
    std::vector<double> topo((size_t)(n + 2) * (m + 2));
    for (int b = 0; b < nb; b++) {
      qgwh_topo(&ps[b], topo.data());
      if (member_upload(c, b, QGW_FIELD_TOPO, topo.data()) != QGW_OK) {
        release(c);
        return nullptr;
      }
    }
  }
  // time state (init_vars sign rule, check_timestep, init_timestep)
  c->nu_hyper = qgwh_nu_hyper_signed(p);
  qgwh_check_timestep(p, c->nu_hyper, &c->dt, &c->DT_max);
  if (p->dt_fixed > 0) c->dt = p->dt_fixed;
  c->dt0 = c->dt;
  c->dt1 = c->dt;
  c->t = 0;
  c->t_out = 0;
  c->it = 0;
  c->qcur = 0;
  c->p4_tile = qgw::kP4TileCount;  // the register-streaming form
  c->p4_waves = 40;
  // the faster dealing order measured: layers first at 2048^2 x 10, chunks
  // first at 4096^2 x 32 (the autotuner tries both)
  c->p4_order = (c->g.nl <= 16) ? 1 : 0;
  c->fft_e = 16;
  c->fft_e_col = (precision == QGW_FP64) ? 8 : 16;
  c->per = per ? 1 : 0;
  // the pipelined row passes by default, except FP64 rows of a length other than 2^k:
  // there the full LDS twiddle table of the Stockham leaves one block per CU,
  // and the one-block form was measured faster (1536^2 x 10: 0.849 against
  // 0.870 ms per step; FP32 the other way, 0.471 against 0.482)
  const bool rows_pow2 = ((Lx / 2) & (Lx / 2 - 1)) == 0;
  c->dst_half = (!per && !dense && qgw::launch_has_half_length(Lx / 2) && qgw::launch_has_half_length(Ly / 2))
                    ? ((precision == QGW_FP64 && !rows_pow2) ? 1 : 2)
                    : -1;
  c->fuse = (c->dst_half > 0) ? 0 : -1;  // four passes measured slower at 2048^2 x 10; the autotuner tries both
  // column pairs per block of the four-pass column pass: the fastest measured
  // at 2048^2 x 10 (FP32 4: 180 us against 191 for 8 and 194 for 2; FP64 2:
  // 357 us against 390 for 4)
  c->p2m_g = (precision == QGW_FP64) ? 2 : 4;
  c->psi_nt = 1;
  c->p4_rev = 1;
  c->modes_set = false;
  // the zero fills above ran on the null stream; the step runs on c->stream
  if (hipDeviceSynchronize() != hipSuccess) {
    release(c);
    return nullptr;
  }
  *status = QGW_OK;
  return c;
}

// Device copy of a group's state (fields, rotation, time), so that timing
// runs can be undone.
struct Snapshot {
  void *buf = nullptr;
  qgwh_rng *rng = nullptr;  // the forcing's random stream
  int qcur = 0;
  void *fa[3] = {nullptr, nullptr, nullptr};
  double t = 0, dt = 0, dt0 = 0, dt1 = 0;
  int it = 0;
};

std::vector<void *> state_fields(qgw_ctx *c) {
  std::vector<void *> f = {c->q[0], c->q[1], c->psi, c->fa[0], c->fa[1], c->fa[2]};
  if (c->comp) f.push_back(c->comp);
  return f;
}

int snapshot_take(qgw_ctx *c, Snapshot &s) {
  const size_t field = c->esz * (size_t)c->g.LS * c->g.nl;
  auto f = state_fields(c);
  HIPCHK(c, hipMalloc(&s.buf, field * f.size()));
  for (size_t i = 0; i < f.size(); i++)
    HIPCHK(c, hipMemcpyAsync((char *)s.buf + i * field, f[i], field, hipMemcpyDeviceToDevice, c->stream));
  if (c->rng && !(s.rng = qgwh_rng_clone(c->rng))) return QGW_ERR_NOMEM;
  s.qcur = c->qcur;
  for (int i = 0; i < 3; i++) s.fa[i] = c->fa[i];
  s.t = c->t;
  s.dt = c->dt;
  s.dt0 = c->dt0;
  s.dt1 = c->dt1;
  s.it = c->it;
  return QGW_OK;
}

int snapshot_restore(qgw_ctx *c, Snapshot &s) {
  const size_t field = c->esz * (size_t)c->g.LS * c->g.nl;
  c->qcur = s.qcur;
  for (int i = 0; i < 3; i++) c->fa[i] = s.fa[i];
  c->t = s.t;
  c->dt = s.dt;
  c->dt0 = s.dt0;
  c->dt1 = s.dt1;
  c->it = s.it;
  if (s.rng) {
    qgwh_rng_assign(c->rng, s.rng);
    qgwh_rng_destroy(s.rng);
    s.rng = nullptr;
  }
  auto f = state_fields(c);
  for (size_t i = 0; i < f.size(); i++)
    HIPCHK(c, hipMemcpyAsync(f[i], (char *)s.buf + i * field, field, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipFree(s.buf));
  s.buf = nullptr;
  return QGW_OK;
}

int time_steps(qgw_ctx *c, int nsteps, double *ms);

// Best of two timings of `reps` steps after one untimed step (the autotuner's
// measurement: single timings of a few steps are too noisy to rank variants
// a few percent apart).
int time_candidate(qgw_ctx *c, int reps, double *ms) {
  double a = 0, b = 0;
  int st = time_steps(c, 1, &a);
  if (!st) st = time_steps(c, reps, &a);
  if (!st) st = time_steps(c, reps, &b);
  *ms = a < b ? a : b;
  return st;
}

int time_steps(qgw_ctx *c, int nsteps, double *ms) {
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  int st = QGW_OK;
  if (hipEventRecord(a, c->stream) != hipSuccess) st = QGW_ERR_HIP;
  for (int i = 0; i < nsteps && !st; i++) st = one_step(c);
  if (!st && hipEventRecord(b, c->stream) != hipSuccess) st = QGW_ERR_HIP;
  if (!st && hipEventSynchronize(b) != hipSuccess) st = QGW_ERR_HIP;
  float f = 0.f;
  if (!st && hipEventElapsedTime(&f, a, b) != hipSuccess) st = QGW_ERR_HIP;
  *ms = f;
  hipEventDestroy(a);
  hipEventDestroy(b);
  return st;
}

// Time `reps` steps with each P4 variant and keep the fastest; the state is
// restored afterwards.
int autotune_group(qgw_ctx *c, int reps, int *p4_tile, double *best_ms) {
  Snapshot snap;
  int st = snapshot_take(c, snap);
  if (st) return st;
  double best = 1e300;
  int arg = c->p4_tile, warg = c->p4_waves;
  // streaming P4: the row chunk follows from the wavefronts per SIMD (at least
  // 16 rows); 2048^2 x 10 takes 40-48 (18-16 rows), 4096^2 x 32 the most
  // (384: 26 rows, 14% faster than 48's 203 rows)
  static const int kWaves[] = {16, 32, 40, 48, 96, 192, 384};
  constexpr int kNW = sizeof(kWaves) / sizeof(kWaves[0]);
  for (int v = 0; v < (c->per ? 0 : qgw::kP4Variants) && !st; v++) {  // periodic: one P4 kernel
    const bool stream = (v >= qgw::kP4TileCount);  // the streaming forms: also the row chunk
    for (int wi = 0; wi < (stream ? kNW : 1) && !st; wi++) {
      c->p4_tile = v;
      if (stream) c->p4_waves = kWaves[wi];
      double ms = 0;
      st = time_steps(c, 1, &ms);  // first-launch code loading
      if (!st) st = time_candidate(c, reps, &ms);
      if (!st && ms < best) {
        best = ms;
        arg = v;
        warg = c->p4_waves;
      }
    }
  }
  c->p4_waves = warg;
  c->p4_tile = arg;
  if (!st && !c->per && arg >= qgw::kP4TileCount) {
    // the streaming forms: the other dealing order of the walks, timed
    // against the current one with twice the repetitions and kept only when
    // faster by more than 0.5% (the two orders differ by a few percent)
    const int oarg = c->p4_order;
    double ms0 = 0, ms1 = 0;
    st = time_candidate(c, 2 * reps, &ms0);
    c->p4_order = 1 - oarg;
    if (!st) st = time_candidate(c, 2 * reps, &ms1);
    if (!st && ms1 < 0.995 * ms0)
      best = ms1;
    else
      c->p4_order = oarg;
  }
  if (!st && c->fuse >= 0) {
    // the inversion in six passes, or in four with 2, 4 or 8 column pairs per
    // block of the column pass, each with plain and non-temporal psi stores
    // (which change the time of the tendency pass that follows)
    double bf = 1e300;
    int farg = c->fuse, garg = c->p2m_g, narg = c->psi_nt;
    const int forms[4][2] = {{0, 0}, {1, 2}, {1, 4}, {1, 8}};
    for (int v = 0; v < 8 && !st; v++) {
      c->fuse = forms[v / 2][0];
      if (forms[v / 2][0]) c->p2m_g = forms[v / 2][1];
      c->psi_nt = v % 2;
      double ms = 0;
      st = time_steps(c, 1, &ms);  // first-launch code loading
      if (!st) st = time_candidate(c, reps, &ms);
      if (!st && ms < bf) {
        bf = ms;
        farg = c->fuse;
        garg = c->p2m_g;
        narg = c->psi_nt;
      }
    }
    c->fuse = farg;
    c->p2m_g = garg;
    c->psi_nt = narg;
  } else if (!st) {
    // plain or non-temporal psi stores
    double bn = 1e300;
    int narg = c->psi_nt;
    for (int v = 0; v < 2 && !st; v++) {
      c->psi_nt = v;
      double ms = 0;
      st = time_candidate(c, reps, &ms);
      if (!st && ms < bn) {
        bn = ms;
        narg = v;
      }
    }
    c->psi_nt = narg;
  }
  if (!st && c->prec != QGW_FP64 && c->fuse != 1) {
    // projection: VALU or matrix cores
    double bt = 1e300;
    int targ = c->t_kernel;
    for (int v = 1; v <= 3 && !st; v++) {
      c->t_kernel = v;
      double ms = 0;
      st = time_candidate(c, reps, &ms);
      if (!st && ms < bt) {
        bt = ms;
        targ = v;
      }
    }
    c->t_kernel = targ;
  }
  if (!st && c->dst_half >= 0) {
    // sine transforms: odd-extension FFTs, or half-length FFTs (one block per
    // item, or the row passes in pipelined persistent blocks) with each radix schedule
    double bt = 1e300;
    int darg = c->dst_half, earg = c->fft_e;
    const int forms[5][2] = {{0, 16}, {1, 16}, {1, 8}, {2, 16}, {2, 8}};
    for (int v = 0; v < 5 && !st; v++) {
      c->dst_half = forms[v][0];
      c->fft_e = forms[v][1];
      double ms = 0;
      st = time_candidate(c, reps, &ms);
      if (!st && ms < bt) {
        bt = ms;
        darg = forms[v][0];
        earg = forms[v][1];
      }
    }
    c->dst_half = darg;
    c->fft_e = earg;
    // the column pass's radix schedule, with the row passes' chosen
    if (c->dst_half > 0) {
      int carg = c->fft_e_col;
      double bc = 1e300;
      for (int v = 0; v < 2 && !st; v++) {
        c->fft_e_col = v ? 8 : 16;
        double ms = 0;
        st = time_candidate(c, reps, &ms);
        if (!st && ms < bc) {
          bc = ms;
          carg = c->fft_e_col;
        }
      }
      c->fft_e_col = carg;
    }
  }
  // the chosen configuration, timed once more (the figure returned)
  if (!st) st = time_candidate(c, reps, &best);
  int st2 = snapshot_restore(c, snap);
  if (st) return st;
  if (st2) return st2;
  if (p4_tile) *p4_tile = arg;
  if (best_ms) *best_ms = best / reps;
  return QGW_OK;
}

}  // namespace

// An ensemble: groups of members (each a qgw_ctx stacking its members along
// the layer axis) on one stream.  A fixed step is periodic in the buffer
// rotation (q alternates between two buffers, the tendency history cycles
// through three), so kGraphSteps = 6 steps of every group are captured once
// into a hipGraph and replayed; the graph is keyed by the rotation state it
// starts from.
struct qgw_batch {
  int nmem = 0;
  int prec = 0;
  hipStream_t stream = nullptr;
  std::vector<qgw_ctx *> groups;
  std::vector<int> gof, iof;  // member -> group, index inside the group
  bool use_graph = true;
  struct Cached {
    std::vector<const void *> sig;
    hipGraphExec_t exec;
  };
  std::vector<Cached> cache;
  int graph_launches = 0;
  // graph capture with more than one group: each group on a stream of its
  // own, forked from and joined back to `stream` by events, so the captured
  // graph holds one independent branch per group
  std::vector<hipStream_t> gstreams;
  std::vector<hipEvent_t> gjoin;
  hipEvent_t gfork = nullptr;
  char err[256];
};

namespace {

constexpr int kGraphSteps = 6;

int bfail(qgw_batch *b, hipError_t e, const char *what) {
  snprintf(b->err, sizeof b->err, "%s: %s", what, hipGetErrorString(e));
  return QGW_ERR_HIP;
}

int bpass(qgw_batch *b, qgw_ctx *c, int st) {
  if (st && c) snprintf(b->err, sizeof b->err, "%s", c->err);
  return st;
}

// A group whose step takes host work (stochastic draws, 4-D forcing records):
// it is stepped launch by launch beside the replayed graph of the others.
bool host_fed(const qgw_ctx *c) { return c->nst > 1 || c->nrec4 > 0; }

std::vector<const void *> batch_sig(const qgw_batch *b) {
  std::vector<const void *> sig;
  for (auto *c : b->groups) {
    if (host_fed(c)) continue;
    sig.push_back(c->fa[0]);
    sig.push_back(c->fa[1]);
    sig.push_back(c->fa[2]);
    sig.push_back(c->q[c->qcur]);
  }
  return sig;
}

bool graph_ok(const qgw_batch *b) {
  if (!b->use_graph) return false;
  int graphed = 0;
  for (auto *c : b->groups) {
    const double dt = c->p.dt_fixed;
    if (c->prof || !(dt > 0) || c->dt != dt || c->dt0 != dt || c->dt1 != dt || !c->modes_set) return false;
    if (!host_fed(c)) graphed++;
  }
  return graphed > 0;
}

int batch_eager_step(qgw_batch *b) {
  for (auto *c : b->groups) {
    int st = one_step(c);
    if (st) return bpass(b, c, st);
  }
  return QGW_OK;
}

// Destroys the cached graphs; a replay still in flight finishes first.
void batch_release_graphs(qgw_batch *b) {
  if (!b->cache.empty() && b->stream) hipStreamSynchronize(b->stream);
  for (auto &e : b->cache) hipGraphExecDestroy(e.exec);
  b->cache.clear();
}

// Streams and events of a multi-group capture, made on first use.
int batch_branch_streams(qgw_batch *b) {
  const size_t ng = b->groups.size();
  if (ng < 2 || b->gstreams.size() == ng) return QGW_OK;
  hipError_t e = hipSuccess;
  if (!b->gfork) e = hipEventCreateWithFlags(&b->gfork, hipEventDisableTiming);
  while (e == hipSuccess && b->gstreams.size() < ng) {
    hipStream_t st = nullptr;
    e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) break;
    b->gstreams.push_back(st);
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) break;
    b->gjoin.push_back(ev);
  }
  return (e == hipSuccess) ? QGW_OK : bfail(b, e, "batch branch streams");
}

// Capture kGraphSteps steps of every group from the current rotation state.
// With several groups, each group's steps are captured on a stream of its own
// (forked from the batch stream by an event and joined back by another), so
// the groups' kernels are independent branches of the graph and may overlap.
// Nothing runs during capture, so the host time state is put back after it.
int batch_capture(qgw_batch *b, hipGraphExec_t *exec) {
  struct T {
    double t;
    int it;
  };
  std::vector<T> keep;
  for (auto *c : b->groups) keep.push_back({c->t, c->it});
  const auto sig0 = batch_sig(b);
  int st = batch_branch_streams(b);
  if (st) return st;
  std::vector<qgw_ctx *> gg;  // the groups the graph holds
  for (auto *c : b->groups)
    if (!host_fed(c)) gg.push_back(c);
  const bool branches = gg.size() > 1;
  hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return bfail(b, e, "hipStreamBeginCapture");
  if (!branches) {
    for (int s = 0; s < kGraphSteps && !st; s++) st = bpass(b, gg[0], one_step(gg[0]));
  } else {
    e = hipEventRecord(b->gfork, b->stream);
    if (e != hipSuccess) st = bfail(b, e, "hipEventRecord");
    for (size_t i = 0; i < gg.size() && !st; i++) {
      qgw_ctx *c = gg[i];
      e = hipStreamWaitEvent(b->gstreams[i], b->gfork, 0);
      if (e != hipSuccess) {
        st = bfail(b, e, "hipStreamWaitEvent");
        break;
      }
      hipStream_t own = c->stream;
      c->stream = b->gstreams[i];
      for (int s = 0; s < kGraphSteps && !st; s++) st = bpass(b, c, one_step(c));
      c->stream = own;
      if (st) break;
      e = hipEventRecord(b->gjoin[i], b->gstreams[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->gjoin[i], 0);
      if (e != hipSuccess) st = bfail(b, e, "branch join");
    }
  }
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(b->stream, &graph);
  for (size_t i = 0; i < b->groups.size(); i++) {
    b->groups[i]->t = keep[i].t;
    b->groups[i]->it = keep[i].it;
  }
  if (st) {
    if (graph) hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return bfail(b, e, "hipStreamEndCapture");
  if (batch_sig(b) != sig0) {
    hipGraphDestroy(graph);
    snprintf(b->err, sizeof b->err, "buffer rotation is not periodic over %d steps", kGraphSteps);
    return QGW_ERR_STATE;
  }
  e = hipGraphInstantiateWithFlags(exec, graph, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return bfail(b, e, "hipGraphInstantiate");
  return QGW_OK;
}

// The graph of the current rotation state, captured on first use.
int batch_exec(qgw_batch *b, hipGraphExec_t *exec) {
  const auto sig = batch_sig(b);
  *exec = nullptr;
  for (auto &e : b->cache)
    if (e.sig == sig) *exec = e.exec;
  if (*exec) return QGW_OK;
  int st = batch_capture(b, exec);
  if (st) return st;
  b->cache.push_back({sig, *exec});
  return QGW_OK;
}

int batch_run(qgw_batch *b, int nsteps) {
  while (nsteps >= kGraphSteps && graph_ok(b)) {
    hipGraphExec_t exec = nullptr;
    int st = batch_exec(b, &exec);
    if (st) return st;
    hipError_t e = hipGraphLaunch(exec, b->stream);
    if (e != hipSuccess) return bfail(b, e, "hipGraphLaunch");
    b->graph_launches++;
    for (auto *c : b->groups) {
      if (host_fed(c)) {  // its steps, launch by launch, after the replay on the same stream
        for (int s = 0; s < kGraphSteps; s++) {
          st = one_step(c);
          if (st) return bpass(b, c, st);
        }
        continue;
      }
      for (int s = 0; s < kGraphSteps; s++) {
        c->t = c->t + c->dt;
        c->it++;
      }
    }
    nsteps -= kGraphSteps;
  }
  for (int s = 0; s < nsteps; s++) {
    int st = batch_eager_step(b);
    if (st) return st;
  }
  return QGW_OK;
}

void batch_free(qgw_batch *b) {
  if (!b) return;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_release_graphs(b);
  for (auto *c : b->groups) release(c);
  for (auto st : b->gstreams) hipStreamDestroy(st);
  for (auto ev : b->gjoin) hipEventDestroy(ev);
  if (b->gfork) hipEventDestroy(b->gfork);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

extern "C" {

qgw_ctx *qgw_create(const qgw_params *p, int precision, int *status) {
  int dummy;
  if (!status) status = &dummy;
  return create_group(p, 1, precision, nullptr, status);
}

int qgw_destroy(qgw_ctx *c) {
  if (!c) return QGW_ERR_ARG;
  if (c->stream) hipStreamSynchronize(c->stream);
  release(c);
  return QGW_OK;
}

int qgw_set_modes(qgw_ctx *c, const double *cl2m, const double *cm2l, const double *iRd2) {
  return member_set_modes(c, 0, cl2m, cm2l, iRd2);
}

int qgw_upload(qgw_ctx *c, int field, const double *host) { return member_upload(c, 0, field, host); }

int qgw_download(qgw_ctx *c, int field, double *host) { return member_download(c, 0, field, host); }

int qgw_invert_pv(qgw_ctx *c) {
  if (!c) return QGW_ERR_ARG;
  return run_invert(c);
}

int qgw_rhs(qgw_ctx *c) {
  if (!c) return QGW_ERR_ARG;
  int st = run_invert(c);
  if (st) return st;
  const double z[3] = {0, 0, 0};
  return run_p4(c, z, 0);
}

int qgw_adjust_timestep(qgw_ctx *c, double *dt_out) {
  if (!c) return QGW_ERR_ARG;
  if (c->p.dt_fixed > 0) {
    if (dt_out) *dt_out = c->p.dt_fixed;
    return QGW_OK;
  }
  double nd;
  int st = adaptive_dt(c, &nd);
  if (st) return st;
  c->dt = nd;
  if (dt_out) *dt_out = nd;
  return QGW_OK;
}

int qgw_timestep(qgw_ctx *c) {
  if (!c) return QGW_ERR_ARG;
  return one_step(c);
}

int run_steps(qgw_ctx *c, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    int st = one_step(c);
    if (st) return st;
  }
  return QGW_OK;
}

int qgw_run(qgw_ctx *c, int nsteps) {
  if (!c || nsteps < 0) return QGW_ERR_ARG;
  int st = run_steps(c, nsteps);
  return (st || nsteps == 0) ? st : check_finite(c);
}

int qgw_diagnostics(qgw_ctx *c, double *E_layer, double *Z_layer, double *E_tot, double *Z_tot) {
  if (!c) return QGW_ERR_ARG;
  std::vector<double> h;
  int st = group_diag(c, h);
  if (st) return st;
  if ((st = check_finite(c))) return st;
  member_energy(c, 0, h.data(), E_layer, Z_layer, E_tot, Z_tot);
  return QGW_OK;
}

int qgw_get_time(qgw_ctx *c, double *t, double *dt, double *dt0, double *dt1, int *it) {
  if (!c) return QGW_ERR_ARG;
  if (t) *t = c->t;
  if (dt) *dt = c->dt;
  if (dt0) *dt0 = c->dt0;
  if (dt1) *dt1 = c->dt1;
  if (it) *it = c->it;
  return QGW_OK;
}

int qgw_set_time(qgw_ctx *c, double t, double dt, double dt0, double dt1, int it) {
  if (!c) return QGW_ERR_ARG;
  c->t = t;
  c->dt = dt;
  c->dt0 = dt0;
  c->dt1 = dt1;
  c->it = it;
  return QGW_OK;
}

int qgw_set_next_output(qgw_ctx *c, double t_out) {
  if (!c) return QGW_ERR_ARG;
  c->t_out = t_out;
  return QGW_OK;
}

int qgw_set_stoch_seed(qgw_ctx *c, unsigned seed) {
  if (!c) return QGW_ERR_ARG;
  if (!c->rng) {
    snprintf(c->err, sizeof c->err, "no stochastic forcing in this run (k_f = 0 or at most one forced point)");
    return QGW_ERR_STATE;
  }
  qgwh_rng *r = qgwh_rng_create(seed);
  if (!r) return QGW_ERR_NOMEM;
  qgwh_rng_destroy(c->rng);
  c->rng = r;
  return QGW_OK;
}

int qgw_set_forcing_file(qgw_ctx *c, const char *path) {
  if (!c || !path || !*path) return QGW_ERR_ARG;
  if (c->nrec4 <= 0) {
    snprintf(c->err, sizeof c->err, "no 4-D forcing in this run (dt_forc_period = 0)");
    return QGW_ERR_STATE;
  }
  c->forc_path = path;
  c->irec4[0] = c->irec4[1] = -1;  // read again from the new file
  return QGW_OK;
}

int qgw_synchronize(qgw_ctx *c) {
  if (!c) return QGW_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_finite(c);
}

int qgw_time_run(qgw_ctx *c, int nsteps, double *ms) {
  if (!c || !ms || nsteps < 0) return QGW_ERR_ARG;
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  HIPCHK(c, hipEventRecord(a, c->stream));
  int st = run_steps(c, nsteps);
  if (!st) {
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float f = 0.f;
    HIPCHK(c, hipEventElapsedTime(&f, a, b));
    *ms = f;
    st = check_finite(c);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  return st;
}

int qgw_profile_enable(qgw_ctx *c, int on) {
  if (!c) return QGW_ERR_ARG;
  if (!on) {
    int st = prof_flush(c);
    if (st) return st;
  }
  c->prof = on != 0;
  return QGW_OK;
}

int qgw_profile_read(qgw_ctx *c, double *ms, int *count, int npass) {
  if (!c) return QGW_ERR_ARG;
  int st = prof_flush(c);
  if (st) return st;
  for (int i = 0; i < npass && i < QGW_NPASS; i++) {
    if (ms) ms[i] = c->prof_ms[i];
    if (count) count[i] = c->prof_cnt[i];
    c->prof_ms[i] = 0;
    c->prof_cnt[i] = 0;
  }
  return QGW_OK;
}

int qgw_tune_set(qgw_ctx *c, int knob, int value) {
  if (!c) return QGW_ERR_ARG;
  if (knob == QGW_TUNE_P4_TILE) {
    if (value < 0 || value >= qgw::kP4Variants) return QGW_ERR_ARG;
    c->p4_tile = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FFT_E) {
    if (value != 8 && value != 16) return QGW_ERR_ARG;
    c->fft_e = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FFT_E_COL) {
    if (value != 8 && value != 16) return QGW_ERR_ARG;
    c->fft_e_col = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_DST) {
    if (c->dst_gemm) return value == 3 ? QGW_OK : QGW_ERR_ARG;  // dense transforms only
    if (value < 0 || value > 2 || (value >= 1 && c->dst_half < 0)) return QGW_ERR_ARG;
    if (c->dst_half >= 0) c->dst_half = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P2M_PAIRS) {
    if (value != 0 && value != 1 && value != 2 && value != 4 && value != 8) return QGW_ERR_ARG;
    c->p2m_g = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_PSI_NT) {
    if (value < 0 || value > 1) return QGW_ERR_ARG;
    c->psi_nt = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_REV) {
    if (value < 0 || value > 1) return QGW_ERR_ARG;
    c->p4_rev = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_ORDER) {
    if (value < 0 || value > 1) return QGW_ERR_ARG;
    c->p4_order = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FUSE) {
    if (value < 0 || value > 1 || (value == 1 && c->fuse < 0)) return QGW_ERR_ARG;
    if (c->fuse >= 0) c->fuse = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_WAVES) {
    if (value != 2 && value != 4 && value != 8 && value != 16 && value != 32 && value != 40 && value != 48 &&
        value != 64 && value != 96 && value != 128 && value != 192 && value != 256 && value != 384 && value != 512 &&
        value != 768)
      return QGW_ERR_ARG;
    c->p4_waves = value;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_T_KERNEL) {
    if (value < 0 || value > 3 || (value >= 2 && c->prec == QGW_FP64)) return QGW_ERR_ARG;
    c->t_kernel = value;
    return QGW_OK;
  }
  return QGW_ERR_ARG;
}

int qgw_tune_get(qgw_ctx *c, int knob, int *value) {
  if (!c || !value) return QGW_ERR_ARG;
  if (knob == QGW_TUNE_P4_TILE) {
    *value = c->p4_tile;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_T_KERNEL) {
    *value = c->t_kernel;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_WAVES) {
    *value = c->p4_waves;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_DST) {
    *value = c->dst_gemm ? 3 : (c->dst_half > 0 ? c->dst_half : 0);
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FUSE) {
    *value = c->fuse > 0 ? 1 : 0;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P2M_PAIRS) {
    *value = c->p2m_g;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_PSI_NT) {
    *value = c->psi_nt;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_REV) {
    *value = c->p4_rev;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_P4_ORDER) {
    *value = c->p4_order;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FFT_E) {
    *value = c->fft_e;
    return QGW_OK;
  }
  if (knob == QGW_TUNE_FFT_E_COL) {
    *value = c->fft_e_col;
    return QGW_OK;
  }
  return QGW_ERR_ARG;
}

int qgw_autotune(qgw_ctx *c, int reps, int *p4_tile, double *best_ms) {
  if (!c || reps < 1) return QGW_ERR_ARG;
  return autotune_group(c, reps, p4_tile, best_ms);
}

int qgw_device_bytes(qgw_ctx *c, size_t *bytes) {
  if (!c || !bytes) return QGW_ERR_ARG;
  *bytes = c->bytes;
  return QGW_OK;
}

const char *qgw_strerror(int status) { return qgwh_strerror(status); }

const char *qgw_last_error(qgw_ctx *c) { return c ? c->err : ""; }

int qgw_diagnostics_device(qgw_ctx *c, double *sums) {
  if (!c || !sums) return QGW_ERR_ARG;
  if (c->g.nl > kNoAllocLayers) {
    snprintf(c->err, sizeof c->err, "qgw_diagnostics_device: %d layers, at most %d", c->g.nl, kNoAllocLayers);
    return QGW_ERR_ARG;
  }
  if (!c->scratch) {
    snprintf(c->err, sizeof c->err, "qgw_diagnostics_device: no scratch on this device");
    return QGW_ERR_STATE;
  }
  qgw::LaunchArgs a;
  a.g = c->g;
  a.prec = c->prec;
  a.stream = c->stream;
  a.psi = c->psi;
  a.q = c->q[c->qcur];
  a.comp = (c->prec == QGW_FP32) ? c->comp : nullptr;
  // every partial is written by its block before k_diag_final reads it, so
  // the scratch needs no clearing between calls
  int st = qgw::launch_diag(a, c->scratch, kNoAllocBlocks, sums);
  if (st || (st = check_launch(c, "diag"))) return st;
  return QGW_OK;
}


qgw_batch *qgw_batch_create(int nmembers, const qgw_params *params, int precision, int *status) {
  int dummy;
  if (!status) status = &dummy;
  *status = QGW_ERR_ARG;
  if (nmembers < 1 || !params) return nullptr;
  for (int i = 0; i < nmembers; i++) {
    if (!(params[i].dt_fixed > 0)) return nullptr;  // the graph step needs a fixed dt
    qgw_grid g;
    int lx, ly, dense;
    if (check_params(&params[i], &g, &lx, &ly, &dense) != QGW_OK) return nullptr;  // before any device call
  }
  qgw_batch *b = new qgw_batch();
  memset(b->err, 0, sizeof b->err);
  b->nmem = nmembers;
  b->prec = precision;
  b->gof.assign(nmembers, -1);
  b->iof.assign(nmembers, -1);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    b->stream = nullptr;
    *status = QGW_ERR_HIP;
    batch_free(b);
    return nullptr;
  }
  // group members in order of first appearance
  std::vector<std::vector<int>> members;
  for (int i = 0; i < nmembers; i++) {
    int gi = -1;
    for (size_t k = 0; k < members.size() && gi < 0; k++)
      if (same_group(params[members[k][0]], params[i])) gi = (int)k;
    if (gi < 0) {
      gi = (int)members.size();
      members.push_back({});
    }
    b->gof[i] = gi;
    b->iof[i] = (int)members[gi].size();
    members[gi].push_back(i);
  }
  for (auto &mem : members) {
    std::vector<qgw_params> ps;
    for (int i : mem) ps.push_back(params[i]);
    int st = QGW_OK;
    qgw_ctx *c = create_group(ps.data(), (int)ps.size(), precision, b->stream, &st);
    if (!c) {
      *status = st;
      batch_free(b);
      return nullptr;
    }
    b->groups.push_back(c);
  }
  *status = QGW_OK;
  return b;
}

int qgw_batch_destroy(qgw_batch *b) {
  if (!b) return QGW_ERR_ARG;
  batch_free(b);
  return QGW_OK;
}

int qgw_batch_groups(qgw_batch *b, int *ngroups, int *group_of_member) {
  if (!b) return QGW_ERR_ARG;
  if (ngroups) *ngroups = (int)b->groups.size();
  if (group_of_member)
    for (int i = 0; i < b->nmem; i++) group_of_member[i] = b->gof[i];
  return QGW_OK;
}

int qgw_batch_set_modes(qgw_batch *b, int member, const double *cl2m, const double *cm2l, const double *iRd2) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  return bpass(b, c, member_set_modes(c, b->iof[member], cl2m, cm2l, iRd2));
}

int qgw_batch_upload(qgw_batch *b, int member, int field, const double *host) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  return bpass(b, c, member_upload(c, b->iof[member], field, host));
}

int qgw_batch_download(qgw_batch *b, int member, int field, double *host) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  return bpass(b, c, member_download(c, b->iof[member], field, host));
}

int qgw_batch_invert_pv(qgw_batch *b) {
  if (!b) return QGW_ERR_ARG;
  for (auto *c : b->groups) {
    int st = run_invert(c);
    if (st) return bpass(b, c, st);
  }
  return QGW_OK;
}

// The non-finite flags of every group (waits for the batch stream).
int batch_check_finite(qgw_batch *b) {
  for (auto *c : b->groups) {
    int st = check_finite(c);
    if (st) return bpass(b, c, st);
  }
  return QGW_OK;
}

int qgw_batch_step(qgw_batch *b, int nsteps) {
  if (!b || nsteps < 0) return QGW_ERR_ARG;
  int st = batch_run(b, nsteps);
  return (st || nsteps == 0) ? st : batch_check_finite(b);
}

int qgw_batch_time_step(qgw_batch *b, int nsteps, double *ms) {
  if (!b || !ms || nsteps < 0) return QGW_ERR_ARG;
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return bfail(b, e, "hipEventCreate");
  e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    hipEventDestroy(e0);
    return bfail(b, e, "hipEventCreate");
  }
  int st = QGW_OK;
  // capture (and instantiate) the graph before the timed region: a replay
  // returns to the rotation state it started from, so one graph serves the run
  if (nsteps >= kGraphSteps && graph_ok(b)) {
    hipGraphExec_t exec = nullptr;
    st = batch_exec(b, &exec);
  }
  if (!st && (e = hipEventRecord(e0, b->stream)) != hipSuccess) st = bfail(b, e, "hipEventRecord");
  if (!st) st = batch_run(b, nsteps);
  if (!st && (e = hipEventRecord(e1, b->stream)) != hipSuccess) st = bfail(b, e, "hipEventRecord");
  if (!st && (e = hipEventSynchronize(e1)) != hipSuccess) st = bfail(b, e, "hipEventSynchronize");
  float f = 0.f;
  if (!st && (e = hipEventElapsedTime(&f, e0, e1)) != hipSuccess) st = bfail(b, e, "hipEventElapsedTime");
  *ms = f;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return st ? st : batch_check_finite(b);
}

int qgw_batch_prepare(qgw_batch *b) {
  if (!b) return QGW_ERR_ARG;
  if (!graph_ok(b)) return QGW_OK;
  hipGraphExec_t exec = nullptr;
  return batch_exec(b, &exec);
}

int qgw_batch_set_stoch_seed(qgw_batch *b, int member, unsigned seed) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  int st = qgw_set_stoch_seed(c, seed);
  return st ? bpass(b, c, st) : QGW_OK;
}

int qgw_batch_set_forcing_file(qgw_batch *b, int member, const char *path) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  int st = qgw_set_forcing_file(c, path);
  return st ? bpass(b, c, st) : QGW_OK;
}

int qgw_batch_synchronize(qgw_batch *b) {
  if (!b) return QGW_ERR_ARG;
  hipError_t e = hipStreamSynchronize(b->stream);
  return e == hipSuccess ? batch_check_finite(b) : bfail(b, e, "hipStreamSynchronize");
}

int qgw_batch_diagnostics(qgw_batch *b, int member, double *E_layer, double *Z_layer, double *E_tot,
                          double *Z_tot) {
  if (!b || member < 0 || member >= b->nmem) return QGW_ERR_ARG;
  qgw_ctx *c = b->groups[b->gof[member]];
  std::vector<double> h;
  int st = group_diag(c, h);
  if (st) return bpass(b, c, st);
  if ((st = check_finite(c))) return bpass(b, c, st);
  member_energy(c, b->iof[member], h.data(), E_layer, Z_layer, E_tot, Z_tot);
  return QGW_OK;
}

int qgw_batch_get_time(qgw_batch *b, double *t, int *it) {
  if (!b || b->groups.empty()) return QGW_ERR_ARG;
  if (t) *t = b->groups[0]->t;
  if (it) *it = b->groups[0]->it;
  return QGW_OK;
}

int qgw_batch_use_graph(qgw_batch *b, int on) {
  if (!b) return QGW_ERR_ARG;
  b->use_graph = on != 0;
  return QGW_OK;
}

int qgw_batch_graph_launches(qgw_batch *b, int *count) {
  if (!b || !count) return QGW_ERR_ARG;
  *count = b->graph_launches;
  return QGW_OK;
}

int qgw_batch_autotune(qgw_batch *b, int reps, int *p4_tile_per_group) {
  if (!b || reps < 1) return QGW_ERR_ARG;
  for (size_t k = 0; k < b->groups.size(); k++) {
    qgw_ctx *c = b->groups[k];
    int v = 0;
    double ms = 0;
    int st = autotune_group(c, reps, &v, &ms);
    if (st) return bpass(b, c, st);
    if (p4_tile_per_group) p4_tile_per_group[k] = v;
  }
  batch_release_graphs(b);  // the captured launches used the previous choice
  return QGW_OK;
}

int qgw_batch_tune_set(qgw_batch *b, int group, int knob, int value) {
  if (!b || group < 0 || group >= (int)b->groups.size()) return QGW_ERR_ARG;
  int st = qgw_tune_set(b->groups[group], knob, value);
  if (!st) batch_release_graphs(b);
  return st;
}

int qgw_batch_tune_get(qgw_batch *b, int group, int knob, int *value) {
  if (!b || group < 0 || group >= (int)b->groups.size()) return QGW_ERR_ARG;
  return qgw_tune_get(b->groups[group], knob, value);
}

int qgw_batch_profile_enable(qgw_batch *b, int on) {
  if (!b) return QGW_ERR_ARG;
  for (auto *c : b->groups) {
    int st = qgw_profile_enable(c, on);
    if (st) return bpass(b, c, st);
  }
  return QGW_OK;
}

int qgw_batch_profile_read(qgw_batch *b, double *ms, int *count, int npass) {
  if (!b) return QGW_ERR_ARG;
  std::vector<double> m(QGW_NPASS, 0.0), mm(QGW_NPASS);
  std::vector<int> n(QGW_NPASS, 0), nn(QGW_NPASS);
  for (auto *c : b->groups) {
    int st = qgw_profile_read(c, mm.data(), nn.data(), QGW_NPASS);
    if (st) return bpass(b, c, st);
    for (int i = 0; i < QGW_NPASS; i++) {
      m[i] += mm[i];
      n[i] += nn[i];
    }
  }
  for (int i = 0; i < npass && i < QGW_NPASS; i++) {
    if (ms) ms[i] = m[i];
    if (count) count[i] = n[i];
  }
  return QGW_OK;
}

int qgw_batch_device_bytes(qgw_batch *b, size_t *bytes) {
  if (!b || !bytes) return QGW_ERR_ARG;
  size_t s = 0;
  for (auto *c : b->groups) s += c->bytes;
  *bytes = s;
  return QGW_OK;
}

const char *qgw_batch_last_error(qgw_batch *b) { return b ? b->err : ""; }

}  // extern "C"
