// launch.hpp -- launch descriptors shared by the context code (qgw_ctx.hip)
// and the per-precision kernel instantiation units (qgw_launch_f32.hip,
// qgw_launch_f64.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "qgw_hip.h"
#include "kernels.hpp"
#include "fused.hpp"

namespace qgw {

// Layers per member handled by the layer<->mode projections (the reference's
// nl_max, src/qg.c:45); above 32 the projections run in 32 x 32 blocks.
constexpr int kMaxLayers = 1000;

// P4 tile shapes (columns x rows); the workgroup has 4*columns threads.  Variant
// kP4TileCount is the streaming form (k_p4s), kP4Aligned the streaming form
// with aligned row stores (k_p4a).
constexpr int kP4TileCount = 4;
constexpr int kP4Aligned = kP4TileCount + 1;
constexpr int kP4Variants = kP4TileCount + 2;
constexpr int kP4TileW[kP4TileCount] = {64, 64, 64, 128};
constexpr int kP4TileH[kP4TileCount] = {8, 16, 32, 16};

struct LaunchArgs {
  Geo g{};
  int prec = 0;
  int Lx = 0, Ly = 0;
  hipStream_t stream = nullptr;
  void *q = nullptr, *qn = nullptr, *psi = nullptr, *what = nullptr, *whatT = nullptr;
  void *cl2m = nullptr, *cm2l = nullptr, *twx = nullptr, *twy = nullptr;
  void *twxh = nullptr, *twyh = nullptr;  // length-(L/2) tables of the half-length transforms
  int dst_half = 0;                       // 1: P1/P2/P3 through FFTs of length L/2; 2: and P1/P3 pipelined (powers of two)
  int fft_e = 16;                         // elements per thread of those FFTs (8 or 16, lengths >= 1024)
  int fft_e_col = 16;                     // the same for the column pass P2
  int per = 0;                            // periodic domain (bc_fac = -1): Lx = Ly = N, FFT transforms
  void *fnew = nullptr, *f1 = nullptr, *f2 = nullptr, *comp = nullptr, *topo = nullptr, *wind = nullptr;
  const double *La2 = nullptr, *Kb2 = nullptr, *iRd2 = nullptr;
  double norm = 1.0, ring_c = 0.0;
  double dk = 0.0;  // y wavenumber step of the bounded transforms, pi / Lx
  int cus = 256;    // compute units of the device (persistent grids)
  StepCoef cf{};
  int nlap = 0;
  int p4_tile = kP4TileCount;  // P4 variant: a tile index, kP4TileCount (streaming) or kP4Aligned
  int p4_waves = 40;           // streaming P4: target wavefronts per SIMD (sets the row chunk)
  int t_kernel = 0;            // projection: 0 auto (matrix cores for FP32 at >= 16 layers), 1 VALU, 2 matrix cores
  int p2m_g = 0;               // four-pass column pass: column pairs per block (0: the most that fit)
  int psi_nt = 1;              // the half-length inverse row passes store psi non-temporally
  int p4_rev = 1;              // streaming P4: the interior walks of odd row chunks run up their chunk
  int p4_order = 1;            // streaming P4: 1 the layers of a row chunk before the next chunk, 0 the chunks first
  unsigned long long *cfl_max = nullptr;
};

// Per-precision entry points (defined in qgw_launch_f32.hip / qgw_launch_f64.hip).
template <typename T> int launch_p1_t(const LaunchArgs &a);
template <typename T> int launch_t1_t(const LaunchArgs &a);
template <typename T> int launch_p2_t(const LaunchArgs &a);
template <typename T> int launch_t2_t(const LaunchArgs &a);
template <typename T> int launch_p3_t(const LaunchArgs &a);
template <typename T> int launch_p4_t(const LaunchArgs &a);
// The four-pass inversion (fused.hpp): projections inside the row passes,
// column tiles in place on a.what.
template <typename T> int launch_p1m_t(const LaunchArgs &a);
template <typename T> int launch_p2m_t(const LaunchArgs &a);
template <typename T> int launch_p3m_t(const LaunchArgs &a);
template <typename T> int launch_cfl_t(const LaunchArgs &a);
template <typename T> int launch_diag_t(const LaunchArgs &a, double *part, int nblk, double *out);
template <typename T> int launch_from_f64_t(const LaunchArgs &a, int layers, const double *src, void *dst, void *resid);
template <typename T> int launch_to_f64_t(const LaunchArgs &a, int layers, const void *src, const void *resid, double *dst);
// Stochastic forcing, q += S s_fac on layer 0 of a.q: k_stoch_cols then k_stoch_rows.
struct StochArgs {
  int N = 0, nc = 0;
  const int *col_start = nullptr, *mode_j = nullptr, *mode_k = nullptr;
  const double *coef = nullptr, *tw = nullptr;
  const double *E = nullptr;  // nc x n weighted twiddles of the forced columns
  double *G = nullptr;
};
template <typename T> int launch_stoch_t(const LaunchArgs &a, const StochArgs &s, double s_fac);
// 4-D forcing on a.q (k_forcing), every layer.
template <typename T>
int launch_forcing_t(const LaunchArgs &a, const void *F1, const void *F2, double w1, double w2, double dt);
// Dense sine-transform helpers: Helmholtz division on a.whatT-shaped array
// `w`, and the psi/q wall rings after the dense inverse row transform.
template <typename T> int launch_helm_div_t(const LaunchArgs &a, void *w);
template <typename T> int launch_rings_t(const LaunchArgs &a);
bool launch_has_length(int L);
bool launch_has_half_length(int M);  // lengths M = L/2 of the half-length transforms

// The specialisations live in the two instantiation units.
#define QGW_DECLARE_SPECS(T)                                                                          \
  template <> int launch_p1_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_t1_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_p2_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_t2_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_p3_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_p4_t<T>(const LaunchArgs &a);                                                \
  template <> int launch_p1m_t<T>(const LaunchArgs &a);                                               \
  template <> int launch_p2m_t<T>(const LaunchArgs &a);                                               \
  template <> int launch_p3m_t<T>(const LaunchArgs &a);                                               \
  template <> int launch_cfl_t<T>(const LaunchArgs &a);                                               \
  template <> int launch_diag_t<T>(const LaunchArgs &a, double *part, int nblk, double *out);         \
  template <> int launch_from_f64_t<T>(const LaunchArgs &a, int layers, const double *src, void *dst, \
                                       void *resid);                                                  \
  template <> int launch_to_f64_t<T>(const LaunchArgs &a, int layers, const void *src, const void *resid, \
                                     double *dst);                                                        \
  template <> int launch_forcing_t<T>(const LaunchArgs &a, const void *F1, const void *F2, double w1,      \
                                      double w2, double dt);                                              \
  template <> int launch_stoch_t<T>(const LaunchArgs &a, const StochArgs &s, double s_fac);              \
  template <> int launch_helm_div_t<T>(const LaunchArgs &a, void *w);                                   \
  template <> int launch_rings_t<T>(const LaunchArgs &a);
QGW_DECLARE_SPECS(float)
QGW_DECLARE_SPECS(double)
#undef QGW_DECLARE_SPECS

#define QGW_DISPATCH(fn, a, ...) \
  ((a).prec == QGW_FP64 ? fn##_t<double>((a), ##__VA_ARGS__) : fn##_t<float>((a), ##__VA_ARGS__))

inline int launch_p1(const LaunchArgs &a) { return QGW_DISPATCH(launch_p1, a); }
inline int launch_t1(const LaunchArgs &a) { return QGW_DISPATCH(launch_t1, a); }
inline int launch_p2(const LaunchArgs &a) { return QGW_DISPATCH(launch_p2, a); }
inline int launch_t2(const LaunchArgs &a) { return QGW_DISPATCH(launch_t2, a); }
inline int launch_p3(const LaunchArgs &a) { return QGW_DISPATCH(launch_p3, a); }
inline int launch_p4(const LaunchArgs &a) { return QGW_DISPATCH(launch_p4, a); }
inline int launch_p1m(const LaunchArgs &a) { return QGW_DISPATCH(launch_p1m, a); }
inline int launch_p2m(const LaunchArgs &a) { return QGW_DISPATCH(launch_p2m, a); }
inline int launch_p3m(const LaunchArgs &a) { return QGW_DISPATCH(launch_p3m, a); }
inline int launch_cfl(const LaunchArgs &a) { return QGW_DISPATCH(launch_cfl, a); }
inline int launch_diag(const LaunchArgs &a, double *part, int nblk, double *out) {
  return QGW_DISPATCH(launch_diag, a, part, nblk, out);
}
inline int launch_from_f64(const LaunchArgs &a, int layers, const double *src, void *dst, void *resid) {
  return QGW_DISPATCH(launch_from_f64, a, layers, src, dst, resid);
}
inline int launch_to_f64(const LaunchArgs &a, int layers, const void *src, const void *resid, double *dst) {
  return QGW_DISPATCH(launch_to_f64, a, layers, src, resid, dst);
}
inline int launch_stoch(const LaunchArgs &a, const StochArgs &s, double s_fac) {
  return QGW_DISPATCH(launch_stoch, a, s, s_fac);
}
inline int launch_forcing(const LaunchArgs &a, const void *F1, const void *F2, double w1, double w2, double dt) {
  return QGW_DISPATCH(launch_forcing, a, F1, F2, w1, w2, dt);
}
inline int launch_helm_div(const LaunchArgs &a, void *w) { return QGW_DISPATCH(launch_helm_div, a, w); }
inline int launch_rings(const LaunchArgs &a) { return QGW_DISPATCH(launch_rings, a); }

}  // namespace qgw
