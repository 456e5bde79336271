// launch_impl.hpp -- kernel instantiation and launch for one arithmetic type.
// Included exactly once by qgw_launch_f32.hip (QGW_T = float) and by
// qgw_launch_f64.hip (QGW_T = double).
#pragma once
#ifndef QGW_T
#error "define QGW_T before including launch_impl.hpp"
#endif
#include "launch.hpp"

namespace qgw {

using T = QGW_T;

// Supported transform lengths L = 2(n+1) (bounded) or n (periodic): 2^a
// (8..8192), 3*2^a (24..6144), the 5^b 2^a lengths of the grids NX = 100,
// 200, 500, 1000 and 2000 (L = 200, 400, 1000, 2000, 4000; periodic NX = 100
// is L = 100) and 7*2^6 (NX = 224).
#define QGW_LENGTHS(X)                                                           \
  X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192) \
  X(24) X(48) X(96) X(192) X(384) X(768) X(1536) X(3072) X(6144)              \
  X(100) X(200) X(400) X(1000) X(2000) X(4000) X(448)

// Lengths M of the half-length transforms (M = L/2 >= 64; M even).
#define QGW_HLENGTHS(X)                                                                                  \
  X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(96) X(192) X(384) X(768) X(1536) X(3072) \
  X(100) X(200) X(500) X(1000) X(2000) X(224)
// Radix schedule of a half-length transform: powers of two from 1024 run with
// 16 (radix-16 stages) or 8 (radix-8) elements per thread: a.fft_e for the
// row passes, a.fft_e_col for the column pass.
#define QGW_HALF_EP(M) QGW_HALF_EP_E(M, a.fft_e)
#define QGW_HALF_EP_E(M, EV)                              \
  if constexpr (is_pow2(M) && (M) >= 1024) {              \
    if ((EV) == 8) {                                      \
      QGW_GO(M, 8)                                        \
    } else {                                              \
      QGW_GO(M, 16)                                       \
    }                                                     \
  } else {                                                \
    QGW_GO(M, 0)                                          \
  }

// Persistent grid of a pipelined kernel: its resident blocks on every compute
// unit, at most one block per item.
template <class K> static unsigned persist_grid(K kern, int nt, long nitems, int cus, int &occ) {
  if (occ <= 0) {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, reinterpret_cast<const void *>(kern), nt, 0) != hipSuccess ||
        o < 1)
      o = 1;
    occ = o;
  }
  const long g = (long)occ * (cus > 0 ? cus : 1);
  return (unsigned)(g < nitems ? g : nitems);
}

// Half-length transforms of length M: for the row passes P1/P3 the pipelined
// form under dst_half == 2, else one block per item.
template <int M, int EP> static void p1h_go(const LaunchArgs &a, dim3 grid) {
  if (a.dst_half == 2) {
    static int occ = 0;
    const unsigned pg = persist_grid(k_p1hp<T, M, EP>, DstHalf<M, EP>::NT, (long)grid.x, a.cus, occ);
    hipLaunchKernelGGL((k_p1hp<T, M, EP>), dim3(pg), dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (const T *)a.q,
                       (T *)a.what, (const cx<T> *)a.twxh, (const cx<T> *)a.twx);
    return;
  }
  hipLaunchKernelGGL((k_p1h<T, M, EP>), grid, dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (const T *)a.q,
                     (T *)a.what, (const cx<T> *)a.twxh, (const cx<T> *)a.twx);
}

template <int M, int EP> static void p3h_go(const LaunchArgs &a, dim3 grid) {
  if (a.dst_half == 2) {
    static int occ = 0;
    const unsigned pg = persist_grid(k_p3hp<T, M, EP>, DstHalf<M, EP>::NT, (long)grid.x, a.cus, occ);
    hipLaunchKernelGGL((k_p3hp<T, M, EP>), dim3(pg), dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g,
                       (const T *)a.what, (T *)a.psi, (T *)a.q, (T)a.ring_c, (const cx<T> *)a.twxh,
                       (const cx<T> *)a.twx, a.psi_nt);
    return;
  }
  hipLaunchKernelGGL((k_p3h<T, M, EP>), grid, dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (const T *)a.what,
                     (T *)a.psi, (T *)a.q, (T)a.ring_c, (const cx<T> *)a.twxh, (const cx<T> *)a.twx, a.psi_nt);
}

template <int M, int EP> static void p2h_go(const LaunchArgs &a, dim3 grid) {
  hipLaunchKernelGGL((k_p2h<T, M, EP>), grid, dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (T *)a.whatT, a.La2,
                     a.iRd2, a.dk, a.norm, (const cx<T> *)a.twyh, (const cx<T> *)a.twy);
}

template <> int launch_p1_t<T>(const LaunchArgs &a) {
  const dim3 grid(a.g.m * ((a.g.nl + 1) / 2));
  if (a.per) {
    switch (a.Lx) {
#define QGW_CASE(L)                                                                                  \
  case L:                                                                                            \
    hipLaunchKernelGGL((k_pf1<T, L>), grid, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (const T *)a.q, \
                       (T *)a.what, (const cx<T> *)a.twx);                                           \
    return 0;
      QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
    }
    return QGW_ERR_ARG;
  }
  if (a.dst_half) {
    switch (a.Lx / 2) {
#define QGW_GO(M, EP) p1h_go<M, EP>(a, grid);
#define QGW_CASE(M) \
  case M:           \
    QGW_HALF_EP(M)  \
    return 0;
      QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
    }
    return QGW_ERR_ARG;
  }
  switch (a.Lx) {
#define QGW_CASE(L)                                                                                 \
  case L:                                                                                           \
    hipLaunchKernelGGL((k_p1<T, L>), grid, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (const T *)a.q, \
                       (T *)a.what, (const cx<T> *)a.twx);                                          \
    return 0;
    QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
  }
  return QGW_ERR_ARG;
}

template <> int launch_p3_t<T>(const LaunchArgs &a) {
  const dim3 grid(a.g.m * ((a.g.nl + 1) / 2));
  if (a.per) {
    switch (a.Lx) {
#define QGW_CASE(L)                                                                                     \
  case L:                                                                                               \
    hipLaunchKernelGGL((k_pf3<T, L>), grid, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (const T *)a.what, \
                       (T *)a.psi, (T *)a.q, (const cx<T> *)a.twx);                                     \
    return 0;
      QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
    }
    return QGW_ERR_ARG;
  }
  if (a.dst_half) {
    switch (a.Lx / 2) {
#define QGW_GO(M, EP) p3h_go<M, EP>(a, grid);
#define QGW_CASE(M) \
  case M:           \
    QGW_HALF_EP(M)  \
    return 0;
      QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
    }
    return QGW_ERR_ARG;
  }
  switch (a.Lx) {
#define QGW_CASE(L)                                                                                    \
  case L:                                                                                              \
    hipLaunchKernelGGL((k_p3<T, L>), grid, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (const T *)a.what, \
                       (T *)a.psi, (T *)a.q, (T)a.ring_c, (const cx<T> *)a.twx);                       \
    return 0;
    QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
  }
  return QGW_ERR_ARG;
}

template <> int launch_p2_t<T>(const LaunchArgs &a) {
  const long rows = (long)a.g.nl * a.g.n;
  const dim3 grid((unsigned)((rows + 1) / 2));
  if (a.per) {
    const dim3 gp((unsigned)((long)a.g.nl * (a.g.n / 2 + 1)));
    switch (a.Ly) {
#define QGW_CASE(L)                                                                                 \
  case L:                                                                                           \
    hipLaunchKernelGGL((k_pf2<T, L>), gp, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (T *)a.whatT,    \
                       a.La2, a.Kb2, a.iRd2, a.norm, (const cx<T> *)a.twy);                         \
    return 0;
      QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
    }
    return QGW_ERR_ARG;
  }
  if (a.dst_half) {
    switch (a.Ly / 2) {
#define QGW_GO(M, EP) p2h_go<M, EP>(a, grid);
#define QGW_CASE(M) \
  case M:           \
    QGW_HALF_EP_E(M, a.fft_e_col)  \
    return 0;
      QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
    }
    return QGW_ERR_ARG;
  }
  switch (a.Ly) {
#define QGW_CASE(L)                                                                               \
  case L:                                                                                         \
    hipLaunchKernelGGL((k_p2<T, L>), grid, dim3(RegPlan<L>::NT), 0, a.stream, a.g, (T *)a.whatT, \
                       a.La2, a.Kb2, a.iRd2, a.norm, (const cx<T> *)a.twy);                       \
    return 0;
    QGW_LENGTHS(QGW_CASE)
#undef QGW_CASE
  }
  return QGW_ERR_ARG;
}

// ---- four-pass inversion ---------------------------------------------------
template <int M, int EP> static void p1m_go(const LaunchArgs &a) {
  static int occ = 0;
  const long items = (long)a.g.m * (a.g.nl / a.g.nlm) * ((a.g.nlm + 1) / 2);
  const unsigned pg = persist_grid(k_p1m<T, M, EP>, DstHalf<M, EP>::NT, items, a.cus, occ);
  hipLaunchKernelGGL((k_p1m<T, M, EP>), dim3(pg), dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (const T *)a.q,
                     (T *)a.what, (const T *)a.cl2m, (const cx<T> *)a.twxh, (const cx<T> *)a.twx);
}
template <int M, int EP> static void p3m_go(const LaunchArgs &a) {
  static int occ = 0;
  const long items = (long)a.g.m * (a.g.nl / a.g.nlm) * ((a.g.nlm + 1) / 2);
  const unsigned pg = persist_grid(k_p3m<T, M, EP>, DstHalf<M, EP>::NT, items, a.cus, occ);
  hipLaunchKernelGGL((k_p3m<T, M, EP>), dim3(pg), dim3(DstHalf<M, EP>::NT), 0, a.stream, a.g, (const T *)a.what,
                     (T *)a.psi, (T *)a.q, (const T *)a.cm2l, (T)a.ring_c, (const cx<T> *)a.twxh,
                     (const cx<T> *)a.twx, a.psi_nt);
}
template <int M, int EP, int G> static void p2m_launch(const LaunchArgs &a) {
  const int ncg = (a.g.n + 2 * G - 1) / (2 * G);
  const long tiles = (long)ncg * a.g.nl;
  hipLaunchKernelGGL((k_p2m<T, M, EP, G>), dim3((unsigned)tiles), dim3(G * DstHalf<M, EP>::NT), 0, a.stream, a.g,
                     (T *)a.what, a.La2, a.iRd2, a.dk, a.norm, (const cx<T> *)a.twyh, (const cx<T> *)a.twy, ncg);
}
// a.p2m_g column pairs per block, capped by what the block size and LDS allow
template <int M, int EP> static void p2m_go(const LaunchArgs &a) {
  constexpr int GM = p2m_groups<T, M, EP>();
  const int want = a.p2m_g > 0 ? a.p2m_g : GM;
  if (want >= 8 && GM >= 8) p2m_launch<M, EP, (GM >= 8 ? 8 : 1)>(a);
  else if (want >= 4 && GM >= 4) p2m_launch<M, EP, (GM >= 4 ? 4 : 1)>(a);
  else if (want >= 2 && GM >= 2) p2m_launch<M, EP, (GM >= 2 ? 2 : 1)>(a);
  else p2m_launch<M, EP, 1>(a);
}

template <> int launch_p1m_t<T>(const LaunchArgs &a) {
  if (a.per || a.g.nlm < 1 || a.g.nl % a.g.nlm) return QGW_ERR_ARG;
  switch (a.Lx / 2) {
#define QGW_GO(M, EP) p1m_go<M, EP>(a);
#define QGW_CASE(M) \
  case M:           \
    QGW_HALF_EP(M)  \
    return 0;
    QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
  }
  return QGW_ERR_ARG;
}

template <> int launch_p3m_t<T>(const LaunchArgs &a) {
  if (a.per || a.g.nlm < 1 || a.g.nl % a.g.nlm) return QGW_ERR_ARG;
  switch (a.Lx / 2) {
#define QGW_GO(M, EP) p3m_go<M, EP>(a);
#define QGW_CASE(M) \
  case M:           \
    QGW_HALF_EP(M)  \
    return 0;
    QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
  }
  return QGW_ERR_ARG;
}

template <> int launch_p2m_t<T>(const LaunchArgs &a) {
  if (a.per) return QGW_ERR_ARG;
  switch (a.Ly / 2) {
#define QGW_GO(M, EP) p2m_go<M, EP>(a);
#define QGW_CASE(M)                \
  case M:                          \
    QGW_HALF_EP_E(M, a.fft_e_col)  \
    return 0;
    QGW_HLENGTHS(QGW_CASE)
#undef QGW_CASE
#undef QGW_GO
  }
  return QGW_ERR_ARG;
}

// Transpose tiles: 32 columns (contiguous in the source) x kTR rows
// (contiguous in the destination): 128-byte segments on both sides.
constexpr int kTC = 32;
constexpr int kTR = (sizeof(T) == 8) ? 16 : 32;
constexpr int kTKB = 1;  // modes per barrier pair

template <int NLC, bool ACC = false>
static void tproj_go(const LaunchArgs &a, const void *src, void *dst, const void *M, int R, int C, int Ps, long Ss,
                     int Pd, long Sd, int nlc = 0, int nk = 0) {
  const int nl = a.g.nlm;
  const dim3 grid((C + kTC - 1) / kTC, (R + kTR - 1) / kTR, a.g.nl / nl);
  hipLaunchKernelGGL((k_tproj<T, kTC, kTR, NLC, kTKB, ACC>), grid, dim3(256), 0, a.stream, (const T *)src, (T *)dst,
                     (const T *)M, nlc ? nlc : nl, nk ? nk : nl, nl, R, C, Ps, Ss, Pd, Sd, (long)nl * Ss,
                     (long)nl * Sd, (long)nl * nl);
}

// More than 32 layers: the contraction in blocks of 32 layers x 32 modes, each
// mode block stored by its first layer block and accumulated by the others.
static int tproj_blocks(const LaunchArgs &a, const void *src, void *dst, const void *M, int R, int C, int Ps,
                        long Ss, int Pd, long Sd) {
  constexpr int B = 32;
  const int nl = a.g.nlm;
  for (int k0 = 0; k0 < nl; k0 += B) {
    const int nk = (nl - k0 < B) ? nl - k0 : B;
    for (int l0 = 0; l0 < nl; l0 += B) {
      const int nlc = (nl - l0 < B) ? nl - l0 : B;
      const T *s0 = (const T *)src + (long)l0 * Ss;
      T *d0 = (T *)dst + (long)k0 * Sd;
      const T *m0 = (const T *)M + (long)k0 * nl + l0;
      if (l0 == 0)
        tproj_go<B, false>(a, s0, d0, m0, R, C, Ps, Ss, Pd, Sd, nlc, nk);
      else
        tproj_go<B, true>(a, s0, d0, m0, R, C, Ps, Ss, Pd, Sd, nlc, nk);
    }
  }
  return 0;
}

static int tproj(const LaunchArgs &a, const void *src, void *dst, const void *M, int R, int C, int Ps, long Ss,
                 int Pd, long Sd) {
  const int nl = a.g.nlm;
  if (nl < 1 || a.g.nl % nl != 0 || nl > kMaxLayers) return QGW_ERR_ARG;
  if (nl > 32) return tproj_blocks(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  const bool mfma = sizeof(T) == 4 && (a.t_kernel >= 2 || (a.t_kernel == 0 && nl >= 16));
  if (mfma) {
    const int tcb = (a.t_kernel == 3) ? 16 : 32;
    const int ntc = (C + tcb - 1) / tcb, ntr = (R + 31) / 32;
    const long blocks = (long)ntc * ntr * (a.g.nl / nl);
    if (blocks > 0x7fffffffL) return QGW_ERR_ARG;
    if (tcb == 16)
      hipLaunchKernelGGL((k_tproj_mfma<8, 16>), dim3((unsigned)blocks), dim3(256), 0, a.stream, (const float *)src,
                         (float *)dst, (const float *)M, nl, R, C, Ps, Ss, Pd, Sd, ntc, ntr);
    else
      hipLaunchKernelGGL((k_tproj_mfma<8, 32>), dim3((unsigned)blocks), dim3(256), 0, a.stream, (const float *)src,
                         (float *)dst, (const float *)M, nl, R, C, Ps, Ss, Pd, Sd, ntc, ntr);
  } else if (nl <= 4) {
    tproj_go<4>(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  } else if (nl <= 8) {
    tproj_go<8>(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  } else if (nl <= 12) {
    tproj_go<12>(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  } else if (nl <= 16) {
    tproj_go<16>(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  } else {
    tproj_go<32>(a, src, dst, M, R, C, Ps, Ss, Pd, Sd);
  }
  return 0;
}

// Columns of a what row: n spectral values, or n + 2 reals (N/2 + 1 complex,
// interleaved) in the periodic form.
static int what_cols(const LaunchArgs &a) { return a.per ? a.g.n + 2 : a.g.n; }

template <> int launch_t1_t<T>(const LaunchArgs &a) {
  // what (layers) [l][j][a] -> whatT (modes) [k][a][j], w_k = sum_l cl2m[k,l] w_l
  return tproj(a, a.what, a.whatT, a.cl2m, a.g.m, what_cols(a), a.g.Pn, a.g.WS, a.g.Pm, a.g.WTS);
}

template <> int launch_t2_t<T>(const LaunchArgs &a) {
  // whatT (modes) [k][a][j] -> what (layers) [l][j][a], w_l = sum_k cm2l[l,k] w_k
  return tproj(a, a.whatT, a.what, a.cm2l, what_cols(a), a.g.m, a.g.Pm, a.g.WTS, a.g.Pn, a.g.WS);
}

template <int NLAP, int KAHAN, int TX, int TY, bool PER = false> static void p4_go(const LaunchArgs &a) {
  const dim3 grid((a.g.n + TX - 1) / TX, (a.g.m + TY - 1) / TY, a.g.nl);
  hipLaunchKernelGGL((k_p4<T, NLAP, KAHAN, TX, TY, PER>), grid, dim3(TX * 4), 0, a.stream, a.g, (const T *)a.psi,
                     (const T *)a.q, (T *)a.qn, (T *)a.fnew, (const T *)a.f1, (const T *)a.f2, (comp_t<T> *)a.comp,
                     (const T *)a.topo, (const T *)a.wind, a.cf);
}

template <int NLAP, int KAHAN> static void p4s_go(const LaunchArgs &a) {
  constexpr int H = (NLAP > 1) ? NLAP : 1;
  const int ow = 64 - 2 * H;
  const int nstrip = (a.g.n + ow - 1) / ow;
  // rows per wavefront: about p4_waves wavefronts per SIMD (256 CUs x 4 SIMDs)
  const long lanes = (long)nstrip * a.g.nl;
  const long target = 1024L * (a.p4_waves > 0 ? a.p4_waves : 40);
  int chunk = (int)(((long)a.g.m * lanes + target - 1) / target);
  chunk = chunk < 16 ? 16 : (chunk > 512 ? 512 : chunk);
  const int nchunk = (a.g.m + chunk - 1) / chunk;
  const long waves = lanes * nchunk;
  const dim3 grid((unsigned)((waves + 3) / 4));
  hipLaunchKernelGGL((k_p4s<T, NLAP, KAHAN>), grid, dim3(256), 0, a.stream, a.g, (const T *)a.psi, (const T *)a.q,
                     (T *)a.qn, (T *)a.fnew, (const T *)a.f1, (const T *)a.f2, (comp_t<T> *)a.comp, (const T *)a.topo,
                     (const T *)a.wind, a.cf, nstrip, chunk, nchunk, a.p4_rev, a.p4_order);
}

// the streaming form over aligned column blocks: blocks of 4 wavefronts over
// p4a_cols output columns, the row chunk set as for k_p4s
template <int NLAP, int KAHAN> static void p4a_go(const LaunchArgs &a) {
  constexpr int BW = p4a_cols<NLAP>();
  const int nbx = (a.g.n + BW - 1) / BW;
  const long lanes = (long)nbx * 4 * a.g.nl;
  const long target = 1024L * (a.p4_waves > 0 ? a.p4_waves : 40);
  int chunk = (int)(((long)a.g.m * lanes + target - 1) / target);
  chunk = chunk < 16 ? 16 : (chunk > 512 ? 512 : chunk);
  const int nchunk = (a.g.m + chunk - 1) / chunk;
  const dim3 grid((unsigned)((long)nbx * nchunk * a.g.nl));
  hipLaunchKernelGGL((k_p4a<T, NLAP, KAHAN>), grid, dim3(256), 0, a.stream, a.g, (const T *)a.psi,
                     (const T *)a.q, (T *)a.qn, (T *)a.fnew, (const T *)a.f1, (const T *)a.f2, (comp_t<T> *)a.comp,
                     (const T *)a.topo, (const T *)a.wind, a.cf, nbx, chunk, nchunk, a.p4_rev, a.p4_order);
}

template <int NLAP, int KAHAN> static int p4_tile(const LaunchArgs &a) {
  static_assert(kP4TileCount == 4, "tile table");
  if (a.per) {  // periodic: the 64x16 tile with wrapped halo loads
    p4_go<NLAP, KAHAN, kP4TileW[1], kP4TileH[1], true>(a);
    return 0;
  }
  if (a.p4_tile == kP4Aligned) {
    p4a_go<NLAP, KAHAN>(a);
    return 0;
  }
  if constexpr (NLAP > 5) {  // deep hyperviscosity (n_hyper > 8): the streaming forms only
    p4s_go<NLAP, KAHAN>(a);
    return 0;
  } else {
    switch (a.p4_tile) {
      case kP4TileCount: p4s_go<NLAP, KAHAN>(a); return 0;
      case 0: p4_go<NLAP, KAHAN, kP4TileW[0], kP4TileH[0]>(a); return 0;
      case 1: p4_go<NLAP, KAHAN, kP4TileW[1], kP4TileH[1]>(a); return 0;
      case 2: p4_go<NLAP, KAHAN, kP4TileW[2], kP4TileH[2]>(a); return 0;
      case 3: p4_go<NLAP, KAHAN, kP4TileW[3], kP4TileH[3]>(a); return 0;
    }
    return QGW_ERR_ARG;
  }
}

template <int KAHAN> static int p4_nlap(const LaunchArgs &a) {
  switch (a.nlap) {
    case 0: return p4_tile<0, KAHAN>(a);
    case 1: return p4_tile<1, KAHAN>(a);
    case 2: return p4_tile<2, KAHAN>(a);
    case 3: return p4_tile<3, KAHAN>(a);
    case 4: return p4_tile<4, KAHAN>(a);
    case 5: return p4_tile<5, KAHAN>(a);
    case 6: return p4_tile<6, KAHAN>(a);
    case 7: return p4_tile<7, KAHAN>(a);
    case 8: return p4_tile<8, KAHAN>(a);
    case 9: return p4_tile<9, KAHAN>(a);
    case 10: return p4_tile<10, KAHAN>(a);
  }
  return QGW_ERR_ARG;
}

template <> int launch_p4_t<T>(const LaunchArgs &a) {
  if (sizeof(T) == 4 && a.comp) return p4_nlap<1>(a);
  return p4_nlap<0>(a);
}

template <> int launch_cfl_t<T>(const LaunchArgs &a) {
  long I = (long)a.g.n * a.g.m * a.g.nl;
  long blocks = (I + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL((k_cfl<T>), dim3((unsigned)blocks), dim3(256), 0, a.stream, a.g, (const T *)a.psi,
                     a.cfl_max);
  return 0;
}

template <> int launch_diag_t<T>(const LaunchArgs &a, double *part, int nblk, double *out) {
  hipLaunchKernelGGL((k_diag_partial<T>), dim3(nblk, a.g.nl), dim3(256), 0, a.stream, a.g, (const T *)a.psi,
                     (const T *)a.q, (const comp_t<T> *)a.comp, part, nblk);
  hipLaunchKernelGGL((k_diag_final<256>), dim3(a.g.nl), dim3(256), 0, a.stream, (const double *)part, nblk, out);
  return 0;
}

template <>
int launch_from_f64_t<T>(const LaunchArgs &a, int layers, const double *src, void *dst, void *resid) {
  long tot = (long)(a.g.n + 2) * (a.g.m + 2) * layers;
  long blocks = (tot + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_from_f64<T>), dim3((unsigned)blocks), dim3(256), 0, a.stream, a.g, layers, src, (T *)dst,
                     (comp_t<T> *)resid);
  return 0;
}

template <> int launch_stoch_t<T>(const LaunchArgs &a, const StochArgs &s, double s_fac) {
  // the N x N field covers the m x n interior (N = n+1 bounded, n periodic)
  if (s.N < 2 || s.nc < 1 || a.g.n > s.N || a.g.m > s.N) return QGW_ERR_ARG;
  const long cells = (long)s.N * s.nc;
  hipLaunchKernelGGL((k_stoch_cols<T>), dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, a.stream, s.N, s.nc,
                     s.col_start, s.mode_j, s.mode_k, (const double2 *)s.coef, (const double2 *)s.tw,
                     (double2 *)s.G);
  constexpr int SR = 8;
  const dim3 grid((unsigned)((a.g.m + SR - 1) / SR), (unsigned)((a.g.n + 255) / 256));
  if (sizeof(T) == 4 && a.comp)
    hipLaunchKernelGGL((k_stoch_rows<T, SR, true>), grid, dim3(256), 0, a.stream, a.g, s.nc, (const double2 *)s.G,
                       (const double2 *)s.E, s_fac, (T *)a.q, (comp_t<T> *)a.comp);
  else
    hipLaunchKernelGGL((k_stoch_rows<T, SR, false>), grid, dim3(256), 0, a.stream, a.g, s.nc, (const double2 *)s.G,
                       (const double2 *)s.E, s_fac, (T *)a.q, (comp_t<T> *)nullptr);
  return 0;
}

template <>
int launch_forcing_t<T>(const LaunchArgs &a, const void *F1, const void *F2, double w1, double w2, double dt) {
  if (!F1 || !F2 || a.g.n < 1 || a.g.m < 1) return QGW_ERR_ARG;
  const dim3 grid((unsigned)(a.g.m * a.g.nl), (unsigned)((a.g.n + 255) / 256));
  if (sizeof(T) == 4 && a.comp)
    hipLaunchKernelGGL((k_forcing<T, true>), grid, dim3(256), 0, a.stream, a.g, (T *)a.q, (comp_t<T> *)a.comp,
                       (const T *)F1, (const T *)F2, w1, w2, dt);
  else
    hipLaunchKernelGGL((k_forcing<T, false>), grid, dim3(256), 0, a.stream, a.g, (T *)a.q, (comp_t<T> *)nullptr,
                       (const T *)F1, (const T *)F2, w1, w2, dt);
  return 0;
}

template <> int launch_helm_div_t<T>(const LaunchArgs &a, void *w) {
  if (a.g.n < 1 || a.g.m < 1 || a.g.nl < 1 || a.g.n > 65535 || a.g.nl > 65535) return QGW_ERR_ARG;
  const dim3 grid((unsigned)((a.g.m + 255) / 256), (unsigned)a.g.n, (unsigned)a.g.nl);
  hipLaunchKernelGGL((k_helm_div<T>), grid, dim3(256), 0, a.stream, a.g, (T *)w, a.La2, a.Kb2, a.iRd2, a.norm);
  return 0;
}

template <> int launch_rings_t<T>(const LaunchArgs &a) {
  if (a.g.nl < 1 || a.g.nl > 65535) return QGW_ERR_ARG;
  const int per = 2 * (a.g.n + 2) + 2 * a.g.m;
  const dim3 grid((unsigned)((per + 255) / 256), (unsigned)a.g.nl);
  hipLaunchKernelGGL((k_rings<T>), grid, dim3(256), 0, a.stream, a.g, (T *)a.psi, (T *)a.q, (T)a.ring_c);
  return 0;
}

template <>
int launch_to_f64_t<T>(const LaunchArgs &a, int layers, const void *src, const void *resid, double *dst) {
  long tot = (long)(a.g.n + 2) * (a.g.m + 2) * layers;
  long blocks = (tot + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((k_to_f64<T>), dim3((unsigned)blocks), dim3(256), 0, a.stream, a.g, layers, (const T *)src,
                     (const comp_t<T> *)resid, dst);
  return 0;
}

}  // namespace qgw
