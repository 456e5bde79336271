// dst_half.hpp -- DST-I of two real rows through one complex FFT of length
// M = N + 1 (half the length of the odd extension).
//
// RODFT00 (the transform of src/elliptic.h:63-66):
//   S_k = 2 sum_{j=1}^{N} x_j sin(pi j k / M),  k = 1..N,  M = N + 1.
// Two real rows a, b are packed as z = a + i b (z_0 = z_M = 0) and
//   y_j = sin(pi j / M) (z_j + z_{M-j}) + (z_j - z_{M-j}) / 2,   j = 0..M-1,
//   Y   = FFT_M(y)                       (e^{-2 pi i j k / M}).
// With Y_k = P + iQ and Y_{M-k} = U + iV, the two rows' transforms are
//   even:  S^a_{2k} = V - Q,            S^b_{2k} = P - U          (k = 1..M/2-1)
//   odd:   S_{2k+1} = sum_{k'=0}^{k} r_{k'},  r^a_k = P + U, r^b_k = Q + V,
//          with r_0 halved                                      (k = 0..M/2-1)
// -- the sine-transform recurrence of Press et al., Numerical Recipes, sec.
// 12.3 ("sinft"), applied to both halves of the complex FFT at once.  The
// running sum is a block-wide prefix sum: each thread owns M/(2 NT)
// consecutive k.  It is accumulated in FP64 in every form: its partial sums
// are of the size of the largest coefficient, so an FP32 sum would leave an
// error of ~eps |S_1| sqrt(k) on the small high-wavenumber coefficients,
// which the hyperviscous term amplifies.
#pragma once
#include "fft_reg.hpp"

namespace qgw {

template <int M, int EP = 0> struct DstHalf {
  static constexpr int E = RegPlan<M, EP>::E;
  static constexpr int NT = RegPlan<M, EP>::NT;
  static constexpr int KPT = M / (2 * NT);  // consecutive k per thread in the post pass
  // LDS: the FFT's own exchange buffer, and the natural-order buffer (rslot)
  static constexpr int slots = (home_slots<M, EP>() > M + M / 16 + 1) ? home_slots<M, EP>() : M + M / 16 + 1;
  static_assert(KPT >= 1 && M % (2 * NT) == 0, "post-pass split");
};

// Exclusive prefix sum over the block of one complex value per thread
// (the real and imaginary parts are independent sums).
template <typename T, int NT> __device__ __forceinline__ cx<T> block_excl_scan(cx<T> v, int t) {
  constexpr int WW = (NT < 64) ? NT : 64;  // threads per wavefront in use
  constexpr int NW = (NT + 63) / 64;
  __shared__ cx<T> wtot[NW];
  const int lane = t & 63, w = t >> 6;
  cx<T> s = v;
#pragma unroll
  for (int d = 1; d < WW; d <<= 1) {
    const T x = __shfl_up(s.x, d, 64);
    const T y = __shfl_up(s.y, d, 64);
    if (lane >= d) {
      s.x += x;
      s.y += y;
    }
  }
  if constexpr (NW > 1) {
    if (lane == 63) wtot[w] = s;
    __syncthreads();
    cx<T> off = mkc<T>(T(0), T(0));
#pragma unroll
    for (int i = 0; i < NW - 1; i++)
      if (i < w) off = cadd<T>(off, wtot[i]);
    s = cadd<T>(s, off);
  }
  return csub<T>(s, v);
}

// Pre pass: h[m] = y_p at the home positions p = t + m NT.  ld(j) returns
// z_j = (a_j, b_j) for j = 1..N (the caller's row access); z_0 = z_M = 0.
// tw2: the length-2M table e^{-2 pi i k / 2M}, so sin(pi p / M) = -Im tw2[p].
template <typename T, int M, int EP, class Ld>
__device__ __forceinline__ void dst_pre(cx<T> *h, int t, const cx<T> *__restrict__ tw2, Ld ld) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    if (p == 0) {
      h[m] = mkc<T>(T(0), T(0));
    } else {
      const cx<T> z = ld(p), zm = ld(M - p);
      const T s = -tw2[p].y;
      const T hx = T(0.5) * (z.x - zm.x), hy = T(0.5) * (z.y - zm.y);
      h[m] = mkc<T>(s * (z.x + zm.x) + hx, s * (z.y + zm.y) + hy);
    }
  }
}

// Post pass: from Y at the home positions to S^a + i S^b in buf, natural
// order (entry k-1 holds S_k, k = 1..N, at rslot(k-1)).  Ends with a barrier.
template <typename T, int M, int EP> __device__ __forceinline__ void dst_post(cx<T> *h, cx<T> *buf, int t) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT, KPT = DstHalf<M, EP>::KPT;
  __syncthreads();  // the FFT's last readers of buf are done
#pragma unroll
  for (int m = 0; m < E; m++) buf[rslot(t + m * NT)] = h[m];
  __syncthreads();
  cx<T> ev[KPT];
  cx<double> cum[KPT];
  cx<double> run = mkc<double>(0.0, 0.0);
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = t * KPT + i;
    const cx<T> Yk = buf[rslot(k)];
    const cx<T> Ym = buf[rslot(k == 0 ? 0 : M - k)];
    ev[i] = mkc<T>(Ym.y - Yk.y, Yk.x - Ym.x);
    cx<double> r = mkc<double>((double)Yk.x + (double)Ym.x, (double)Yk.y + (double)Ym.y);
    if (k == 0) r = mkc<double>(0.5 * r.x, 0.5 * r.y);
    run = cadd<double>(run, r);
    cum[i] = run;
  }
  const cx<double> off = block_excl_scan<double, NT>(run, t);
  __syncthreads();  // every read of Y is done
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = t * KPT + i;
    if (k >= 1) buf[rslot(2 * k - 1)] = ev[i];
    const cx<double> o = cadd<double>(off, cum[i]);
    buf[rslot(2 * k)] = mkc<T>((T)o.x, (T)o.y);
  }
  __syncthreads();
}

}  // namespace qgw
