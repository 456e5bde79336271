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

// DPP permutation of a double (two 32-bit moves); lanes with no source lane,
// and rows outside ROWS, receive 0.
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, ROWS, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
// Inclusive sum over the 64 lanes of a wavefront: row_shr 1, 2, 4, 8 within
// rows of 16 lanes, then row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3).
__device__ __forceinline__ double wave_incl_scan(double s) {
  s += dpp_d<0x111, 0xF>(s);
  s += dpp_d<0x112, 0xF>(s);
  s += dpp_d<0x114, 0xF>(s);
  s += dpp_d<0x118, 0xF>(s);
  s += dpp_d<0x142, 0xA>(s);
  s += dpp_d<0x143, 0xC>(s);
  return s;
}
// Exclusive prefix sum over the block of one complex value per thread
// (the real and imaginary parts are independent sums).  The DPP moves carry
// values from lower to higher lanes only, so a partial last wavefront (NT not
// a multiple of 64) sums correctly in its active lanes.
// wtot: NT/64 entries of LDS for the wavefront totals (its own per transform
// when a block runs several transforms side by side); null: a static array.
template <int NT> __device__ __forceinline__ cx<double> block_excl_scan(cx<double> v, int t, cx<double> *wext) {
  constexpr int NW = (NT + 63) / 64;
  __shared__ cx<double> wown[NW];
  cx<double> *wtot = wext ? wext : wown;
  const int lane = t & 63, w = t >> 6;
  const double sx = wave_incl_scan(v.x), sy = wave_incl_scan(v.y);
  cx<double> e = mkc<double>(dpp_d<0x138, 0xF>(sx), dpp_d<0x138, 0xF>(sy));  // wave_shr:1: lane i - 1's sum
  if constexpr (NW > 1) {
    if (lane == 63) wtot[w] = mkc<double>(sx, sy);
    __syncthreads();
    cx<double> off = mkc<double>(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < NW - 1; i++)
      if (i < w) off = cadd<double>(off, wtot[i]);
    e = cadd<double>(off, e);
  }
  return e;
}

// Pre pass: h[m] = y_p at the home positions p = t + m NT.  ld(j) returns
// z_j = (a_j, b_j) for j = 1..N (the caller's row access); z_0 = z_M = 0.
// tw2: the length-2M table e^{-2 pi i k / 2M}, so sin(pi p / M) = -Im tw2[p].
// sn(p) returns sin(pi p / M).
template <typename T, int M, int EP, class Sn, class Ld>
__device__ __forceinline__ void dst_pre_s(cx<T> *h, int t, const Sn &sn, Ld ld) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
#pragma unroll
  for (int m = 0; m < E; m++) {
    // branch-free: p = 0 (y_0 = 0) reads the valid pair of p = 1 and is zeroed
    // after, so no wait is forced inside a divergent branch
    const int p = t + m * NT;
    const int pc = (p == 0) ? 1 : p;
    const cx<T> z = ld(pc), zm = ld(M - pc);
    const T s = sn(pc);
    const T hx = T(0.5) * (z.x - zm.x), hy = T(0.5) * (z.y - zm.y);
    const cx<T> y = mkc<T>(s * (z.x + zm.x) + hx, s * (z.y + zm.y) + hy);
    h[m] = (p == 0) ? mkc<T>(T(0), T(0)) : y;
  }
}
template <typename T, int M, int EP, class Ld>
__device__ __forceinline__ void dst_pre(cx<T> *h, int t, const cx<T> *__restrict__ tw2, Ld ld) {
  dst_pre_s<T, M, EP>(h, t, [&](int p) { return -tw2[p].y; }, ld);
}
// sin(pi p / M) = -Im e^{-2 pi i p / 2M} from the two-level table of length 2M
template <typename T, int M> struct SinLds {
  const TwTab<T, 2 * M> *tb;
  __device__ __forceinline__ T operator()(int p) const {
    const cx<T> a = tb->hi[p >> 6], b = tb->lo[p & 63];
    return -(a.x * b.y + a.y * b.x);
  }
};

// Raw buffer loads of one element: the hardware range check returns 0 for a
// byte offset at or beyond the record size, which also covers an offset of -1
// element (it wraps to 2^32 - size).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void *p, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, int off) {
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
// Pre pass from two rows in global memory, element x - 1 holding z_x (x = 1..N),
// through buffer loads: the range check gives z_0 = z_M = 0 (offsets of -1 and
// N elements), so p = 0 needs no clamp; b1 has a record size of 0 when the
// second row does not exist.  The sines are the imaginary parts of the
// length-2M table tw2, loaded alone.
template <typename T, int M, int EP>
__device__ __forceinline__ void dst_pre_rows(cx<T> *h, int t, __amdgpu_buffer_rsrc_t b0, __amdgpu_buffer_rsrc_t b1,
                                             const cx<T> *__restrict__ tw2) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT, S = (int)sizeof(T);
  const __amdgpu_buffer_rsrc_t bs = row_rsrc(tw2, 2 * M * 2 * S);
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    const cx<T> z = mkc<T>(buf_ld<T>(b0, (p - 1) * S), buf_ld<T>(b1, (p - 1) * S));
    const cx<T> zm = mkc<T>(buf_ld<T>(b0, (M - 1 - p) * S), buf_ld<T>(b1, (M - 1 - p) * S));
    const T s = -buf_ld<T>(bs, p * 2 * S + S);
    const T hx = T(0.5) * (z.x - zm.x), hy = T(0.5) * (z.y - zm.y);
    const cx<T> y = mkc<T>(s * (z.x + zm.x) + hx, s * (z.y + zm.y) + hy);
    h[m] = (m == 0 && t == 0) ? mkc<T>(T(0), T(0)) : y;
  }
}

// Pre pass from LDS: z_x at entry rslot(x) (x = 1..N), the sines from the
// two-level table of length 2M.  With NT a multiple of 64 every address is a
// per-thread base plus a compile-time offset: for u a multiple of 16,
// rslot(u + t) = rslot(u) + rslot(t) and rslot(u - t) = rslot(u) - (t +
// ceil(t / 16)); and sin(pi (t + m NT) / M) = -Im(hi[t/64 + m NT/64] lo[t mod 64])
// takes the same lo entry for every m.  Lane t = 0 (p = 0) reads entries 0 and
// rslot(M), whatever they hold, and y_0 = 0 is selected after.
template <typename T, int M, int EP>
__device__ __forceinline__ void dst_pre_lds(cx<T> *h, int t, const cx<T> *buf, const TwTab<T, 2 * M> &tb) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  if constexpr (NT % 64 == 0) {
    const int rp = t + (t >> 4);
    const int rn = t + ((t + 15) >> 4);
    const cx<T> lo = tb.lo[t & 63];
    const cx<T> *hib = tb.hi + (t >> 6);
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int up = m * NT, um = M - m * NT;
      const cx<T> z = buf[rp + up + (up >> 4)];
      const cx<T> zm = buf[um + (um >> 4) - rn];
      const cx<T> a = hib[m * (NT / 64)];
      const T s = -(a.x * lo.y + a.y * lo.x);
      const T hx = T(0.5) * (z.x - zm.x), hy = T(0.5) * (z.y - zm.y);
      const cx<T> y = mkc<T>(s * (z.x + zm.x) + hx, s * (z.y + zm.y) + hy);
      h[m] = (m == 0 && t == 0) ? mkc<T>(T(0), T(0)) : y;
    }
  } else {
    dst_pre_s<T, M, EP>(h, t, SinLds<T, M>{&tb}, [&](int x) { return buf[rslot(x)]; });
  }
}

// Post pass: from Y at the home positions to S^a + i S^b in buf, natural
// order: entry k - 1 + SH holds S_k (k = 1..N) at rslot(k - 1 + SH); with SH = 1
// entry 0 holds no coefficient (zero or unwritten).  Ends with a barrier.
template <typename T, int M, int EP, int SH = 0>
__device__ __forceinline__ void dst_post(cx<T> *h, cx<T> *buf, int t, cx<double> *wtot = nullptr) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT, KPT = DstHalf<M, EP>::KPT;
  // Addresses as a per-thread base plus compile-time offsets where the pad
  // allows (rslot(B + c) = rslot(B) + c + c/16 while B mod 16 + c mod 16 < 16):
  // the home positions t + m NT (NT a multiple of 16), the thread's KPT
  // consecutive k and its 2 KPT consecutive outputs (KPT and 2 KPT dividing 16
  // or multiples of it).
  constexpr bool fold = NT % 16 == 0 && M % 16 == 0 && (16 % KPT == 0 || KPT % 16 == 0) &&
                        (16 % (2 * KPT) == 0 || (2 * KPT) % 16 == 0);
  const int rt = rslot(t);
  __syncthreads();  // the FFT's last readers of buf are done
#pragma unroll
  for (int m = 0; m < E; m++) {
    if constexpr (fold)
      buf[rt + m * NT + ((m * NT) >> 4)] = h[m];
    else
      buf[rslot(t + m * NT)] = h[m];
  }
  __syncthreads();
  cx<T> ev[KPT];
  cx<double> cum[KPT];
  cx<double> run = mkc<double>(0.0, 0.0);
  const int rk = rslot(t * KPT);                            // entry of k = KPT t
  const int rm = rslot(M - KPT * (t + 1));                  // entry of M - KPT (t + 1)
  const int rm0 = (t == 0) ? 0 : rslot(M - KPT * t);        // Y_{M-k} of i = 0 (Y_0 for k = 0)
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = t * KPT + i;
    cx<T> Yk, Ym;
    if constexpr (fold) {
      Yk = buf[rk + i + (i >> 4)];
      Ym = buf[i == 0 ? rm0 : rm + (KPT - i) + ((KPT - i) >> 4)];
    } else {
      Yk = buf[rslot(k)];
      Ym = buf[rslot(k == 0 ? 0 : M - k)];
    }
    ev[i] = mkc<T>(Ym.y - Yk.y, Yk.x - Ym.x);
    cx<double> r = mkc<double>((double)Yk.x + (double)Ym.x, (double)Yk.y + (double)Ym.y);
    if (k == 0) r = mkc<double>(0.5 * r.x, 0.5 * r.y);
    run = cadd<double>(run, r);
    cum[i] = run;
  }
  const cx<double> off = block_excl_scan<NT>(run, t, wtot);
  const int r2 = rslot(2 * KPT * t);  // entry of 2k for i = 0
  __syncthreads();  // every read of Y is done
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = t * KPT + i;
    if constexpr (fold) {
      if (i > 0 || SH == 1)
        buf[r2 + (2 * i - 1 + SH) + ((2 * i - 1 + SH) >> 4)] = ev[i];  // entry 2k - 1 + SH
      else if (k >= 1)
        buf[rslot(2 * k - 1)] = ev[i];
      buf[r2 + (2 * i + SH) + ((2 * i + SH) >> 4)] = mkc<T>((T)(off.x + cum[i].x), (T)(off.y + cum[i].y));
    } else {
      if (k >= 1) buf[rslot(2 * k - 1 + SH)] = ev[i];
      const cx<double> o = cadd<double>(off, cum[i]);
      buf[rslot(2 * k + SH)] = mkc<T>((T)o.x, (T)o.y);
    }
  }
  __syncthreads();
}

}  // namespace qgw
