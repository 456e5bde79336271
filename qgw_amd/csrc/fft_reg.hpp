// fft_reg.hpp -- register-blocked workgroup FFT for gfx950.
//
// Each of NT = L/E threads holds E elements of the length-L sequence in
// registers, always at its "home" positions p = t + m*NT (m = 0..E-1), both
// before the first stage and after the last one.  Stages are Stockham radix-R
// butterflies (R divides E, so a thread runs E/R butterflies per stage); only
// the exchanges between stages go through LDS, with one pad slot per 16
// entries, which keeps the strided writes of the early stages on distinct
// banks.  Because input and output live at the home positions, callers load
// their sequence straight from global memory and store the result straight
// back, with no extra LDS pass at either end.
//
// Lengths 2^a use E = 16 (L >= 1024) or E = 8 (radices 16/8/4/2); other
// lengths (factors 2, 3, 5, 7) use E = 8 (E = 4 where 8 does not divide L) with
// the LDS Stockham of fft.hpp between a write and a read of the home
// positions.
#pragma once
#include "fft.hpp"

namespace qgw {

constexpr bool is_pow2(int L) { return L > 0 && (L & (L - 1)) == 0; }

// EP: elements per thread (0: 16 for powers of two from 1024, else 8).  A
// radix schedule of the transform: E = 16 runs radix-16 stages, E = 8 radix-8.
template <int L, int EP = 0> struct RegPlan {
  static constexpr int E = (EP > 0 && is_pow2(L)) ? EP : (is_pow2(L) ? ((L >= 1024) ? 16 : 8) : (L % 8 == 0 ? 8 : 4));
  static_assert(L % E == 0, "home positions");
  static constexpr int NT = L / E;
  // radix of stage s (powers of two): E while the remaining length allows, then the rest
  static constexpr int radix(int s) {
    int rem = L, r = 0;
    for (int i = 0; i <= s; i++) {
      r = (rem >= E) ? E : rem;
      rem /= r;
    }
    return r;
  }
  static constexpr int count() {
    int rem = L, c = 0;
    while (rem > 1) {
      rem /= (rem >= E) ? E : rem;
      c++;
    }
    return c;
  }
  static constexpr int ns(int s) {
    int p = 1;
    for (int i = 0; i < s; i++) p *= radix(i);
    return p;
  }
  // LDS slots: one pad per 16 entries
  static constexpr int slots = L + L / 16 + 1;
};

__device__ __forceinline__ int rslot(int i) { return i + (i >> 4); }

// ---- radix-16 butterfly: 4 x 4 with internal twiddles W16^(n2 k1) --------
template <typename T> __device__ __forceinline__ void bfly16(cx<T> *v) {
  const T c1 = T(0.92387953251128675612818318939678829);  // cos(pi/8)
  const T s1 = T(0.38268343236508977172845998403039887);  // sin(pi/8)
  const T r2 = T(0.70710678118654752440084436210484904);
  cx<T> y[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; n2++) {
    cx<T> a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
    Bfly<T, 4>::run(a);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) y[n2][k1] = a[k1];
  }
  // y[n2][k1] *= W16^(n2*k1), W16 = exp(-2 pi i/16)
  // n2*k1 = 1: (c1, -s1); 2: (r2, -r2); 3: (s1, -c1); 4: -i; 6: (-r2, -r2); 9: (-c1, s1)
  auto mulw = [&](cx<T> a, T wr, T wi) { return mkc<T>(a.x * wr - a.y * wi, a.x * wi + a.y * wr); };
  y[1][1] = mulw(y[1][1], c1, -s1);
  y[1][2] = mulw(y[1][2], r2, -r2);
  y[1][3] = mulw(y[1][3], s1, -c1);
  y[2][1] = mulw(y[2][1], r2, -r2);
  y[2][2] = cnegi<T>(y[2][2]);
  y[2][3] = mulw(y[2][3], -r2, -r2);
  y[3][1] = mulw(y[3][1], s1, -c1);
  y[3][2] = mulw(y[3][2], -r2, -r2);
  y[3][3] = mulw(y[3][3], -c1, s1);
#pragma unroll
  for (int k1 = 0; k1 < 4; k1++) {
    cx<T> a[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
    Bfly<T, 4>::run(a);
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = a[k2];
  }
}

template <typename T, int R> __device__ __forceinline__ void bfly(cx<T> *v) {
  if constexpr (R == 16)
    bfly16<T>(v);
  else
    Bfly<T, R>::run(v);
}

// Twiddles of one butterfly: v[r] *= W^(r k), W = tw[base] (index into the
// size-L table).  The loads fetch W^(2^b), b < log2 R; the other powers are
// products.  Loads and products are split so that a stage's table loads can
// be issued a stage early, ahead of the LDS exchange that precedes their use,
// instead of stalling the stage on an L2 round trip after its barrier.
template <int R> constexpr int tw_loads() { return R == 2 ? 1 : (R == 4 ? 2 : (R == 8 ? 3 : 4)); }

// Twiddle table accessors: entry i of the size-L table e^{-2 pi i k / L}.
template <typename T> struct TwGlobal {  // the table in global memory
  const cx<T> *__restrict__ p;
  __device__ __forceinline__ cx<T> operator()(int i) const { return p[i]; }
};
// Two-level table in LDS: W^i = hi[i / 64] * lo[i % 64], hi[h] = W^(64 h),
// lo[l] = W^l (L >= 64; exact where either factor is 1).
template <typename T, int L> struct TwTab {
  static_assert(L >= 64, "two-level table");
  static constexpr int NHI = (L + 63) / 64;
  cx<T> lo[64], hi[NHI];
};
template <typename T, int L>
__device__ __forceinline__ void twtab_fill(TwTab<T, L> &tb, const cx<T> *__restrict__ tw, int t, int nt) {
  for (int i = t; i < 64 + TwTab<T, L>::NHI; i += nt) {
    if (i < 64)
      tb.lo[i] = tw[i];
    else
      tb.hi[i - 64] = tw[(i - 64) * 64];
  }
}
// The same fill split in two: the table loads into registers first, the LDS
// stores later, so that the stores wait only for the table loads and work
// issued in between (a transform's row loads) is not held up behind them.
template <typename T, int L, int NT> struct TwTabFill {
  static constexpr int CNT = 64 + TwTab<T, L>::NHI;
  static constexpr int PER = (CNT + NT - 1) / NT;
  cx<T> v[PER];
  __device__ __forceinline__ void load(const cx<T> *__restrict__ tw, int t) {
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int i = t + r * NT;
      const int e = (i < 64) ? i : ((i < CNT) ? (i - 64) * 64 : 0);
      v[r] = tw[e];
    }
  }
  __device__ __forceinline__ void store(TwTab<T, L> &tb, int t) const {
#pragma unroll
    for (int r = 0; r < PER; r++) {
      const int i = t + r * NT;
      if (i < 64)
        tb.lo[i] = v[r];
      else if (i < CNT)
        tb.hi[i - 64] = v[r];
    }
  }
};
template <typename T, int L> struct TwLds {
  const TwTab<T, L> *tb;
  __device__ __forceinline__ cx<T> operator()(int i) const { return cmul<T>(tb->hi[i >> 6], tb->lo[i & 63]); }
};

template <typename T, int R, class TW>
__device__ __forceinline__ void twiddle_load(cx<T> *wl, const TW &tw, int base) {
#pragma unroll
  for (int b = 0; b < tw_loads<R>(); b++) wl[b] = tw(base << b);
}

template <typename T, int R> __device__ __forceinline__ void twiddle_apply(cx<T> *v, const cx<T> *wl) {
  if constexpr (R == 2) {
    v[1] = cmul<T>(v[1], wl[0]);
  } else if constexpr (R == 4) {
    const cx<T> w1 = wl[0], w2 = wl[1];
    v[1] = cmul<T>(v[1], w1);
    v[2] = cmul<T>(v[2], w2);
    v[3] = cmul<T>(v[3], cmul<T>(w1, w2));
  } else if constexpr (R == 8) {
    const cx<T> w1 = wl[0], w2 = wl[1], w4 = wl[2];
    const cx<T> w3 = cmul<T>(w1, w2);
    v[1] = cmul<T>(v[1], w1);
    v[2] = cmul<T>(v[2], w2);
    v[3] = cmul<T>(v[3], w3);
    v[4] = cmul<T>(v[4], w4);
    v[5] = cmul<T>(v[5], cmul<T>(w1, w4));
    v[6] = cmul<T>(v[6], cmul<T>(w2, w4));
    v[7] = cmul<T>(v[7], cmul<T>(w3, w4));
  } else if constexpr (R == 16) {
    const cx<T> w1 = wl[0], w2 = wl[1], w4 = wl[2], w8 = wl[3];
    const cx<T> w3 = cmul<T>(w1, w2), w5 = cmul<T>(w1, w4), w6 = cmul<T>(w2, w4);
    const cx<T> w7 = cmul<T>(w3, w4);
    v[1] = cmul<T>(v[1], w1);
    v[2] = cmul<T>(v[2], w2);
    v[3] = cmul<T>(v[3], w3);
    v[4] = cmul<T>(v[4], w4);
    v[5] = cmul<T>(v[5], w5);
    v[6] = cmul<T>(v[6], w6);
    v[7] = cmul<T>(v[7], w7);
    v[8] = cmul<T>(v[8], w8);
    v[9] = cmul<T>(v[9], cmul<T>(w1, w8));
    v[10] = cmul<T>(v[10], cmul<T>(w2, w8));
    v[11] = cmul<T>(v[11], cmul<T>(w3, w8));
    v[12] = cmul<T>(v[12], cmul<T>(w4, w8));
    v[13] = cmul<T>(v[13], cmul<T>(w5, w8));
    v[14] = cmul<T>(v[14], cmul<T>(w6, w8));
    v[15] = cmul<T>(v[15], cmul<T>(w7, w8));
  }
}

// Table loads of every twiddle stage S of the plan needs (none for S = 0):
// butterfly w of the thread uses wl[w * log2(R) ...], W = tw[k L/(Ns R)] with
// k = (t + w NT) mod Ns.  At most E/2 entries.
template <typename T, int L, int S, int EP, class TW>
__device__ __forceinline__ void stage_tw_load(cx<T> *wl, const TW &tw, int t) {
  using Pl = RegPlan<L, EP>;
  constexpr int R = Pl::radix(S), Ns = Pl::ns(S), NB = Pl::E / R, LG = tw_loads<R>();
  if constexpr (Ns > 1) {
#pragma unroll
    for (int w = 0; w < NB; w++) twiddle_load<T, R>(wl + w * LG, tw, ((t + w * Pl::NT) % Ns) * (L / (Ns * R)));
  }
}

// Stage S of the power-of-two plan.  On entry v[] holds this stage's inputs:
// v[w*R + r] = X[j_w + r*L/R], j_w = t + w*NT, and wcur this stage's twiddle
// loads (stage_tw_load).  On exit (S not last) v[] holds the next stage's
// inputs; for the last stage v[w*R + r] is the output at position
// t + (w + r*E/R)*NT.  The next stage's twiddle loads are issued before this
// stage's butterflies.
template <typename T, int L, int S, int EP, class TW>
__device__ __forceinline__ void reg_stage(cx<T> *v, cx<T> *buf, const TW &tw, int t, const cx<T> *wcur) {
  using Pl = RegPlan<L, EP>;
  constexpr int E = Pl::E, NT = Pl::NT;
  constexpr int R = Pl::radix(S);
  constexpr int Ns = Pl::ns(S);
  constexpr int NB = E / R;  // butterflies per thread
  constexpr int LG = tw_loads<R>();
  constexpr bool more = S + 1 < Pl::count();
  cx<T> wnext[E / 2];
  if constexpr (more) stage_tw_load<T, L, S + 1, EP>(wnext, tw, t);
#pragma unroll
  for (int w = 0; w < NB; w++) {
    if constexpr (Ns > 1) twiddle_apply<T, R>(v + w * R, wcur + w * LG);
    bfly<T, R>(v + w * R);
  }
  if constexpr (more) {
    constexpr int R2 = Pl::radix(S + 1);
    constexpr int NB2 = E / R2;
    // With the one-per-16 pad, rslot(B + c) = rslot(B) + c + c/16 whenever B mod 16
    // plus c mod 16 stays below 16: so each butterfly's R writes are one
    // per-thread base plus compile-time offsets when Ns = 1 (B = jR, R <= 16)
    // or Ns R >= 16 (B mod 16 = j mod Ns, c = r Ns), and the reads are when NT
    // and L/R2 are multiples of 16.
    constexpr bool wfold = (Ns == 1 && R <= 16) || (Ns * R >= 16 && Ns <= 16) || Ns % 16 == 0;
    constexpr bool rfold = NT % 16 == 0 && (L / R2) % 16 == 0;
    __syncthreads();  // previous readers of buf are done
#pragma unroll
    for (int w = 0; w < NB; w++) {
      const int j = t + w * NT;
      const int base = (j / Ns) * Ns * R + (j % Ns);
      const int rb = rslot(base);
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (wfold)
          buf[rb + r * Ns + ((r * Ns) >> 4)] = v[w * R + r];
        else
          buf[rslot(base + r * Ns)] = v[w * R + r];
      }
    }
    __syncthreads();
    const int rt = rslot(t);
#pragma unroll
    for (int w = 0; w < NB2; w++) {
#pragma unroll
      for (int r = 0; r < R2; r++) {
        const int c = w * NT + r * (L / R2);
        if constexpr (rfold)
          v[w * R2 + r] = buf[rt + c + (c >> 4)];
        else
          v[w * R2 + r] = buf[rslot(t + c)];
      }
    }
    reg_stage<T, L, S + 1, EP, TW>(v, buf, tw, t, wnext);
  }
}

// Home-position index helpers: element m of a thread sits at position
// t + m*NT.  Stage inputs of the first stage and outputs of the last stage are
// permutations of the home set; these map register slots to m.
template <int L, int EP = 0> constexpr int in_m(int slot) {
  // first stage radix R0: v[w*R0 + r] = X[t + w*NT + r*(E/R0)*NT]
  constexpr int E = RegPlan<L, EP>::E, R0 = RegPlan<L, EP>::radix(0);
  return (slot / R0) + (slot % R0) * (E / R0);
}
template <int L, int EP = 0> constexpr int out_m(int slot) {
  constexpr int E = RegPlan<L, EP>::E, RL = RegPlan<L, EP>::radix(RegPlan<L, EP>::count() - 1);
  return (slot / RL) + (slot % RL) * (E / RL);
}

// Forward FFT of the sequence held at the home positions.  h[m] is the value
// at position t + m*NT, in and out.  buf: RegPlan<L>::slots entries (powers of
// two) or lds_slots(L) (other lengths).
// Power-of-two lengths with any table accessor (TwGlobal, TwLds).
template <typename T, int L, int EP, class TW>
__device__ __forceinline__ void fft_home_tw(cx<T> *h, cx<T> *buf, const TW &tw, int t) {
  static_assert(is_pow2(L), "register plan");
  constexpr int E = RegPlan<L, EP>::E;
  cx<T> v[E];
#pragma unroll
  for (int s = 0; s < E; s++) v[s] = h[in_m<L, EP>(s)];
  reg_stage<T, L, 0, EP, TW>(v, buf, tw, t, v);  // stage 0 has no twiddles
#pragma unroll
  for (int s = 0; s < E; s++) h[out_m<L, EP>(s)] = v[s];
}

template <typename T, int L, int EP = 0>
__device__ __forceinline__ void fft_home(cx<T> *h, cx<T> *buf, const cx<T> *__restrict__ tw, int t) {
  constexpr int E = RegPlan<L, EP>::E, NT = RegPlan<L, EP>::NT;
  if constexpr (is_pow2(L)) {
    fft_home_tw<T, L, EP>(h, buf, TwGlobal<T>{tw}, t);
  } else {
    __syncthreads();
#pragma unroll
    for (int m = 0; m < E; m++) buf[lds_slot(t + m * NT)] = h[m];
    __syncthreads();
    Stages<T, L, NT, 0>::run(buf, tw, t);
#pragma unroll
    for (int m = 0; m < E; m++) h[m] = buf[lds_slot(t + m * NT)];
  }
}

template <int L, int EP = 0> constexpr int home_slots() {
  return is_pow2(L) ? RegPlan<L, EP>::slots : lds_slots(L);
}

}  // namespace qgw
