// fft.hpp -- workgroup-wide complex FFT in LDS for gfx950 (CDNA4).
//
// Stockham autosort formulation (Stockham 1966; the GPU form of Govindaraju
// et al., SC 2008): every stage reads its R inputs with stride L/R and writes
// its R outputs with stride Ns, so input and output are both in natural order
// and no bit reversal is needed.  Between stages the data sits in one LDS
// buffer (read all -> barrier -> write all -> barrier), with one pad slot per
// 32 entries so that the stride-R writes of the early stages spread over the
// banks.  Radices are 8 (as many as possible), then 4, 2 and 3, so that the
// transform lengths 2(n+1) of the model's grids (powers of two and 3*2^k)
// are covered.  Twiddles come from a table tw[m] = exp(-2 pi i m / L) built
// on the host in FP64 and stored in the arithmetic type of the kernel.
//
// One workgroup of NT threads transforms one sequence of length L; the
// sequence is forward (e^{-2 pi i jk/L}) and unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace qgw {

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using cx = typename vec2<T>::type;

template <typename T> __device__ __forceinline__ cx<T> mkc(T x, T y) {
  cx<T> r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by -i
template <typename T> __device__ __forceinline__ cx<T> cnegi(cx<T> a) { return mkc<T>(a.y, -a.x); }
template <typename T> __device__ __forceinline__ cx<T> cneg(cx<T> a) { return mkc<T>(-a.x, -a.y); }

// LDS slot of logical index i: one pad entry after every 32.
__device__ __forceinline__ int lds_slot(int i) { return i + (i >> 5); }
constexpr int lds_slots(int L) { return L + (L >> 5) + 1; }

// ---- radix plan -----------------------------------------------------------
constexpr int plan_count(int L) {
  int c = 0;
  while (L % 8 == 0) { L /= 8; c++; }
  while (L % 4 == 0) { L /= 4; c++; }
  while (L % 2 == 0) { L /= 2; c++; }
  while (L % 3 == 0) { L /= 3; c++; }
  return c;
}
constexpr int plan_rest(int L) {
  while (L % 8 == 0) L /= 8;
  while (L % 4 == 0) L /= 4;
  while (L % 2 == 0) L /= 2;
  while (L % 3 == 0) L /= 3;
  return L;
}
constexpr int plan_radix(int L, int s) {
  int r = 0;
  int c = 0;
  while (L % 8 == 0) { if (c == s) return 8; L /= 8; c++; }
  while (L % 4 == 0) { if (c == s) return 4; L /= 4; c++; }
  while (L % 2 == 0) { if (c == s) return 2; L /= 2; c++; }
  while (L % 3 == 0) { if (c == s) return 3; L /= 3; c++; }
  return r;
}
constexpr int plan_ns(int L, int s) {  // product of the radices before stage s
  int p = 1;
  for (int i = 0; i < s; i++) p *= plan_radix(L, i);
  return p;
}

// ---- butterflies (forward) -----------------------------------------------
template <typename T, int R> struct Bfly;

template <typename T> struct Bfly<T, 2> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    cx<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
  }
};

template <typename T> struct Bfly<T, 3> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    // W3 = -1/2 - i sqrt(3)/2
    const T h = T(0.5);
    const T s3 = T(0.86602540378443864676372317075293618);
    cx<T> s = cadd<T>(v[1], v[2]);
    cx<T> d = csub<T>(v[1], v[2]);
    cx<T> m = mkc<T>(v[0].x - h * s.x, v[0].y - h * s.y);
    cx<T> r = mkc<T>(s3 * d.y, -s3 * d.x);  // -i sqrt(3)/2 * d
    v[0] = cadd<T>(v[0], s);
    v[1] = cadd<T>(m, r);
    v[2] = csub<T>(m, r);
  }
};

template <typename T> struct Bfly<T, 4> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    cx<T> a0 = cadd<T>(v[0], v[2]), a1 = csub<T>(v[0], v[2]);
    cx<T> a2 = cadd<T>(v[1], v[3]), a3 = cnegi<T>(csub<T>(v[1], v[3]));
    v[0] = cadd<T>(a0, a2);
    v[2] = csub<T>(a0, a2);
    v[1] = cadd<T>(a1, a3);
    v[3] = csub<T>(a1, a3);
  }
};

template <typename T> struct Bfly<T, 8> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    const T r2 = T(0.70710678118654752440084436210484904);
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Bfly<T, 4>::run(e);
    Bfly<T, 4>::run(o);
    // o[k] *= W8^k: W8 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = -(1 + i)/sqrt2
    cx<T> o1 = mkc<T>(r2 * (o[1].x + o[1].y), r2 * (o[1].y - o[1].x));
    cx<T> o2 = cnegi<T>(o[2]);
    cx<T> o3 = mkc<T>(r2 * (o[3].y - o[3].x), -r2 * (o[3].x + o[3].y));
    v[0] = cadd<T>(e[0], o[0]);
    v[4] = csub<T>(e[0], o[0]);
    v[1] = cadd<T>(e[1], o1);
    v[5] = csub<T>(e[1], o1);
    v[2] = cadd<T>(e[2], o2);
    v[6] = csub<T>(e[2], o2);
    v[3] = cadd<T>(e[3], o3);
    v[7] = csub<T>(e[3], o3);
  }
};

// ---- one Stockham stage, then the rest -----------------------------------
template <typename T, int L, int NT, int S> struct Stages {
  static constexpr int R = plan_radix(L, S);
  static constexpr int Ns = plan_ns(L, S);
  static constexpr int NW = L / R;                 // butterflies in this stage
  static constexpr int PER = (NW + NT - 1) / NT;   // butterflies per thread
  static constexpr int TWS = L / (Ns * R);         // twiddle index step

  static __device__ __forceinline__ void run(cx<T> *buf, const cx<T> *__restrict__ tw, int t) {
    cx<T> v[PER][R];
#pragma unroll
    for (int w = 0; w < PER; w++) {
      const int j = t + w * NT;
      if ((NW % NT) == 0 || j < NW) {
#pragma unroll
        for (int r = 0; r < R; r++) v[w][r] = buf[lds_slot(j + r * NW)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < PER; w++) {
      const int j = t + w * NT;
      if ((NW % NT) == 0 || j < NW) {
        const int k = j % Ns;
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; r++) v[w][r] = cmul<T>(v[w][r], tw[r * k * TWS]);
        }
        Bfly<T, R>::run(v[w]);
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) buf[lds_slot(base + r * Ns)] = v[w][r];
      }
    }
    __syncthreads();
    if constexpr (S + 1 < plan_count(L)) Stages<T, L, NT, S + 1>::run(buf, tw, t);
  }
};

// Forward FFT of buf[0..L) (natural order in and out).  The caller fills the
// buffer and issues a barrier first; on return the result is visible to all
// threads of the workgroup.
template <typename T, int L, int NT>
__device__ __forceinline__ void fft_lds(cx<T> *buf, const cx<T> *__restrict__ tw, int t) {
  static_assert(plan_rest(L) == 1, "transform length must be 2^a 3^b");
  Stages<T, L, NT, 0>::run(buf, tw, t);
}

// Threads per workgroup for a transform of length L: one radix-8 butterfly
// per thread in the radix-8 stages, at least one wavefront.
constexpr int fft_threads(int L) { return (L / 8 < 64) ? 64 : ((L / 8 > 1024) ? 1024 : L / 8); }

}  // namespace qgw
