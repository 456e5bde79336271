// fft.hpp -- workgroup-wide complex FFT in LDS for gfx950 (CDNA4).
//
// Stockham autosort formulation (Stockham 1966; the GPU form of Govindaraju
// et al., SC 2008): every stage reads its R inputs with stride L/R and writes
// its R outputs with stride Ns, so input and output are both in natural order
// and no bit reversal is needed.  Between stages the data sits in one LDS
// buffer (read all -> barrier -> write all -> barrier), with one pad slot per
// 32 entries so that the stride-R writes of the early stages spread over the
// banks.  Radices are 8 (as many as possible), then 4, 2, 3, 5 and 7, so that
// the transform lengths 2(n+1) of the model's grids of the form 2^a 3^b 5^c
// 7^d are covered (powers of two, 3*2^k, and grids such as NX = 100,
// 200, 1000, 2000).  Twiddles come from a table tw[m] = exp(-2 pi i m / L) built
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
constexpr int kPlanRadices[6] = {8, 4, 2, 3, 5, 7};
constexpr int plan_count(int L) {
  int c = 0;
  for (int r : kPlanRadices)
    while (L % r == 0) { L /= r; c++; }
  return c;
}
constexpr int plan_rest(int L) {
  for (int r : kPlanRadices)
    while (L % r == 0) L /= r;
  return L;
}
constexpr int plan_radix(int L, int s) {
  int c = 0;
  for (int r : kPlanRadices)
    while (L % r == 0) {
      if (c == s) return r;
      L /= r;
      c++;
    }
  return 0;
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

// X_k = sum_j x_j W5^(jk), W5 = exp(-2 pi i/5): with t = x1 + x4, x2 + x3 and
// u = x1 - x4, x2 - x3, X_1,4 = x0 + c1 t1 + c2 t2 -/+ i (s1 u1 + s2 u2) and
// X_2,3 = x0 + c2 t1 + c1 t2 -/+ i (s2 u1 - s1 u2), c_k = cos(2 pi k/5), s_k = sin(2 pi k/5).
template <typename T> struct Bfly<T, 5> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    const T c1 = T(0.30901699437494742410229341718281906), c2 = T(-0.80901699437494742410229341718281906);
    const T s1 = T(0.95105651629515357211643933337938214), s2 = T(0.58778525229247312916870595463907277);
    const cx<T> t1 = cadd<T>(v[1], v[4]), t2 = cadd<T>(v[2], v[3]);
    const cx<T> u1 = csub<T>(v[1], v[4]), u2 = csub<T>(v[2], v[3]);
    const cx<T> a1 = mkc<T>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const cx<T> a2 = mkc<T>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const cx<T> b1 = mkc<T>(s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y);
    const cx<T> b2 = mkc<T>(s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y);
    v[0] = cadd<T>(v[0], cadd<T>(t1, t2));
    v[1] = cadd<T>(a1, cnegi<T>(b1));
    v[4] = csub<T>(a1, cnegi<T>(b1));
    v[2] = cadd<T>(a2, cnegi<T>(b2));
    v[3] = csub<T>(a2, cnegi<T>(b2));
  }
};

// The same for W7 = exp(-2 pi i/7): t_j = x_j + x_(7-j), u_j = x_j - x_(7-j)
// (j = 1..3), X_k and X_(7-k) = x0 + sum_j cos(2 pi jk/7) t_j -/+ i sum_j
// sin(2 pi jk/7) u_j, the products jk reduced mod 7.
template <typename T> struct Bfly<T, 7> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    const T c1 = T(0.62348980185873353052500488400423981), c2 = T(-0.22252093395631440428890256449679476),
            c3 = T(-0.90096886790241912623610231950744505);
    const T s1 = T(0.78183148246802980870844452667405775), s2 = T(0.97492791218182360701813168299393122),
            s3 = T(0.43388373911755812047576833284835875);
    const cx<T> t1 = cadd<T>(v[1], v[6]), t2 = cadd<T>(v[2], v[5]), t3 = cadd<T>(v[3], v[4]);
    const cx<T> u1 = csub<T>(v[1], v[6]), u2 = csub<T>(v[2], v[5]), u3 = csub<T>(v[3], v[4]);
    auto lin3 = [](const cx<T> &x0, T a, const cx<T> &p, T b, const cx<T> &q, T c, const cx<T> &r) {
      return mkc<T>(x0.x + a * p.x + b * q.x + c * r.x, x0.y + a * p.y + b * q.y + c * r.y);
    };
    const cx<T> z = mkc<T>(T(0), T(0));
    const cx<T> a1 = lin3(v[0], c1, t1, c2, t2, c3, t3), b1 = lin3(z, s1, u1, s2, u2, s3, u3);
    const cx<T> a2 = lin3(v[0], c2, t1, c3, t2, c1, t3), b2 = lin3(z, s2, u1, -s3, u2, -s1, u3);
    const cx<T> a3 = lin3(v[0], c3, t1, c1, t2, c2, t3), b3 = lin3(z, s3, u1, -s1, u2, s2, u3);
    v[0] = cadd<T>(v[0], cadd<T>(t1, cadd<T>(t2, t3)));
    v[1] = cadd<T>(a1, cnegi<T>(b1));
    v[6] = csub<T>(a1, cnegi<T>(b1));
    v[2] = cadd<T>(a2, cnegi<T>(b2));
    v[5] = csub<T>(a2, cnegi<T>(b2));
    v[3] = cadd<T>(a3, cnegi<T>(b3));
    v[4] = csub<T>(a3, cnegi<T>(b3));
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
  static_assert(plan_rest(L) == 1, "transform length must be 2^a 3^b 5^c 7^d");
  Stages<T, L, NT, 0>::run(buf, tw, t);
}

// Threads per workgroup for a transform of length L: one radix-8 butterfly
// per thread in the radix-8 stages, at least one wavefront.
constexpr int fft_threads(int L) { return (L / 8 < 64) ? 64 : ((L / 8 > 1024) ? 1024 : L / 8); }

}  // namespace qgw
