// kernels.hpp -- the per-step kernels of the QG timestep on gfx950.
//
// Device layout (per precision T):
//   q[2], psi, f[3], comp   nl layers of (m+2) rows, row pitch P >= n+2; the
//                           one-point wall ring is stored (reference idx
//                           layout, src/domain.h:62, with a padded pitch);
//                           comp (FP32 forms) holds 16-bit elements, comp_t
//   what                    nl layers of m rows (y) x n columns (x frequency), pitch Pn
//   whatT                   nl modes of n rows (x frequency) x m columns (y), pitch Pm
//
// Passes of one step (bounded branch of invert_pv, src/elliptic.h:76-155, and
// rhs/timestep, src/dynamics.h:47-73, src/timestep.h:114-167):
//   P1  per grid row and layer pair: forward sine transform along x
//   T1  what (layers) -> whatT (modes): tile transpose + layer->mode projection
//   P2  per x-frequency and mode: forward sine transform along y, division by
//       -(L^2 + K^2 + 1/Rd^2) * 4(n+1)(m+1), inverse sine transform along y
//   T2  whatT (modes) -> what (layers): tile transpose + mode->layer projection
//   P3  per grid row and layer pair: inverse sine transform along x, psi row
//       and the wall ring of psi and q
//   P4  per 2-D tile and layer: Arakawa Jacobian, beta, nu lap q,
//       nu_hyper lap^(n_hyper+2) psi, wind, drag, topography, AB3 update
//
// Each sine transform of a real sequence of length n is computed together
// with a second one: the complex sequence a + i b, odd-extended to length
// L = 2(n+1), is Fourier transformed, and DST(a) = -Im, DST(b) = Re (the DFT of
// an odd real sequence is -i times its RODFT00 transform).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft.hpp"
#include "fft_reg.hpp"
#include "dst_half.hpp"

namespace qgw {

// Workgroup id -> work item such that the ids the hardware deals round-robin
// to one of the 8 XCDs cover a contiguous range of work items.
__device__ __forceinline__ int xcd_remap(int b, int total) {
  const int x = b & 7, i = b >> 3;
  const int q = total >> 3, r = total & 7;
  return x * q + (x < r ? x : r) + i;
}

// Kahan compensation of the FP32 forms: the rounding error c of the stored q
// (q - c is the compensated value, |c| <= ulp(q)/2) kept as a 16-bit integer
// multiple of ulp(q)/2^15, i.e. 15 bits below q's last place: q - c carries 39
// significant bits, more than the FP32 increments added to it hold, in half
// the bytes of an FP32 compensation.  The FP64 form keeps no compensation.
template <typename T> struct CompOf {
  using type = T;
};
template <> struct CompOf<float> {
  using type = short;
};
template <typename T> using comp_t = typename CompOf<T>::type;
__device__ __forceinline__ int comp_exp(float q) { return (int)((__float_as_uint(q) >> 23) & 0xffu); }
// c in units of ulp(q)/2^15 = 2^(e - 165), e the biased exponent of q; clamped
// to +-32767 (|c| above 2^15 ulp happens only under cancellation, where q's own
// rounding dominates); q = 0 or subnormal decodes to 0
template <typename T> __device__ __forceinline__ comp_t<T> comp_make(T c, T q) {
  if constexpr (sizeof(T) == 4) {
    const float v = __builtin_rintf(__builtin_ldexpf(c, 165 - comp_exp(q)));
    return (short)__builtin_fminf(__builtin_fmaxf(v, -32767.f), 32767.f);
  } else {
    return c;
  }
}
template <typename T> __device__ __forceinline__ T comp_val(comp_t<T> v, T q) {
  if constexpr (sizeof(T) == 4)
    return __builtin_ldexpf((float)v, comp_exp(q) - 165);
  else
    return v;
}

// Sizes shared by every kernel.
struct Geo {
  int n, m, nl;          // interior points in x, y; layers (all members)
  int nlm;               // layers per member: ensemble members are stacked along the layer axis
  int P;                 // row pitch of the storage fields (elements)
  long LS;               // layer stride of the storage fields (= (m+2)*P)
  int Pn;                // pitch of what (row = y index, column = x frequency)
  long WS;               // mode stride of what (= m*Pn)
  int Pm;                // pitch of whatT (row = x frequency, column = y)
  long WTS;              // mode stride of whatT (= n*Pm)
};

// ===========================================================================
// P1: forward sine transform along x of two layers of one grid row.  The
// pair (layer l0, layer l1) is packed as a + i b and odd-extended; each thread
// loads its home positions straight from q.
// ===========================================================================
template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_p1(Geo g, const T *__restrict__ q, T *__restrict__ what, const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[home_slots<L>()];
  const int t = threadIdx.x;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *r0 = q + l0 * g.LS + (long)j * g.P + 1;
  const T *r1 = q + (has1 ? l1 : l0) * g.LS + (long)j * g.P + 1;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    T a = T(0), b = T(0);
    if (p != 0 && p != L / 2) {
      const int x = (p < L / 2) ? p - 1 : L - p - 1;
      a = r0[x];
      b = has1 ? r1[x] : T(0);
      if (p > L / 2) {
        a = -a;
        b = -b;
      }
    }
    h[m] = mkc<T>(a, b);
  }
  fft_home<T, L>(h, buf, tw, t);
  T *o0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  T *o1 = what + l1 * g.WS + (long)(j - 1) * g.Pn;
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    if (p >= 1 && p < L / 2) {
      o0[p - 1] = -h[m].y;
      if (has1) o1[p - 1] = h[m].x;
    }
  }
}

// ===========================================================================
// Transpose with a layer/mode contraction:
//   dst[k][c][r] = sum_l M[k*nl + l] * src[l][r][c],   r < R, c < C.
// T1 uses M = cl2m (layers -> modes, w_k = sum_l cl2m[k,l] q_l,
// src/elliptic.h:90-98); T2 uses M = cm2l (modes -> layers, psi_l +=
// cm2l[l,k] w_k, src/elliptic.h:140-147).  Both projections commute with the
// sine transforms, so they sit where the data is transposed anyway.
// ===========================================================================
// Runs over more than 32 layers take the matrix in blocks of at most NLC
// layers and 32 modes: nl is the number of layers of this block, nk its
// modes, ms the matrix row stride (the member's layer count); ACC adds the
// block's contribution to the destination instead of storing it (every
// layer block after the first).  zs, zd, zm: the member strides of src, dst
// and M (member blockIdx.z).
template <typename T, int TC, int TR, int NLC, int KB, bool ACC = false>
__global__ void __launch_bounds__(256)
k_tproj(const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ M, int nl, int nk, int ms, int R,
        int C, int Ps, long Ss, int Pd, long Sd, long zs, long zd, long zm) {
  static_assert(TC * 8 == 256 && TR % 8 == 0, "tile shape");
  constexpr int U = TR / 8;  // source rows per thread
  // KB output modes per barrier pair: their tiles are formed together, then stored together
  __shared__ T tile[KB][TR][TC + 1];
  // member blockIdx.z: its layers and its own matrix
  src += (long)blockIdx.z * zs;
  dst += (long)blockIdx.z * zd;
  M += (long)blockIdx.z * zm;
  const int c0 = blockIdx.x * TC, r0 = blockIdx.y * TR;
  const int tid = threadIdx.x;
  // every layer of the thread's points, lanes along c (contiguous in src):
  // a uniform branch on the layer, then loads at clamped (valid) in-layer
  // offsets from the layer's scalar base, zeroed outside the tile's R x C
  const int ci = tid % TC, rg = tid / TC;
  int off[U];
  bool ok[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int r = r0 + rg + 8 * u, c = c0 + ci;
    ok[u] = r < R && c < C;
    off[u] = ((r < R) ? r : R - 1) * Ps + ((c < C) ? c : C - 1);
  }
  T v[NLC][U];
#pragma unroll
  for (int l = 0; l < NLC; l++) {
    if (l < nl) {
      const T *lb = src + (long)l * Ss;
#pragma unroll
      for (int u = 0; u < U; u++) {
        const T x = lb[off[u]];
        v[l][u] = ok[u] ? x : T(0);
      }
    } else {
#pragma unroll
      for (int u = 0; u < U; u++) v[l][u] = T(0);
    }
  }
  // store mapping: lanes along r (contiguous in dst)
  const int rl = tid % TR, cg = tid / TR;
  constexpr int CS = 256 / TR;  // columns per store pass
  for (int k0 = 0; k0 < nk; k0 += KB) {
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
      const int k = k0 + kb;
      if (k >= nk) break;  // uniform
      // row k of the matrix: NLC uniform (scalar) loads issued together and
      // waited on once; entries past nl are clamped to a valid address and
      // meet zero layers (v[l][u] = 0 for l >= nl), so their products vanish
      T mk[NLC];
#pragma unroll
      for (int l = 0; l < NLC; l++) mk[l] = M[(long)k * ms + ((l < nl) ? l : 0)];
#pragma unroll
      for (int u = 0; u < U; u++) {
        T acc = T(0);
#pragma unroll
        for (int l = 0; l < NLC; l++) acc += mk[l] * v[l][u];
        tile[kb][rg + 8 * u][ci] = acc;
      }
    }
    __syncthreads();
    const int r = r0 + rl;
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
      const int k = k0 + kb;
      if (k >= nk) break;
      T *kd = dst + (long)k * Sd;
#pragma unroll
      for (int cc = cg; cc < TC; cc += CS) {
        const int c = c0 + cc;
        if (r < R && c < C) {
          if constexpr (ACC)
            kd[c * Pd + r] += tile[kb][rl][cc];
          else
            kd[c * Pd + r] = tile[kb][rl][cc];
        }
      }
    }
    __syncthreads();
  }
}

// ===========================================================================
// Transpose + layer<->mode projection on the matrix cores (FP32, up to 32
// layers): out_k = sum_l M[k][l] in_l as D = M (32 x 32) * V (32 layers x
// points) with v_mfma_f32_32x32x2f32.  A block takes 16 source columns x 32
// source rows; a point group is 32 consecutive source rows of one column, so
// the accumulator rows land contiguous in the destination (128-byte stores)
// and no LDS pass is needed on the way out.  Layers stream through LDS in
// chunks of 8; layers and modes beyond nl are zero.
//   A (32 x 2):  lane i holds M[i % 32][2s + i / 32]
//   B (2 x 32):  lane i holds V[2s + i / 32][point i % 32]
//   D (32 x 32): register j of lane i holds D[8 (j / 4) + 4 (i / 32) + j % 4][i % 32]
// ===========================================================================
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int KC, int TCB>
__global__ void __launch_bounds__(256)
k_tproj_mfma(const float *__restrict__ src, float *__restrict__ dst, const float *__restrict__ M, int nl, int R,
             int C, int Ps, long Ss, int Pd, long Sd, int ntc, int ntr) {
  constexpr int TRB = 32, NCH = 32 / KC;
  constexpr int GPW = TCB / 4;        // point groups (source columns) per wavefront
  constexpr int LR = 256 / TCB;       // source rows per load pass
  constexpr int LU = TRB / LR;        // load passes
  static_assert(TCB % 4 == 0 && TRB % LR == 0, "tile shape");
  __shared__ float lds[KC][TRB][TCB + 1];
  // tiles in raster order per XCD, so that tiles sharing cache lines meet in one L2
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int z = tile / (ntc * ntr), cy = tile % (ntc * ntr);
  src += (long)z * nl * Ss;
  dst += (long)z * nl * Sd;
  M += (long)z * nl * nl;
  const int c0 = (cy % ntc) * TCB, r0 = (cy / ntc) * TRB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int half = lane >> 5, l32 = lane & 31;
  float a[16];
#pragma unroll
  for (int s = 0; s < 16; s++) {
    const int l = 2 * s + half;
    a[s] = (l32 < nl && l < nl) ? M[l32 * nl + l] : 0.f;
  }
  f32x16 acc[GPW];
#pragma unroll
  for (int g = 0; g < GPW; g++)
#pragma unroll
    for (int j = 0; j < 16; j++) acc[g][j] = 0.f;
  const int lc = tid % TCB, lr = tid / TCB;
  // Loads go to clamped (always valid) addresses and are zeroed by a select,
  // so that a chunk's loads are all in flight before the first LDS write.
  const int cc = (c0 + lc < C) ? c0 + lc : C - 1;
  const bool cok = c0 + lc < C;
  long roff[LU];
  bool rok[LU];
#pragma unroll
  for (int u = 0; u < LU; u++) {
    const int r = r0 + lr + LR * u;
    rok[u] = cok && r < R;
    roff[u] = (long)(r < R ? r : R - 1) * Ps + cc;
  }
  // chunk ch+1's loads are issued before chunk ch's matrix products
  float v[KC][LU];
  auto load_chunk = [&](int l0) {
#pragma unroll
    for (int kk = 0; kk < KC; kk++) {
      const int l = (l0 + kk < nl) ? l0 + kk : nl - 1;
#pragma unroll
      for (int u = 0; u < LU; u++) v[kk][u] = src[l * Ss + roff[u]];
    }
  };
  load_chunk(0);
#pragma unroll
  for (int ch = 0; ch < NCH; ch++) {
    const int l0 = ch * KC;
    if (l0 >= nl) break;  // uniform
#pragma unroll
    for (int kk = 0; kk < KC; kk++)
#pragma unroll
      for (int u = 0; u < LU; u++) lds[kk][lr + LR * u][lc] = (l0 + kk < nl && rok[u]) ? v[kk][u] : 0.f;
    __syncthreads();
    if (ch + 1 < NCH && l0 + KC < nl) load_chunk(l0 + KC);
#pragma unroll
    for (int g = 0; g < GPW; g++) {
      const int cl = w * GPW + g;
#pragma unroll
      for (int s = 0; s < KC / 2; s++) {
        const float b = lds[2 * s + half][l32][cl];
        acc[g] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[ch * (KC / 2) + s], b, acc[g], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  const int r = r0 + l32;
#pragma unroll
  for (int g = 0; g < GPW; g++) {
    const int c = c0 + w * GPW + g;
    if (c >= C || r >= R) continue;
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int k = 8 * (j >> 2) + 4 * half + (j & 3);
      if (k < nl) dst[k * Sd + (long)c * Pd + r] = acc[g][j];
    }
  }
}

// ===========================================================================
// P2: per pair of whatT rows (mode k, x frequency a): forward transform along
// y, Helmholtz division, inverse transform along y.  The division is applied
// at the home positions in registers: position p holds C_p, and the input of
// the inverse transform at p is (-Im C_p) f_0(|p|) + i Re(C_p) f_1(|p|) with
// |p| = p or L - p, which is the odd extension of the divided spectra.
// f = 1 / (-(L_a^2 + K_b^2 + 1/Rd_k^2) * 4(n+1)(m+1)) (src/elliptic.h:120-138);
// the denominator is summed in FP64.
// ===========================================================================
template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_p2(Geo g, T *__restrict__ whatT, const double *__restrict__ La2, const double *__restrict__ Kb2,
     const double *__restrict__ iRd2, double norm, const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[home_slots<L>()];
  const int t = threadIdx.x;
  const int n = g.n;
  const long rows = (long)g.nl * n;
  const long r0 = 2L * blockIdx.x, r1 = r0 + 1;
  const bool has1 = r1 < rows;
  const int k0 = (int)(r0 / n), a0 = (int)(r0 % n) + 1;
  const int k1 = has1 ? (int)(r1 / n) : k0, a1 = has1 ? (int)(r1 % n) + 1 : a0;
  T *row0 = whatT + k0 * g.WTS + (long)(a0 - 1) * g.Pm;
  T *row1 = whatT + k1 * g.WTS + (long)(a1 - 1) * g.Pm;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    T a = T(0), b = T(0);
    if (p != 0 && p != L / 2) {
      const int y = (p < L / 2) ? p - 1 : L - p - 1;
      a = row0[y];
      b = has1 ? row1[y] : T(0);
      if (p > L / 2) {
        a = -a;
        b = -b;
      }
    }
    h[m] = mkc<T>(a, b);
  }
  fft_home<T, L>(h, buf, tw, t);
  const double d0 = La2[a0], d1 = La2[a1];
  const double ir0 = iRd2[k0], ir1 = iRd2[k1];
  // FP32 forms: the denominator in FP32 except for the lowest barotropic
  // wavenumbers (mode 0, a, b <= 16), where it is smallest and psi largest
  const bool lo0 = (k0 == 0 && a0 <= 16), lo1 = (k1 == 0 && a1 <= 16);
  const T d0f = (T)(d0 + ir0), d1f = (T)(d1 + ir1), normf = (T)norm;
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    if (p == 0 || p == L / 2) {
      h[m] = mkc<T>(T(0), T(0));
    } else {
      const int b = (p < L / 2) ? p : L - p;
      const double kb = Kb2[b];
      T f0, f1;
      if constexpr (sizeof(T) == 8) {
        f0 = (T)(1.0 / (-(d0 + kb + ir0) * norm));
        f1 = (T)(1.0 / (-(d1 + kb + ir1) * norm));
      } else {
        const T kbf = (T)kb;
        f0 = (lo0 && b <= 16) ? (T)(1.0 / (-(d0 + kb + ir0) * norm)) : T(1) / (-(d0f + kbf) * normf);
        f1 = (lo1 && b <= 16) ? (T)(1.0 / (-(d1 + kb + ir1) * norm)) : T(1) / (-(d1f + kbf) * normf);
      }
      h[m] = mkc<T>(-h[m].y * f0, h[m].x * f1);
    }
  }
  fft_home<T, L>(h, buf, tw, t);
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    if (p >= 1 && p < L / 2) {
      row0[p - 1] = -h[m].y;
      if (has1) row1[p - 1] = h[m].x;
    }
  }
}

// ===========================================================================
// Half-length forms of P1, P2 and P3 (dst_half.hpp): the same transforms and
// outputs, through complex FFTs of length M = N + 1 instead of 2(N + 1).
//   twh: e^{-2 pi i k / M} (length M)    tw2: e^{-2 pi i k / 2M} (length 2M)
// EP: elements per thread of the FFT (its radix schedule; 0 = default).
// ===========================================================================
template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p1h(Geo g, const T *__restrict__ q, T *__restrict__ what, const cx<T> *__restrict__ twh,
      const cx<T> *__restrict__ tw2) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  const int t = threadIdx.x;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *r0 = q + l0 * g.LS + (long)j * g.P;  // storage column x = 1..n holds x_1..x_N
  const T *r1 = q + (has1 ? l1 : l0) * g.LS + (long)j * g.P;
  cx<T> h[E];
  dst_pre<T, M, EP>(h, t, tw2, [&](int x) { return mkc<T>(r0[x], has1 ? r1[x] : T(0)); });
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP>(h, buf, t);
  T *o0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  T *o1 = what + l1 * g.WS + (long)(j - 1) * g.Pn;
  for (int i = t; i < M - 1; i += NT) {
    const cx<T> v = buf[rslot(i)];
    o0[i] = v.x;
    if (has1) o1[i] = v.y;
  }
}

// FP32 with 8 elements per thread: four wavefronts per SIMD (128 registers;
// the transform pair is issue-bound, and the fourth wavefront hides more of
// its LDS and load latency than the few spilled registers cost).
template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
__attribute__((amdgpu_waves_per_eu((DstHalf<M, EP>::E == 8) ? (sizeof(T) == 4 ? 4 : 3) : 1)))
k_p2h(Geo g, T *__restrict__ whatT, const double *__restrict__ La2, const double *__restrict__ iRd2, double dk,
      double norm, const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  __shared__ TwTab<T, 2 * M> tb2;  // sines of the inverse transform's pre pass
  const int t = threadIdx.x;
  const int n = g.n;
  // rows q0 = 2 blockIdx.x and q0 + 1 of nl n (< 2^31): mode k, column a (1-based);
  // one 32-bit division, the second row following from the first
  const unsigned q0 = 2u * blockIdx.x, un = (unsigned)n;
  const bool has1 = q0 + 1 < (unsigned)g.nl * un;
  const int k0 = (int)(q0 / un), a0 = (int)(q0 - (unsigned)k0 * un) + 1;
  const int k1 = (has1 && a0 == n) ? k0 + 1 : k0, a1 = has1 ? (a0 == n ? 1 : a0 + 1) : a0;
  T *row0 = whatT + k0 * g.WTS + (long)(a0 - 1) * g.Pm;
  T *row1 = whatT + k1 * g.WTS + (long)(a1 - 1) * g.Pm;
  // the sine table's loads go first and its LDS stores after the pre pass, so
  // that neither waits behind the row loads; published by the transform's barriers
  TwTabFill<T, 2 * M, NT> tfill;
  tfill.load(tw2, t);
  cx<T> h[E];
  // forward transform along y (row element y-1 holds x_y)
  dst_pre_rows<T, M, EP>(h, t, row_rsrc(row0, (M - 1) * (int)sizeof(T)),
                         row_rsrc(row1, has1 ? (M - 1) * (int)sizeof(T) : 0), tw2);
  tfill.store(tb2, t);
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP, 1>(h, buf, t);  // entry b holds S_b
  // Helmholtz division, f = 1 / (-(L_a^2 + K_b^2 + 1/Rd_k^2) * 4(n+1)(m+1)) (src/elliptic.h:120-138).
  // K_b = pi b / Lx (src/domain.h:170-173) is formed here as b dk, dk = pi / Lx,
  // within 2 ulp of the quotient, so that no table load stalls the division.
  const double d0 = La2[a0], d1 = La2[a1];
  const double ir0 = iRd2[k0], ir1 = iRd2[k1];
  const bool lo0 = (k0 == 0 && a0 <= 16), lo1 = (k1 == 0 && a1 <= 16);
  // FP32: (d + K_b^2) norm = d norm + b^2 (dk^2 norm), b^2 exact in FP32 (b < 2^12)
  const T d0n = (T)((d0 + ir0) * norm), d1n = (T)((d1 + ir1) * norm), c2 = (T)(dk * dk * norm);
  // all of the thread's entries are read before any is written back, so that
  // the LDS reads are in flight together
  constexpr int NS = (M + NT - 1) / NT;
  cx<T> hv[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int b = t + s * NT;
    if (b < M) hv[s] = buf[rslot(b)];
  }
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int b = t + s * NT;
    T f0, f1;
    if constexpr (sizeof(T) == 8) {
      const double K = dk * (double)b, kb = K * K;
      f0 = (T)(1.0 / (-(d0 + kb + ir0) * norm));
      f1 = (T)(1.0 / (-(d1 + kb + ir1) * norm));
    } else {
      // FP32 forms: FP64 only for the lowest barotropic wavenumbers (mode 0, a, b <= 16);
      // elsewhere FP32 and the hardware reciprocal (1 ulp), well inside the
      // transforms' FP32 rounding
      const T bf = (T)b, bb = bf * bf;
      f0 = -__builtin_amdgcn_rcpf(__builtin_fmaf(bb, c2, d0n));
      f1 = -__builtin_amdgcn_rcpf(__builtin_fmaf(bb, c2, d1n));
      if (b <= 16) {
        const double K = dk * (double)b, kb = K * K;
        if (lo0) f0 = (T)(1.0 / (-(d0 + kb + ir0) * norm));
        if (lo1) f1 = (T)(1.0 / (-(d1 + kb + ir1) * norm));
      }
    }
    if (b >= 1 && b < M) buf[rslot(b)] = mkc<T>(hv[s].x * f0, hv[s].y * f1);
  }
  __syncthreads();
  // inverse transform (the same DST-I) of the divided spectra
  dst_pre_lds<T, M, EP>(h, t, buf, tb2);
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP>(h, buf, t);
  for (int i = t; i < M - 1; i += NT) {
    const cx<T> v = buf[rslot(i)];
    row0[i] = v.x;
    if (has1) row1[i] = v.y;
  }
}

// Element x of a psi row: a plain store, or (nt) a non-temporal one, written
// as a buffer store with the nt cache-policy bit (a plain store and a
// non-temporal one to the same address under a uniform branch would be merged
// by the compiler into one plain store).  At 2048^2 x 10 the non-temporal psi
// stores of the inverse row pass make the tendency pass that reads psi next
// faster by more than they slow the row pass itself (QGW_TUNE_PSI_NT).
template <typename T> __device__ __forceinline__ void psi_st(T *row, int x, T v, int nt) {
  if (nt) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(row, (short)0, 0x7ffffff0, 0x00020000);
    if constexpr (sizeof(T) == 8) {
      typedef unsigned u2 __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, x * 8, 0, 2);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, x * 4, 0, 2);
    }
  } else {
    row[x] = v;
  }
}

template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p3h(Geo g, const T *__restrict__ what, T *__restrict__ psi, T *__restrict__ q, T ring_c,
      const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2, int psi_nt) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  const int t = threadIdx.x;
  const int n = g.n, m_ = g.m;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *i0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  const T *i1 = what + (has1 ? l1 : l0) * g.WS + (long)(j - 1) * g.Pn;
  cx<T> h[E];
  dst_pre<T, M, EP>(h, t, tw2, [&](int x) { return mkc<T>(i0[x - 1], has1 ? i1[x - 1] : T(0)); });
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP>(h, buf, t);
  T *p0 = psi + l0 * g.LS + (long)j * g.P;
  T *p1 = psi + l1 * g.LS + (long)j * g.P;
  T *q0r = q + l0 * g.LS;
  T *q1r = q + l1 * g.LS;
  // both layers in one pass over buf, the wall-ring values of q beside psi
  for (int i = t; i < n; i += NT) {
    const int x = i + 1;
    const cx<T> c = buf[rslot(i)];
    psi_st(p0, x, c.x, psi_nt);
    if (has1) psi_st(p1, x, c.y, psi_nt);
    if (x == 1 || x == n) {
      const long o = (long)j * g.P + (x == 1 ? 0 : n + 1);
      q0r[o] = ring_c * c.x;
      if (has1) q1r[o] = ring_c * c.y;
    }
    if (j == 1 || j == m_) {
      const long o = (long)(j == 1 ? 0 : m_ + 1) * g.P + x;
      q0r[o] = ring_c * c.x;
      if (has1) q1r[o] = ring_c * c.y;
    }
  }
  if (t == 0) {
    p0[0] = T(0);
    p0[n + 1] = T(0);
    if (has1) {
      p1[0] = T(0);
      p1[n + 1] = T(0);
    }
  }
  if (j == 1 || j == m_) {
    const long rr = (long)(j == 1 ? 0 : m_ + 1) * g.P;
    for (int x = t; x < n + 2; x += NT) {
      psi[l0 * g.LS + rr + x] = T(0);
      if (has1) psi[l1 * g.LS + rr + x] = T(0);
    }
  }
}

// ===========================================================================
// Pipelined half-length row transforms (every half length M): the same outputs as
// k_p1h / k_p3h from persistent blocks.  A block walks the items
// it = blockIdx.x, blockIdx.x + gridDim.x, ... (a row and layer pair).  The
// rows of item it + gridDim.x are loaded into registers while item it is
// transformed; the twiddles (two-level tables for powers of two, the full
// table for other lengths) and the pre-pass sines come from LDS tables filled once per
// block, so the global traffic inside the loop is the row loads and stores
// alone and no wait on a table load also waits on the next item's rows.  Each
// item's row pair is staged in LDS (entry x holds z_x, entry 0 the zero z_0),
// from where the pre pass reads z_p and z_{M-p}: both runs of
// 16 lanes then start on a multiple of 16 entries, where the one-per-16 pad
// keeps them on distinct banks.  (The column pass P2
// stays one block per item: its two transforms fill the register file, and
// the prefetch registers would cost it a wavefront per SIMD.)
// The thread index passes through an empty asm at the top of each item, so
// that the index arithmetic of the item is formed inside the loop instead of
// being hoisted out of it into registers held across all items; its range
// 0 <= t < NT is restated after the asm.
// ===========================================================================
template <int NT> __device__ __forceinline__ int loop_tid(int t) {
  asm volatile("" : "+v"(t));
  __builtin_assume(t >= 0 && t < NT);  // the range survives, for the shifts and trip counts
  return t;
}
template <typename T, int M, int EP> struct PipeRows {
  static constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  T za[E], zb[E];
  // z_p (p = t + m NT) of two rows of N = M - 1 elements, element p - 1 holding
  // z_p: buffer loads, whose range check gives z_0 = 0 and, when has1 is
  // false, zeros for the whole second row, with no select after the loads
  __device__ __forceinline__ void fetch(const T *__restrict__ r0, const T *__restrict__ r1, bool has1, int t) {
    const __amdgpu_buffer_rsrc_t b0 = row_rsrc(r0, (M - 1) * (int)sizeof(T));
    const __amdgpu_buffer_rsrc_t b1 = row_rsrc(r1, has1 ? (M - 1) * (int)sizeof(T) : 0);
    const int o = (t - 1) * (int)sizeof(T);
#pragma unroll
    for (int m = 0; m < E; m++) {
      za[m] = buf_ld<T>(b0, o + m * NT * (int)sizeof(T));
      zb[m] = buf_ld<T>(b1, o + m * NT * (int)sizeof(T));
    }
  }
  // publish the fetched rows in buf (entry p holds z_p, entry 0 the zero z_0);
  // the caller's barrier before this call retires the previous item's readers of buf
  __device__ __forceinline__ void stage(cx<T> *buf, int t) const {
    const int rp = t + (t >> 4);
#pragma unroll
    for (int m = 0; m < E; m++) {
      if constexpr (NT % 16 == 0)
        buf[rp + m * NT + ((m * NT) >> 4)] = mkc<T>(za[m], zb[m]);  // = rslot(t + m NT)
      else
        buf[rslot(t + m * NT)] = mkc<T>(za[m], zb[m]);
    }
  }
};

// Transform of the pipelined row passes: the register plan with the two-level
// LDS table for powers of two; for other lengths the LDS Stockham of fft.hpp on buf,
// with its full twiddle table in LDS (twf), so that no global twiddle load
// waits behind the prefetched rows.
template <typename T, int M, int EP>
__device__ __forceinline__ void fft_pipe(cx<T> *h, cx<T> *buf, const TwLds<T, M> &twa, const cx<T> *twf, int t) {
  if constexpr (is_pow2(M)) {
    fft_home_tw<T, M, EP>(h, buf, twa, t);
  } else {
    constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
    __syncthreads();  // the pre pass's readers of buf are done
#pragma unroll
    for (int m = 0; m < E; m++) buf[lds_slot(t + m * NT)] = h[m];
    __syncthreads();
    Stages<T, M, NT, 0>::run(buf, twf, t);
#pragma unroll
    for (int m = 0; m < E; m++) h[m] = buf[lds_slot(t + m * NT)];
  }
}
// the full length-M table of the LDS Stockham (lengths other than 2^k; one entry otherwise)
template <typename T, int M> constexpr int twf_len() { return is_pow2(M) ? 1 : M; }
template <typename T, int M>
__device__ __forceinline__ void twf_fill(cx<T> *twf, const cx<T> *__restrict__ tw, int t, int nt) {
  if constexpr (!is_pow2(M))
    for (int i = t; i < M; i += nt) twf[i] = tw[i];
}

template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p1hp(Geo g, const T *__restrict__ q, T *__restrict__ what, const cx<T> *__restrict__ twh,
       const cx<T> *__restrict__ tw2) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  __shared__ TwTab<T, M> tb;
  __shared__ cx<T> twf[twf_len<T, M>()];
  __shared__ TwTab<T, 2 * M> tb2;
  const int t = threadIdx.x;
  const int npair = (g.nl + 1) / 2;
  const int nitems = g.m * npair;
  twtab_fill(tb, twh, t, NT);
  twf_fill<T, M>(twf, twh, t, NT);
  twtab_fill(tb2, tw2, t, NT);
  const TwLds<T, M> twa{&tb};
  PipeRows<T, M, EP> rows;
  // storage column x = 1..n of row j holds x_1..x_N
  auto fetch = [&](int it, int tf) {
    const int lp = it % npair, j = it / npair + 1, l0 = 2 * lp;
    const bool has1 = l0 + 1 < g.nl;
    const T *r0 = q + l0 * g.LS + (long)j * g.P + 1;
    const T *r1 = q + (has1 ? l0 + 1 : l0) * g.LS + (long)j * g.P + 1;
    rows.fetch(r0, r1, has1, tf);
  };
  int it = blockIdx.x;
  if (it < nitems) fetch(it, t);
  for (; it < nitems; it += gridDim.x) {
    const int tl = loop_tid<NT>(t);
    __syncthreads();
    rows.stage(buf, tl);
    __syncthreads();
    if (it + (int)gridDim.x < nitems) fetch(it + gridDim.x, tl);
    cx<T> h[E];
    dst_pre_lds<T, M, EP>(h, tl, buf, tb2);
    fft_pipe<T, M, EP>(h, buf, twa, twf, tl);
    dst_post<T, M, EP>(h, buf, tl);
    const int lp = it % npair, j = it / npair + 1, l0 = 2 * lp, l1 = l0 + 1;
    const bool has1 = l1 < g.nl;
    T *o0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
    T *o1 = what + l1 * g.WS + (long)(j - 1) * g.Pn;
    // unrolled by two only: the full unroll keeps eight LDS reads in flight
    // and costs the kernel a wavefront per SIMD (88 against 76 VGPRs in FP32)
#pragma unroll 2
    for (int i = tl; i < M - 1; i += NT) {
      const cx<T> v = buf[rslot(i)];
      o0[i] = v.x;
      if (has1) o1[i] = v.y;
    }
  }
}

template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p3hp(Geo g, const T *__restrict__ what, T *__restrict__ psi, T *__restrict__ q, T ring_c,
       const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2, int psi_nt) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  __shared__ TwTab<T, M> tb;
  __shared__ cx<T> twf[twf_len<T, M>()];
  __shared__ TwTab<T, 2 * M> tb2;
  const int t = threadIdx.x;
  const int n = g.n, m_ = g.m;
  const int npair = (g.nl + 1) / 2;
  const int nitems = g.m * npair;
  twtab_fill(tb, twh, t, NT);
  twf_fill<T, M>(twf, twh, t, NT);
  twtab_fill(tb2, tw2, t, NT);
  const TwLds<T, M> twa{&tb};
  PipeRows<T, M, EP> rows;
  auto fetch = [&](int it, int tf) {
    const int lp = it % npair, j = it / npair + 1, l0 = 2 * lp;
    const bool has1 = l0 + 1 < g.nl;
    const T *i0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
    const T *i1 = what + (has1 ? l0 + 1 : l0) * g.WS + (long)(j - 1) * g.Pn;
    rows.fetch(i0, i1, has1, tf);
  };
  int it = blockIdx.x;
  if (it < nitems) fetch(it, t);
  for (; it < nitems; it += gridDim.x) {
    const int tl = loop_tid<NT>(t);
    __syncthreads();
    rows.stage(buf, tl);
    __syncthreads();
    if (it + (int)gridDim.x < nitems) fetch(it + gridDim.x, tl);
    cx<T> h[E];
    dst_pre_lds<T, M, EP>(h, tl, buf, tb2);
    fft_pipe<T, M, EP>(h, buf, twa, twf, tl);
    dst_post<T, M, EP>(h, buf, tl);
    const int lp = it % npair, j = it / npair + 1, l0 = 2 * lp, l1 = l0 + 1;
    const bool has1 = l1 < g.nl;
    T *p0 = psi + l0 * g.LS + (long)j * g.P;
    T *p1 = psi + l1 * g.LS + (long)j * g.P;
    T *q0r = q + l0 * g.LS;
    T *q1r = q + l1 * g.LS;
    // both layers in one pass over buf (one LDS read per entry), the wall-ring
    // values of q beside psi
#pragma unroll 2
    for (int i = tl; i < n; i += NT) {
      const int x = i + 1;
      const cx<T> c = buf[rslot(i)];
      psi_st(p0, x, c.x, psi_nt);
      if (has1) psi_st(p1, x, c.y, psi_nt);
      if (x == 1 || x == n) {
        const long o = (long)j * g.P + (x == 1 ? 0 : n + 1);
        q0r[o] = ring_c * c.x;
        if (has1) q1r[o] = ring_c * c.y;
      }
      if (j == 1 || j == m_) {
        const long o = (long)(j == 1 ? 0 : m_ + 1) * g.P + x;
        q0r[o] = ring_c * c.x;
        if (has1) q1r[o] = ring_c * c.y;
      }
    }
    if (tl == 0) {
      p0[0] = T(0);
      p0[n + 1] = T(0);
      if (has1) {
        p1[0] = T(0);
        p1[n + 1] = T(0);
      }
    }
    if (j == 1 || j == m_) {
      const long rr = (long)(j == 1 ? 0 : m_ + 1) * g.P;
      for (int x = tl; x < n + 2; x += NT) {
        psi[l0 * g.LS + rr + x] = T(0);
        if (has1) psi[l1 * g.LS + rr + x] = T(0);
      }
    }
  }
}

// ===========================================================================
// Periodic branch (bc_fac = -1, src/elliptic.h:32-47,103-118): the reference's
// in-place r2c / c2r 2-D transform pair, as
//   P1 (k_pf1): per (row, layer pair): FFT_N of a + i b, split into the two
//       half spectra A_k = (Z_k + conj Z_{N-k})/2, B_k = (Z_k - conj Z_{N-k})/(2i),
//       k = 0..N/2, stored interleaved (re, im) in what rows of N + 2 reals;
//   T1: the transpose puts the (re, im) pair of each x wavenumber in two
//       adjacent whatT rows;
//   P2 (k_pf2): per (mode, x wavenumber): FFT_N along y of re + i im, division
//       by -(L_j^2 + K_i^2 + 1/Rd_k^2) * N^2 where non-zero, the (0,0)
//       coefficient zeroed, inverse FFT;
//   P3 (k_pf3): per (row, layer pair): Z_p = A_p + i B_p (p <= N/2) and
//       conj A_{N-p} + i conj B_{N-p} (p > N/2), inverse FFT_N, a = Re, b = Im;
//       psi row, and the periodic rings of psi and q (src/domain.h:225-256).
// The imaginary parts of the k = 0 and k = N/2 coefficients are ignored in
// P3, as a c2r transform does.  Twiddles: length N.
// ===========================================================================
template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_pf1(Geo g, const T *__restrict__ q, T *__restrict__ what, const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[DstHalf<L>::slots];
  const int t = threadIdx.x;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *r0 = q + l0 * g.LS + (long)j * g.P + 1;  // x = p + 1
  const T *r1 = q + (has1 ? l1 : l0) * g.LS + (long)j * g.P + 1;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    h[m] = mkc<T>(r0[p], has1 ? r1[p] : T(0));
  }
  fft_home<T, L>(h, buf, tw, t);
  __syncthreads();
#pragma unroll
  for (int m = 0; m < E; m++) buf[rslot(t + m * NT)] = h[m];
  __syncthreads();
  T *o0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  T *o1 = what + l1 * g.WS + (long)(j - 1) * g.Pn;
  for (int k = t; k <= L / 2; k += NT) {
    const cx<T> Z = buf[rslot(k)], Zm = buf[rslot(k == 0 ? 0 : L - k)];
    o0[2 * k] = T(0.5) * (Z.x + Zm.x);
    o0[2 * k + 1] = T(0.5) * (Z.y - Zm.y);
    if (has1) {
      o1[2 * k] = T(0.5) * (Z.y + Zm.y);
      o1[2 * k + 1] = T(0.5) * (Zm.x - Z.x);
    }
  }
}

template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_pf2(Geo g, T *__restrict__ whatT, const double *__restrict__ Kx2, const double *__restrict__ Ly2,
      const double *__restrict__ iRd2, double norm, const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[home_slots<L>()];
  const int t = threadIdx.x;
  const int nk = g.n / 2 + 1;  // x wavenumbers
  const int k = blockIdx.x / nk, a = blockIdx.x % nk;
  T *re = whatT + k * g.WTS + (long)(2 * a) * g.Pm;
  T *im = re + g.Pm;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    h[m] = mkc<T>(re[p], im[p]);
  }
  fft_home<T, L>(h, buf, tw, t);
  const double kx2 = Kx2[a], ir = iRd2[k];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    const double fact = -(Ly2[p] + kx2 + ir);
    if (a == 0 && p == 0) {
      h[m] = mkc<T>(T(0), T(0));
    } else if (fact != 0.0) {
      const T f = (T)(1.0 / (fact * norm));
      h[m] = mkc<T>(h[m].x * f, h[m].y * f);
    }
    h[m] = mkc<T>(h[m].x, -h[m].y);  // inverse by conjugation
  }
  fft_home<T, L>(h, buf, tw, t);
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    re[p] = h[m].x;
    im[p] = -h[m].y;
  }
}

template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_pf3(Geo g, const T *__restrict__ what, T *__restrict__ psi, T *__restrict__ q, const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[home_slots<L>()];
  const int t = threadIdx.x;
  const int n = g.n, m_ = g.m;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *i0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  const T *i1 = what + (has1 ? l1 : l0) * g.WS + (long)(j - 1) * g.Pn;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    const int kk = (p <= L / 2) ? p : L - p;
    const bool edge = (kk == 0 || kk == L / 2);
    const T ar = i0[2 * kk], ai = edge ? T(0) : i0[2 * kk + 1];
    const T br = has1 ? i1[2 * kk] : T(0), bi = (has1 && !edge) ? i1[2 * kk + 1] : T(0);
    // Z = A + i B (p <= N/2) or conj A + i conj B (p > N/2); conjugated for the inverse
    const T sa = (p <= L / 2) ? T(1) : T(-1);
    const cx<T> z = mkc<T>(ar - sa * bi, sa * ai + br);
    h[m] = mkc<T>(z.x, -z.y);
  }
  fft_home<T, L>(h, buf, tw, t);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int l = side ? l1 : l0;
    if (side && !has1) break;
    T *pr = psi + l * g.LS;
    T *qr = q + l * g.LS;
    const long rj = (long)j * g.P;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int p = t + m * NT;
      const T v = side ? -h[m].y : h[m].x;  // conj(FFT(conj Z)): a = Re, b = Im
      const int x = p + 1;
      pr[rj + x] = v;
      if (x == n) pr[rj] = v;
      if (x == 1) pr[rj + n + 1] = v;
      if (j == m_) {  // row 0 is the copy of row N
        pr[x] = v;
        if (x == n) pr[0] = v;
        if (x == 1) pr[n + 1] = v;
      }
      if (j == 1) {  // row N+1 is the copy of row 1
        const long rt = (long)(m_ + 1) * g.P;
        pr[rt + x] = v;
        if (x == n) pr[rt] = v;
        if (x == 1) pr[rt + n + 1] = v;
      }
    }
    // q: periodic copies of its current interior
    if (t == 0) {
      qr[rj] = qr[rj + n];
      qr[rj + n + 1] = qr[rj + 1];
    }
    if (j == m_ || j == 1) {
      const long rs = rj, rt = (j == m_) ? 0 : (long)(m_ + 1) * g.P;
      for (int x = t; x < n + 2; x += NT) {
        const int xs = (x == 0) ? n : (x == n + 1) ? 1 : x;
        qr[rt + x] = qr[rs + xs];
      }
    }
  }
}

// ===========================================================================
// P3: inverse transform along x of two layers of one grid row, the psi row
// with its zero wall values, and the wall ring of q (src/domain.h:308-348:
// q ring = 2 bc_fac/Delta^2 * adjacent psi; corners are zero because psi's
// ring is).
// ===========================================================================
template <typename T, int L>
__global__ void __launch_bounds__(RegPlan<L>::NT)
k_p3(Geo g, const T *__restrict__ what, T *__restrict__ psi, T *__restrict__ q, T ring_c,
     const cx<T> *__restrict__ tw) {
  constexpr int E = RegPlan<L>::E, NT = RegPlan<L>::NT;
  __shared__ cx<T> buf[home_slots<L>()];
  const int t = threadIdx.x;
  const int n = g.n, m_ = g.m;
  const int npair = (g.nl + 1) / 2;
  const int lp = blockIdx.x % npair;
  const int j = blockIdx.x / npair + 1;
  const int l0 = 2 * lp, l1 = l0 + 1;
  const bool has1 = l1 < g.nl;
  const T *i0 = what + l0 * g.WS + (long)(j - 1) * g.Pn;
  const T *i1 = what + (has1 ? l1 : l0) * g.WS + (long)(j - 1) * g.Pn;
  cx<T> h[E];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    T a = T(0), b = T(0);
    if (p != 0 && p != L / 2) {
      const int x = (p < L / 2) ? p - 1 : L - p - 1;
      a = i0[x];
      b = has1 ? i1[x] : T(0);
      if (p > L / 2) {
        a = -a;
        b = -b;
      }
    }
    h[m] = mkc<T>(a, b);
  }
  fft_home<T, L>(h, buf, tw, t);
#pragma unroll
  for (int side = 0; side < 2; side++) {
    const int l = side ? l1 : l0;
    if (side && !has1) break;
    T *pr = psi + l * g.LS + (long)j * g.P;
    T *qr = q + l * g.LS;
#pragma unroll
    for (int m = 0; m < E; m++) {
      const int p = t + m * NT;
      if (p >= 1 && p < L / 2) {
        const T v = side ? h[m].x : -h[m].y;
        pr[p] = v;
        if (p == 1) qr[(long)j * g.P] = ring_c * v;
        if (p == n) qr[(long)j * g.P + n + 1] = ring_c * v;
        if (j == 1) qr[p] = ring_c * v;
        if (j == m_) qr[(long)(m_ + 1) * g.P + p] = ring_c * v;
      }
    }
    if (t == 0) {
      pr[0] = T(0);
      pr[n + 1] = T(0);
    }
    if (j == 1 || j == m_) {
      const long rr = (long)(j == 1 ? 0 : m_ + 1) * g.P;
      T *pw = psi + l * g.LS + rr;
      for (int x = t; x < n + 2; x += NT) pw[x] = T(0);
      if (t == 0) {
        qr[rr] = T(0);
        qr[rr + n + 1] = T(0);
      }
    }
  }
}

// ===========================================================================
// Scalars of the tendency/AB3 kernel (P4).
// ===========================================================================
struct StepCoef {
  double c0, c1, c2;     // AB3 coefficients
  double beta, nu, nu_hyper, inv_d2, inv_12d2, inv_2d;
  double drag;           // -hEkb*f0/(2*dh_bottom)
  double topo_fac;       // f0/dh_bottom
  double ring_c;         // 2*bc_fac/Delta^2
  int update;            // 0: write the tendency only (rhs), 1: also the AB3 update
  unsigned *nonfinite;   // set to 1 when an output (new q, or the tendency) is NaN or Inf; may be null
};

// Non-finite detection at one FMA per output: v * 0 is +-0 for finite v and
// NaN for NaN or Inf, so the running sum stays zero until a non-finite value
// enters it.  The flag is raised once per thread at the end of its work.
template <typename T> __device__ __forceinline__ T nf_acc(T acc, T v) { return __builtin_fma(v, T(0), acc); }
__device__ __forceinline__ float nf_acc(float acc, float v) { return __builtin_fmaf(v, 0.f, acc); }
template <typename T> __device__ __forceinline__ void nf_flag(T acc, unsigned *flag) {
  if (flag && acc != acc) atomicOr(flag, 1u);
}

// Visit every point (ly, lx) of a (TX + 2R) x HH region with the 4 x TX
// threads of a P4 workgroup: the TX middle columns one column per thread, the
// 2R side columns as one flattened pass, so no pass runs mostly masked.
template <int TX, int R, int HH, typename F>
__device__ __forceinline__ void p4_region(int tx, int ty, F &&f) {
#pragma unroll
  for (int ly = ty; ly < HH; ly += 4) f(ly, R + tx);
  if constexpr (R > 0) {
    const int tid = ty * TX + tx;
#pragma unroll
    for (int e = tid; e < 2 * R * HH; e += 4 * TX) {
      const int ly = e / (2 * R), c = e % (2 * R);
      f(ly, c < R ? c : TX + c);
    }
  }
}

template <typename T, int TX, int TY, int NLAP> struct P4Smem {
  static constexpr int H = (NLAP > 1) ? NLAP : 1;  // psi halo
  static constexpr int PW = TX + 2 * H, PH = TY + 2 * H;
  static constexpr int AW = (NLAP > 1) ? PW - 2 : 1, AH = (NLAP > 1) ? PH - 2 : 1;
  static constexpr int BW = (NLAP > 2) ? PW - 4 : 1, BH = (NLAP > 2) ? PH - 4 : 1;
  T P[PH][PW];
  T A[AH][AW];
  T B[BH][BW];
  T Q[TY + 2][TX + 2];
};

// Laplacian level of radius RD (tile interior +- RD) from the level below
// (radius RD+1, row width SW) into dst (row width DW).  Interior tiles (FAST)
// need no wall handling; elsewhere a wall-ring point takes the value
// 2 bc_fac/Delta^2 * adjacent psi (the ring set by adjust_bc(...,1,psi) after
// every Laplacian of calc_diff, src/dynamics.h:25-43) and points beyond the
// ring are never used.
template <typename T, int TX, int TY, int H, bool FAST, int RD, int SW, int DW>
__device__ __forceinline__ void p4_level(const T *src, T *dst, const T *sP, int x0, int y0, int n, int m, T inv_d2,
                                         T ring_c, int tx, int ty) {
  constexpr int HH = TY + 2 * RD, PW = TX + 2 * H;
  p4_region<TX, RD, HH>(tx, ty, [&](int ly, int lx) {
    const T *c = src + (ly + 1) * SW + (lx + 1);
    T v;
    if constexpr (FAST) {
      v = (c[1] + c[-1] + c[SW] + c[-SW] - T(4) * c[0]) * inv_d2;
    } else {
      const int gx = x0 - RD + lx, gy = y0 - RD + ly;
      if (gx >= 1 && gx <= n && gy >= 1 && gy <= m) {
        v = (c[1] + c[-1] + c[SW] + c[-SW] - T(4) * c[0]) * inv_d2;
      } else if (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) {
        const T *pp = sP + (ly + H - RD) * PW + (lx + H - RD);  // psi at this point
        if (gx == 0)
          v = (gy == 0 || gy == m + 1) ? T(0) : ring_c * pp[1];
        else if (gx == n + 1)
          v = (gy == 0 || gy == m + 1) ? T(0) : ring_c * pp[-1];
        else if (gy == 0)
          v = ring_c * pp[PW];
        else
          v = ring_c * pp[-PW];
      } else {
        v = T(0);
      }
    }
    dst[ly * DW + lx] = v;
  });
}

// PER: periodic domain (bc_fac = -1): the psi and q tiles are loaded with
// wrapped indices (the stored rings are periodic copies, src/domain.h:225-256,
// and so is every Laplacian level), the levels use the plain stencil
// everywhere, and the whole q ring (corners included) is carried.
// Index v of any sign mapped to 1..N (a tile with its halo may span more
// than one period on small grids).
__device__ __forceinline__ int wrap1(int v, int N) { return ((v - 1) % N + N) % N + 1; }

// Levels LEV..NLAP-1 of the nested Laplacians: odd levels land in A, even in
// B (A holds the largest odd level, B the largest even one).
template <typename T, int TX, int TY, int H, bool FAST, int NLAP, int LEV, class SM>
__device__ __forceinline__ void p4_levels(SM &sm, const T *sPp, int x0, int y0, int n, int m, T inv_d2, T ring_c,
                                          int tx, int ty) {
  if constexpr (LEV <= NLAP - 1) {
    constexpr int PW = TX + 2 * H;
    constexpr int SW = (LEV == 1) ? PW : ((LEV % 2 == 0) ? SM::AW : SM::BW);
    constexpr int DW = (LEV % 2 == 1) ? SM::AW : SM::BW;
    const T *src = (LEV == 1) ? sPp : ((LEV % 2 == 0) ? &sm.A[0][0] : &sm.B[0][0]);
    T *dst = (LEV % 2 == 1) ? &sm.A[0][0] : &sm.B[0][0];
    p4_level<T, TX, TY, H, FAST, H - LEV, SW, DW>(src, dst, sPp, x0, y0, n, m, inv_d2, ring_c, tx, ty);
    __syncthreads();
    p4_levels<T, TX, TY, H, FAST, NLAP, LEV + 1>(sm, sPp, x0, y0, n, m, inv_d2, ring_c, tx, ty);
  }
}

template <typename T, int NLAP, int KAHAN, int TX, int TY, bool FAST, bool PER = false>
__device__ __forceinline__ void p4_body(P4Smem<T, TX, TY, NLAP> &sm, const Geo &g, const T *__restrict__ P,
                                        const T *__restrict__ Q, const T *__restrict__ topo, int k, int x0, int y0,
                                        int tx, int ty, const T (&hf1)[TY / 4], const T (&hf2)[TY / 4],
                                        const comp_t<T> (&hcp)[TY / 4], T *__restrict__ qn, T *__restrict__ fnew,
                                        comp_t<T> *__restrict__ comp, const T *__restrict__ wind, const StepCoef &cf) {
  using SM = P4Smem<T, TX, TY, NLAP>;
  constexpr int H = SM::H, PW = SM::PW, PH = SM::PH;
  (void)PW;
  const int n = g.n, m = g.m;
  const bool bottom = (k == g.nlm - 1);  // k: layer within the member
  // ---- psi (halo H) and q (halo 1) tiles; zero outside the storage grid
  p4_region<TX, H, PH>(tx, ty, [&](int ly, int lx) {
    const int gx = x0 - H + lx, gy = y0 - H + ly;
    if constexpr (PER)
      sm.P[ly][lx] = P[wrap1(gy, m) * g.P + wrap1(gx, n)];
    else if constexpr (FAST)
      sm.P[ly][lx] = P[gy * g.P + gx];
    else
      sm.P[ly][lx] = (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) ? P[gy * g.P + gx] : T(0);
  });
  p4_region<TX, 1, TY + 2>(tx, ty, [&](int ly, int lx) {
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    if constexpr (PER) {
      sm.Q[ly][lx] = Q[wrap1(gy, m) * g.P + wrap1(gx, n)];
    } else {
      const bool in = FAST || (gx <= n + 1 && gy <= m + 1);
      sm.Q[ly][lx] = in ? Q[gy * g.P + gx] : T(0);
    }
  });
  __syncthreads();
  const T inv_d2 = (T)cf.inv_d2;
  const T ring_c = (T)cf.ring_c;
  const T *sPp = &sm.P[0][0];
  // ---- nested Laplacians of psi down to radius 1 (calc_diff, src/dynamics.h:25-43):
  // level 1 from psi into A, then alternately A -> B and B -> A (radius H - level)
  p4_levels<T, TX, TY, H, FAST || PER, NLAP, 1>(sm, sPp, x0, y0, n, m, inv_d2, ring_c, tx, ty);
  // the radius-1 level feeding the last Laplacian: A for even NLAP, B for odd
  constexpr bool lastA = (NLAP % 2) == 0;
  constexpr int LW = lastA ? SM::AW : SM::BW;
  const T *last = lastA ? &sm.A[0][0] : &sm.B[0][0];

  const T beta = (T)cf.beta, nu = (T)cf.nu, nuh = (T)cf.nu_hyper;
  const T i12 = (T)cf.inv_12d2, i2d = (T)cf.inv_2d;
  const T c0 = (T)cf.c0, c1 = (T)cf.c1, c2 = (T)cf.c2;
  const T drag = (T)cf.drag, tfac = (T)cf.topo_fac;
  const bool use_nu = cf.nu != 0.0;
  const int gx = x0 + tx;
  T nfa = T(0);
#pragma unroll
  for (int r = 0; r < TY / 4; r++) {
    const int tyr = ty + 4 * r;
    const int gy = y0 + tyr;
    if (!FAST && (gx > n || gy > m)) continue;
    const int px = tx + H, py = tyr + H;  // sP index
    const int qx = tx + 1, qy = tyr + 1;  // sQ index
#define PS(a, b) sm.P[py + (b)][px + (a)]
#define QS(a, b) sm.Q[qy + (b)][qx + (a)]
#define TS(a, b) tp[(b)*g.P + (a)]
    // Arakawa Jacobian (src/domain.h:17-27)
    const T jpq = ((PS(1, 0) - PS(-1, 0)) * (QS(0, 1) - QS(0, -1)) + (PS(0, -1) - PS(0, 1)) * (QS(1, 0) - QS(-1, 0)) +
                   PS(1, 0) * (QS(1, 1) - QS(1, -1)) - PS(-1, 0) * (QS(-1, 1) - QS(-1, -1)) -
                   PS(0, 1) * (QS(1, 1) - QS(-1, 1)) + PS(0, -1) * (QS(1, -1) - QS(-1, -1)) +
                   QS(0, 1) * (PS(1, 1) - PS(-1, 1)) - QS(0, -1) * (PS(1, -1) - PS(-1, -1)) -
                   QS(1, 0) * (PS(1, 1) - PS(1, -1)) + QS(-1, 0) * (PS(-1, 1) - PS(-1, -1))) *
                  i12;
    T f = -jpq - beta * (PS(1, 0) - PS(-1, 0)) * i2d;
    if (use_nu) f += nu * ((QS(1, 0) + QS(-1, 0) + QS(0, 1) + QS(0, -1) - T(4) * QS(0, 0)) * inv_d2);
    const T lap1 = (PS(1, 0) + PS(-1, 0) + PS(0, 1) + PS(0, -1) - T(4) * PS(0, 0)) * inv_d2;
    if constexpr (NLAP == 1) {
      f += nuh * lap1;
    } else if constexpr (NLAP >= 2) {
      const T *c = last + (tyr + 1) * LW + (tx + 1);
      f += nuh * ((c[1] + c[-1] + c[LW] + c[-LW] - T(4) * c[0]) * inv_d2);
    }
    if (k == 0) f += wind[gy];
    if (bottom) {
      const T *tp = topo + gy * g.P + gx;  // bottom topography, read through the cache
      const T jt = ((PS(1, 0) - PS(-1, 0)) * (TS(0, 1) - TS(0, -1)) + (PS(0, -1) - PS(0, 1)) * (TS(1, 0) - TS(-1, 0)) +
                    PS(1, 0) * (TS(1, 1) - TS(1, -1)) - PS(-1, 0) * (TS(-1, 1) - TS(-1, -1)) -
                    PS(0, 1) * (TS(1, 1) - TS(-1, 1)) + PS(0, -1) * (TS(1, -1) - TS(-1, -1)) +
                    TS(0, 1) * (PS(1, 1) - PS(-1, 1)) - TS(0, -1) * (PS(1, -1) - PS(-1, -1)) -
                    TS(1, 0) * (PS(1, 1) - PS(1, -1)) + TS(-1, 0) * (PS(-1, 1) - PS(-1, -1))) *
                   i12;
      f += drag * lap1 - jt * tfac;
    }
    const T qo = QS(0, 0);
#undef PS
#undef QS
#undef TS
    const int o = gy * g.P + gx;  // offset inside layer k (the pointers carry the layer)
    fnew[o] = f;
    if (!cf.update) {
      nfa = nf_acc(nfa, f);
      continue;
    }
    const T inc = c0 * f + c1 * hf1[r] + c2 * hf2[r];
    if constexpr (KAHAN) {
      const T y = inc - comp_val<T>(hcp[r], qo);
      const T tq = qo + y;
      comp[o] = comp_make<T>((tq - qo) - y, tq);
      qn[o] = tq;
      nfa = nf_acc(nfa, tq);
    } else {
      const T tq = qo + inc;
      qn[o] = tq;
      nfa = nf_acc(nfa, tq);
    }
    // carry q's wall ring into the next buffer (bounded: corners stay zero)
    if (gy == 1) qn[gx] = sm.Q[qy - 1][qx];
    if (gy == m) qn[(m + 1) * g.P + gx] = sm.Q[qy + 1][qx];
    if (gx == 1) qn[gy * g.P] = sm.Q[qy][qx - 1];
    if (gx == n) qn[gy * g.P + n + 1] = sm.Q[qy][qx + 1];
    if constexpr (PER) {
      if (gy == 1 && gx == 1) qn[0] = sm.Q[qy - 1][qx - 1];
      if (gy == 1 && gx == n) qn[n + 1] = sm.Q[qy - 1][qx + 1];
      if (gy == m && gx == 1) qn[(m + 1) * g.P] = sm.Q[qy + 1][qx - 1];
      if (gy == m && gx == n) qn[(m + 1) * g.P + n + 1] = sm.Q[qy + 1][qx + 1];
    }
  }
  nf_flag(nfa, cf.nonfinite);
}

// ===========================================================================
// P4: tendency and AB3 update on a TX x TY tile of one layer.
//   f = -J(psi,q) - beta psi_x + nu lap q + nu_hyper lap^NLAP psi
//   layer 0:    += wind(y_j)
//   layer nl-1: += -hEkb f0/(2 dh) lap psi - J(psi, topo) f0/dh
//   q_next = q + c0 f + c1 f1 + c2 f2 (Kahan-compensated when KAHAN)
// (rhs, src/dynamics.h:47-73; timestep, src/timestep.h:135-141).  NLAP =
// n_hyper/2 + 1 nested 5-point Laplacians (calc_diff); NLAP = 0 skips the
// term (nu_hyper = 0).  Tiles whose psi halo lies inside the storage grid and
// whose Laplacian levels lie inside the interior take a branch-free path.
// ===========================================================================
template <typename T, int NLAP, int KAHAN, int TX, int TY, bool PER = false>
__global__ void __launch_bounds__(TX * 4)
k_p4(Geo g, const T *__restrict__ psi, const T *__restrict__ q, T *__restrict__ qn, T *__restrict__ fnew,
     const T *__restrict__ f1, const T *__restrict__ f2, comp_t<T> *__restrict__ comp, const T *__restrict__ topo,
     const T *__restrict__ wind, StepCoef cf) {
  using SM = P4Smem<T, TX, TY, NLAP>;
  constexpr int H = SM::H;
  __shared__ SM sm;
  const int n = g.n, m = g.m;
  // Tiles are taken in raster order per XCD: workgroup ids are dealt to the
  // 8 XCDs round-robin, so each XCD gets one contiguous run of tiles and the
  // halo rows shared by vertically adjacent tiles stay in its L2.
  const int ntx = gridDim.x, nty = gridDim.y;
  const int tile = xcd_remap(blockIdx.x + ntx * (blockIdx.y + nty * blockIdx.z), ntx * nty * gridDim.z);
  const int k = tile / (ntx * nty);
  const int tyi = (tile / ntx) % nty, txi = tile % ntx;
  const int x0 = 1 + txi * TX, y0 = 1 + tyi * TY;  // first interior point of the tile
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  // AB3 history and compensation of this thread's outputs, fetched before the tiles
  const long lo = k * g.LS;
  f1 += lo;
  f2 += lo;
  qn += lo;
  fnew += lo;
  if (KAHAN) comp += lo;
  T hf1[TY / 4], hf2[TY / 4];
  comp_t<T> hcp[TY / 4];
#pragma unroll
  for (int r = 0; r < TY / 4; r++) {
    const int gx = x0 + tx, gy = y0 + ty + 4 * r;
    const bool ok = cf.update && gx <= n && gy <= m;
    const int o = gy * g.P + gx;
    hf1[r] = ok ? f1[o] : T(0);
    hf2[r] = ok ? f2[o] : T(0);
    hcp[r] = (ok && KAHAN) ? comp[o] : comp_t<T>(0);
  }
  const T *P = psi + lo;
  const T *Q = q + lo;
  const int mb = k / g.nlm, km = k - mb * g.nlm;  // member and its layer
  topo += mb * g.LS;
  wind += (long)mb * (m + 2);
  const bool fast = (x0 - H >= 0) && (x0 + TX - 1 + H <= n + 1) && (y0 - H >= 0) && (y0 + TY - 1 + H <= m + 1);
  if constexpr (PER)
    p4_body<T, NLAP, KAHAN, TX, TY, false, true>(sm, g, P, Q, topo, km, x0, y0, tx, ty, hf1, hf2, hcp, qn, fnew,
                                                 comp, wind, cf);
  else if (fast)
    p4_body<T, NLAP, KAHAN, TX, TY, true>(sm, g, P, Q, topo, km, x0, y0, tx, ty, hf1, hf2, hcp, qn, fnew, comp,
                                          wind, cf);
  else
    p4_body<T, NLAP, KAHAN, TX, TY, false>(sm, g, P, Q, topo, km, x0, y0, tx, ty, hf1, hf2, hcp, qn, fnew, comp,
                                           wind, cf);
}

// ===========================================================================
// P4, streaming form.  One wavefront owns a strip of 64 storage columns (its
// 64 - 2H middle columns are outputs, H = max(NLAP, 1)) of one layer and
// walks a chunk of rows.  psi, q, the nested Laplacian levels and the
// Jacobian's 3x3 neighbourhoods live in rolling register windows; the
// horizontal neighbours come from DPP wavefront shifts (wave_shl:1 gives lane
// i the value of lane i+1, wave_shr:1 that of lane i-1); there is no LDS and
// no barrier.  Each psi row is read once per chunk (plus 2H+2 warm-up rows).
// Same arithmetic and wall rules as the tiled form above.
// ===========================================================================
// bound_ctrl set: a lane with no source lane (63 for the right shift, 0 for
// the left) receives 0, and no old value has to be staged in the destination.
__device__ __forceinline__ float dpp_right(float v) {  // value of lane i+1
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_left(float v) {  // value of lane i-1
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double dpp_right(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double dpp_left(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Element at byte offset xb of a row: a wavefront-uniform row pointer plus an
// unsigned per-lane offset, which maps to the scalar-base form of the memory
// instructions (no 64-bit address arithmetic per lane).
// The row pointer passes through readfirstlane (free for a value already in
// scalar registers) so that the compiler keeps it scalar instead of hoisting a
// per-lane "base + offset" pointer out of the row loop; it is rebuilt as a
// global-address-space pointer so that the scalar-base load/store forms apply.
typedef __attribute__((address_space(1))) char gchar;
__device__ __forceinline__ gchar *uniform_gptr(const void *p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (gchar *)(((unsigned long long)hi << 32) | lo);
}
template <typename T> __device__ __forceinline__ T ld_at(const T *rowp, unsigned xb) {
  return *(const __attribute__((address_space(1))) T *)(uniform_gptr(rowp) + xb);
}

// Raw buffer view of one layer of a field: a scalar resource, a per-lane byte
// offset (VGPR) and a per-row byte offset (SGPR) -- the loads carry no
// per-lane address arithmetic.  Resource word 3 = 0x00020000 (32-bit data
// format, the gfx9 family); stride 0; num_records = the bytes of the view.
struct BufView {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ BufView buf_view(const void *base, unsigned bytes) {
  return BufView{__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000)};
}
// AUX: the cache-policy bits of the instruction (0: default caching).
template <typename T, int AUX = 0> __device__ __forceinline__ T buf_ld(const BufView &b, unsigned xb, unsigned rowb) {
  if constexpr (sizeof(T) == 2)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(b.r, (int)xb, (int)rowb, AUX));
  else if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(b.r, (int)xb, (int)rowb, AUX));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(b.r, (int)xb, (int)rowb, AUX));
}
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
template <int AUX = 0> __device__ __forceinline__ void buf_st(const BufView &b, unsigned xb, unsigned rowb, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), b.r, (int)xb, (int)rowb, AUX);
}
template <int AUX = 0> __device__ __forceinline__ void buf_st(const BufView &b, unsigned xb, unsigned rowb, double v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), b.r, (int)xb, (int)rowb, AUX);
}
template <typename T> __device__ __forceinline__ void st_at(T *rowp, unsigned xb, T v) {
  asm volatile("" : "+v"(xb));
  *(__attribute__((address_space(1))) T *)(uniform_gptr(rowp) + xb) = v;
}
__device__ __forceinline__ void buf_st(const BufView &b, unsigned xb, unsigned rowb, short v) {
  __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), b.r, (int)xb, (int)rowb, 0);
}

// 12 Delta^2 J(p, q) in the 12-term form of src/domain.h:17-27, from the 3x3
// neighbourhoods (c: row j, u: row j+1, d: row j-1; l: column i-1, r: i+1).
template <typename T>
__device__ __forceinline__ T arakawa12(T pcl, T pcr, T pu, T pd, T pul, T pur, T pdl, T pdr, T qcl, T qcr, T qu,
                                       T qd, T qul, T qur, T qdl, T qdr) {
  return (pcr - pcl) * (qu - qd) + (pd - pu) * (qcr - qcl) + pcr * (qur - qdr) - pcl * (qul - qdl) -
         pu * (qur - qul) + pd * (qdr - qdl) + qu * (pur - pul) - qd * (pdr - pdl) - qcr * (pur - pdr) +
         qcl * (pul - pdl);
}

template <int I> struct IC {
  static constexpr int value = I;
};
// slot of row r - i in a circular window of W rows written at phase ph (= r mod W)
template <int W> __device__ constexpr int wslot(int ph, int i) { return ((ph - i) % W + W) % W; }

// The row walk of one wavefront.  The loop is unrolled by W (a multiple of 2
// and 3 holding the psi rows in use), so every window slot and prefetch buffer
// is a fixed register: no moves shift the windows.  EDGE = false: the strip
// and its rows touch no wall, so every load is in bounds and no wall rule or
// ring carry applies.
// OWN: the walk stores only the output columns X0 .. X0 + ownw - 1 (the
// block's own columns in k_p4a); the other outputs of its strip belong to a
// neighbouring block, which computes them as well.
template <typename T, int NLAP, int KAHAN, bool EDGE, bool OWN = false, bool REV = false>
__device__ __forceinline__ void p4s_walk(const Geo &g, const T *__restrict__ P, const T *__restrict__ Qf,
                                         T *__restrict__ qn, T *__restrict__ fnew, const T *__restrict__ f1,
                                         const T *__restrict__ f2, comp_t<T> *__restrict__ comp,
                                         const T *__restrict__ topo, const T *__restrict__ wind, const StepCoef &cf,
                                         int x, int lane, int km,
                                         int y0, int y1, bool ringstrip, int X0 = 0, int ownw = 0) {
  constexpr int H = (NLAP > 1) ? NLAP : 1;       // halo, and the delay of the output row
  constexpr int NP = H + 2;                      // psi rows r, r-1, ..., r-H-1
  constexpr int D = 3;                           // prefetch depth (iterations)
  constexpr int W = (NP <= 6) ? 6 : 12;          // unroll and psi window length (a multiple of 3)
  constexpr int RING = W;                        // prefetch slots: row r uses slot (r - rbeg) mod W
  static_assert(D < RING && W % 3 == 0, "unroll");
  constexpr int NL = (NLAP > 1) ? NLAP - 1 : 1;  // stored levels 1..NLAP-1
  static_assert(NP <= W, "psi window");
  const int n = g.n, m = g.m;
  const bool xin = !EDGE || (x >= 0 && x <= n + 1);
  const bool xring = EDGE && (x == 0 || x == n + 1);
  const bool outl = lane >= H && lane < 64 - H && (!EDGE || (x >= 1 && x <= n));
  const int xc = xin ? x : 0;  // clamped column for addresses
  const bool bottom = (km == g.nlm - 1);
  const bool upd = cf.update != 0;
  const T inv_d2 = (T)cf.inv_d2, ring_c = (T)cf.ring_c;
  const T beta = (T)cf.beta, nu = (T)cf.nu, nuh = (T)cf.nu_hyper;
  const T i12 = (T)cf.inv_12d2, i2d = (T)cf.inv_2d;
  const T c0 = (T)cf.c0, c1 = (T)cf.c1, c2 = (T)cf.c2;
  const T drag = (T)cf.drag, tfac = (T)cf.topo_fac;
  const bool use_nu = cf.nu != 0.0;
  // REV: the walk runs up its chunk, virtual row v being grid row y0 + y1 - 1 - v
  // (an interior walk only); every stencil below is written for grid rows u =
  // y + 1 and d = y - 1, so both directions give the same bits
  static_assert(!(REV && EDGE), "reversed walks are interior walks");
  const int ymir = y0 + y1 - 1;
  auto prow = [&](int v) -> unsigned { return (unsigned)(REV ? ymir - v : v); };

  T pw[W], nbw[W];                                    // psi rows; psi of the ring columns' neighbour
  T lv[NL][3];                                        // level a: rows r-a, r-a-1, r-a-2 (circular)
  T jpc[3], jpl[3], jpr[3], jqc[3], jql[3], jqr[3];  // rows y+1, y, y-1 of psi and q (circular)
  T jtc[3], jtl[3], jtr[3];                           // topography rows (bottom layer)
  T bp[RING], bq[RING], bt[RING], b1[RING], b2[RING], bw[RING];  // loads of r .. r+D (D + 1 slots live)
  comp_t<T> bc[RING];
#pragma unroll
  for (int i = 0; i < W; i++) pw[i] = nbw[i] = T(0);
#pragma unroll
  for (int a = 0; a < NL; a++)
#pragma unroll
    for (int i = 0; i < 3; i++) lv[a][i] = T(0);
#pragma unroll
  for (int i = 0; i < 3; i++) {
    jpc[i] = jpl[i] = jpr[i] = T(0);
    jqc[i] = jql[i] = jqr[i] = T(0);
    jtc[i] = jtl[i] = jtr[i] = T(0);
  }
#pragma unroll
  for (int i = 0; i < RING; i++) {
    bp[i] = bq[i] = bt[i] = b1[i] = b2[i] = bw[i] = T(0);
    bc[i] = comp_t<T>(0);
  }

  const int rbeg = y0 - H, rend = y1 + H;
  const unsigned xb = (unsigned)xc * sizeof(T);  // byte offset of this lane's column
  const unsigned lbytes = (unsigned)(g.LS * sizeof(T));
  const BufView vP = buf_view(P, lbytes), vQ = buf_view(Qf, lbytes), vT = buf_view(topo, lbytes);
  const BufView v1 = buf_view(f1, lbytes), v2 = buf_view(f2, lbytes);
  // the compensation's element is sizeof(comp_t<T>) bytes: its offsets scale by CS / sizeof(T)
  constexpr unsigned CS = sizeof(comp_t<T>);
  const unsigned xbc = (unsigned)xc * CS, rowbc = (unsigned)g.P * CS;
  const BufView vC = buf_view(comp, (unsigned)(g.LS * CS));
  const unsigned rowb = (unsigned)(g.P * sizeof(T));
  const BufView vW = buf_view(wind, (unsigned)((m + 2) * sizeof(T)));
  // loads of iteration r: psi row r, q and topography row r-H+1, AB3 history
  // and wind of row r-H; issued D iterations ahead
  auto fetch = [&](auto Sc, int r) {
    constexpr int s = decltype(Sc)::value;
    // wind of row r-H: a vector load (in-order completion), so that waiting
    // for it does not also wait for the scalar loads issued since
    if (km == 0) {
      if constexpr (EDGE)
        bw[s] = (r - H >= 0 && r - H <= m + 1) ? wind[r - H] : T(0);
      else
        bw[s] = buf_ld<T>(vW, 0u, prow(r - H) * (unsigned)sizeof(T));
    }
    if constexpr (EDGE) {
      auto ldrow = [&](const T *base, int row) -> T {
        return (xin && row >= 0 && row <= m + 1) ? ld_at(base + row * g.P, xb) : T(0);
      };
      auto ldout = [&](const T *base, int row) -> T {
        return (outl && row >= y0 && row < y1) ? ld_at(base + row * g.P, xb) : T(0);
      };
      bp[s] = ldrow(P, r);
      if (r + 1 - H >= y0 - 1) {  // q and topography rows y0 - 1 .. y1 only (wavefront-uniform)
        bq[s] = ldrow(Qf, r + 1 - H);
        if (bottom) bt[s] = ldrow(topo, r + 1 - H);
      }
      if (upd) {
        b1[s] = ldout(f1, r - H);
        b2[s] = ldout(f2, r - H);
        if (KAHAN) bc[s] = (outl && r - H >= y0 && r - H < y1) ? ld_at(comp + (r - H) * g.P, xbc) : comp_t<T>(0);
      }
    } else {
      // every row an interior walk touches lies inside the layer
      if (r >= rend) return;  // past the walk (wavefront-uniform)
      bp[s] = buf_ld<T>(vP, xb, prow(r) * rowb);
      // the Jacobians and nu lap q use q (and the topography) at rows y0 - 1
      // .. y1 only: the first 2H - 2 warm-up rows load psi alone (wavefront-uniform)
      if (r + 1 - H >= y0 - 1) {
        bq[s] = buf_ld<T>(vQ, xb, prow(r + 1 - H) * rowb);
        if (bottom) bt[s] = buf_ld<T>(vT, xb, prow(r + 1 - H) * rowb);
      }
      if (r - H >= y0) {  // the AB3 history only for output rows (wavefront-uniform)
        b1[s] = buf_ld<T>(v1, xb, prow(r - H) * rowb);
        b2[s] = buf_ld<T>(v2, xb, prow(r - H) * rowb);
        if (KAHAN) bc[s] = buf_ld<comp_t<T>>(vC, xbc, prow(r - H) * rowbc);
      }
    }
  };
  const BufView vF = buf_view(fnew, lbytes), vN = buf_view(qn, lbytes);
  T nfa = T(0);  // non-finite detection (nf_acc)
  auto body = [&](auto PHc, int r) {
    constexpr int ph = decltype(PHc)::value;
    constexpr int cur = ph % RING;
    constexpr int j0 = ph % 3, j1 = (ph + 2) % 3, j2 = (ph + 1) % 3;  // virtual rows y+1, y, y-1
    constexpr int jU = REV ? j2 : j0, jD = REV ? j0 : j2;              // grid rows y+1, y-1
    const int y = r - H;  // output row of this iteration
    if (r + D < rend) fetch(IC<(ph + D) % RING>{}, r + D);
    const T p0 = bp[cur];
    pw[wslot<W>(ph, 0)] = p0;
    if constexpr (EDGE) {
      if (ringstrip) {
        const T pr = dpp_right(p0), pl = dpp_left(p0);
        nbw[wslot<W>(ph, 0)] = (x == 0) ? pr : pl;
      }
    }
    // ---- nested Laplacians: level a at row r - a (calc_diff, src/dynamics.h:25-43)
    T hyp = T(0);
    if constexpr (NLAP >= 1) {
      T s0 = pw[wslot<W>(ph, 0)], s1 = pw[wslot<W>(ph, 1)], s2 = pw[wslot<W>(ph, 2)];
#pragma unroll
      for (int a = 1; a <= NLAP; a++) {
        const T su = REV ? s2 : s0, sd = REV ? s0 : s2;  // grid rows above and below
        T v = (dpp_left(s1) + dpp_right(s1) + su + sd - T(4) * s1) * inv_d2;
        if constexpr (EDGE) {
          const int yl = r - a;
          if (yl == 0 || yl == m + 1) {
            // wall row: 2 bc_fac/Delta^2 * psi of the adjacent interior row; corners 0
            v = xring ? T(0) : ring_c * ((yl == 0) ? pw[wslot<W>(ph, a - 1)] : pw[wslot<W>(ph, a + 1)]);
          } else if (xring) {
            v = ring_c * nbw[wslot<W>(ph, a)];  // wall column: its interior neighbour's psi
          }
        }
        if (a < NLAP) {
          lv[a - 1][j0] = v;
          s0 = v;
          s1 = lv[a - 1][j1];
          s2 = lv[a - 1][j2];
        } else {
          hyp = v;
        }
      }
    }
    // ---- 3x3 windows for the Jacobians at row y
    const T pn = pw[wslot<W>(ph, H - 1)];  // psi row y+1
    const T qnew = bq[cur];
    jpc[j0] = pn;
    jpl[j0] = dpp_left(pn);
    jpr[j0] = dpp_right(pn);
    jqc[j0] = qnew;
    jql[j0] = dpp_left(qnew);
    jqr[j0] = dpp_right(qnew);
    if (bottom) {
      const T tnew = bt[cur];
      jtc[j0] = tnew;
      jtl[j0] = dpp_left(tnew);
      jtr[j0] = dpp_right(tnew);
    }
    if (y < y0) return;  // warm-up row (wavefront-uniform)
    // ---- tendency at row y (rhs, src/dynamics.h:52-71)
    const T jpq = arakawa12<T>(jpl[j1], jpr[j1], jpc[jU], jpc[jD], jpl[jU], jpr[jU], jpl[jD], jpr[jD], jql[j1],
                               jqr[j1], jqc[jU], jqc[jD], jql[jU], jqr[jU], jql[jD], jqr[jD]) *
                  i12;
    T f = -jpq - beta * (jpr[j1] - jpl[j1]) * i2d;
    if (use_nu) f += nu * ((jql[j1] + jqr[j1] + jqc[jU] + jqc[jD] - T(4) * jqc[j1]) * inv_d2);
    if constexpr (NLAP >= 1) f += nuh * hyp;
    if (km == 0) f += bw[cur];
    if (bottom) {
      const T lap1 = (jpl[j1] + jpr[j1] + jpc[jU] + jpc[jD] - T(4) * jpc[j1]) * inv_d2;
      const T jt = arakawa12<T>(jpl[j1], jpr[j1], jpc[jU], jpc[jD], jpl[jU], jpr[jU], jpl[jD], jpr[jD], jtl[j1],
                                jtr[j1], jtc[jU], jtc[jD], jtl[jU], jtr[jU], jtl[jD], jtr[jD]) *
                   i12;
      f += drag * lap1 - jt * tfac;
    }
    if (!outl) return;
    const int ro = y * g.P;  // output row offset (wavefront-uniform; edge walks)
    const unsigned rb = prow(y) * rowb;
    if constexpr (OWN) {
      if ((unsigned)(x - X0) >= (unsigned)ownw) return;  // a neighbouring block's column
    }
    if constexpr (EDGE)
      st_at(fnew + ro, xb, f);
    else
      buf_st(vF, xb, rb, f);
    if (!upd) {
      nfa = nf_acc(nfa, f);
      return;
    }
    const T qo = jqc[j1];
    const T inc = c0 * f + c1 * b1[cur] + c2 * b2[cur];
    T qv, cv = T(0);
    comp_t<T> cs = comp_t<T>(0);
    if constexpr (KAHAN) {
      const T yk = inc - comp_val<T>(bc[cur], qo);
      qv = qo + yk;
      cv = (qv - qo) - yk;
      cs = comp_make<T>(cv, qv);
    } else {
      qv = qo + inc;
    }
    nfa = nf_acc(nfa, qv);
    if constexpr (EDGE) {
      if constexpr (KAHAN) st_at(comp + ro, xbc, cs);
      st_at(qn + ro, xb, qv);
    } else {
      if constexpr (KAHAN) buf_st(vC, xbc, prow(y) * rowbc, cs);
      buf_st(vN, xb, rb, qv);
    }
    if constexpr (EDGE) {
      // carry q's wall ring into the next buffer (corners stay zero)
      if (y == 1) qn[x] = jqc[j2];
      if (y == m) qn[(m + 1) * g.P + x] = jqc[j0];
      if (x == 1) qn[y * g.P] = jql[j1];
      if (x == n) qn[y * g.P + n + 1] = jqr[j1];
    }
  };
#define QGW_P4S_PRE(k)                          \
  if constexpr ((k) < D) {                      \
    if (rbeg + (k) < rend) fetch(IC<(k)>{}, rbeg + (k)); \
  }
  QGW_P4S_PRE(0)
  QGW_P4S_PRE(1)
  QGW_P4S_PRE(2)
  QGW_P4S_PRE(3)
  QGW_P4S_PRE(4)
#undef QGW_P4S_PRE
  for (int r = rbeg; r < rend; r += W) {
#define QGW_P4S_PHASE(k)                   \
  if constexpr ((k) < W) {                 \
    if (r + (k) >= rend) break;            \
    body(IC<(k)>{}, r + (k));              \
  }
    QGW_P4S_PHASE(0)
    QGW_P4S_PHASE(1)
    QGW_P4S_PHASE(2)
    QGW_P4S_PHASE(3)
    QGW_P4S_PHASE(4)
    QGW_P4S_PHASE(5)
    QGW_P4S_PHASE(6)
    QGW_P4S_PHASE(7)
    QGW_P4S_PHASE(8)
    QGW_P4S_PHASE(9)
    QGW_P4S_PHASE(10)
    QGW_P4S_PHASE(11)
#undef QGW_P4S_PHASE
  }
  nf_flag(nfa, cf.nonfinite);
}

template <typename T, int NLAP, int KAHAN>
__global__ void __launch_bounds__(256)
k_p4s(Geo g, const T *__restrict__ psi, const T *__restrict__ q, T *__restrict__ qn, T *__restrict__ fnew,
      const T *__restrict__ f1, const T *__restrict__ f2, comp_t<T> *__restrict__ comp, const T *__restrict__ topo,
      const T *__restrict__ wind, StepCoef cf, int nstrip, int chunk, int nchunk, int rev, int lfirst) {
  constexpr int H = (NLAP > 1) ? NLAP : 1;  // halo
  constexpr int OW = 64 - 2 * H;            // output columns per strip
  const int n = g.n, m = g.m;
  const int lane = threadIdx.x & 63;
  // Blocks are dealt to the 8 XCDs round-robin; the remap gives each XCD one
  // contiguous run of walks, strips fastest, so the columns shared by
  // neighbouring strips and the warm-up rows shared by consecutive chunks are
  // fetched into one L2.  lfirst: the layers of a chunk before the next chunk
  // (the walks in flight cover the same rows of every layer; 4% faster at
  // 2048^2 x 10), else the chunks of a layer before the next layer (3-4%
  // faster at 4096^2 x 32).
  const int wave = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
  if (wave >= nstrip * nchunk * g.nl) return;  // whole wavefronts only; no barrier follows
  const int sidx = wave % nstrip;
  const int k = lfirst ? (wave / nstrip) % g.nl : wave / (nstrip * nchunk);  // layer (all members)
  const int cidx = lfirst ? wave / (nstrip * g.nl) : (wave / nstrip) % nchunk;
  const int mb = k / g.nlm, km = k - mb * g.nlm;
  topo += mb * g.LS;
  wind += (long)mb * (m + 2);
  const int xlo = 1 + sidx * OW - H;  // storage column of lane 0
  const int x = xlo + lane;
  const int y0 = 1 + cidx * chunk;
  const int y1 = (y0 + chunk < m + 1) ? y0 + chunk : m + 1;
  const long lo = k * g.LS;
  const T *P = psi + lo, *Qf = q + lo;
  f1 += lo;
  f2 += lo;
  qn += lo;
  fnew += lo;
  if (KAHAN) comp += lo;
  const bool ringstrip = (xlo <= 0 || xlo + 63 >= n + 1);
  // interior: no wall column or row in reach of the stencils, so every row the
  // walk loads lies inside the layer
  const bool interior = xlo >= 1 && xlo + 63 <= n && y0 - H - NLAP >= 1 && y1 + H - 1 <= m;
  // rev: the interior walks of odd chunks run up their chunk, so that the rows
  // two neighbouring chunks both load (the psi and q rows around their common
  // boundary) are loaded by both at about the same time, at the start or the
  // end of both walks, and the second load meets L2
  if (interior && rev && (cidx & 1))
    p4s_walk<T, NLAP, KAHAN, false, false, true>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0,
                                                 y1, false);
  else if (interior)
    p4s_walk<T, NLAP, KAHAN, false>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0, y1, false);
  else
    p4s_walk<T, NLAP, KAHAN, true>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0, y1,
                                   ringstrip);
}

// Streaming form over aligned column blocks: a block of 4 wavefronts owns
// BW = 32 floor(4 OW / 32) output columns X0 .. X0 + BW - 1 (224 at n_hyper
// 4; X0 = 1 + BW bx) of one row chunk and layer, wavefront w the BW / 4 of
// them from X0 + (BW / 4) w, walking the strip of 64 columns from there minus
// the halo H as in k_p4s (OW = 64 - 2H outputs possible) and storing only its
// own.  Block boundaries then fall on 128-byte line boundaries (32 columns
// of FP32), so every line of the outputs is written by the wavefronts of one
// block, about together, on one CU; in k_p4s (blocks of 4 OW columns) the
// lines across a block boundary are completed by a wavefront of another block
// at another time.  Measured on the access pattern alone (five loaded and
// three stored streams per row, 109 VGPRs, FP32, 10 layers of 2050 rows): 235
// against 244 us at 2048 columns, 440 against 484 us at 4096.  In the stencil
// kernel, which computes 4% more strip columns here: faster at 4096^2 x 32
// (FP32 3983 against 4095 us, FP64 5766 against 6170 us), slower at 2048^2 x
// 10 (311 against 302 us); the autotuner picks.
template <int NLAP> constexpr int p4a_cols() {
  constexpr int H = (NLAP > 1) ? NLAP : 1, OW = 64 - 2 * H;
  return 32 * ((4 * OW) / 32);
}
template <typename T, int NLAP, int KAHAN>
__global__ void __launch_bounds__(256)
k_p4a(Geo g, const T *__restrict__ psi, const T *__restrict__ q, T *__restrict__ qn, T *__restrict__ fnew,
      const T *__restrict__ f1, const T *__restrict__ f2, comp_t<T> *__restrict__ comp, const T *__restrict__ topo,
      const T *__restrict__ wind, StepCoef cf, int nbx, int chunk, int nchunk, int rev, int lfirst) {
  constexpr int H = (NLAP > 1) ? NLAP : 1;  // halo
  constexpr int BW = p4a_cols<NLAP>();      // output columns per block
  constexpr int OWU = BW / 4;               // output columns per wavefront
  const int n = g.n, m = g.m;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // consecutive column blocks of a chunk and layer on one XCD; chunks and
  // layers in the order lfirst sets, as in k_p4s
  const int bi = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = bi % nbx;
  const int k = lfirst ? (bi / nbx) % g.nl : bi / (nbx * nchunk);  // layer (all members)
  const int cidx = lfirst ? bi / (nbx * g.nl) : (bi / nbx) % nchunk;
  const int X0 = 1 + BW * bx + OWU * w;  // the wavefront's first output column
  if (X0 > n) return;  // a strip past the grid (whole wavefronts; no barrier follows)
  const int mb = k / g.nlm, km = k - mb * g.nlm;
  topo += mb * g.LS;
  wind += (long)mb * (m + 2);
  const int xlo = X0 - H;  // storage column of lane 0
  const int x = xlo + lane;
  const int y0 = 1 + cidx * chunk;
  const int y1 = (y0 + chunk < m + 1) ? y0 + chunk : m + 1;
  const long lo = k * g.LS;
  const T *P = psi + lo, *Qf = q + lo;
  f1 += lo;
  f2 += lo;
  qn += lo;
  fnew += lo;
  if (KAHAN) comp += lo;
  const bool ringstrip = (xlo <= 0 || xlo + 63 >= n + 1);
  const bool interior = xlo >= 1 && xlo + 63 <= n && y0 - H - NLAP >= 1 && y1 + H - 1 <= m;
  if (interior && rev && (cidx & 1))  // as in k_p4s
    p4s_walk<T, NLAP, KAHAN, false, true, true>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0, y1,
                                                false, X0, OWU);
  else if (interior)
    p4s_walk<T, NLAP, KAHAN, false, true>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0, y1,
                                          false, X0, OWU);
  else
    p4s_walk<T, NLAP, KAHAN, true, true>(g, P, Qf, qn, fnew, f1, f2, comp, topo, wind, cf, x, lane, km, y0, y1,
                                         ringstrip, X0, OWU);
}

// ===========================================================================
// CFL reduction (adaptive dt): max over interior points and layers of
// |psi(i+1,j) - psi(i,j)| and |psi(i,j+1) - psi(i,j)| (src/timestep.h:74-87).
// The maximum of a non-negative double is reduced through its bit pattern.
// ===========================================================================
// A maximum that keeps NaN (fmax drops it): a NaN anywhere in psi reaches
// the reduced value, whose bits (above those of +inf as unsigned integers)
// then win the atomicMax, so the host sees a non-finite maximum.
__device__ inline double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }

template <typename T>
__global__ void __launch_bounds__(256) k_cfl(Geo g, const T *__restrict__ psi, unsigned long long *out) {
  const long I = (long)g.n * g.m * g.nl;
  double mx = 0.0;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < I; e += (long)gridDim.x * 256L) {
    const int k = (int)(e / ((long)g.n * g.m));
    const long r = e % ((long)g.n * g.m);
    const int j = (int)(r / g.n) + 1, i = (int)(r % g.n) + 1;
    const T *p = psi + k * g.LS + (long)j * g.P + i;
    const double c = (double)p[0];
    const double du = fabs((double)p[1] - c);
    const double dv = fabs((double)p[g.P] - c);
    mx = nan_max(mx, nan_max(du, dv));
  }
  __shared__ double red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = nan_max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

// ===========================================================================
// Diagnostics: per layer sum of psi*q and q^2 over all stored points, FP64,
// in a fixed order (per-block partials, then one block per layer).
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(256)
k_diag_partial(Geo g, const T *__restrict__ psi, const T *__restrict__ q, const comp_t<T> *__restrict__ comp,
               double *part, int nblk) {
  const int k = blockIdx.y;
  const long S = (long)(g.m + 2) * (g.n + 2);
  double e = 0.0, z = 0.0;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < S; idx += (long)nblk * 256L) {
    const int j = (int)(idx / (g.n + 2)), i = (int)(idx % (g.n + 2));
    const long o = k * g.LS + (long)j * g.P + i;
    double qv = (double)q[o];
    if (comp) qv -= (double)comp_val<T>(comp[o], q[o]);
    const double pv = (double)psi[o];
    e += pv * qv;
    z += qv * qv;
  }
  __shared__ double se[256], sz[256];
  se[threadIdx.x] = e;
  sz[threadIdx.x] = z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sz[threadIdx.x] += sz[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ((long)k * nblk + blockIdx.x)] = se[0];
    part[2 * ((long)k * nblk + blockIdx.x) + 1] = sz[0];
  }
}

template <int NB>
__global__ void __launch_bounds__(NB) k_diag_final(const double *part, int nblk, double *out) {
  const int k = blockIdx.x;
  double e = 0.0, z = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    e += part[2 * ((long)k * nblk + b)];
    z += part[2 * ((long)k * nblk + b) + 1];
  }
  __shared__ double se[256], sz[256];
  se[threadIdx.x] = e;
  sz[threadIdx.x] = z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sz[threadIdx.x] += sz[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * k] = se[0];
    out[2 * k + 1] = sz[0];
  }
}

// ===========================================================================
// Stochastic ring forcing field (calc_forc, src/forcing.h:168-208): the
// unnormalised c2r of a sparse N x (N/2+1) half spectrum, synthesised
// directly.  The nc forced columns have their points
// mode_j/mode_k[col_start[c] .. col_start[c+1]) (row j, index into coef).
//   G[j'][c] = sum_p coef[k_p] e^{+2 pi i j_p j'/N}                (k_stoch_cols)
//   S[j'][i'] = sum_c Re(G[j'][c] E[c][i'])                        (k_stoch_rows)
// and k_stoch_rows adds S sqrt(dt) to q at once (src/timestep.h:143-150).
// with E[c][i'] = w_c e^{+2 pi i col_i[c] i'/N}, w = 1 for the zero and
// Nyquist columns (only their real parts enter a real inverse), 2 otherwise:
// a real product of [Re G, -Im G] (m x 2nc) and [Re E; Im E] (2nc x n).
// tw[t] = e^{+2 pi i t/N}.  S point (i', j') is storage point (i'+1, j'+1)
// of layer 0 of q; rows j' < m and columns i' < n.
// ===========================================================================
template <typename T>  // the field type of the launch unit (the sums are FP64 either way)
__global__ void __launch_bounds__(256)
k_stoch_cols(int N, int nc, const int *__restrict__ col_start, const int *__restrict__ mode_j,
             const int *__restrict__ mode_k, const double2 *__restrict__ coef, const double2 *__restrict__ tw,
             double2 *__restrict__ G) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e >= (long)N * nc) return;
  const int jp = (int)(e / nc), c = (int)(e % nc);
  double re = 0, im = 0;
  for (int p = col_start[c]; p < col_start[c + 1]; p++) {
    const double2 a = coef[mode_k[p]];
    const double2 w = tw[(int)(((long)mode_j[p] * jp) % N)];
    re += a.x * w.x - a.y * w.y;
    im += a.x * w.y + a.y * w.x;
  }
  G[e] = make_double2(re, im);
}

// One block: 256 columns i' x SR rows j'.  The block's rows of G pass through
// LDS 64 columns at a time; E streams from L2, one coalesced row per c, with
// UE rows in flight before their products; q (and the compensation) of the
// thread's points are fetched before the sums.
// q += S s_fac on layer 0, Kahan-compensated when KAHAN.
template <typename T, int SR, bool KAHAN>
__global__ void __launch_bounds__(256)
k_stoch_rows(Geo g, int nc, const double2 *__restrict__ G, const double2 *__restrict__ E, double s_fac,
             T *__restrict__ q, comp_t<T> *__restrict__ comp) {
  constexpr int CC = 64, UE = 8;
  __shared__ double2 sg[SR][CC + UE];  // zero past the chunk: the unrolled steps need no branch
  const int j0 = blockIdx.x * SR;
  const int ip = blockIdx.y * 256 + threadIdx.x;
  const bool act = ip < g.n;
  const int ipc = act ? ip : g.n - 1;  // clamped column: every lane loads, uniform control flow
  const int rows = min(SR, g.m - j0);
  T qv[SR], cv[SR];
#pragma unroll
  for (int r = 0; r < SR; r++) {
    const long o = (long)(j0 + (r < rows ? r : 0) + 1) * g.P + ipc + 1;
    qv[r] = q[o];
    cv[r] = KAHAN ? comp_val<T>(comp[o], qv[r]) : T(0);
  }
  double acc[SR];
#pragma unroll
  for (int r = 0; r < SR; r++) acc[r] = 0;
  for (int c0 = 0; c0 < nc; c0 += CC) {
    const int cn = min(CC, nc - c0);
    __syncthreads();
    for (int e = threadIdx.x; e < SR * (CC + UE); e += 256) {
      const int r = e / (CC + UE), cc = e % (CC + UE);
      sg[r][cc] = (cc < cn && r < rows) ? G[(long)(j0 + r) * nc + c0 + cc] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    for (int cc = 0; cc < cn; cc += UE) {
      double2 w[UE];
#pragma unroll
      for (int u = 0; u < UE; u++) w[u] = E[(long)(c0 + min(cc + u, cn - 1)) * g.n + ipc];
      __builtin_amdgcn_sched_barrier(0);  // the UE loads stay ahead of their products
#pragma unroll
      for (int u = 0; u < UE; u++) {
#pragma unroll
        for (int r = 0; r < SR; r++) {
          const double2 gv = sg[r][cc + u];
          acc[r] = fma(gv.x, w[u].x, acc[r]);
          acc[r] = fma(-gv.y, w[u].y, acc[r]);
        }
      }
    }
  }
  if (!act) return;
#pragma unroll
  for (int r = 0; r < SR; r++) {
    if (r >= rows) break;
    const long o = (long)(j0 + r + 1) * g.P + ip + 1;
    const T inc = (T)(acc[r] * s_fac);
    if constexpr (KAHAN) {
      const T y = inc - cv[r], t = qv[r] + y;
      comp[o] = comp_make<T>((t - qv[r]) - y, t);
      q[o] = t;
    } else {
      q[o] = qv[r] + inc;
    }
  }
}

// ===========================================================================
// 4-D forcing added after the AB3 update (src/timestep.h:152-163):
// (w1 F1 + w2 F2) dt, records around the time before the step, on the
// interior of every layer.  One block row per (layer, grid row); 256 columns
// per block.
// ===========================================================================
template <typename T, bool KAHAN>
__global__ void __launch_bounds__(256)
k_forcing(Geo g, T *__restrict__ q, comp_t<T> *__restrict__ comp, const T *__restrict__ F1, const T *__restrict__ F2,
          double w1, double w2, double dt) {
  const int k = blockIdx.x / g.m, j = blockIdx.x % g.m + 1;
  const int i = blockIdx.y * 256 + threadIdx.x + 1;
  if (i > g.n) return;
  const long o = (long)k * g.LS + (long)j * g.P + i;
  T v = q[o];
  T c = KAHAN ? comp_val<T>(comp[o], v) : T(0);
  auto add = [&](T inc) {
    if constexpr (KAHAN) {
      const T y = inc - c;
      const T t = v + y;
      c = (t - v) - y;
      v = t;
    } else {
      v = v + inc;
    }
  };
  const double qf = w1 * (double)F1[o] + w2 * (double)F2[o];
  add((T)(qf * dt));
  q[o] = v;
  if constexpr (KAHAN) comp[o] = comp_make<T>(c, v);
}

// ===========================================================================
// Dense sine transforms (grids whose transform lengths the FFT kernels do not
// cover): the row and column DST-I are matrix products with S_ij =
// 2 sin(pi i j / (n+1)) on the matrix cores (library GEMM, see qgw_ctx.hip);
// these two kernels are the parts between the products.
//
// Helmholtz division on a whatT-shaped array (mode k, row a = x frequency
// 1..n, column b = y frequency 1..m): w /= -(L_a^2 + K_b^2 + 1/Rd_k^2) * norm
// (src/elliptic.h:120-138), the factor formed in FP64.
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(256)
k_helm_div(Geo g, T *__restrict__ w, const double *__restrict__ La2, const double *__restrict__ Kb2,
           const double *__restrict__ iRd2, double norm) {
  const int b = blockIdx.x * 256 + threadIdx.x;  // column (y frequency b + 1)
  const int a = blockIdx.y;                      // row (x frequency a + 1)
  const int k = blockIdx.z;                      // mode (all members)
  if (b >= g.m) return;
  const long o = k * g.WTS + (long)a * g.Pm + b;
  const double f = 1.0 / (-(La2[a + 1] + Kb2[b + 1] + iRd2[k]) * norm);
  w[o] = (T)((double)w[o] * f);
}

// The wall rings after the dense inverse row transform has written the psi
// interior: psi's ring is zero (adjust_bc(psi, 0)), q's ring (corners left as
// they are) is 2 bc_fac/Delta^2 times the adjacent interior psi
// (adjust_bc(q, 1, psi), src/domain.h:308-348).  One thread per ring point.
template <typename T>
__global__ void __launch_bounds__(256) k_rings(Geo g, T *__restrict__ psi, T *__restrict__ q, T ring_c) {
  const int n = g.n, m = g.m;
  const int per = 2 * (n + 2) + 2 * m;  // rows 0 and m+1 in full, then columns 0 and n+1 of rows 1..m
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (e >= per) return;
  T *P = psi + k * g.LS, *Q = q + k * g.LS;
  int j, i;
  if (e < 2 * (n + 2)) {
    j = (e < n + 2) ? 0 : m + 1;
    i = (e < n + 2) ? e : e - (n + 2);
  } else {
    const int r = e - 2 * (n + 2);
    j = 1 + (r >> 1);
    i = (r & 1) ? n + 1 : 0;
  }
  P[(long)j * g.P + i] = T(0);
  if (j == 0 && i >= 1 && i <= n) Q[i] = ring_c * P[(long)g.P + i];
  if (j == m + 1 && i >= 1 && i <= n) Q[(long)(m + 1) * g.P + i] = ring_c * P[(long)m * g.P + i];
  if (i == 0 && j >= 1 && j <= m) Q[(long)j * g.P] = ring_c * P[(long)j * g.P + 1];
  if (i == n + 1 && j >= 1 && j <= m) Q[(long)j * g.P + n + 1] = ring_c * P[(long)j * g.P + n];
}

// ===========================================================================
// Conversions between the host FP64 layout and the device arrays.
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(256)
k_from_f64(Geo g, int layers, const double *__restrict__ src, T *__restrict__ dst, comp_t<T> *__restrict__ resid) {
  const long W = g.n + 2;
  const long S = W * (g.m + 2);
  const long tot = S * layers;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += (long)gridDim.x * 256L) {
    const long k = e / S, r = e % S;
    const long j = r / W, i = r % W;
    const double v = src[e];
    const T tv = (T)v;
    const long o = k * g.LS + j * g.P + i;
    dst[o] = tv;
    if (resid) resid[o] = comp_make<T>((T)((double)tv - v), tv);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_to_f64(Geo g, int layers, const T *__restrict__ src, const comp_t<T> *__restrict__ resid, double *__restrict__ dst) {
  const long W = g.n + 2;
  const long S = W * (g.m + 2);
  const long tot = S * layers;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += (long)gridDim.x * 256L) {
    const long k = e / S, r = e % S;
    const long j = r / W, i = r % W;
    const long o = k * g.LS + j * g.P + i;
    double v = (double)src[o];
    if (resid) v -= (double)comp_val<T>(resid[o], src[o]);
    dst[e] = v;
  }
}

}  // namespace qgw
