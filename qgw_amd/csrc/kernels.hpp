// kernels.hpp -- the per-step kernels of the QG timestep on gfx950.
//
// Device layout (per precision T):
//   q[2], psi, f[3], comp   nl layers of (m+2) rows, row pitch P >= n+2; the
//                           one-point wall ring is stored (reference idx
//                           layout, src/domain.h:62, with a padded pitch)
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

namespace qgw {

// Workgroup id -> work item such that the ids the hardware deals round-robin
// to one of the 8 XCDs cover a contiguous range of work items.
__device__ __forceinline__ int xcd_remap(int b, int total) {
  const int x = b & 7, i = b >> 3;
  const int q = total >> 3, r = total & 7;
  return x * q + (x < r ? x : r) + i;
}

// Sizes shared by every kernel.
struct Geo {
  int n, m, nl;          // interior points in x, y; layers
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
template <typename T, int TC, int TR, int NLC>
__global__ void __launch_bounds__(256)
k_tproj(const T *__restrict__ src, T *__restrict__ dst, const T *__restrict__ M, int nl, int R, int C, int Ps,
        long Ss, int Pd, long Sd) {
  static_assert(TC * 8 == 256 && TR % 8 == 0, "tile shape");
  constexpr int U = TR / 8;  // source rows per thread
  __shared__ T tile[TR][TC + 1];
  const int c0 = blockIdx.x * TC, r0 = blockIdx.y * TR;
  const int tid = threadIdx.x;
  // every layer of the thread's points, lanes along c (contiguous in src)
  const int ci = tid % TC, rg = tid / TC;
  T v[NLC][U];
#pragma unroll
  for (int l = 0; l < NLC; l++)
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int r = r0 + rg + 8 * u, c = c0 + ci;
      v[l][u] = (l < nl && r < R && c < C) ? src[l * Ss + (long)r * Ps + c] : T(0);
    }
  // store mapping: lanes along r (contiguous in dst)
  const int rl = tid % TR, cg = tid / TR;
  constexpr int CS = 256 / TR;  // columns per store pass
  for (int k = 0; k < nl; k++) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      T acc = T(0);
#pragma unroll
      for (int l = 0; l < NLC; l++)
        if (l < nl) acc += M[k * nl + l] * v[l][u];  // uniform address: scalar loads
      tile[rg + 8 * u][ci] = acc;
    }
    __syncthreads();
    const int r = r0 + rl;
#pragma unroll
    for (int cc = cg; cc < TC; cc += CS) {
      const int c = c0 + cc;
      if (r < R && c < C) dst[k * Sd + (long)c * Pd + r] = tile[rl][cc];
    }
    __syncthreads();
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
  const double d0 = La2[a0] + 0.0, d1 = La2[a1] + 0.0;
  const double ir0 = iRd2[k0], ir1 = iRd2[k1];
#pragma unroll
  for (int m = 0; m < E; m++) {
    const int p = t + m * NT;
    if (p == 0 || p == L / 2) {
      h[m] = mkc<T>(T(0), T(0));
    } else {
      const int b = (p < L / 2) ? p : L - p;
      const double kb = Kb2[b];
      const double den0 = -(d0 + kb + ir0) * norm;
      const double den1 = -(d1 + kb + ir1) * norm;
      T f0, f1;
      if constexpr (sizeof(T) == 8) {
        f0 = (T)(1.0 / den0);
        f1 = (T)(1.0 / den1);
      } else {
        f0 = T(1) / (T)den0;
        f1 = T(1) / (T)den1;
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
// P4: tendency and AB3 update on a TX x TY tile of one layer.
//   f = -J(psi,q) - beta psi_x + nu lap q + nu_hyper lap^NLAP psi
//   layer 0:    += wind(y_j)
//   layer nl-1: += -hEkb f0/(2 dh) lap psi - J(psi, topo) f0/dh
//   q_next = q + c0 f + c1 f1 + c2 f2 (Kahan-compensated when KAHAN)
// NLAP = n_hyper/2 + 1 nested 5-point Laplacians (calc_diff,
// src/dynamics.h:13-44), the intermediate ones carrying the wall value
// 2 bc_fac/Delta^2 * adjacent psi on the ring; NLAP = 0 skips the term
// (nu_hyper = 0).
// ===========================================================================
struct StepCoef {
  double c0, c1, c2;     // AB3 coefficients
  double beta, nu, nu_hyper, inv_d2, inv_12d2, inv_2d;
  double drag;           // -hEkb*f0/(2*dh_bottom)
  double topo_fac;       // f0/dh_bottom
  double ring_c;         // 2*bc_fac/Delta^2
  int update;            // 0: write the tendency only (rhs), 1: also the AB3 update
};

template <int TX, int TY, int H> struct P4Tile {
  static constexpr int PW = TX + 2 * H, PH = TY + 2 * H;
};

template <typename T, int NLAP, int KAHAN, int TX, int TY>
__global__ void __launch_bounds__(TX * 4)
k_p4(Geo g, const T *__restrict__ psi, const T *__restrict__ q, T *__restrict__ qn, T *__restrict__ fnew,
     const T *__restrict__ f1, const T *__restrict__ f2, T *__restrict__ comp, const T *__restrict__ topo,
     const T *__restrict__ wind, StepCoef cf) {
  constexpr int H = (NLAP > 1) ? NLAP : 1;  // psi halo
  constexpr int PW = TX + 2 * H, PH = TY + 2 * H;
  constexpr int NTH = TX * 4;
  __shared__ T sP[PH][PW];
  __shared__ T sA[(NLAP > 1) ? (PH - 2) : 1][(NLAP > 1) ? (PW - 2) : 1];
  __shared__ T sB[(NLAP > 2) ? (PH - 4) : 1][(NLAP > 2) ? (PW - 4) : 1];
  __shared__ T sQ[TY + 2][TX + 2];
  __shared__ T sT[TY + 2][TX + 2];

  const int n = g.n, m = g.m, nl = g.nl;
  // Tiles are taken in raster order per XCD: workgroup ids are dealt to the
  // 8 XCDs round-robin, so each XCD gets one contiguous run of tiles and the
  // halo rows shared by vertically adjacent tiles stay in its L2.
  const int ntx = gridDim.x, nty = gridDim.y;
  const int tile = xcd_remap(blockIdx.x + ntx * (blockIdx.y + nty * blockIdx.z), ntx * nty * gridDim.z);
  const int k = tile / (ntx * nty);
  const int tyi = (tile / ntx) % nty, txi = tile % ntx;
  const int x0 = 1 + txi * TX, y0 = 1 + tyi * TY;  // first interior point of the tile
  const int tid = threadIdx.x;
  // AB3 history and compensation of this thread's outputs, fetched before the tiles
  const int ptx = tid % TX, pty0 = tid / TX;
  T hf1[TY / 4], hf2[TY / 4], hcp[TY / 4];
#pragma unroll
  for (int r = 0; r < TY / 4; r++) {
    const int gx = x0 + ptx, gy = y0 + pty0 + 4 * r;
    const bool ok = cf.update && gx <= n && gy <= m;
    const long o = k * g.LS + (long)gy * g.P + gx;
    hf1[r] = ok ? f1[o] : T(0);
    hf2[r] = ok ? f2[o] : T(0);
    hcp[r] = (ok && KAHAN) ? comp[o] : T(0);
  }
  const T *P = psi + k * g.LS;
  const T *Q = q + k * g.LS;
  const bool bottom = (k == nl - 1);

  // ---- load psi (halo H) and q (halo 1), zero outside the storage grid
  for (int e = tid; e < PW * PH; e += NTH) {
    const int ly = e / PW, lx = e % PW;
    const int gx = x0 - H + lx, gy = y0 - H + ly;
    sP[ly][lx] = (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) ? P[(long)gy * g.P + gx] : T(0);
  }
  for (int e = tid; e < (TX + 2) * (TY + 2); e += NTH) {
    const int ly = e / (TX + 2), lx = e % (TX + 2);
    const int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
    const bool in = (gx <= n + 1 && gy <= m + 1);
    sQ[ly][lx] = in ? Q[(long)gy * g.P + gx] : T(0);
    if (bottom) sT[ly][lx] = in ? topo[(long)gy * g.P + gx] : T(0);
  }
  __syncthreads();

  const T inv_d2 = (T)cf.inv_d2;
  const T ring_c = (T)cf.ring_c;
  // value of a Laplacian level at storage point (gx,gy): interior -> computed,
  // ring -> 2 bc_fac/Delta^2 * adjacent psi (psi from sP), outside -> 0
  auto ring_val = [&](int gx, int gy, int lx, int ly) -> T {
    // (lx,ly) index sP at the same point
    if (gx == 0) return (gy == 0 || gy == m + 1) ? T(0) : ring_c * sP[ly][lx + 1];
    if (gx == n + 1) return (gy == 0 || gy == m + 1) ? T(0) : ring_c * sP[ly][lx - 1];
    if (gy == 0) return ring_c * sP[ly + 1][lx];
    return ring_c * sP[ly - 1][lx];  // gy == m+1
  };

  // ---- nested Laplacians of psi down to radius 1 (kept in sA / sB)
  if constexpr (NLAP > 1) {
    // level 1 on radius H-1 -> sA
    for (int e = tid; e < (PW - 2) * (PH - 2); e += NTH) {
      const int ly = e / (PW - 2), lx = e % (PW - 2);
      const int gx = x0 - H + 1 + lx, gy = y0 - H + 1 + ly;
      T v;
      if (gx >= 1 && gx <= n && gy >= 1 && gy <= m) {
        const int sy = ly + 1, sx = lx + 1;
        v = (sP[sy][sx + 1] + sP[sy][sx - 1] + sP[sy + 1][sx] + sP[sy - 1][sx] - T(4) * sP[sy][sx]) * inv_d2;
      } else if (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) {
        v = ring_val(gx, gy, lx + 1, ly + 1);
      } else {
        v = T(0);
      }
      sA[ly][lx] = v;
    }
    __syncthreads();
  }
  if constexpr (NLAP > 2) {
    // levels 2..NLAP-1 alternate between sA (radius H-1) and sB (radius H-2) in place
    // level 2 -> sB (radius H-2)
    for (int e = tid; e < (PW - 4) * (PH - 4); e += NTH) {
      const int ly = e / (PW - 4), lx = e % (PW - 4);
      const int gx = x0 - H + 2 + lx, gy = y0 - H + 2 + ly;
      T v;
      if (gx >= 1 && gx <= n && gy >= 1 && gy <= m) {
        const int sy = ly + 1, sx = lx + 1;
        v = (sA[sy][sx + 1] + sA[sy][sx - 1] + sA[sy + 1][sx] + sA[sy - 1][sx] - T(4) * sA[sy][sx]) * inv_d2;
      } else if (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) {
        v = ring_val(gx, gy, lx + 2, ly + 2);
      } else {
        v = T(0);
      }
      sB[ly][lx] = v;
    }
    __syncthreads();
    // further levels (NLAP >= 4): radius shrinks by one each time; reuse sA then sB
    int rad = H - 2;  // radius held in sB
    for (int lev = 3; lev < NLAP; lev++) {
      const int nr = rad - 1;
      const int wid = TX + 2 * nr, hei = TY + 2 * nr;
      // source is sB (radius rad, width TX+2rad) when lev is odd, sA otherwise
      const bool fromB = (lev % 2) == 1;
      for (int e = tid; e < wid * hei; e += NTH) {
        const int ly = e / wid, lx = e % wid;
        const int gx = x0 - nr + lx, gy = y0 - nr + ly;
        T v;
        if (gx >= 1 && gx <= n && gy >= 1 && gy <= m) {
          const int sy = ly + 1, sx = lx + 1;
          if (fromB)
            v = (sB[sy][sx + 1] + sB[sy][sx - 1] + sB[sy + 1][sx] + sB[sy - 1][sx] - T(4) * sB[sy][sx]) * inv_d2;
          else
            v = (sA[sy][sx + 1] + sA[sy][sx - 1] + sA[sy + 1][sx] + sA[sy - 1][sx] - T(4) * sA[sy][sx]) * inv_d2;
        } else if (gx >= 0 && gx <= n + 1 && gy >= 0 && gy <= m + 1) {
          v = ring_val(gx, gy, lx + H - nr, ly + H - nr);
        } else {
          v = T(0);
        }
        // source and destination differ, so no barrier is needed inside the loop
        if (fromB)
          sA[ly][lx] = v;
        else
          sB[ly][lx] = v;
      }
      __syncthreads();
      rad = nr;
    }
  }

  // ---- outputs: each thread owns column tx and rows ty, ty+4, ...
  const int tx = tid % TX;
  const int ty0 = tid / TX;
  const int gx = x0 + tx;
  const T beta = (T)cf.beta, nu = (T)cf.nu, nuh = (T)cf.nu_hyper;
  const T i12 = (T)cf.inv_12d2, i2d = (T)cf.inv_2d;
  const T c0 = (T)cf.c0, c1 = (T)cf.c1, c2 = (T)cf.c2;
#pragma unroll
  for (int r = 0; r < TY / 4; r++) {
    const int ty = ty0 + 4 * r;
    const int gy = y0 + ty;
    if (gx > n || gy > m) continue;
    const int px = tx + H, py = ty + H;  // sP index
    const int qx = tx + 1, qy = ty + 1;  // sQ index
#define PS(a, b) sP[py + (b)][px + (a)]
#define QS(a, b) sQ[qy + (b)][qx + (a)]
#define TS(a, b) sT[qy + (b)][qx + (a)]
    // Arakawa Jacobian (src/domain.h:17-27)
    const T jpq = ((PS(1, 0) - PS(-1, 0)) * (QS(0, 1) - QS(0, -1)) + (PS(0, -1) - PS(0, 1)) * (QS(1, 0) - QS(-1, 0)) +
                   PS(1, 0) * (QS(1, 1) - QS(1, -1)) - PS(-1, 0) * (QS(-1, 1) - QS(-1, -1)) -
                   PS(0, 1) * (QS(1, 1) - QS(-1, 1)) + PS(0, -1) * (QS(1, -1) - QS(-1, -1)) +
                   QS(0, 1) * (PS(1, 1) - PS(-1, 1)) - QS(0, -1) * (PS(1, -1) - PS(-1, -1)) -
                   QS(1, 0) * (PS(1, 1) - PS(1, -1)) + QS(-1, 0) * (PS(-1, 1) - PS(-1, -1))) *
                  i12;
    T f = -jpq - beta * (PS(1, 0) - PS(-1, 0)) * i2d;
    if (cf.nu != 0.0)
      f += nu * ((QS(1, 0) + QS(-1, 0) + QS(0, 1) + QS(0, -1) - T(4) * QS(0, 0)) * inv_d2);
    T lap1 = (PS(1, 0) + PS(-1, 0) + PS(0, 1) + PS(0, -1) - T(4) * PS(0, 0)) * inv_d2;
    if constexpr (NLAP == 1) {
      f += nuh * lap1;
    } else if constexpr (NLAP == 2) {
      const int ax = tx + H - 1, ay = ty + H - 1;
      f += nuh * ((sA[ay][ax + 1] + sA[ay][ax - 1] + sA[ay + 1][ax] + sA[ay - 1][ax] - T(4) * sA[ay][ax]) * inv_d2);
    } else if constexpr (NLAP > 2) {
      // the last stored level is in sB when NLAP is odd, in sA when even
      if constexpr ((NLAP % 2) == 1) {
        const int bx = tx + 1, by = ty + 1;
        f += nuh * ((sB[by][bx + 1] + sB[by][bx - 1] + sB[by + 1][bx] + sB[by - 1][bx] - T(4) * sB[by][bx]) * inv_d2);
      } else {
        const int ax = tx + 1, ay = ty + 1;
        f += nuh * ((sA[ay][ax + 1] + sA[ay][ax - 1] + sA[ay + 1][ax] + sA[ay - 1][ax] - T(4) * sA[ay][ax]) * inv_d2);
      }
    }
    if (k == 0) f += wind[gy];
    if (bottom) {
      const T jt = ((PS(1, 0) - PS(-1, 0)) * (TS(0, 1) - TS(0, -1)) + (PS(0, -1) - PS(0, 1)) * (TS(1, 0) - TS(-1, 0)) +
                    PS(1, 0) * (TS(1, 1) - TS(1, -1)) - PS(-1, 0) * (TS(-1, 1) - TS(-1, -1)) -
                    PS(0, 1) * (TS(1, 1) - TS(-1, 1)) + PS(0, -1) * (TS(1, -1) - TS(-1, -1)) +
                    TS(0, 1) * (PS(1, 1) - PS(-1, 1)) - TS(0, -1) * (PS(1, -1) - PS(-1, -1)) -
                    TS(1, 0) * (PS(1, 1) - PS(1, -1)) + TS(-1, 0) * (PS(-1, 1) - PS(-1, -1))) *
                   i12;
      f += (T)cf.drag * lap1 - jt * (T)cf.topo_fac;
    }
#undef PS
#undef QS
#undef TS
    const long o = k * g.LS + (long)gy * g.P + gx;
    fnew[o] = f;
    if (!cf.update) continue;
    const T qo = sQ[qy][qx];
    const T inc = c0 * f + c1 * hf1[r] + c2 * hf2[r];
    if constexpr (KAHAN) {
      const T cc = hcp[r];
      const T y = inc - cc;
      const T tq = qo + y;
      comp[o] = (tq - qo) - y;
      qn[o] = tq;
    } else {
      qn[o] = qo + inc;
    }
    // carry q's wall ring into the next buffer (corners stay zero)
    if (gy == 1) qn[k * g.LS + gx] = sQ[qy - 1][qx];
    if (gy == m) qn[k * g.LS + (long)(m + 1) * g.P + gx] = sQ[qy + 1][qx];
    if (gx == 1) qn[k * g.LS + (long)gy * g.P] = sQ[qy][qx - 1];
    if (gx == n) qn[k * g.LS + (long)gy * g.P + n + 1] = sQ[qy][qx + 1];
  }
}

// ===========================================================================
// CFL reduction (adaptive dt): max over interior points and layers of
// |psi(i+1,j) - psi(i,j)| and |psi(i,j+1) - psi(i,j)| (src/timestep.h:74-87).
// The maximum of a non-negative double is reduced through its bit pattern.
// ===========================================================================
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
    mx = fmax(mx, fmax(du, dv));
  }
  __shared__ double red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
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
k_diag_partial(Geo g, const T *__restrict__ psi, const T *__restrict__ q, const T *__restrict__ comp,
               double *part, int nblk) {
  const int k = blockIdx.y;
  const long S = (long)(g.m + 2) * (g.n + 2);
  double e = 0.0, z = 0.0;
  for (long idx = blockIdx.x * 256L + threadIdx.x; idx < S; idx += (long)nblk * 256L) {
    const int j = (int)(idx / (g.n + 2)), i = (int)(idx % (g.n + 2));
    const long o = k * g.LS + (long)j * g.P + i;
    double qv = (double)q[o];
    if (comp) qv -= (double)comp[o];
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
// Conversions between the host FP64 layout and the device arrays.
// ===========================================================================
template <typename T>
__global__ void __launch_bounds__(256)
k_from_f64(Geo g, int layers, const double *__restrict__ src, T *__restrict__ dst, T *__restrict__ resid) {
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
    if (resid) resid[o] = (T)((double)tv - v);
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_to_f64(Geo g, int layers, const T *__restrict__ src, const T *__restrict__ resid, double *__restrict__ dst) {
  const long W = g.n + 2;
  const long S = W * (g.m + 2);
  const long tot = S * layers;
  for (long e = blockIdx.x * 256L + threadIdx.x; e < tot; e += (long)gridDim.x * 256L) {
    const long k = e / S, r = e % S;
    const long j = r / W, i = r % W;
    const long o = k * g.LS + j * g.P + i;
    double v = (double)src[o];
    if (resid) v -= (double)resid[o];
    dst[e] = v;
  }
}

}  // namespace qgw
