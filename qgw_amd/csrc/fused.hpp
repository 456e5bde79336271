// fused.hpp -- the inversion in four passes over HBM (bounded branch of
// invert_pv, src/elliptic.h:76-155), for runs of few layers per member:
//
//   P1m  per (grid row j, member, mode pair): the layer -> mode projection
//        w_k = sum_l cl2m[k,l] q_l (src/elliptic.h:90-98) of the row's layers,
//        formed at the thread's positions as the rows stream in, then the
//        forward sine transform along x; writes what[k][j][a] (mode-major,
//        x frequency fastest).
//   P2m  per (mode, tile of 2G consecutive x frequencies): the tile's columns
//        (64-byte row segments) are staged in LDS, G column pairs are
//        transformed side by side along y, divided by
//        -(L_a^2 + K_b^2 + 1/Rd_k^2) * 4(n+1)(m+1) and transformed back, and
//        the tile is stored in place.
//   P3m  per (grid row j, member, layer pair): the mode -> layer projection
//        psi_l = sum_k cm2l[l,k] w_k (src/elliptic.h:140-147) of the row's
//        spectral rows, then the inverse sine transform along x; psi row and
//        the wall rings of psi and q.
//
// Both projections commute with the sine transforms, so projecting before the
// row transform (P1m) and after the inverse row transform's input is formed
// (P3m) gives the reference's result; no transposed copy of the spectral
// array is made.  Each row of the layers (modes) is read by every mode (layer)
// pair of its member: the persistent blocks take a row's items together and
// the block ids are remapped so that one XCD holds them, so the repeated reads
// of a row meet its L2.
#pragma once
#include "kernels.hpp"

namespace qgw {

// Rows of nl layers (l = 0..nl-1, stride S elements), one element each at the
// thread's positions p = t + m NT (element p - 1 of a row holds z_p; the
// buffer range check gives z_0 = 0), contracted with two rows of a matrix:
//   za[m] = sum_l ca[l] row_l[p - 1],   zb[m] = sum_l cb[l] row_l[p - 1].
// Two layers' loads are issued before their products.
template <typename T, int M, int EP>
__device__ __forceinline__ void proj_fetch(T *za, T *zb, const T *__restrict__ row0, long S, int nl,
                                           const T *__restrict__ ca, const T *__restrict__ cb, int t) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT, SZ = (int)sizeof(T);
#pragma unroll
  for (int m = 0; m < E; m++) za[m] = zb[m] = T(0);
  const int o = (t - 1) * SZ;
  int l = 0;
  for (; l + 1 < nl; l += 2) {
    const __amdgpu_buffer_rsrc_t r0 = row_rsrc(row0 + (long)l * S, (M - 1) * SZ);
    const __amdgpu_buffer_rsrc_t r1 = row_rsrc(row0 + (long)(l + 1) * S, (M - 1) * SZ);
    T v0[E], v1[E];
#pragma unroll
    for (int m = 0; m < E; m++) {
      v0[m] = buf_ld<T>(r0, o + m * NT * SZ);
      v1[m] = buf_ld<T>(r1, o + m * NT * SZ);
    }
    const T a0 = ca[l], a1 = ca[l + 1], b0 = cb[l], b1 = cb[l + 1];
#pragma unroll
    for (int m = 0; m < E; m++) {
      za[m] += a0 * v0[m] + a1 * v1[m];
      zb[m] += b0 * v0[m] + b1 * v1[m];
    }
  }
  if (l < nl) {
    const __amdgpu_buffer_rsrc_t r0 = row_rsrc(row0 + (long)l * S, (M - 1) * SZ);
    const T a0 = ca[l], b0 = cb[l];
#pragma unroll
    for (int m = 0; m < E; m++) {
      const T v = buf_ld<T>(r0, o + m * NT * SZ);
      za[m] += a0 * v;
      zb[m] += b0 * v;
    }
  }
}

// Publish the projected pair in buf (entry p holds z_p, entry 0 the zero z_0),
// as PipeRows::stage does; the caller's barrier before it retires buf's
// previous readers.
template <typename T, int M, int EP>
__device__ __forceinline__ void stage_pair(cx<T> *buf, const T *za, const T *zb, int t) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  const int rp = t + (t >> 4);
#pragma unroll
  for (int m = 0; m < E; m++) {
    if constexpr (NT % 16 == 0)
      buf[rp + m * NT + ((m * NT) >> 4)] = mkc<T>(za[m], zb[m]);
    else
      buf[rslot(t + m * NT)] = mkc<T>(za[m], zb[m]);
  }
}

// Item -> (row j, member mb, pair pr): the pairs of one row and member are
// consecutive items.
struct RowItem {
  int j, mb, pr;
};
__device__ __forceinline__ RowItem row_item(int it, int npair, int nmem) {
  RowItem r;
  r.pr = it % npair;
  const int rest = it / npair;
  r.mb = rest % nmem;
  r.j = rest / nmem + 1;
  return r;
}

template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p1m(Geo g, const T *__restrict__ q, T *__restrict__ what, const T *__restrict__ cl2m,
      const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  __shared__ TwTab<T, M> tb;
  __shared__ cx<T> twf[twf_len<T, M>()];
  __shared__ TwTab<T, 2 * M> tb2;
  const int t = threadIdx.x;
  const int nlm = g.nlm, nmem = g.nl / nlm, npair = (nlm + 1) / 2;
  const int nitems = g.m * nmem * npair;
  twtab_fill(tb, twh, t, NT);
  twf_fill<T, M>(twf, twh, t, NT);
  twtab_fill(tb2, tw2, t, NT);
  const TwLds<T, M> twa{&tb};
  // consecutive slots on one XCD: the items of a row run there together
  const int slot = xcd_remap(blockIdx.x, gridDim.x);
  for (int it = slot; it < nitems; it += gridDim.x) {
    const int tl = loop_tid<NT>(t);
    const RowItem ri = row_item(it, npair, nmem);
    const int k0 = 2 * ri.pr, k1 = k0 + 1;
    const bool has1 = k1 < nlm;
    const T *Mm = cl2m + (long)ri.mb * nlm * nlm;
    // rows k0 and k1 of the member's matrix (k0's again for a missing k1, whose result is not stored)
    const T *ca = Mm + (long)k0 * nlm, *cb = has1 ? Mm + (long)k1 * nlm : ca;
    T za[E], zb[E];
    // storage column x = 1..n of row j holds x_1..x_N
    proj_fetch<T, M, EP>(za, zb, q + (long)ri.mb * nlm * g.LS + (long)ri.j * g.P + 1, g.LS, nlm, ca, cb, tl);
    __syncthreads();  // the previous item's readers of buf are done
    stage_pair<T, M, EP>(buf, za, zb, tl);
    __syncthreads();
    cx<T> h[E];
    dst_pre_lds<T, M, EP>(h, tl, buf, tb2);
    fft_pipe<T, M, EP>(h, buf, twa, twf, tl);
    dst_post<T, M, EP>(h, buf, tl);
    const int kb = ri.mb * nlm;
    T *o0 = what + (long)(kb + k0) * g.WS + (long)(ri.j - 1) * g.Pn;
    T *o1 = what + (long)(kb + k1) * g.WS + (long)(ri.j - 1) * g.Pn;
#pragma unroll 2
    for (int i = tl; i < M - 1; i += NT) {
      const cx<T> v = buf[rslot(i)];
      o0[i] = v.x;
      if (has1) o1[i] = v.y;
    }
  }
}

// Store of a (psi row j, layer pair) from the natural-order buffer of the
// inverse row transform (entry i holds x = i + 1): the psi rows with their
// zero wall columns, the wall rings of psi (zero) and of q (ring_c psi of the
// adjacent interior point, src/domain.h:308-348).
template <typename T, int M, int EP>
__device__ __forceinline__ void p3_store_pair(const Geo &g, const cx<T> *buf, T *__restrict__ psi, T *__restrict__ q,
                                              T ring_c, int j, long lb, int l0, bool has1, int tl, int psi_nt) {
  constexpr int NT = DstHalf<M, EP>::NT;
  const int n = g.n, m_ = g.m, l1 = l0 + 1;
  T *p0 = psi + (lb + l0) * g.LS + (long)j * g.P;
  T *p1 = psi + (lb + l1) * g.LS + (long)j * g.P;
  T *q0r = q + (lb + l0) * g.LS;
  T *q1r = q + (lb + l1) * g.LS;
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
      p0[rr - (long)j * g.P + x] = T(0);
      if (has1) p1[rr - (long)j * g.P + x] = T(0);
    }
  }
}

template <typename T, int M, int EP>
__global__ void __launch_bounds__((DstHalf<M, EP>::NT))
k_p3m(Geo g, const T *__restrict__ what, T *__restrict__ psi, T *__restrict__ q, const T *__restrict__ cm2l,
      T ring_c, const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2, int psi_nt) {
  constexpr int E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  __shared__ cx<T> buf[DstHalf<M, EP>::slots];
  __shared__ TwTab<T, M> tb;
  __shared__ cx<T> twf[twf_len<T, M>()];
  __shared__ TwTab<T, 2 * M> tb2;
  const int t = threadIdx.x;
  const int nlm = g.nlm, nmem = g.nl / nlm, npair = (nlm + 1) / 2;
  const int nitems = g.m * nmem * npair;
  twtab_fill(tb, twh, t, NT);
  twf_fill<T, M>(twf, twh, t, NT);
  twtab_fill(tb2, tw2, t, NT);
  const TwLds<T, M> twa{&tb};
  const int slot = xcd_remap(blockIdx.x, gridDim.x);
  for (int it = slot; it < nitems; it += gridDim.x) {
    const int tl = loop_tid<NT>(t);
    const RowItem ri = row_item(it, npair, nmem);
    const int l0 = 2 * ri.pr, l1 = l0 + 1;
    const bool has1 = l1 < nlm;
    const T *Mm = cm2l + (long)ri.mb * nlm * nlm;
    const T *ca = Mm + (long)l0 * nlm, *cb = has1 ? Mm + (long)l1 * nlm : ca;
    T za[E], zb[E];
    // spectral row j - 1 of the member's modes: element x - 1 holds z_x
    proj_fetch<T, M, EP>(za, zb, what + (long)ri.mb * nlm * g.WS + (long)(ri.j - 1) * g.Pn, g.WS, nlm, ca, cb,
                         tl);
    __syncthreads();
    stage_pair<T, M, EP>(buf, za, zb, tl);
    __syncthreads();
    cx<T> h[E];
    dst_pre_lds<T, M, EP>(h, tl, buf, tb2);
    fft_pipe<T, M, EP>(h, buf, twa, twf, tl);
    dst_post<T, M, EP>(h, buf, tl);
    p3_store_pair<T, M, EP>(g, buf, psi, q, ring_c, ri.j, (long)ri.mb * nlm, l0, has1, tl, psi_nt);
  }
}

// ---- P2m: column tiles ------------------------------------------------------
// Entries per pair buffer: at least S, and 4 mod 32 so that the G buffers
// start 8 (FP32) or 16 (FP64) banks apart: the staging writes of one row's
// 2G columns then fall on distinct banks.
constexpr int p2m_slots(int S) { return ((S - 4 + 31) / 32) * 32 + 4; }
// G column pairs per block: 2G columns of T = 64 bytes per row where the
// block (G NT threads, at most 1024) and its pair buffers (at most 144 KiB of
// LDS) allow, else the largest power of two that fits.  Each pair's transform
// must span whole wavefronts (the post pass's prefix sum moves values across
// the lanes of a wavefront), so with NT not a multiple of 64 G is 1.
template <typename T, int M, int EP> constexpr int p2m_groups() {
  const int NT = DstHalf<M, EP>::NT;
  if (NT % 64 != 0) return 1;
  int G = 64 / (2 * (int)sizeof(T));
  const int bytes = p2m_slots(DstHalf<M, EP>::slots) * 2 * (int)sizeof(T);
  while (G > 1 && (G * NT > 1024 || G * bytes > 144 * 1024)) G /= 2;
  return G;
}

// GG: column pairs per block (at most p2m_groups; fewer pairs leave room for
// more resident blocks per CU, whose load and transform phases overlap).
template <typename T, int M, int EP, int GG>
__global__ void __launch_bounds__((GG * DstHalf<M, EP>::NT))
k_p2m(Geo g, T *__restrict__ w, const double *__restrict__ La2, const double *__restrict__ iRd2, double dk,
      double norm, const cx<T> *__restrict__ twh, const cx<T> *__restrict__ tw2, int ncg) {
  constexpr int G = GG, E = DstHalf<M, EP>::E, NT = DstHalf<M, EP>::NT;
  static_assert(GG >= 1 && GG <= p2m_groups<T, M, EP>(), "pairs per block");
  constexpr int BT = G * NT, W = 2 * G, SL = p2m_slots(DstHalf<M, EP>::slots);
  constexpr int NWV = (NT + 63) / 64;
  static_assert(BT % G == 0, "tile rows per pass");
  __shared__ cx<T> bufs[G][SL];
  __shared__ TwTab<T, 2 * M> tb2;  // sines of the pre passes
  __shared__ cx<double> wtot[G][NWV];
  const int tid = threadIdx.x;
  const int n = g.n, m = g.m;  // m = M - 1 rows
  // tiles of one mode in raster order per XCD: neighbouring tiles share 128-byte lines in one L2
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int k = tile / ncg, a0 = (tile - k * ncg) * W;  // mode, first column (0-based x frequency)
  T *base = w + (long)k * g.WS;
  TwTabFill<T, 2 * M, BT> tfill;
  tfill.load(tw2, tid);
  // stage the tile: lanes along the G column pairs of a row, one pair (both
  // columns, 2 sizeof(T) bytes) per lane, BT / G rows per pass; a pair's two
  // columns land in one complex entry of its buffer
  const int cp = tid % G, r0 = tid / G;
  constexpr int RS = BT / G;
  const bool cok = a0 + 2 * cp < n, cok1 = a0 + 2 * cp + 1 < n;
  cx<T> *bpair = bufs[cp];
  {
    const T *src = base + a0 + 2 * cp;
#pragma unroll 8
    for (int j = r0; j < m; j += RS) {
      const T *e = src + (long)j * g.Pn;
      cx<T> v = mkc<T>(T(0), T(0));
      if (cok1)
        v = *reinterpret_cast<const cx<T> *>(e);
      else if (cok)
        v.x = e[0];
      bpair[rslot(j + 1)] = v;  // entry y = j + 1 holds z_y
    }
  }
  tfill.store(tb2, tid);
  __syncthreads();
  const int gi = tid / NT, t = tid - gi * NT;
  cx<T> *buf = bufs[gi];
  cx<T> h[E];
  // forward transform along y: entry b holds S_b afterwards
  dst_pre_lds<T, M, EP>(h, t, buf, tb2);
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP, 1>(h, buf, t, wtot[gi]);
  // Helmholtz division, f = 1 / (-(L_a^2 + K_b^2 + 1/Rd_k^2) * 4(n+1)(m+1))
  // (src/elliptic.h:120-138); columns a and a + 1 of mode k, K_b = b dk
  const int a = a0 + 2 * gi + 1;               // 1-based x frequency of .x
  const int a1 = (a + 1 <= n) ? a + 1 : a;     // of .y (a past-the-end column is discarded)
  const double d0 = La2[a], d1 = La2[a1];
  const double ir = iRd2[k];
  const bool lo0 = (k % g.nlm == 0 && a <= 16), lo1 = (k % g.nlm == 0 && a1 <= 16);
  const T d0n = (T)((d0 + ir) * norm), d1n = (T)((d1 + ir) * norm), c2 = (T)(dk * dk * norm);
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
      f0 = (T)(1.0 / (-(d0 + kb + ir) * norm));
      f1 = (T)(1.0 / (-(d1 + kb + ir) * norm));
    } else {
      // FP32: FP64 only for the lowest barotropic wavenumbers (mode 0 of a
      // member, a, b <= 16), elsewhere FP32 and the hardware reciprocal
      const T bf = (T)b, bb = bf * bf;
      f0 = -__builtin_amdgcn_rcpf(__builtin_fmaf(bb, c2, d0n));
      f1 = -__builtin_amdgcn_rcpf(__builtin_fmaf(bb, c2, d1n));
      if (b <= 16) {
        const double K = dk * (double)b, kb = K * K;
        if (lo0) f0 = (T)(1.0 / (-(d0 + kb + ir) * norm));
        if (lo1) f1 = (T)(1.0 / (-(d1 + kb + ir) * norm));
      }
    }
    if (b >= 1 && b < M) buf[rslot(b)] = mkc<T>(hv[s].x * f0, hv[s].y * f1);
  }
  __syncthreads();
  // inverse transform (the same DST-I): entry i holds row i afterwards
  dst_pre_lds<T, M, EP>(h, t, buf, tb2);
  fft_home<T, M, EP>(h, buf, twh, t);
  dst_post<T, M, EP>(h, buf, t, wtot[gi]);
  {
    T *dst = base + a0 + 2 * cp;
#pragma unroll 8
    for (int j = r0; j < m; j += RS) {
      const cx<T> v = bpair[rslot(j)];
      T *e = dst + (long)j * g.Pn;
      if (cok1)
        *reinterpret_cast<cx<T> *>(e) = v;
      else if (cok)
        e[0] = v.x;
    }
  }
}

}  // namespace qgw
