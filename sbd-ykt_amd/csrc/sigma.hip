// sigma = H c on the alpha x beta product basis (replaces mult, tpb/mult.h:376-632).
//
// In the separated phase order H splits into
//   sigma[ia,ib] = hii*c                                                   (diagonal)
//                + sum_{ja in S(ia)+D(ia)} va(ia,ja) c[ja,ib]             (alpha product)
//                + sum_{jb in S(ib)+D(ib)} vb(ib,jb) c[ia,jb]             (beta product)
//                + sum_{ja in S(ia)} sa [ sum_{k in occ(ib)} (ai|kk) c[ja,ib]
//                                        + sum_{jb in S(ib)} sb (ai|bj) c[ja,jb] ]
//                + sum_{jb in S(ib)} sb [ sum_{k in occ(ia)} (kk|bj) ] c[ia,jb]   (alpha-beta)
// where va, vb are the string-pair values of the excitation tables.  The alpha product
// streams whole rows c[ja,:] (coalesced, one workgroup per row tile, no atomics); the
// beta product is the same kernel on the transposed vector; the alpha-beta kernel owns
// one bra row per workgroup, stages the ket row c[ja,:] in LDS chunk by chunk together
// with the (ai|..) slice, and gathers from LDS.  Every output element has one owner.
#include <algorithm>

#include "sbdg_internal.h"

namespace sbdg {

namespace {

template <class T>
struct V16;
template <>
struct V16<double> {
  using t = double2;
  static constexpr int n = 2;
};
template <>
struct V16<float> {
  using t = float4;
  static constexpr int n = 4;
};

constexpr int TS = 256;  // threads of the row-streaming product
constexpr int SPMM_U = 4;  // table entries gathered per step of the row product

// rows per workgroup of the row product: FP64 one row of 512 columns, FP32 four rows (one
// wave each) of 256 columns
template <class T>
constexpr int spmm_rows() {
  return sizeof(T) == 8 ? 1 : 4;
}

// Column range [cb, ce) of one row product:
//   out[r, c] = (DIAG ? hii*X : 0)[r, c] + sum_q (val[q] + sgn_q J[p_q][c]) X[col[q], c]
//             + (ZG ? sum_{s in slots(r)} Z[slot[s], c - cb] : 0)
// jai[q] = +-(p_q) selects the row of the column strings' occupation table J that carries
// the other-spin occupation part of a single (row 0: zero, for doubles).  Z holds the
// core-grouped doubles of the same column range (k_core_prod).  A work item is one row of
// one column tile; tiles are outer, rows inner.  Workgroups take spmm_rows consecutive
// items, are dealt to the 8 XCDs in turn, and workgroup L takes items from
// (L % 8) * per + L / 8 (per = ceil(workgroups / 8)): each XCD walks a contiguous run of
// rows, whose strings share many excitation partners, so the rows of X they read meet in
// that XCD's L2.  Measured on one GPU against one 1,024-column row per FP32 workgroup, four
// rows of 256 columns take 3-4 % less time per pass; for FP64 four rows of 128 columns take
// 4-11 % more than one row of 512, and 2 or 8 rows, or 8 entries per step, were slower.
template <class T, bool DIAG, bool VECTOR, bool ZG>
__global__ void __launch_bounds__(TS) k_spmm(int64_t nrows, int64_t ncols, int64_t cb, int64_t ce,
                                             const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                             const T* __restrict__ val, const int32_t* __restrict__ jai,
                                             const T* __restrict__ J, const T* __restrict__ X,
                                             const T* __restrict__ hii, T* __restrict__ out,
                                             const int32_t* __restrict__ zptr, const int32_t* __restrict__ zslot,
                                             const T* __restrict__ Z, int64_t zcols) {
  constexpr int VN = V16<T>::n;
  using VT = typename V16<T>::t;
  constexpr int RW = spmm_rows<T>();
  constexpr int RT = TS / RW;  // threads per row
  const int64_t ntl = (ce - cb + RT * VN - 1) / (RT * VN);
  const int64_t items = ntl * nrows;
  const int64_t per = ((items + RW - 1) / RW + 7) / 8;
  const int64_t bid = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;  // each XCD: a contiguous run of rows
  const int64_t item = bid * RW + (RW == 1 ? 0 : threadIdx.x / RT);
  if (item >= items) return;  // no barrier in this kernel: a row group may leave alone
  const int64_t tile = item / nrows;
  const int64_t r = item - tile * nrows;
  const int64_t c0 = cb + tile * (int64_t)(RT * VN) + (int64_t)(RW == 1 ? threadIdx.x : threadIdx.x % RT) * VN;
  if (c0 >= ce) return;
  T acc[VN];
  const bool full = VECTOR && (c0 + VN <= ce);
  if (DIAG) {
    const T* h = hii + r * ncols + c0;
    const T* x = X + r * ncols + c0;
    if (full) {
      T hv[VN], xv[VN];
      *reinterpret_cast<VT*>(hv) = *reinterpret_cast<const VT*>(h);
      *reinterpret_cast<VT*>(xv) = *reinterpret_cast<const VT*>(x);
#pragma unroll
      for (int k = 0; k < VN; ++k) acc[k] = hv[k] * xv[k];
    } else {
#pragma unroll
      for (int k = 0; k < VN; ++k) acc[k] = (c0 + k < ce) ? h[k] * x[k] : T(0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < VN; ++k) acc[k] = T(0);
  }
  int64_t q = rowptr[r];
  const int64_t q1 = rowptr[r + 1];
  if (full) {
    for (; q + SPMM_U <= q1; q += SPMM_U) {
      int32_t j[SPMM_U];
      T v[SPMM_U], sg[SPMM_U], x[SPMM_U][VN], jv[SPMM_U][VN];
#pragma unroll
      for (int u = 0; u < SPMM_U; ++u) {
        j[u] = col[q + u];
        v[u] = val[q + u];
        const int32_t p = jai[q + u];
        sg[u] = p < 0 ? T(-1) : T(1);
        *reinterpret_cast<VT*>(x[u]) = *reinterpret_cast<const VT*>(X + (int64_t)j[u] * ncols + c0);
        *reinterpret_cast<VT*>(jv[u]) = *reinterpret_cast<const VT*>(J + (int64_t)(p < 0 ? -p : p) * ncols + c0);
      }
#pragma unroll
      for (int u = 0; u < SPMM_U; ++u)
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[k] += (v[u] + sg[u] * jv[u][k]) * x[u][k];
    }
    for (; q < q1; ++q) {
      const T v = val[q];
      const int32_t p = jai[q];
      const T sg = p < 0 ? T(-1) : T(1);
      T x[VN], jv[VN];
      *reinterpret_cast<VT*>(x) = *reinterpret_cast<const VT*>(X + (int64_t)col[q] * ncols + c0);
      *reinterpret_cast<VT*>(jv) = *reinterpret_cast<const VT*>(J + (int64_t)(p < 0 ? -p : p) * ncols + c0);
#pragma unroll
      for (int k = 0; k < VN; ++k) acc[k] += (v + sg * jv[k]) * x[k];
    }
    if (ZG) {
      const int64_t zc = c0 - cb;
      int32_t s = zptr[r];
      const int32_t s1 = zptr[r + 1];
      for (; s + 2 <= s1; s += 2) {
        T z0[VN], z1[VN];
        *reinterpret_cast<VT*>(z0) = *reinterpret_cast<const VT*>(Z + (int64_t)zslot[s] * zcols + zc);
        *reinterpret_cast<VT*>(z1) = *reinterpret_cast<const VT*>(Z + (int64_t)zslot[s + 1] * zcols + zc);
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[k] += z0[k] + z1[k];
      }
      if (s < s1) {
        T z0[VN];
        *reinterpret_cast<VT*>(z0) = *reinterpret_cast<const VT*>(Z + (int64_t)zslot[s] * zcols + zc);
#pragma unroll
        for (int k = 0; k < VN; ++k) acc[k] += z0[k];
      }
    }
    *reinterpret_cast<VT*>(out + r * ncols + c0) = *reinterpret_cast<VT*>(acc);
  } else {
    for (; q < q1; ++q) {
      const T v = val[q];
      const int32_t p = jai[q];
      const T sg = p < 0 ? T(-1) : T(1);
      const T* x = X + (int64_t)col[q] * ncols + c0;
      const T* jr = J + (int64_t)(p < 0 ? -p : p) * ncols + c0;
#pragma unroll
      for (int k = 0; k < VN; ++k)
        if (c0 + k < ce) acc[k] += (v + sg * jr[k]) * x[k];
    }
    if (ZG) {
      for (int32_t s = zptr[r]; s < zptr[r + 1]; ++s) {
        const T* z = Z + (int64_t)zslot[s] * zcols + (c0 - cb);
#pragma unroll
        for (int k = 0; k < VN; ++k)
          if (c0 + k < ce) acc[k] += z[k];
      }
    }
#pragma unroll
    for (int k = 0; k < VN; ++k)
      if (c0 + k < ce) out[r * ncols + c0 + k] = acc[k];
  }
}

// Same-spin doubles grouped by core, column range [cb, ce):
//   Z[x, c - cb] = sum_{y in core(x)} eps_x eps_y W[P_x][P_y] X[row_y, c]
// for the members x of one product block (up to CORE_XB members of one core).  Members
// whose pairs overlap have W = 0 (they differ by a single excitation, or are equal), so
// every double pair of the string set is counted exactly once, in the core it shares.
// The block's signed coefficients and member rows are staged in LDS, CORE_YC members at a
// time (every lane reads the same entry: broadcast), and four member rows are loaded
// before they are used.  Each lane owns VN columns.
// Workgroups are dealt to the 8 XCDs in turn; workgroup L takes work item
// w = (L % 8) * per + L / 8 (per = ceil(items / 8)), so each XCD walks one contiguous range of
// items, and item w is column tile w / ngroups, block group w % ngroups: the blocks of one
// core are neighbours in that range and find the core's member rows in their XCD's L2.
constexpr int CORE_YC = 128;  // members staged per step

template <class T, bool VECTOR>
__global__ void __launch_bounds__(TS) k_core_prod(int64_t ncols, int64_t cb, int64_t ce,
                                                  const int32_t* __restrict__ start,
                                                  const int32_t* __restrict__ row,
                                                  const int32_t* __restrict__ pair,
                                                  const int32_t* __restrict__ blk,
                                                  const int32_t* __restrict__ blk_x0,
                                                  const int32_t* __restrict__ grp, int64_t ngroups,
                                                  const T* __restrict__ W, int64_t npair, const T* __restrict__ X,
                                                  T* __restrict__ Z, int64_t zcols) {
  constexpr int VN = V16<T>::n;
  constexpr int XB = CORE_XB;
  using VT = typename V16<T>::t;
  __shared__ __align__(16) T cf[CORE_YC][XB];  // signed coefficients of the staged members
  __shared__ int32_t rw[CORE_YC];              // their rows
  const int64_t ntiles = (ce - cb + TS * VN - 1) / (TS * VN);
  const int64_t ngrp = ngroups;
  const int64_t per = (ntiles * ngrp + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= ntiles * ngrp) return;  // whole workgroup: no barrier is skipped by part of it
  const int64_t tile = w / ngrp;
  const int64_t g = w - tile * ngrp;
  const int64_t c0 = cb + tile * (TS * VN) + (int64_t)threadIdx.x * VN;
  const bool active = c0 < ce;  // inactive lanes still take part in the barriers
  const bool full = VECTOR && (c0 + VN <= ce);
  const int64_t zc = c0 - cb;
  for (int b = grp[g]; b < grp[g + 1]; ++b) {
    const int core = blk[b];
    const int s0 = start[core];
    const int m = start[core + 1] - s0;
    const int x0 = blk_x0[b];
    const int nx = min(XB, m - x0);
    T acc[XB][VN];
#pragma unroll
    for (int xi = 0; xi < XB; ++xi)
#pragma unroll
      for (int k = 0; k < VN; ++k) acc[xi][k] = T(0);
    for (int y0 = 0; y0 < m; y0 += CORE_YC) {
      const int ny = min(CORE_YC, m - y0);
      __syncthreads();  // previous step done with the staged entries
      for (int e = threadIdx.x; e < ny * XB; e += TS) {
        const int yy = e / XB, xi = e - yy * XB;
        T v = T(0);
        if (xi < nx) {
          const int px = pair[s0 + x0 + xi], py = pair[s0 + y0 + yy];
          v = W[(int64_t)(px & 0x7FFFFFFF) * npair + (py & 0x7FFFFFFF)];
          if ((px < 0) != (py < 0)) v = -v;
        }
        cf[yy][xi] = v;
        if (xi == 0) rw[yy] = row[s0 + y0 + yy];
      }
      __syncthreads();
      if (active) {
        int yy = 0;
        for (; yy + 4 <= ny; yy += 4) {
          T xv[4][VN];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const T* x = X + (int64_t)rw[yy + u] * ncols + c0;
            if (full) {
              *reinterpret_cast<VT*>(xv[u]) = *reinterpret_cast<const VT*>(x);
            } else {
#pragma unroll
              for (int k = 0; k < VN; ++k) xv[u][k] = (c0 + k < ce) ? x[k] : T(0);
            }
          }
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int xi = 0; xi < XB; ++xi) {
              const T w = cf[yy + u][xi];
#pragma unroll
              for (int k = 0; k < VN; ++k) acc[xi][k] += w * xv[u][k];
            }
        }
        for (; yy < ny; ++yy) {
          T xv[VN];
          const T* x = X + (int64_t)rw[yy] * ncols + c0;
          if (full) {
            *reinterpret_cast<VT*>(xv) = *reinterpret_cast<const VT*>(x);
          } else {
#pragma unroll
            for (int k = 0; k < VN; ++k) xv[k] = (c0 + k < ce) ? x[k] : T(0);
          }
#pragma unroll
          for (int xi = 0; xi < XB; ++xi) {
            const T w = cf[yy][xi];
#pragma unroll
            for (int k = 0; k < VN; ++k) acc[xi][k] += w * xv[k];
          }
        }
      }
    }
    if (active) {
#pragma unroll
      for (int xi = 0; xi < XB; ++xi) {
        if (xi < nx) {
          T* z = Z + (int64_t)(s0 + x0 + xi) * zcols + zc;
          if (full) {
            *reinterpret_cast<VT*>(z) = *reinterpret_cast<VT*>(acc[xi]);
          } else {
#pragma unroll
            for (int k = 0; k < VN; ++k)
              if (c0 + k < ce) z[k] = acc[xi][k];
          }
        }
      }
    }
  }
}

constexpr int TT = 64;  // transpose tile

// out[c][r] = in[r][c]   (ADD: out[c][r] += in[r][c])
template <class T, bool ADD>
__global__ void __launch_bounds__(256) k_transpose(int64_t R, int64_t C, const T* __restrict__ in,
                                                   T* __restrict__ out, const T* __restrict__ add) {
  __shared__ T tile[TT][TT + 1];
  const int64_t r0 = (int64_t)blockIdx.y * TT, c0 = (int64_t)blockIdx.x * TT;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < TT; k += 4) {
    const int64_t r = r0 + k, c = c0 + tx;
    if (r < R && c < C) tile[k][tx] = in[r * C + c];
  }
  __syncthreads();
  for (int k = ty; k < TT; k += 4) {
    const int64_t c = c0 + k, r = r0 + tx;
    if (r < R && c < C) {
      if (ADD) out[c * R + r] += tile[tx][k] + (add ? add[c * R + r] : T(0));
      else out[c * R + r] = tile[tx][k];
    }
  }
}

// alpha-beta and cross-spin single terms.  One workgroup per (bra row ia, tile of lane
// positions); thread t owns positions col0 + t + m*blockDim.x, m < M (every wave owns
// whole 64-lane groups of the ELL table), i.e. columns colperm[position].  The ket rows
// ja in S(ia) are taken G at a time (G = 2 in FP64, 4 in FP32: 16 bytes either way):
// the G (ai|..) slices and the G rows' current column chunk sit in LDS interleaved, so one
// 16-byte LDS read serves all G rows, and every ELL entry read from the L2-resident table
// serves them too.  LDS: [slice G-vectors of the pairs b > j] [chunk + 1 ket G-vectors
// from kbase, the last one zero].
template <class T>
struct AbRows {
  static constexpr int G = 16 / (int)sizeof(T);
  struct __align__(16) V {
    T v[G];
  };
  // x with its sign flipped where bit 31 of sbit is set
  __device__ static inline T flip(T x, uint32_t sbit) {
    if constexpr (sizeof(T) == 8)
      return __longlong_as_double(__double_as_longlong(x) ^ ((unsigned long long)(sbit & 0x80000000u) << 32));
    else
      return __int_as_float(__float_as_int(x) ^ (int)(sbit & 0x80000000u));
  }
};

// The LDS vector at byte address a of the workgroup's LDS.  The alpha-beta kernel has no
// static LDS, so its dynamic LDS starts at address 0 (checked at launch) and entries carry
// absolute byte addresses.
typedef unsigned int lds_u32x4 __attribute__((ext_vector_type(4)));
template <class V>
__device__ inline V lds_at(uint32_t a) {
  static_assert(sizeof(V) == 16, "16-byte LDS vectors");
  const lds_u32x4 w = *reinterpret_cast<const __attribute__((address_space(3))) lds_u32x4*>(static_cast<size_t>(a));
  V r;
  __builtin_memcpy(&r, &w, 16);
  return r;
}

// One batch of four stream rows of one lane: s[g] += sign_e * V[pair_e][g] * cb[jb_e][g] for
// the G staged rows at once (all eight LDS reads issued before the products).
template <class T>
__device__ inline void ab_batch(const uint32_t (&p)[4], T (&s)[AbRows<T>::G]) {
  using R = AbRows<T>;
  using RV = typename R::V;
  constexpr int G = R::G;
  RV v[4], x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    v[u] = lds_at<RV>((p[u] >> 14) & 0xFFF0u);  // slice vector of pair (b, j)
    x[u] = lds_at<RV>(p[u] & 0x3FFFFu);         // ket vector
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
#pragma unroll
    for (int g = 0; g < G; ++g) s[g] += R::flip(v[u].v[g], p[u]) * x[u].v[g];
  }
}

// Alpha-beta terms of one bra row ia over one column tile:
//   sigma[ia, ib] += sum_{ja in S(ia)} sa * sum_{entries (jb, b<-j) of ib} sb (ai|bj) c[ja, jb]
// (the other-spin occupation parts of the singles are in the row products).  Chunk by chunk of
// ket columns: G ket rows of the chunk and their slices sit in LDS as interleaved
// G-vectors, each lane walks its ELL entries for that chunk (the lane -> column map is per
// chunk, so lanes of a group carry lists of similar length), and the chunk's sums are
// added into sigma before the next chunk (one owner per element: a column belongs to one
// tile; with one tile per row, through LDS in column order).  A wave's entries of a chunk are one stream (its groups one after another, four
// rows per batch); the next batch is loaded while the current one is used, across group
// boundaries, and after the last batch the stream's first batch is loaded again for the
// next ket rows.
// Workgroup w runs on XCD w % 8: consecutive workgroups of one XCD take the column tiles of
// one bra row (rows wider than one tile), so the tiles' staged ket rows and slices meet in
// that XCD's L2.  NTMAX bounds the workgroup (1024 threads; 768 for the widest FP64 rows,
// whose accumulators need more than 128 VGPRs).
template <class T, int M, int NTMAX>
__global__ void __launch_bounds__(NTMAX) k_ab(int64_t na, int64_t nb, int L, int64_t ntiles,
                                             const int64_t* __restrict__ rpSa,
                                             const int32_t* __restrict__ colSa,
                                             const uint32_t* __restrict__ excSa,
                                             const int64_t* __restrict__ soff, const int32_t* __restrict__ nq,
                                             const uint32_t* __restrict__ ent,
                                             const int32_t* __restrict__ colperm, int nch, int64_t chunk,
                                             int64_t kbase,
                                             int64_t groups, const T* __restrict__ eri, const T* __restrict__ c,
                                             T* __restrict__ sigma, int assign) {
  using R = AbRows<T>;
  using RV = typename R::V;
  constexpr int G = R::G;
  // 16-byte vectors per thread per staging round: SR ket vectors (4*SR VGPRs beside the
  // accumulators; eight only where they fit without spills) and SV slice vectors (at most
  // 2,016 pairs: one round at 1,024 threads)
  constexpr int SR = (sizeof(T) == 8 && M <= 20) ? 8 : 4;
  constexpr int SV = 2;
  extern __shared__ __align__(16) unsigned char smem[];
  RV* Vv = reinterpret_cast<RV*>(smem);  // slice G-vectors of the pairs b > j
  RV* cbv = Vv + kbase;                   // chunk + 1 G-vectors (the last one stays zero: padding)
  const int64_t LL = (int64_t)L * L;
  const int NP = L * (L - 1) / 2, NPS = max(NP, 1);  // slice vectors: pairs b > j (at least one)
  const int NT = blockDim.x;
  const int64_t npos = groups * 64;
  const int64_t xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int64_t tile = slot % ntiles;
  const int64_t ia = (slot / ntiles) * 8 + xcd;
  if (ia >= na) return;  // the grid rounds the rows up to a multiple of 8 (whole workgroup)
  const int64_t col0 = tile * (int64_t)NT * M;
  const int lane = threadIdx.x & 63;
  const int64_t s0 = (tile * (NT >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * nch;  // wave-uniform
  if (threadIdx.x == 0)
    for (int g = 0; g < G; ++g) cbv[chunk].v[g] = T(0);
  const int64_t q0 = rpSa[ia], q1 = rpSa[ia + 1];
  for (int k = 0; k < nch; ++k) {
    const int64_t j0 = (int64_t)k * chunk;
    const int nj = (int)(min(nb, j0 + chunk) - j0);
    // this wave's entry stream of chunk k: `rows` rows (a multiple of four); pn holds the
    // next batch, pf the row after it (wrapping to the start)
    const int64_t so = soff[s0 + k];
    const int rows = (int)((soff[s0 + k + 1] - so) >> 6);
    const uint32_t* sp = ent + so;  // wave-uniform; lanes add their index
    uint32_t pn[4];
    int pf = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) pn[u] = (uint32_t)chunk;
    if (rows > 0) {
#pragma unroll
      for (int u = 0; u < 4; ++u) pn[u] = sp[u * 64 + lane];
      pf = rows > 4 ? 4 : 0;
    }
    T acc[M];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = T(0);
    for (int64_t t = q0; t < q1; t += G) {
      // rows t..t+G-1; rows at or past q1 are absent (zero)
      T coef[G];
      const T* crow[G];
      const T* srow[G];
      bool has[G];
      __syncthreads();  // previous rows done with LDS (and the previous chunk's sums are stored)
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int64_t q = t + g;
        has[g] = q < q1;
        coef[g] = T(0);
        crow[g] = c + ia * nb + j0;
        srow[g] = eri;
        if (has[g]) {
          const uint32_t e = excSa[q];
          coef[g] = ((e >> 24) & 1u) ? T(-1) : T(1);
          crow[g] = c + (int64_t)colSa[q] * nb + j0;
          srow[g] = eri + (int64_t)((int)((e >> 6) & 63u) * L + (int)(e & 63u)) * LL;  // (ai|..), a*L + i
        }
      }
      // the G slices and the G ket rows' chunk into LDS: in rounds of SR ket and SV slice
      // vectors per thread, every global load of a round issued before its LDS stores (the loads do not depend on
      // each other), so a round costs one memory latency.  The thread index is opaque here so
      // that the addresses are formed per step rather than held in registers through the
      // products.
      int xt = (int)threadIdx.x;
      asm volatile("" : "+v"(xt));
      for (int x0 = xt, y0 = xt; x0 < nj || y0 < NPS; x0 += SR * NT, y0 += SV * NT) {
        RV sv[SV], kv[SR];
#pragma unroll
        for (int u = 0; u < SV; ++u) {
          // slice vector y is orbital pair (b, j), b > j, y = b(b-1)/2 + j: element b*L + j
          // of the slice (the float root is exact to within one, corrected below)
          const int y = y0 + u * NT;
          const bool sl = y < NP;
          int src = 0;
          if (sl) {
            int b = (int)((1.0f + __builtin_sqrtf(8.0f * (float)y + 1.0f)) * 0.5f);
            if (b * (b - 1) / 2 > y) --b;
            else if (b * (b + 1) / 2 <= y) ++b;
            src = b * L + (y - b * (b - 1) / 2);
          }
#pragma unroll
          for (int g = 0; g < G; ++g) sv[u].v[g] = (has[g] && sl) ? srow[g][src] : T(0);
        }
#pragma unroll
        for (int u = 0; u < SR; ++u) {
          const int x = x0 + u * NT;
#pragma unroll
          for (int g = 0; g < G; ++g)
            kv[u].v[g] = (has[g] && x < nj) ? __builtin_nontemporal_load(crow[g] + x) : T(0);
        }
#pragma unroll
        for (int u = 0; u < SV; ++u) {
          const int y = y0 + u * NT;
          if (y < NPS) Vv[y] = sv[u];
        }
#pragma unroll
        for (int u = 0; u < SR; ++u) {
          const int x = x0 + u * NT;
          if (x < nj) cbv[x] = kv[u];
        }
      }
      __syncthreads();
      // the stream's batch counts, opaque to the compiler here: they are then reloaded for
      // every step instead of being held in M registers
      int qi = (int)((s0 + k) * M);
      asm volatile("" : "+s"(qi));
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int nb4 = nq[qi + m];  // batches of this wave's m-th group
        T sg[G];
#pragma unroll
        for (int r = 0; r < G; ++r) sg[r] = T(0);
        for (int b = 0; b < nb4; ++b) {
          uint32_t p[4];
          const uint32_t* rp = sp + pf * 64;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            p[u] = pn[u];
            pn[u] = rp[u * 64 + lane];
          }
          pf = pf + 4 == rows ? 0 : pf + 4;
          ab_batch<T>(p, sg);
        }
        T a = acc[m];
#pragma unroll
        for (int r = 0; r < G; ++r) a += coef[r] * sg[r];
        asm volatile("" : "+v"(a));  // formed here, so the group's partial sums die here
        acc[m] = a;
      }
    }
    // opaque again: the column map is read here, not before the products
    int ft = (int)threadIdx.x;
    asm volatile("" : "+v"(ft));
    const bool first = assign && k == 0;  // every column has a position in chunk 0
    if (ntiles == 1) {
      // One tile holds the whole row, each column at exactly one position of this chunk: the
      // sums go through LDS (the slices and ket vectors, idle now and staged again for the
      // next rows; the zero slot after them stays untouched) one block of columns at a time,
      // then into sigma with coalesced
      // read-modify-writes.  Scattered 8-byte ones fetch and write back whole cache lines
      // once the rows in flight outgrow the L2 (at 40,000 columns they moved 4.7x the
      // kernel's algorithmic bytes).  Each element is read and written by the same thread in
      // every chunk.
      T* rowbuf = reinterpret_cast<T*>(Vv);
      const int64_t cap = (kbase + chunk) * G;
      T* out = sigma + ia * nb;
      for (int64_t b0 = 0; b0 < nb; b0 += cap) {
        const int64_t b1 = min(nb, b0 + cap);
        __syncthreads();  // the products, or the previous block's readers, are done with LDS
        int fb = ft;  // opaque per block: the column map is read per block, not held in M registers
        asm volatile("" : "+v"(fb));
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int64_t pos = (int64_t)m * NT + fb;
          if ((pos >> 6) < groups) {
            const int32_t ib = colperm[(int64_t)k * npos + pos];
            if (ib >= b0 && ib < b1) rowbuf[ib - b0] = acc[m];
          }
        }
        __syncthreads();
        for (int64_t x = threadIdx.x; x < b1 - b0; x += NT) {
          if (first) out[b0 + x] = rowbuf[x];
          else out[b0 + x] += rowbuf[x];
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int64_t pos = col0 + (int64_t)m * NT + ft;
        if ((pos >> 6) < groups) {
          const int32_t ib = colperm[(int64_t)k * npos + pos];
          if (ib >= 0) {
            if (first) sigma[ia * nb + ib] = acc[m];
            else sigma[ia * nb + ib] += acc[m];
          }
        }
      }
    }
    // with several tiles a column's owner thread changes from chunk to chunk: the next
    // chunk's read-modify-write of an element must see this chunk's store.  The barriers
    // inside the row loop order them only when the bra row has alpha singles; this one also
    // covers rows without (and the LDS reads of the last flush block).
    if (nch > 1) __syncthreads();
  }
}

template <class T>
void launch_spmm(bool diag, int64_t nrows, int64_t ncols, const SpmmTable& tb, const T* J, const T* X,
                 const T* hii, T* out, hipStream_t st, int64_t cb = 0, int64_t ce = -1,
                 const CoreTable* ct = nullptr, const T* Z = nullptr, int64_t zcols = 0) {
  if (ce < 0) ce = ncols;
  if (nrows == 0 || ce <= cb) return;
  constexpr int VN = V16<T>::n;
  constexpr int RW = spmm_rows<T>();
  const int64_t ntiles = (ce - cb + TS / RW * VN - 1) / (TS / RW * VN);
  const int64_t nblk = ((ntiles * nrows + RW - 1) / RW + 7) / 8 * 8;
  if (nblk > (int64_t)INT32_MAX) throw Error(SBDG_ERR_LIMIT, "grid too large");
  const T* val = nullptr;
  if constexpr (sizeof(T) == 8) val = reinterpret_cast<const T*>(tb.val.p);
  else val = reinterpret_cast<const T*>(tb.val32.p);
  const bool vec = (ncols % VN) == 0 && (cb % VN) == 0;
  const int32_t* zp = ct ? ct->slot_ptr.p : nullptr;
  const int32_t* zs = ct ? ct->slot.p : nullptr;
  dim3 g((unsigned)nblk), b(TS);
#define SBDG_SPMM(D, V, ZG)                                                                                   \
  hipLaunchKernelGGL((k_spmm<T, D, V, ZG>), g, b, 0, st, nrows, ncols, cb, ce, tb.rowptr.p, tb.col.p, val, \
                     tb.jai.p, J, X, hii, out, zp, zs, Z, zcols)
  if (ct) {
    if (diag) {
      if (vec) SBDG_SPMM(true, true, true);
      else SBDG_SPMM(true, false, true);
    } else {
      if (vec) SBDG_SPMM(false, true, true);
      else SBDG_SPMM(false, false, true);
    }
  } else {
    if (diag) {
      if (vec) SBDG_SPMM(true, true, false);
      else SBDG_SPMM(true, false, false);
    } else {
      if (vec) SBDG_SPMM(false, true, false);
      else SBDG_SPMM(false, false, false);
    }
  }
#undef SBDG_SPMM
}

// Z columns per pass over the range of one spin: a multiple of CORE_COLS that fits the
// workspace, at most the whole row length rounded up
template <class T>
int64_t core_zcols(const SigmaPlan& P, const CoreTable& ct, int64_t ncols) {
  const int64_t whole = (ncols + CORE_COLS - 1) / CORE_COLS * CORE_COLS;
  if (ct.nmem == 0) return whole;
  const int64_t fit = (int64_t)(P.zbytes / (sizeof(T) * (size_t)ct.nmem)) / CORE_COLS * CORE_COLS;
  if (fit < CORE_COLS) throw Error(SBDG_ERR_LIMIT, "core workspace too small");
  return std::min(fit, whole);
}

// products of one spin (rows of X are that spin's strings): singles + diagonal by the
// excitation table, doubles by core, one pass per column range of Z
template <class T>
void launch_spin(bool diag, int64_t nrows, int64_t ncols, const SpmmTable& tb, const T* J, const CoreTable* ct,
                 const T* W, int64_t npair, const T* X, const T* hii, T* out, T* Z, const SigmaPlan& P,
                 hipStream_t st) {
  if (!ct) {
    launch_spmm<T>(diag, nrows, ncols, tb, J, X, hii, out, st);
    return;
  }
  if (nrows == 0 || ncols == 0) return;
  constexpr int VN = V16<T>::n;
  const int64_t zcols = core_zcols<T>(P, *ct, ncols);
  const bool vec = (ncols % VN) == 0;
  for (int64_t cb = 0; cb < ncols; cb += zcols) {
    const int64_t ce = std::min(ncols, cb + zcols);
    if (ct->ngrp > 0) {
      const int64_t ntiles = (ce - cb + TS * VN - 1) / (TS * VN);
      const int64_t items = ntiles * ct->ngrp;
      const int64_t nwg = (items + 7) / 8 * 8;
      if (nwg > (int64_t)INT32_MAX) throw Error(SBDG_ERR_LIMIT, "grid too large");
      dim3 g((unsigned)nwg), b(TS);
      if (vec)
        hipLaunchKernelGGL((k_core_prod<T, true>), g, b, 0, st, ncols, cb, ce, ct->start.p, ct->row.p, ct->pair.p,
                           ct->blk.p, ct->blk_x0.p, ct->grp.p, ct->ngrp, W, npair, X, Z, zcols);
      else
        hipLaunchKernelGGL((k_core_prod<T, false>), g, b, 0, st, ncols, cb, ce, ct->start.p, ct->row.p,
                           ct->pair.p, ct->blk.p, ct->blk_x0.p, ct->grp.p, ct->ngrp, W, npair, X, Z, zcols);
    }
    launch_spmm<T>(diag, nrows, ncols, tb, J, X, hii, out, st, cb, ce, ct, Z, zcols);
  }
}

template <class T, bool ADD>
void launch_transpose(int64_t R, int64_t C, const T* in, T* out, hipStream_t st, const T* add = nullptr) {
  if (R == 0 || C == 0) return;
  dim3 g((unsigned)((C + TT - 1) / TT), (unsigned)((R + TT - 1) / TT));
  hipLaunchKernelGGL((k_transpose<T, ADD>), g, dim3(256), 0, st, R, C, in, out, add);
}

template <class T, int M, int NTMAX = 1024>
void launch_ab_m(const SigmaPlan& P, size_t lds, const T* eri, const T* c, T* sigma, hipStream_t st, int assign) {
  const EllTable& E = *P.ell;
  if (E.nt > NTMAX || E.nt % 64) throw Error(SBDG_ERR_STATE, "alpha-beta workgroup size");
  const int64_t nblk = (P.na + 7) / 8 * 8 * E.ntiles;
  if (nblk > (int64_t)INT32_MAX) throw Error(SBDG_ERR_LIMIT, "grid too large");
  static const bool lds_at_zero = [] {  // entries hold absolute LDS byte addresses
    hipFuncAttributes fa;
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_ab<T, M, NTMAX>)) == hipSuccess &&
           fa.sharedSizeBytes == 0;
  }();
  if (!lds_at_zero) throw Error(SBDG_ERR_STATE, "alpha-beta kernel has static LDS: entry addresses invalid");
  hipLaunchKernelGGL((k_ab<T, M, NTMAX>), dim3((unsigned)nblk), dim3(E.nt), lds, st, P.na, P.nb, P.L, E.ntiles,
                     P.sa->rowptr.p, P.sa->col.p, P.sa->exc.p, E.soff.p, E.nq.p, E.ent.p, E.colperm.p, E.nch, E.chunk,
                     E.kbase, E.groups, eri, c, sigma, assign);
}

template <class T>
void launch_ab(const SigmaPlan& P, const T* eri, const T* c, T* sigma, hipStream_t st, int assign) {
  if (P.na == 0 || P.nb == 0) return;
  const EllTable& E = *P.ell;
  if ((int64_t)E.nt * E.m * E.ntiles != E.groups * 64) throw Error(SBDG_ERR_STATE, "alpha-beta table geometry");
  const size_t lds = ab_lds_bytes(E.kbase, E.chunk);
  switch (E.m) {  // columns per thread, fixed when the table was built
    case 1: launch_ab_m<T, 1>(P, lds, eri, c, sigma, st, assign); break;
    case 2: launch_ab_m<T, 2>(P, lds, eri, c, sigma, st, assign); break;
    case 4: launch_ab_m<T, 4>(P, lds, eri, c, sigma, st, assign); break;
    case 8: launch_ab_m<T, 8>(P, lds, eri, c, sigma, st, assign); break;
    case 12: launch_ab_m<T, 12>(P, lds, eri, c, sigma, st, assign); break;
    case 16: launch_ab_m<T, 16>(P, lds, eri, c, sigma, st, assign); break;
    case 20: launch_ab_m<T, 20>(P, lds, eri, c, sigma, st, assign); break;
    case 24: launch_ab_m<T, 24>(P, lds, eri, c, sigma, st, assign); break;
    default:
      if constexpr (sizeof(T) == 8) {  // 768 threads (ab_geometry)
        switch (E.m) {
          case 32: launch_ab_m<T, 32, 768>(P, lds, eri, c, sigma, st, assign); return;
          case 40: launch_ab_m<T, 40, 768>(P, lds, eri, c, sigma, st, assign); return;
          case 48: launch_ab_m<T, 48, 768>(P, lds, eri, c, sigma, st, assign); return;
          case 53: launch_ab_m<T, 53, 768>(P, lds, eri, c, sigma, st, assign); return;
        }
      } else {
        switch (E.m) {
          case 40: launch_ab_m<T, 40>(P, lds, eri, c, sigma, st, assign); return;
          case 48: launch_ab_m<T, 48>(P, lds, eri, c, sigma, st, assign); return;
        }
      }
      throw Error(SBDG_ERR_STATE, "alpha-beta columns per thread");
  }
}

}  // namespace

void transpose_d(int64_t R, int64_t C, const double* in, double* out, hipStream_t st) {
  launch_transpose<double, false>(R, C, in, out, st);
  SBDG_CHECK(hipGetLastError());
}

size_t ab_lds_bytes(int64_t kbase, int64_t chunk) {
  return (size_t)16 * ((size_t)chunk + 1 + (size_t)kbase);  // 16-byte row vectors, any precision
}

template <class T>
void sigma_apply(const SigmaPlan& P, const T* c, T* sigma, const SigmaScratch<T>& scr, hipStream_t st,
                 hipEvent_t* ev) {
  T* cT = scr.cT;
  T* sT = scr.sT;
  const T* hii;
  const T* eri;
  if constexpr (sizeof(T) == 8) {
    hii = reinterpret_cast<const T*>(P.hii);
    eri = reinterpret_cast<const T*>(P.eri);
  } else {
    hii = reinterpret_cast<const T*>(P.hii32);
    eri = reinterpret_cast<const T*>(P.eri32);
  }
  if (ev) SBDG_CHECK(hipEventRecord(ev[0], st));
  // the alpha-beta terms (LDS-bound) into their own buffer on the second stream, while the
  // spin phases (HBM-bound) run on this one; the transpose-add joins them
  T* sab = scr.sab;
  if (P.st2) {
    SBDG_CHECK(hipEventRecord(P.fork, st));
    SBDG_CHECK(hipStreamWaitEvent(P.st2, P.fork, 0));
    launch_ab<T>(P, eri, c, sab, P.st2, 1);
    SBDG_CHECK(hipEventRecord(P.join, P.st2));
  }
  launch_transpose<T, false>(P.na, P.nb, c, cT, st);
  if (ev) SBDG_CHECK(hipEventRecord(ev[1], st));
  const T* W;
  if constexpr (sizeof(T) == 8) W = reinterpret_cast<const T*>(P.W);
  else W = reinterpret_cast<const T*>(P.W32);
  T* Z = reinterpret_cast<T*>(P.Z);
  const T *Ja, *Jb;
  if constexpr (sizeof(T) == 8) {
    Ja = reinterpret_cast<const T*>(P.Ja);
    Jb = reinterpret_cast<const T*>(P.Jb);
  } else {
    Ja = reinterpret_cast<const T*>(P.Ja32);
    Jb = reinterpret_cast<const T*>(P.Jb32);
  }
  launch_spin<T>(false, P.nb, P.na, *P.mb, Ja, P.cb, W, P.npair, cT, nullptr, sT, Z, P, st);
  if (ev) SBDG_CHECK(hipEventRecord(ev[2], st));
  launch_spin<T>(true, P.na, P.nb, *P.ma, Jb, P.ca, W, P.npair, c, hii, sigma, Z, P, st);
  if (ev) SBDG_CHECK(hipEventRecord(ev[3], st));
  if (P.st2) {
    SBDG_CHECK(hipStreamWaitEvent(st, P.join, 0));
    if (ev) SBDG_CHECK(hipEventRecord(ev[4], st));
    launch_transpose<T, true>(P.nb, P.na, sT, sigma, st, sab);
    if (ev) SBDG_CHECK(hipEventRecord(ev[5], st));
  } else {
    launch_transpose<T, true>(P.nb, P.na, sT, sigma, st);
    if (ev) SBDG_CHECK(hipEventRecord(ev[4], st));
    launch_ab<T>(P, eri, c, sigma, st, 0);
    if (ev) SBDG_CHECK(hipEventRecord(ev[5], st));
  }
  SBDG_CHECK(hipGetLastError());
}

template void sigma_apply<double>(const SigmaPlan&, const double*, double*, const SigmaScratch<double>&,
                                  hipStream_t, hipEvent_t*);
template void sigma_apply<float>(const SigmaPlan&, const float*, float*, const SigmaScratch<float>&, hipStream_t,
                                 hipEvent_t*);

}  // namespace sbdg
