// 1- and 2-particle RDMs of a real vector (replaces Correlation, tpb/correlation.h:194-444,
// and feeds the Hellmann-Feynman backward).
//
// Output layout is the reference's: one[s][p + L q] = <c+_{p s} c_{q s}>,
// two[s + 2t][i + L j + L^2 k + L^3 l] = <c+_{i s} c+_{j t} c_{l t} c_{k s}>.
// Mixed-spin blocks are the physical ones.
//
// Work is organised by which string pairs connect bra and ket (separated phase order,
// W real, tables symmetric so both orders of every pair are visited):
//  * alpha pairs (ia,ja), beta fixed: g = sum_ib W[ia,ib] W[ja,ib]; singles also need the
//    beta-occupation-resolved sums gb[k] = sum_ib W[ia,ib] W[ja,ib] n_k(ib).  Beta pairs
//    are the same on the transposed vector.  (row-pair Gram products, one workgroup per
//    bra row, LDS accumulators)
//  * alpha-beta pairs: T_(a,i)[b,j] = sum sa sb W[ia,ib] W[ja,jb] over alpha singles with
//    orbitals (a<-i) and beta singles (b<-j); one workgroup per group of alpha pairs with
//    one key (a,i), the ket row staged in LDS, accumulators in LDS.
//  * the diagonal: density correlators sum_D W_D^2 n_p n_q.
// Every RDM slot is written by exactly one owner (singles by their key (a,i), doubles by
// their orbital quadruple, alpha-beta by (a,i), the diagonal by (p,q)): no global atomics.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sbdg_internal.h"

namespace sbdg {

void transpose_d(int64_t R, int64_t C, const double* in, double* out, hipStream_t st);

namespace {

constexpr int RT = 256;

// Sums that many lanes of a workgroup add into one LDS slot are kept as exact 64-bit
// integer adds of each term's fixed-point image (term * 2^56, rounded to nearest), so their
// value does not depend on the order in which the lanes arrive.  W has unit norm, so every
// such sum is bounded by Cauchy-Schwarz: |sum over columns of W[r,c] W[q,c] n_k(c)| <= 1 for
// the row-pair sums, and <= PCH (32) for the alpha-beta partials of one workgroup, whose
// beta excitation (b <- j) maps each bra column to at most one ket column.  2^56 leaves a
// range of +-128 and resolves 1.4e-17 per term.
constexpr double FX_SCALE = 72057594037927936.0;  // 2^56
__device__ inline void fx_add(unsigned long long* a, double v) {
  atomicAdd(a, (unsigned long long)__double2ll_rn(v * FX_SCALE));
}
__device__ inline double fx_get(unsigned long long a) { return (double)(long long)a * (1.0 / FX_SCALE); }

__device__ inline double block_sum_d(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

// first entry of [q0, q1) (partners in ascending order) whose partner lies above r
__device__ inline int64_t first_above(int64_t q0, int64_t q1, const int32_t* __restrict__ col, int64_t r) {
  while (q0 < q1) {
    const int64_t mid = (q0 + q1) >> 1;
    if ((int64_t)col[mid] <= r) q0 = mid + 1;
    else q1 = mid;
  }
  return q0;
}

// The row-pair sums are symmetric in the two rows (gocc too: its weights belong to the
// columns), and every pair (r, c) of an excitation table has its mirror (c, r), so the
// Gram kernels form only the pairs with c > r; this copies them to the pairs with c < r
// (the same products in the same order: the same bits).
__global__ void k_gram_mirror(int L, const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                              double* __restrict__ g, double* __restrict__ gocc) {
  const int64_t r = blockIdx.x;
  const int64_t q1 = first_above(rowptr[r], rowptr[r + 1], col, r);
  for (int64_t q = rowptr[r] + threadIdx.x; q < q1; q += blockDim.x) {
    const int64_t c = col[q];
    const int64_t m = first_above(rowptr[c], rowptr[c + 1], col, r - 1);  // the entry of r in row c
    g[q] = g[m];
    if (gocc)
      for (int k = 0; k < L; ++k) gocc[q * L + k] = gocc[m * L + k];
  }
}

// entry of partner c in row r of an excitation table (partners ascending); -1 if absent
__device__ inline int64_t find_entry(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                     int64_t r, int64_t c) {
  const int64_t e = first_above(rowptr[r], rowptr[r + 1], col, c - 1);
  return e < rowptr[r + 1] && col[e] == c ? e : -1;
}

// Gram products of the members of one large core of the same-spin doubles (members
// x = core + {A, B}, as sigma groups them): every member row pairs with every other, so a
// workgroup takes a block pair (bi, bj) of one core's members, BM rows each, and streams
// them through LDS in tiles of GT columns once for all BM^2 products (thread (ti, tj) owns
// a TM x TM sub-block).  A member pair whose strings differ in four orbitals is a double
// pair, and that pair shares no other core: its sum goes to both entries of the
// excitation table, (x, y) and (y, x), and marks them done.  Fixed summation order.
template <int BM>
__global__ void __launch_bounds__(RT) k_core_gram(int64_t ncols, const int4* __restrict__ items,
                                                  const int32_t* __restrict__ cstart,
                                                  const int32_t* __restrict__ crow, const uint64_t* __restrict__ str,
                                                  const int64_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ col, const double* __restrict__ X,
                                                  double* __restrict__ g, unsigned char* __restrict__ done) {
  constexpr int TM = BM / 16, GT = 2048 / BM;
  __shared__ double A[GT][BM + 1], B[GT][BM + 1];
  const int4 it = items[blockIdx.x];  // core, block bi, block bj
  const int32_t s0 = cstart[it.x], m = cstart[it.x + 1] - s0;
  const int i0 = it.y * BM, j0 = it.z * BM;
  const bool same = it.y == it.z;
  const int ti = threadIdx.x / 16, tj = threadIdx.x % 16;
  double acc[TM][TM];
#pragma unroll
  for (int u = 0; u < TM; ++u)
#pragma unroll
    for (int v = 0; v < TM; ++v) acc[u][v] = 0.0;
  for (int64_t c0 = 0; c0 < ncols; c0 += GT) {
    __syncthreads();  // the previous tile's readers are done
    for (int x = threadIdx.x; x < BM * GT; x += RT) {
      const int r = x / GT, cc = x % GT;
      const int64_t c = c0 + cc;
      A[cc][r] = (i0 + r < m && c < ncols) ? X[(int64_t)crow[s0 + i0 + r] * ncols + c] : 0.0;
      if (!same) B[cc][r] = (j0 + r < m && c < ncols) ? X[(int64_t)crow[s0 + j0 + r] * ncols + c] : 0.0;
    }
    __syncthreads();
    const double(*Bt)[BM + 1] = same ? A : B;
    for (int cc = 0; cc < GT; ++cc) {
      double a[TM], b[TM];
#pragma unroll
      for (int u = 0; u < TM; ++u) a[u] = A[cc][ti * TM + u];
#pragma unroll
      for (int v = 0; v < TM; ++v) b[v] = Bt[cc][tj * TM + v];
#pragma unroll
      for (int u = 0; u < TM; ++u)
#pragma unroll
        for (int v = 0; v < TM; ++v) acc[u][v] += a[u] * b[v];
    }
  }
#pragma unroll
  for (int u = 0; u < TM; ++u)
#pragma unroll
    for (int v = 0; v < TM; ++v) {
      const int i = i0 + ti * TM + u, j = j0 + tj * TM + v;
      if (i >= m || j >= m || i >= j) continue;
      const int64_t ri = crow[s0 + i], rj = crow[s0 + j];
      if (__popcll(str[ri] ^ str[rj]) != 4) continue;  // pairs sharing an orbital: a single
      const int64_t q1 = find_entry(rowptr, col, ri, rj), q2 = find_entry(rowptr, col, rj, ri);
      if (q1 < 0 || q2 < 0) continue;
      g[q1] = acc[u][v];
      g[q2] = acc[u][v];
      done[q1] = 1;
      done[q2] = 1;
    }
}

// the entries above the diagonal that no core product formed, as a list per row
__global__ void k_pair_flags(const int64_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                             const unsigned char* __restrict__ done, int32_t* __restrict__ flag) {
  const int64_t r = blockIdx.x;
  for (int64_t q = rowptr[r] + threadIdx.x; q < rowptr[r + 1]; q += blockDim.x)
    flag[q] = (col[q] > r && !done[q]) ? 1 : 0;
}
__global__ void k_pair_list(int64_t nnz, const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                            int32_t* __restrict__ qlist) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nnz && flag[q]) qlist[pos[q]] = (int32_t)q;
}
__global__ void k_pair_lptr(int64_t n, int64_t nnz, const int64_t* __restrict__ rowptr,
                            const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                            int64_t* __restrict__ lptr) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  const int64_t q = rowptr[r];
  lptr[r] = q < nnz ? pos[q] : (nnz ? (int64_t)pos[nnz - 1] + flag[nnz - 1] : 0);
}

// g[q] = <X[r,:], X[col[q],:]> for the pairs q = qlist[e], e in [lptr[r], lptr[r+1]), of
// row r, GQ pairs per pass: row r is read once per pass for all of them and their sums are
// reduced together (one barrier pair per pass instead of per pair).  Each sum's order is fixed (a thread's columns in order, then
// a shuffle tree, then the waves in order), so g does not depend on timing.  VEC: rows of
// an even length, read two columns at a time.
constexpr int GQ = 8;
template <bool VEC>
__global__ void __launch_bounds__(RT) k_gram_pairs(int64_t ncols, const int64_t* __restrict__ lptr,
                                                   const int32_t* __restrict__ qlist,
                                                   const int32_t* __restrict__ col, const double* __restrict__ X,
                                                   double* __restrict__ g) {
  __shared__ double sh[RT / 64][GQ];
  const int64_t r = blockIdx.x;
  const double* xr = X + r * ncols;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t q1 = lptr[r + 1];
  for (int64_t q0 = lptr[r]; q0 < q1; q0 += GQ) {
    const int nq = (int)min<int64_t>(GQ, q1 - q0);
    const double* xq[GQ];
    double s[GQ];
#pragma unroll
    for (int u = 0; u < GQ; ++u) {
      xq[u] = X + (int64_t)col[qlist[q0 + min(u, nq - 1)]] * ncols;
      s[u] = 0.0;
    }
    if (VEC) {
      for (int64_t cc = 2 * (int64_t)threadIdx.x; cc < ncols; cc += 2 * RT) {
        const double2 x = *reinterpret_cast<const double2*>(xr + cc);
#pragma unroll
        for (int u = 0; u < GQ; ++u)
          if (u < nq) {
            const double2 y = *reinterpret_cast<const double2*>(xq[u] + cc);
            s[u] += x.x * y.x;
            s[u] += x.y * y.y;
          }
      }
    } else {
      for (int64_t cc = threadIdx.x; cc < ncols; cc += RT) {
        const double x = xr[cc];
#pragma unroll
        for (int u = 0; u < GQ; ++u)
          if (u < nq) s[u] += x * xq[u][cc];
      }
    }
#pragma unroll
    for (int u = 0; u < GQ; ++u)
      for (int o = 32; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o, 64);
    __syncthreads();  // the previous pass's readers of sh are done
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < GQ; ++u) sh[wid][u] = s[u];
    __syncthreads();
    if ((int)threadIdx.x < nq) {
      double t = 0.0;
      for (int w = 0; w < RT / 64; ++w) t += sh[w][threadIdx.x];
      g[qlist[q0 + threadIdx.x]] = t;
    }
  }
}

// g[q] as k_gram_pairs and gocc[q*L + k] = sum_c X[r,c] X[col[q],c] n_k(cstr[c]) for the
// pairs of row r, GQO pairs per pass.  The occupations enter relative to one string of the
// column set, ref = cstr[0]: n_k(c) = n_k(ref) + d_k(c), d_k = +-1 only where c's
// occupation differs from ref's (twice c's excitation level against ref), so
// gocc = n_k(ref) g + sum_c p_c d_k(c).  The second sum takes fixed-point LDS adds (exact,
// so their order does not matter), a few per column for strings near ref instead of one
// per occupied orbital, and fewer of them on one address.
constexpr int GQO = 4;
template <bool VEC>
__global__ void __launch_bounds__(RT) k_gram_occ(int64_t ncols, int L, const int64_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ col, const double* __restrict__ X,
                                                 const uint64_t* __restrict__ cstr, double* __restrict__ g,
                                                 double* __restrict__ gocc) {
  __shared__ double sh[RT / 64][GQO];
  __shared__ unsigned long long acc[GQO][64];
  const int64_t r = blockIdx.x;
  const double* xr = X + r * ncols;
  const uint64_t ref = cstr[0];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t q1 = rowptr[r + 1];
  for (int64_t q0 = first_above(rowptr[r], q1, col, r); q0 < q1; q0 += GQO) {
    const int nq = (int)min<int64_t>(GQO, q1 - q0);
    const double* xq[GQO];
    double s[GQO];
#pragma unroll
    for (int u = 0; u < GQO; ++u) {
      xq[u] = X + (int64_t)col[q0 + min(u, nq - 1)] * ncols;
      s[u] = 0.0;
    }
    __syncthreads();  // the previous pass's readers of acc and sh are done
    for (int x = threadIdx.x; x < GQO * 64; x += RT) acc[x / 64][x % 64] = 0ull;
    __syncthreads();
    auto column = [&](double x, const double (&y)[GQO], uint64_t occ) {
      const uint64_t d = occ ^ ref;
#pragma unroll
      for (int u = 0; u < GQO; ++u)
        if (u < nq) {
          const double p = x * y[u];
          s[u] += p;
          if (p != 0.0)
            for (uint64_t z = d; z; z &= z - 1) {
              const int k = __builtin_ctzll(z);
              fx_add(&acc[u][k], ((occ >> k) & 1ull) ? p : -p);
            }
        }
    };
    if (VEC) {
      for (int64_t cc = 2 * (int64_t)threadIdx.x; cc < ncols; cc += 2 * RT) {
        const double2 x = *reinterpret_cast<const double2*>(xr + cc);
        double y0[GQO], y1[GQO];
#pragma unroll
        for (int u = 0; u < GQO; ++u) {
          const double2 y = u < nq ? *reinterpret_cast<const double2*>(xq[u] + cc) : make_double2(0.0, 0.0);
          y0[u] = y.x;
          y1[u] = y.y;
        }
        column(x.x, y0, cstr[cc]);
        column(x.y, y1, cstr[cc + 1]);
      }
    } else {
      for (int64_t cc = threadIdx.x; cc < ncols; cc += RT) {
        double y[GQO];
#pragma unroll
        for (int u = 0; u < GQO; ++u) y[u] = u < nq ? xq[u][cc] : 0.0;
        column(xr[cc], y, cstr[cc]);
      }
    }
#pragma unroll
    for (int u = 0; u < GQO; ++u)
      for (int o = 32; o > 0; o >>= 1) s[u] += __shfl_xor(s[u], o, 64);
    if (lane == 0)
#pragma unroll
      for (int u = 0; u < GQO; ++u) sh[wid][u] = s[u];
    __syncthreads();
    for (int x = threadIdx.x; x < nq * L; x += RT) {
      const int u = x / L, k = x % L;
      double t = 0.0;
      for (int w = 0; w < RT / 64; ++w) t += sh[w][u];
      if (k == 0) g[q0 + u] = t;
      gocc[(q0 + u) * L + k] = (((ref >> k) & 1ull) ? t : 0.0) + fx_get(acc[u][k]);
    }
  }
}

// per row: rsq[r] = sum_c X^2, P[r*L+k] = sum_c X[r,c]^2 n_k(cstr[c])
__global__ void __launch_bounds__(RT) k_rowocc(int64_t ncols, int L, const double* __restrict__ X,
                                               const uint64_t* __restrict__ cstr, double* __restrict__ rsq,
                                               double* __restrict__ P) {
  __shared__ double sh[RT / 64];
  __shared__ unsigned long long acc[64];
  const int64_t r = blockIdx.x;
  if (threadIdx.x < 64) acc[threadIdx.x] = 0ull;
  __syncthreads();
  double s = 0.0;
  for (int64_t cc = threadIdx.x; cc < ncols; cc += RT) {
    const double x = X[r * ncols + cc];
    const double p = x * x;
    s += p;
    if (p != 0.0) {
      uint64_t y = cstr[cc];
      while (y) {
        const int k = __builtin_ctzll(y);
        y &= y - 1;
        fx_add(&acc[k], p);
      }
    }
  }
  s = block_sum_d(s, sh);
  __syncthreads();
  if (threadIdx.x == 0) rsq[r] = s;
  if (threadIdx.x < L) P[r * L + threadIdx.x] = fx_get(acc[threadIdx.x]);
}

// D[i][j] = sum_r w[r] n_i(rstr[r]) * (Y ? Y[r*L+j] : n_j(rstr[r]))  (one thread per (i,j))
__global__ void k_corr(int64_t nrows, int L, const uint64_t* __restrict__ rstr, const double* __restrict__ w,
                       const double* __restrict__ Y, double* __restrict__ D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L * L) return;
  const int i = t / L, j = t % L;
  double v = 0.0;
  for (int64_t r = 0; r < nrows; ++r) {
    const uint64_t s = rstr[r];
    if (!((s >> i) & 1ull)) continue;
    if (Y) v += Y[r * L + j];
    else if ((s >> j) & 1ull) v += w[r];
  }
  D[t] = v;
}

__device__ inline int64_t idx4(int L, int i, int j, int k, int l) {
  return (int64_t)i + (int64_t)L * j + (int64_t)L * L * k + (int64_t)L * L * L * l;
}

// diagonal slots: gamma_s[p,p], Gamma_st[i,j,i,j], Gamma_ss[i,j,j,i]
__global__ void k_diag_slots(int L, const double* __restrict__ dens, const double* __restrict__ Daa,
                             const double* __restrict__ Dbb, const double* __restrict__ Dab,
                             double* __restrict__ one, double* __restrict__ two) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L * L) return;
  const int i = t / L, j = t % L;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  if (i == j) {
    one[i + L * i] = dens[2 * i];
    one[L2 + i + L * i] = dens[2 * i + 1];
  }
  // alpha-beta (block 2) and beta-alpha (block 1): every (i,j)
  two[2 * L4 + idx4(L, i, j, i, j)] = Dab[i * L + j];
  two[1 * L4 + idx4(L, j, i, j, i)] = Dab[i * L + j];
  if (i != j) {
    two[0 * L4 + idx4(L, i, j, i, j)] = Daa[i * L + j];
    two[0 * L4 + idx4(L, i, j, j, i)] = -Daa[i * L + j];
    two[3 * L4 + idx4(L, i, j, i, j)] = Dbb[i * L + j];
    two[3 * L4 + idx4(L, i, j, j, i)] = -Dbb[i * L + j];
  }
}

// singles owners: one thread per key (a,i); its pairs are perm[seg[key]..seg[key+1]).
// spin s (0 alpha / 1 beta); spectators of the same spin from the ket string, of the
// other spin from gocc.  blk(s,t) = s + 2t.
__global__ void k_single_slots(int L, int s, const int32_t* __restrict__ seg, const int32_t* __restrict__ perm,
                               const int64_t* __restrict__ rowof, const int32_t* __restrict__ col,
                               const uint32_t* __restrict__ exc, const uint64_t* __restrict__ str,
                               const double* __restrict__ g, const double* __restrict__ gocc,
                               double* __restrict__ one, double* __restrict__ two) {
  const int key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key >= L * L) return;
  const int a = key / L, i = key % L;
  if (a == i) return;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  double accS[64], accO[64];
  for (int k = 0; k < L; ++k) accS[k] = accO[k] = 0.0;
  double gam = 0.0;
  for (int32_t e = seg[key]; e < seg[key + 1]; ++e) {
    const int32_t q = perm[e];
    const double sg = ((exc[q] >> 24) & 1u) ? -1.0 : 1.0;
    const double v = sg * g[q];
    gam += v;
    uint64_t sp = str[col[q]] & ~(1ull << i);  // ket occupation minus i = common occupation
    while (sp) {
      const int k = __builtin_ctzll(sp);
      sp &= sp - 1;
      accS[k] += v;
    }
    for (int k = 0; k < L; ++k) accO[k] += sg * gocc[(int64_t)q * L + k];
  }
  (void)rowof;
  const int ss = s + 2 * s;          // same-spin block
  const int so = s + 2 * (1 - s);    // (s, other)
  const int os = (1 - s) + 2 * s;    // (other, s)
  one[s * L2 + a + L * i] = gam;
  for (int k = 0; k < L; ++k) {
    if (k != a && k != i) {
      const double v = accS[k];
      two[ss * L4 + idx4(L, a, k, i, k)] = v;
      two[ss * L4 + idx4(L, k, a, k, i)] = v;
      two[ss * L4 + idx4(L, a, k, k, i)] = -v;
      two[ss * L4 + idx4(L, k, a, i, k)] = -v;
    }
    const double w = accO[k];
    two[so * L4 + idx4(L, a, k, i, k)] = w;
    two[os * L4 + idx4(L, k, a, k, i)] = w;
  }
}

// doubles owners: one thread per distinct quadruple (segment of sorted keys)
__global__ void k_double_slots(int L, int s, int64_t nseg, const int32_t* __restrict__ seg,
                               const int32_t* __restrict__ perm, const uint32_t* __restrict__ keys,
                               const uint32_t* __restrict__ exc, const double* __restrict__ g,
                               double* __restrict__ two) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg) return;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  double v = 0.0;
  for (int32_t e = seg[t]; e < seg[t + 1]; ++e) {
    const int32_t q = perm[e];
    v += (((exc[q] >> 24) & 1u) ? -1.0 : 1.0) * g[q];
  }
  const uint32_t key = keys[seg[t]];
  const int J = key % L, I = (key / L) % L, B = (key / L / L) % L, A = key / L / L / L;
  const int ss = s + 2 * s;
  two[ss * L4 + idx4(L, A, B, I, J)] = v;
  two[ss * L4 + idx4(L, B, A, J, I)] = v;
  two[ss * L4 + idx4(L, A, B, J, I)] = -v;
  two[ss * L4 + idx4(L, B, A, I, J)] = -v;
}

__global__ void k_single_keys(int64_t nnz, int L, const uint32_t* __restrict__ exc, uint32_t* __restrict__ keys,
                              int32_t* __restrict__ idx) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const uint32_t e = exc[q];
  keys[q] = ((e >> 6) & 63u) * (uint32_t)L + (e & 63u);  // a*L + i
  idx[q] = (int32_t)q;
}

__global__ void k_double_keys(int64_t nnz, int L, const uint32_t* __restrict__ exc, uint32_t* __restrict__ keys,
                              int32_t* __restrict__ idx) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const uint32_t e = exc[q];
  const uint32_t I = e & 63u, J = (e >> 6) & 63u, A = (e >> 12) & 63u, B = (e >> 18) & 63u;
  keys[q] = ((A * L + B) * L + I) * L + J;
  idx[q] = (int32_t)q;
}

// segment starts of sorted keys: flag[e] = (e==0 || key[e] != key[e-1])
__global__ void k_flag(int64_t n, const uint32_t* __restrict__ keys, int32_t* __restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  flag[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

__global__ void k_segstart(int64_t n, const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                           int32_t* __restrict__ seg, int32_t nseg) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n && flag[e]) seg[pos[e]] = (int32_t)e;
  if (e == 0) seg[nseg] = (int32_t)n;
}

// per key (a,i) dense segment table: kseg[key] = first sorted entry with key >= key
__global__ void k_keyseg(int64_t n, int nkeys, const uint32_t* __restrict__ keys, int32_t* __restrict__ kseg) {
  const int key = blockIdx.x * blockDim.x + threadIdx.x;
  if (key > nkeys) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int)keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  kseg[key] = (int32_t)lo;
}

// alpha-beta: workgroup w per (key, chunk of <= PCH pairs), the sorted pairs
// [wseg[2w], wseg[2w+1]); partial[T][L*L] in part[w*L2].  LDS: ket row chunk + L*L
// accumulators.
constexpr int PCH = 32;
__global__ void __launch_bounds__(RT) k_ab_rdm(int64_t nb, int L, const int32_t* __restrict__ wseg,
                                               const int32_t* __restrict__ wkey, const int32_t* __restrict__ perm,
                                               const int64_t* __restrict__ rowof, const int32_t* __restrict__ colSa,
                                               const uint32_t* __restrict__ excSa,
                                               const uint32_t* __restrict__ pack,
                                               const int32_t* __restrict__ cptr, int nch, int64_t chunk,
                                               const double* __restrict__ W, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* cb = reinterpret_cast<double*>(smem);
  unsigned long long* T = reinterpret_cast<unsigned long long*>(cb + chunk);
  const int LL = L * L;
  const int w = blockIdx.x;
  for (int x = threadIdx.x; x < LL; x += RT) T[x] = 0ull;
  for (int32_t e = wseg[2 * w]; e < wseg[2 * w + 1]; ++e) {
    const int32_t q = perm[e];
    const int64_t ia = rowof[q], ja = colSa[q];
    const double sa = ((excSa[q] >> 24) & 1u) ? -1.0 : 1.0;
    const double* wi = W + ia * nb;
    for (int k = 0; k < nch; ++k) {
      const int64_t j0 = (int64_t)k * chunk, j1 = min(nb, j0 + chunk);
      __syncthreads();
      for (int64_t x = threadIdx.x; x < j1 - j0; x += RT) cb[x] = W[ja * nb + j0 + x];
      __syncthreads();
      for (int64_t ib = threadIdx.x; ib < nb; ib += RT) {
        const double wb = sa * wi[ib];
        if (wb == 0.0) continue;
        const int32_t* cp = cptr + ib * (nch + 1) + k;
        for (int ee = cp[0]; ee < cp[1]; ++ee) {
          const uint32_t p = pack[ee];
          const int jb = (int)(p & 0x3FFFFu) - (int)j0;
          const double v = (((p >> 30) & 1u) ? -wb : wb) * cb[jb];
          fx_add(&T[(p >> 18) & 0xFFFu], v);
        }
      }
    }
  }
  __syncthreads();
  (void)wkey;
  for (int x = threadIdx.x; x < LL; x += RT) part[(int64_t)w * LL + x] = fx_get(T[x]);
}

// The same for rows of at most RTA*R columns: 1024 threads, thread t owning columns
// t + RTA*r, whose bra values W[ia, ib] (signed) stay in registers for the whole alpha pair,
// so each ket chunk costs its own values and the chunk pointers only.  With the chunk sized
// to the whole LDS this reads a pair's bra row once instead of once per chunk.
constexpr int RTA = 1024;
template <int R>
__global__ void __launch_bounds__(RTA) k_ab_rdm_reg(int64_t nb, int L, const int32_t* __restrict__ wseg,
                                                    const int32_t* __restrict__ perm,
                                                    const int64_t* __restrict__ rowof,
                                                    const int32_t* __restrict__ colSa,
                                                    const uint32_t* __restrict__ excSa,
                                                    const uint32_t* __restrict__ pack,
                                                    const int32_t* __restrict__ cptr, int nch, int64_t chunk,
                                                    const double* __restrict__ W, double* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* cb = reinterpret_cast<double*>(smem);
  unsigned long long* T = reinterpret_cast<unsigned long long*>(cb + chunk);
  const int LL = L * L;
  const int w = blockIdx.x;
  for (int x = threadIdx.x; x < LL; x += RTA) T[x] = 0ull;
  for (int32_t e = wseg[2 * w]; e < wseg[2 * w + 1]; ++e) {
    const int32_t q = perm[e];
    const int64_t ia = rowof[q], ja = colSa[q];
    const double sa = ((excSa[q] >> 24) & 1u) ? -1.0 : 1.0;
    double wb[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t ib = (int64_t)r * RTA + threadIdx.x;
      wb[r] = ib < nb ? sa * W[ia * nb + ib] : 0.0;
    }
    for (int k = 0; k < nch; ++k) {
      const int64_t j0 = (int64_t)k * chunk, j1 = min(nb, j0 + chunk);
      __syncthreads();  // the previous chunk's products are done with cb
      for (int64_t x = threadIdx.x; x < j1 - j0; x += RTA) cb[x] = W[ja * nb + j0 + x];
      __syncthreads();
      // opaque per chunk: the chunk-pointer addresses are formed here, not held per column
      int tt = (int)threadIdx.x;
      asm volatile("" : "+v"(tt));
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (wb[r] == 0.0) continue;  // also the columns past nb
        const int32_t* cp = cptr + ((int64_t)r * RTA + tt) * (nch + 1) + k;
        const int e1 = cp[1];
        for (int ee = cp[0]; ee < e1; ++ee) {
          const uint32_t p = pack[ee];
          const int jb = (int)(p & 0x3FFFFu) - (int)j0;
          fx_add(&T[(p >> 18) & 0xFFFu], (((p >> 30) & 1u) ? -wb[r] : wb[r]) * cb[jb]);
        }
      }
    }
  }
  __syncthreads();
  for (int x = threadIdx.x; x < LL; x += RTA) part[(int64_t)w * LL + x] = fx_get(T[x]);
}

// sum the chunk partials of each key (a,i) and write Gamma_ab[a,b,i,j], Gamma_ba[b,a,j,i]
__global__ void k_ab_slots(int L, const int32_t* __restrict__ kfirst, const double* __restrict__ part,
                           double* __restrict__ two) {
  const int key = blockIdx.x;
  const int a = key / L, i = key % L;
  const int LL = L * L;
  const int64_t L4 = (int64_t)LL * LL;
  const int src = a > i ? key : i * L + a;  // keys with a < i: the mirror key, (b, j) swapped
  for (int x = threadIdx.x; x < LL; x += blockDim.x) {
    const int b = x / L, j = x % L;
    const int xs = a > i ? x : j * L + b;
    double v = 0.0;
    for (int32_t c = kfirst[src]; c < kfirst[src + 1]; ++c) v += part[(int64_t)c * LL + xs];
    if (a == i || b == j) continue;
    two[2 * L4 + idx4(L, a, b, i, j)] = v;
    two[1 * L4 + idx4(L, b, a, j, i)] = v;
  }
}

__global__ void k_rowof(int64_t n, const int64_t* __restrict__ rowptr, int64_t* __restrict__ rowof) {
  const int64_t r = blockIdx.x;
  for (int64_t q = rowptr[r] + threadIdx.x; q < rowptr[r + 1]; q += blockDim.x) rowof[q] = r;
  (void)n;
}

__global__ void k_backward(int L, const double* __restrict__ one, const double* __restrict__ two, double g,
                           double* __restrict__ dh, double* __restrict__ deri) {
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L4 && deri) {
    const int s_ = (int)(t % L), r = (int)((t / L) % L), q = (int)((t / L2) % L), p = (int)(t / L2 / L);
    const int64_t x = idx4(L, p, r, q, s_);
    deri[t] = 0.5 * g * (two[x] + two[L4 + x] + two[2 * L4 + x] + two[3 * L4 + x]);
  }
  if (t < L2 && dh) {
    const int q = (int)(t % L), p = (int)(t / L);
    dh[t] = g * (one[p + L * q] + one[L2 + p + L * q]);
  }
}

inline unsigned nb_for(int64_t n, int t) { return (unsigned)std::max<int64_t>(1, (n + t - 1) / t); }

// sort (key, index) pairs; returns sorted keys/indices in the out buffers
void sort_pairs(int64_t n, DBuf<uint32_t>& keys, DBuf<int32_t>& idx, DBuf<uint32_t>& keys_out,
                DBuf<int32_t>& idx_out, int end_bit, hipStream_t st) {
  keys_out.alloc(std::max<int64_t>(n, 1));
  idx_out.alloc(std::max<int64_t>(n, 1));
  if (n == 0) return;
  size_t tmp = 0;
  SBDG_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys.p, keys_out.p, idx.p, idx_out.p, (int)n, 0,
                                                end_bit, st));
  DBuf<unsigned char> t;
  t.alloc(tmp + 1);
  SBDG_CHECK(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys.p, keys_out.p, idx.p, idx_out.p, (int)n, 0, end_bit,
                                                st));
  SBDG_CHECK(hipStreamSynchronize(st));
}

int bits_for(uint64_t v) {
  int b = 1;
  while ((1ull << b) <= v) ++b;
  return b;
}

struct PairBlock {  // g (and gocc for singles) of one table
  DBuf<double> g, gocc;
};

// singles of one spin: Gram products, then key owners
void singles_part(const Table& t, int L, int s, const double* X, int64_t ncols, const uint64_t* rstr,
                  const uint64_t* cstr, double* one, double* two, hipStream_t st) {
  if (t.nnz == 0) return;
  PairBlock pb;
  pb.g.alloc(t.nnz);
  pb.gocc.alloc((size_t)t.nnz * L);
  if (ncols % 2 == 0)
    hipLaunchKernelGGL(k_gram_occ<true>, dim3((unsigned)t.n), dim3(RT), 0, st, ncols, L, t.rowptr.p, t.col.p, X,
                       cstr, pb.g.p, pb.gocc.p);
  else
    hipLaunchKernelGGL(k_gram_occ<false>, dim3((unsigned)t.n), dim3(RT), 0, st, ncols, L, t.rowptr.p, t.col.p, X,
                       cstr, pb.g.p, pb.gocc.p);
  hipLaunchKernelGGL(k_gram_mirror, dim3((unsigned)t.n), dim3(RT), 0, st, L, t.rowptr.p, t.col.p, pb.g.p, pb.gocc.p);
  DBuf<uint32_t> keys, ks;
  DBuf<int32_t> idx, perm, kseg;
  keys.alloc(t.nnz);
  idx.alloc(t.nnz);
  hipLaunchKernelGGL(k_single_keys, dim3(nb_for(t.nnz, 256)), dim3(256), 0, st, t.nnz, L, t.exc.p, keys.p, idx.p);
  sort_pairs(t.nnz, keys, idx, ks, perm, bits_for((uint64_t)L * L), st);
  kseg.alloc(L * L + 1);
  hipLaunchKernelGGL(k_keyseg, dim3(nb_for(L * L + 1, 256)), dim3(256), 0, st, t.nnz, L * L, ks.p, kseg.p);
  hipLaunchKernelGGL(k_single_slots, dim3(nb_for(L * L, 64)), dim3(64), 0, st, L, s, kseg.p, perm.p, nullptr,
                     t.col.p, t.exc.p, rstr, pb.g.p, pb.gocc.p, one, two);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipStreamSynchronize(st));
}

// Row-pair sums of the doubles table t: by core block for the cores of ct (the larger
// cores, when sigma built them), by pair for the rest of the entries above the diagonal,
// then mirrored to the entries below it.
void doubles_gram(const Table& t, const CoreTable* ct, const uint64_t* strs, const double* X, int64_t ncols,
                  double* g, hipStream_t st) {
  DBuf<unsigned char> done;
  done.alloc(t.nnz);
  SBDG_CHECK(hipMemsetAsync(done.p, 0, t.nnz, st));
  if (ct && ct->ncore > 0) {
    std::vector<int32_t> hs(ct->ncore + 1);
    SBDG_CHECK(hipMemcpyAsync(hs.data(), ct->start.p, sizeof(int32_t) * hs.size(), hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipStreamSynchronize(st));
    std::vector<int4> items[3];  // blocks of 16, 32, 64 members
    for (int64_t k = 0; k < ct->ncore; ++k) {
      const int m = hs[k + 1] - hs[k];
      if (m < 2) continue;
      const int w = m <= 16 ? 0 : (m <= 32 ? 1 : 2), bm = 16 << w, nbk = (m + bm - 1) / bm;
      for (int bi = 0; bi < nbk; ++bi)
        for (int bj = bi; bj < nbk; ++bj) items[w].push_back(make_int4((int)k, bi, bj, 0));
    }
    DBuf<int4> ditems[3];
    for (int w = 0; w < 3; ++w) {
      if (items[w].empty()) continue;
      ditems[w].alloc(items[w].size());
      SBDG_CHECK(hipMemcpyAsync(ditems[w].p, items[w].data(), sizeof(int4) * items[w].size(), hipMemcpyHostToDevice,
                                st));
    }
#define SBDG_CG(W, BM)                                                                                      \
  if (!items[W].empty())                                                                                    \
  hipLaunchKernelGGL(k_core_gram<BM>, dim3((unsigned)items[W].size()), dim3(RT), 0, st, ncols, ditems[W].p, \
                     ct->start.p, ct->row.p, strs, t.rowptr.p, t.col.p, X, g, done.p)
    SBDG_CG(0, 16);
    SBDG_CG(1, 32);
    SBDG_CG(2, 64);
#undef SBDG_CG
    SBDG_CHECK(hipGetLastError());
    SBDG_CHECK(hipStreamSynchronize(st));  // the item lists on the host stay alive until here
  }
  DBuf<int32_t> flag, pos, qlist;
  DBuf<int64_t> lptr;
  flag.alloc(t.nnz);
  pos.alloc(t.nnz);
  qlist.alloc(t.nnz);
  lptr.alloc(t.n + 1);
  hipLaunchKernelGGL(k_pair_flags, dim3((unsigned)t.n), dim3(RT), 0, st, t.rowptr.p, t.col.p, done.p, flag.p);
  size_t tmp = 0;
  SBDG_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag.p, pos.p, (int)t.nnz, st));
  DBuf<unsigned char> tb;
  tb.alloc(tmp + 1);
  SBDG_CHECK(hipcub::DeviceScan::ExclusiveSum(tb.p, tmp, flag.p, pos.p, (int)t.nnz, st));
  hipLaunchKernelGGL(k_pair_list, dim3(nb_for(t.nnz, 256)), dim3(256), 0, st, t.nnz, flag.p, pos.p, qlist.p);
  hipLaunchKernelGGL(k_pair_lptr, dim3(nb_for(t.n + 1, 256)), dim3(256), 0, st, t.n, t.nnz, t.rowptr.p, flag.p, pos.p,
                     lptr.p);
  if (ncols % 2 == 0)
    hipLaunchKernelGGL(k_gram_pairs<true>, dim3((unsigned)t.n), dim3(RT), 0, st, ncols, lptr.p, qlist.p, t.col.p, X,
                       g);
  else
    hipLaunchKernelGGL(k_gram_pairs<false>, dim3((unsigned)t.n), dim3(RT), 0, st, ncols, lptr.p, qlist.p, t.col.p,
                       X, g);
  hipLaunchKernelGGL(k_gram_mirror, dim3((unsigned)t.n), dim3(RT), 0, st, 0, t.rowptr.p, t.col.p, g, nullptr);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipStreamSynchronize(st));
}

void doubles_part(const Table& t, int L, int s, const double* X, int64_t ncols, const CoreTable* ct,
                  const uint64_t* strs, double* two, hipStream_t st) {
  if (t.nnz == 0) return;
  DBuf<double> g;
  g.alloc(t.nnz);
  doubles_gram(t, ct, strs, X, ncols, g.p, st);
  DBuf<uint32_t> keys, ks;
  DBuf<int32_t> idx, perm, flag, pos, seg;
  keys.alloc(t.nnz);
  idx.alloc(t.nnz);
  hipLaunchKernelGGL(k_double_keys, dim3(nb_for(t.nnz, 256)), dim3(256), 0, st, t.nnz, L, t.exc.p, keys.p, idx.p);
  sort_pairs(t.nnz, keys, idx, ks, perm, bits_for((uint64_t)L * L * L * L), st);
  flag.alloc(t.nnz);
  pos.alloc(t.nnz);
  hipLaunchKernelGGL(k_flag, dim3(nb_for(t.nnz, 256)), dim3(256), 0, st, t.nnz, ks.p, flag.p);
  size_t tmp = 0;
  SBDG_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, flag.p, pos.p, (int)t.nnz, st));
  DBuf<unsigned char> tb;
  tb.alloc(tmp + 1);
  SBDG_CHECK(hipcub::DeviceScan::ExclusiveSum(tb.p, tmp, flag.p, pos.p, (int)t.nnz, st));
  int32_t lastpos = 0, lastflag = 0;
  SBDG_CHECK(hipMemcpyAsync(&lastpos, pos.p + t.nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  SBDG_CHECK(hipMemcpyAsync(&lastflag, flag.p + t.nnz - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  SBDG_CHECK(hipStreamSynchronize(st));
  const int32_t nseg = lastpos + lastflag;
  seg.alloc(nseg + 1);
  hipLaunchKernelGGL(k_segstart, dim3(nb_for(t.nnz, 256)), dim3(256), 0, st, t.nnz, flag.p, pos.p, seg.p, nseg);
  hipLaunchKernelGGL(k_double_slots, dim3(nb_for(nseg, 256)), dim3(256), 0, st, L, s, (int64_t)nseg, seg.p, perm.p,
                     ks.p, t.exc.p, g.p, two);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipStreamSynchronize(st));
}

}  // namespace

void rdm_compute(const RdmPlan& P, const double* W, double* WT, double* one, double* two, hipStream_t st) {
  const int L = P.L;
  const int64_t na = P.na, nb = P.nb, N = na * nb;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  SBDG_CHECK(hipMemsetAsync(one, 0, sizeof(double) * 2 * L2, st));
  SBDG_CHECK(hipMemsetAsync(two, 0, sizeof(double) * 4 * L4, st));
  if (N == 0) return;
  transpose_d(na, nb, W, WT, st);

  // diagonal: density correlators
  {
    DBuf<double> rsa, Pa, rsb, Pb, Daa, Dbb, Dab, dens;
    rsa.alloc(na);
    Pa.alloc(na * L);
    rsb.alloc(nb);
    Pb.alloc(nb * L);
    Daa.alloc(L2);
    Dbb.alloc(L2);
    Dab.alloc(L2);
    dens.alloc(2 * L);
    // rows of W: rsa[ia] = sum W^2, Pa[ia][j] = sum_ib W^2 n_j(ib)  (beta occupations)
    hipLaunchKernelGGL(k_rowocc, dim3((unsigned)na), dim3(RT), 0, st, nb, L, W, P.b, rsa.p, Pa.p);
    hipLaunchKernelGGL(k_rowocc, dim3((unsigned)nb), dim3(RT), 0, st, na, L, WT, P.a, rsb.p, Pb.p);
    hipLaunchKernelGGL(k_corr, dim3(nb_for(L2, 64)), dim3(64), 0, st, na, L, P.a, rsa.p, nullptr, Daa.p);
    hipLaunchKernelGGL(k_corr, dim3(nb_for(L2, 64)), dim3(64), 0, st, nb, L, P.b, rsb.p, nullptr, Dbb.p);
    hipLaunchKernelGGL(k_corr, dim3(nb_for(L2, 64)), dim3(64), 0, st, na, L, P.a, nullptr, Pa.p, Dab.p);
    // densities: diagonal of Daa/Dbb
    std::vector<double> hdaa(L2), hdbb(L2), hd(2 * L);
    SBDG_CHECK(hipMemcpyAsync(hdaa.data(), Daa.p, sizeof(double) * L2, hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipMemcpyAsync(hdbb.data(), Dbb.p, sizeof(double) * L2, hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < L; ++i) {
      hd[2 * i] = hdaa[i * L + i];
      hd[2 * i + 1] = hdbb[i * L + i];
    }
    SBDG_CHECK(hipMemcpyAsync(dens.p, hd.data(), sizeof(double) * 2 * L, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_diag_slots, dim3(nb_for(L2, 64)), dim3(64), 0, st, L, dens.p, Daa.p, Dbb.p, Dab.p, one, two);
    SBDG_CHECK(hipGetLastError());
    SBDG_CHECK(hipStreamSynchronize(st));
  }
  // alpha pairs (beta spectators from the beta strings of the columns)
  singles_part(*P.sa, L, 0, W, nb, P.a, P.b, one, two, st);
  doubles_part(*P.da, L, 0, W, nb, P.ca, P.a, two, st);
  // beta pairs on the transposed vector (alpha spectators from the alpha strings)
  singles_part(*P.sb, L, 1, WT, na, P.b, P.a, one, two, st);
  doubles_part(*P.db, L, 1, WT, na, P.cb, P.b, two, st);

  // alpha-beta
  const Table& sa = *P.sa;
  if (sa.nnz > 0 && P.sb->nnz > 0) {
    DBuf<int64_t> rowof;
    rowof.alloc(sa.nnz);
    hipLaunchKernelGGL(k_rowof, dim3((unsigned)na), dim3(256), 0, st, na, sa.rowptr.p, rowof.p);
    DBuf<uint32_t> keys, ks;
    DBuf<int32_t> idx, perm, kseg;
    keys.alloc(sa.nnz);
    idx.alloc(sa.nnz);
    hipLaunchKernelGGL(k_single_keys, dim3(nb_for(sa.nnz, 256)), dim3(256), 0, st, sa.nnz, L, sa.exc.p, keys.p,
                       idx.p);
    sort_pairs(sa.nnz, keys, idx, ks, perm, bits_for((uint64_t)L * L), st);
    kseg.alloc(L2 + 1);
    hipLaunchKernelGGL(k_keyseg, dim3(nb_for(L2 + 1, 256)), dim3(256), 0, st, sa.nnz, (int)L2, ks.p, kseg.p);
    std::vector<int32_t> hk(L2 + 1);
    SBDG_CHECK(hipMemcpyAsync(hk.data(), kseg.p, sizeof(int32_t) * (L2 + 1), hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipStreamSynchronize(st));
    // split each key's run into workgroups of <= PCH pairs; only keys (a, i) with a > i: the
    // pairs of key (i, a) are the same pairs read the other way, with the same product and
    // signs, so T_(i,a)[j,b] = T_(a,i)[b,j] (k_ab_slots reads it there)
    std::vector<int32_t> wseg, wkey, kfirst(L2 + 1);
    for (int key = 0; key < (int)L2; ++key) {
      kfirst[key] = (int32_t)wkey.size();
      if (key / L <= key % L) continue;
      for (int32_t e = hk[key]; e < hk[key + 1]; e += PCH) {
        wseg.push_back(e);  // [start, end) of the workgroup's pairs
        wseg.push_back(std::min(e + PCH, hk[key + 1]));
        wkey.push_back(key);
      }
    }
    kfirst[L2] = (int32_t)wkey.size();
    const int nw = (int)wkey.size();
    DBuf<int32_t> dws, dwk, dkf;
    DBuf<double> part;
    dws.alloc(std::max<size_t>(wseg.size(), 1));
    dwk.alloc(std::max(nw, 1));
    dkf.alloc(L2 + 1);
    part.alloc((size_t)std::max(nw, 1) * L2);
    if (nw) SBDG_CHECK(hipMemcpyAsync(dws.p, wseg.data(), sizeof(int32_t) * wseg.size(), hipMemcpyHostToDevice, st));
    if (nw) SBDG_CHECK(hipMemcpyAsync(dwk.p, wkey.data(), sizeof(int32_t) * nw, hipMemcpyHostToDevice, st));
    SBDG_CHECK(hipMemcpyAsync(dkf.p, kfirst.data(), sizeof(int32_t) * (L2 + 1), hipMemcpyHostToDevice, st));
    const size_t lds = sizeof(double) * (size_t)(P.ab->chunk + L2);
    if (nw) {
      const int64_t R = (nb + RTA - 1) / RTA;  // bra values per thread
#define SBDG_ABR(RR)                                                                                          \
  hipLaunchKernelGGL(k_ab_rdm_reg<RR>, dim3(nw), dim3(RTA), lds, st, nb, L, dws.p, perm.p, rowof.p, sa.col.p, \
                     sa.exc.p, P.ab->pack.p, P.ab->cptr.p, P.ab->nch, P.ab->chunk, W, part.p)
      if (R <= 2) SBDG_ABR(2);
      else if (R <= 4) SBDG_ABR(4);
      else if (R <= 8) SBDG_ABR(8);
      else if (R <= 12) SBDG_ABR(12);
      else if (R <= 16) SBDG_ABR(16);
      else if (R <= 20) SBDG_ABR(20);
      else if (R <= 24) SBDG_ABR(24);
      else if (R <= 32) SBDG_ABR(32);
      else if (R <= 40) SBDG_ABR(40);
      else
        hipLaunchKernelGGL(k_ab_rdm, dim3(nw), dim3(RT), lds, st, nb, L, dws.p, dwk.p, perm.p, rowof.p, sa.col.p,
                           sa.exc.p, P.ab->pack.p, P.ab->cptr.p, P.ab->nch, P.ab->chunk, W, part.p);
#undef SBDG_ABR
    }
    hipLaunchKernelGGL(k_ab_slots, dim3((unsigned)L2), dim3(256), 0, st, L, dkf.p, part.p, two);
    SBDG_CHECK(hipGetLastError());
    SBDG_CHECK(hipStreamSynchronize(st));
  }
}

void rdm_backward(const double* d_one, const double* d_two, int L, double g, double* d_dh, double* d_deri,
                  hipStream_t st) {
  const int64_t L4 = (int64_t)L * L * L * L;
  hipLaunchKernelGGL(k_backward, dim3(nb_for(L4, 256)), dim3(256), 0, st, L, d_one, d_two, g, d_dh, d_deri);
  SBDG_CHECK(hipGetLastError());
}

}  // namespace sbdg
