// Excitation tables and the diagonal, built on the device.
//
// Tables (replacing GenerateExcitation + MakeSmartHelper, tpb/helper.h:132-197,367-437):
// for every bra string the kets at XOR-popcount 2 (singles) and 4 (doubles), ascending.
// One workgroup per bra row scans the string set twice (count, then fill with a
// ballot/prefix so the ascending order survives), then a per-entry pass derives the
// orbitals with ctz and the separated-order sign with a masked popcount, and evaluates
// the spin-string-only part of the Slater-Condon element (determinants.h:423-465).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sbdg_internal.h"

namespace sbdg {

namespace {

constexpr int TB = 256;  // threads per row workgroup

__device__ inline uint64_t between_mask(int x, int y) {
  // bits strictly between orbitals x and y (x != y)
  int lo = x < y ? x : y, hi = x < y ? y : x;
  uint64_t below_hi = (hi >= 64) ? ~0ull : ((1ull << hi) - 1ull);
  uint64_t upto_lo = (lo + 1 >= 64) ? ~0ull : ((1ull << (lo + 1)) - 1ull);
  return below_hi & ~upto_lo;
}

__device__ inline int block_sum(int v, int* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
    sh[0] = t;
  }
  __syncthreads();
  t = sh[0];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(TB) k_count(const uint64_t* __restrict__ s, int64_t n,
                                              int64_t* __restrict__ cs, int64_t* __restrict__ cd) {
  __shared__ int sh[TB / 64];
  const int64_t i = blockIdx.x;
  const uint64_t si = s[i];
  int ns = 0, nd = 0;
  for (int64_t j = threadIdx.x; j < n; j += TB) {
    int d = __popcll(si ^ s[j]);
    ns += (d == 2);
    nd += (d == 4);
  }
  ns = block_sum(ns, sh);
  nd = block_sum(nd, sh);
  if (threadIdx.x == 0) {
    cs[i] = ns;
    cd[i] = nd;
  }
}

// exclusive scan of cnt[0..n) into rp[0..n] (one workgroup, fixed order)
__global__ void __launch_bounds__(1024) k_scan(const int64_t* __restrict__ cnt, int64_t n,
                                               int64_t* __restrict__ rp) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t b = t * per, e = min(n, b + per);
  int64_t s = 0;
  for (int64_t k = b; k < e; ++k) s += cnt[k];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    for (int k = 0; k < 1024; ++k) {
      int64_t v = part[k];
      part[k] = acc;
      acc += v;
    }
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int64_t k = b; k < e; ++k) {
    rp[k] = acc;
    acc += cnt[k];
  }
  if (e == n && b < e) rp[n] = acc;
  if (n == 0 && t == 0) rp[0] = 0;
}

__global__ void __launch_bounds__(TB) k_fill(const uint64_t* __restrict__ s, int64_t n,
                                             const int64_t* __restrict__ rps,
                                             const int64_t* __restrict__ rpd,
                                             int32_t* __restrict__ cols, int32_t* __restrict__ cold) {
  __shared__ int wcs[TB / 64], wcd[TB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t i = blockIdx.x;
  const uint64_t si = s[i];
  int64_t ps = rps[i], pd = rpd[i];
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int64_t base = 0; base < n; base += TB) {
    const int64_t j = base + threadIdx.x;
    int d = (j < n) ? __popcll(si ^ s[j]) : 0;
    uint64_t ms = __ballot(d == 2), md = __ballot(d == 4);
    if (lane == 0) {
      wcs[wid] = __popcll(ms);
      wcd[wid] = __popcll(md);
    }
    __syncthreads();
    int offs = 0, offd = 0, tots = 0, totd = 0;
    for (int w = 0; w < TB / 64; ++w) {
      if (w < wid) {
        offs += wcs[w];
        offd += wcd[w];
      }
      tots += wcs[w];
      totd += wcd[w];
    }
    if (d == 2) cols[ps + offs + __popcll(ms & lt)] = (int32_t)j;
    if (d == 4) cold[pd + offd + __popcll(md & lt)] = (int32_t)j;
    ps += tots;
    pd += totd;
    __syncthreads();
  }
}

__device__ inline double eri4(const double* __restrict__ eri, int L, int p, int q, int r, int s) {
  return eri[(((int64_t)p * L + q) * L + r) * L + s];
}

// singles: bra-only a, ket-only i; sign from the ket string between them;
// value sign*(h_ai + sum over occ(ket)\{i} of [(ai|kk) - (ak|ki)])
__global__ void __launch_bounds__(TB) k_single(const uint64_t* __restrict__ s, int L,
                                               const double* __restrict__ h1,
                                               const double* __restrict__ eri,
                                               const int64_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               uint32_t* __restrict__ exc, double* __restrict__ val,
                                               float* __restrict__ val32) {
  const int64_t r = blockIdx.x;
  const uint64_t bra = s[r];
  for (int64_t q = rp[r] + threadIdx.x; q < rp[r + 1]; q += TB) {
    const uint64_t ket = s[col[q]];
    const uint64_t bo = bra & ~ket, ko = ket & ~bra;
    const int a = __builtin_ctzll(bo), i = __builtin_ctzll(ko);
    const int neg = __popcll(ket & between_mask(i, a)) & 1;
    double v = h1[a * L + i];
    uint64_t occ = ket & ~(1ull << i);
    while (occ) {
      int k = __builtin_ctzll(occ);
      occ &= occ - 1;
      v += eri4(eri, L, a, i, k, k) - eri4(eri, L, a, k, k, i);
    }
    v = neg ? -v : v;
    exc[q] = (uint32_t)i | ((uint32_t)a << 6) | ((uint32_t)neg << 24);
    val[q] = v;
    val32[q] = (float)v;
  }
}

// doubles: ket-only I<J, bra-only A<B; sign by the two-pair rule of TwoExcite
// (determinants.h:447-465) applied to the ket string; value sign*((AI|BJ)-(AJ|BI))
__global__ void __launch_bounds__(TB) k_double(const uint64_t* __restrict__ s, int L,
                                               const double* __restrict__ eri,
                                               const int64_t* __restrict__ rp,
                                               const int32_t* __restrict__ col,
                                               uint32_t* __restrict__ exc, double* __restrict__ val,
                                               float* __restrict__ val32) {
  const int64_t r = blockIdx.x;
  const uint64_t bra = s[r];
  for (int64_t q = rp[r] + threadIdx.x; q < rp[r + 1]; q += TB) {
    const uint64_t ket = s[col[q]];
    uint64_t bo = bra & ~ket, ko = ket & ~bra;
    const int A = __builtin_ctzll(bo);
    bo &= bo - 1;
    const int B = __builtin_ctzll(bo);
    const int I = __builtin_ctzll(ko);
    ko &= ko - 1;
    const int J = __builtin_ctzll(ko);
    int neg = (__popcll(ket & between_mask(I, A)) + __popcll(ket & between_mask(J, B))) & 1;
    if (A > J || B < I) neg ^= 1;
    double v = eri4(eri, L, A, I, B, J) - eri4(eri, L, A, J, B, I);
    v = neg ? -v : v;
    exc[q] = (uint32_t)I | ((uint32_t)J << 6) | ((uint32_t)A << 12) | ((uint32_t)B << 18) |
             ((uint32_t)neg << 24);
    val[q] = v;
    val32[q] = (float)v;
  }
}

__device__ inline int32_t single_jai(uint32_t e, int L) {
  const int32_t p = (int32_t)((e >> 6) & 63u) * L + (int32_t)(e & 63u);  // bra-only * L + ket-only
  return ((e >> 24) & 1u) ? -(p + 1) : (p + 1);
}

__global__ void k_jai(int64_t nnz, int L, const uint32_t* __restrict__ exc, int32_t* __restrict__ jai) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nnz) jai[q] = single_jai(exc[q], L);
}

__global__ void k_merge(int64_t n, int L, const int64_t* __restrict__ rps, const int32_t* __restrict__ cs,
                        const double* __restrict__ vs, const float* __restrict__ vs32,
                        const uint32_t* __restrict__ es,
                        const int64_t* __restrict__ rpd, const int32_t* __restrict__ cd,
                        const double* __restrict__ vd, const float* __restrict__ vd32,
                        int64_t* __restrict__ rpm, int32_t* __restrict__ cm,
                        double* __restrict__ vm, float* __restrict__ vm32, int32_t* __restrict__ jm) {
  const int64_t r = blockIdx.x;
  const int64_t b = rps[r] + rpd[r];
  const int64_t ns = rps[r + 1] - rps[r], nd = rpd[r + 1] - rpd[r];
  if (threadIdx.x == 0) {
    rpm[r] = b;
    if (r == n - 1) rpm[n] = rps[n] + rpd[n];
  }
  for (int64_t k = threadIdx.x; k < ns + nd; k += blockDim.x) {
    if (k < ns) {
      cm[b + k] = cs[rps[r] + k];
      vm[b + k] = vs[rps[r] + k];
      vm32[b + k] = vs32[rps[r] + k];
      jm[b + k] = single_jai(es[rps[r] + k], L);
    } else {
      cm[b + k] = cd[rpd[r] + k - ns];
      vm[b + k] = vd[rpd[r] + k - ns];
      vm32[b + k] = vd32[rpd[r] + k - ns];
      jm[b + k] = 0;
    }
  }
}

__global__ void k_abpack(int64_t nnz, int L, const int32_t* __restrict__ col,
                         const uint32_t* __restrict__ exc, uint32_t* __restrict__ pack) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnz) return;
  const uint32_t e = exc[q];
  const uint32_t j = e & 63u, b = (e >> 6) & 63u, neg = (e >> 24) & 1u;
  pack[q] = (uint32_t)col[q] | ((b * (uint32_t)L + j) << 18) | (neg << 30);
}

// cptr[r*(nch+1)+k] = first entry of row r with ket >= k*chunk (k = 0..nch)
__global__ void k_abchunk(int64_t n, int nch, int64_t chunk, const int64_t* __restrict__ rp,
                          const int32_t* __restrict__ col, int32_t* __restrict__ cptr) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * (nch + 1)) return;
  const int64_t r = t / (nch + 1);
  const int k = (int)(t % (nch + 1));
  const int64_t lim = (int64_t)k * chunk;
  int64_t lo = rp[r], hi = rp[r + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)col[mid] < lim) lo = mid + 1;
    else hi = mid;
  }
  cptr[t] = (int32_t)lo;
}


// J[(p+1)*n + r] = sum_{o in occ(s[r])} eri[p*L^2 + o*L + o]; row 0 zero
__global__ void k_occ_table(const uint64_t* __restrict__ s, int64_t n, int L, const double* __restrict__ eri,
                            double* __restrict__ J, float* __restrict__ J32) {
  const int64_t LL = (int64_t)L * L;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (LL + 1) * n) return;
  const int64_t p = t / n, r = t - p * n;
  double v = 0.0;
  if (p > 0) {
    const double* e = eri + (p - 1) * LL;
    uint64_t y = s[r];
    while (y) {
      const int o = __builtin_ctzll(y);
      y &= y - 1;
      v += e[(int64_t)o * L + o];
    }
  }
  J[t] = v;
  J32[t] = (float)v;
}

// memberships: thread t -> (string r, electron pair q); key = core bits, value =
// r | P(a,b) << 32 | neg << 63 with eps = (-1)^(#core below a + #core below b)
__global__ void k_core_enum(const uint64_t* __restrict__ s, int64_t n, int nel, uint64_t* __restrict__ key,
                            uint64_t* __restrict__ val) {
  const int64_t npp = (int64_t)nel * (nel - 1) / 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * npp) return;
  const int64_t r = t / npp;
  int q = (int)(t % npp);
  int j = 1;
  while ((j + 1) * j / 2 <= q) ++j;  // positions i < j of the occupied orbitals
  const int i = q - j * (j - 1) / 2;
  uint64_t x = s[r];
  int a = -1, b = -1;
  for (int pos = 0; x; ++pos) {
    const int o = __builtin_ctzll(x);
    x &= x - 1;
    if (pos == i) a = o;
    if (pos == j) b = o;
  }
  const uint64_t core = s[r] & ~(1ull << a) & ~(1ull << b);
  const int neg = (__popcll(core & ((1ull << a) - 1ull)) + __popcll(core & ((1ull << b) - 1ull))) & 1;
  key[t] = core;
  val[t] = (uint64_t)r | ((uint64_t)(b * (b - 1) / 2 + a) << 32) | ((uint64_t)neg << 63);
}

// W[P*npair+Q] = (AI|BJ) - (AJ|BI) for P=(A<B), Q=(I<J) disjoint, 0 otherwise
__global__ void k_pair_int(const double* __restrict__ eri, int L, int64_t npair, double* __restrict__ W,
                           float* __restrict__ W32) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npair * npair) return;
  const int64_t P = t / npair, Q = t % npair;
  auto unpair = [](int64_t p, int& lo, int& hi) {
    int b = 1;
    while ((int64_t)(b + 1) * b / 2 <= p) ++b;
    hi = b;
    lo = (int)(p - (int64_t)b * (b - 1) / 2);
  };
  int A, B, I, J;
  unpair(P, A, B);
  unpair(Q, I, J);
  double v = 0.0;
  if (A != I && A != J && B != I && B != J) v = eri4(eri, L, A, I, B, J) - eri4(eri, L, A, J, B, I);
  W[t] = v;
  W32[t] = (float)v;
}

// per-string diagonal parts: E[s] = sum_{i in s} h_ii + sum_{i<j in s} [(ii|jj)-(ij|ji)]
__global__ void k_strdiag(const uint64_t* __restrict__ s, int64_t n, int L,
                          const double* __restrict__ h1, const double* __restrict__ eri,
                          double* __restrict__ E) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  uint64_t x = s[t];
  double e = 0.0;
  while (x) {
    int i = __builtin_ctzll(x);
    x &= x - 1;
    e += h1[i * L + i];
    uint64_t y = x;
    while (y) {
      int j = __builtin_ctzll(y);
      y &= y - 1;
      e += eri4(eri, L, i, i, j, j) - eri4(eri, L, i, j, j, i);
    }
  }
  E[t] = e;
}

// hii[ia,ib] = core + Ea[ia] + Eb[ib] + sum_{i in a, j in b} (ii|jj)
__global__ void __launch_bounds__(TB) k_hii(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                            int64_t nb, int L, double core,
                                            const double* __restrict__ Ea,
                                            const double* __restrict__ Eb,
                                            const double* __restrict__ eri, double* __restrict__ hii) {
  __shared__ double J[64];
  const int64_t ia = blockIdx.x;
  const uint64_t A = a[ia];
  if (threadIdx.x < L) {
    double v = 0.0;
    uint64_t x = A;
    while (x) {
      int i = __builtin_ctzll(x);
      x &= x - 1;
      v += eri4(eri, L, i, i, threadIdx.x, threadIdx.x);
    }
    J[threadIdx.x] = v;
  }
  __syncthreads();
  const double base = core + Ea[ia];
  for (int64_t ib = threadIdx.x; ib < nb; ib += TB) {
    uint64_t x = b[ib];
    double v = base + Eb[ib];
    while (x) {
      int j = __builtin_ctzll(x);
      x &= x - 1;
      v += J[j];
    }
    hii[ia * nb + ib] = v;
  }
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void build_table(const uint64_t* d_str, int64_t n, int L, const double* d_h1, const double* d_eri,
                 Table& ts, Table& td, hipStream_t st) {
  ts.n = td.n = n;
  DBuf<int64_t> cs, cd;
  cs.alloc(n);
  cd.alloc(n);
  ts.rowptr.alloc(n + 1);
  td.rowptr.alloc(n + 1);
  if (n > 0) {
    hipLaunchKernelGGL(k_count, dim3((unsigned)n), dim3(TB), 0, st, d_str, n, cs.p, cd.p);
    SBDG_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cs.p, n, ts.rowptr.p);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cd.p, n, td.rowptr.p);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipMemcpyAsync(&ts.nnz, ts.rowptr.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SBDG_CHECK(hipMemcpyAsync(&td.nnz, td.rowptr.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SBDG_CHECK(hipStreamSynchronize(st));
  if (ts.nnz >= (int64_t)INT32_MAX || td.nnz >= (int64_t)INT32_MAX)
    throw Error(SBDG_ERR_LIMIT, "excitation table beyond 2^31 entries");
  for (Table* t : {&ts, &td}) {
    t->col.alloc(std::max<int64_t>(t->nnz, 1));
    t->exc.alloc(std::max<int64_t>(t->nnz, 1));
    t->val.alloc(std::max<int64_t>(t->nnz, 1));
    t->val32.alloc(std::max<int64_t>(t->nnz, 1));
  }
  if (n > 0) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)n), dim3(TB), 0, st, d_str, n, ts.rowptr.p, td.rowptr.p,
                       ts.col.p, td.col.p);
    hipLaunchKernelGGL(k_single, dim3((unsigned)n), dim3(TB), 0, st, d_str, L, d_h1, d_eri, ts.rowptr.p,
                       ts.col.p, ts.exc.p, ts.val.p, ts.val32.p);
    hipLaunchKernelGGL(k_double, dim3((unsigned)n), dim3(TB), 0, st, d_str, L, d_eri, td.rowptr.p, td.col.p,
                       td.exc.p, td.val.p, td.val32.p);
    SBDG_CHECK(hipGetLastError());
  }
  SBDG_CHECK(hipStreamSynchronize(st));
}

void build_spmm(const Table& s, const Table& d, SpmmTable& m, int L, hipStream_t st, bool with_doubles) {
  if (!with_doubles) {  // singles only: copy the singles CSR
    m.n = s.n;
    m.nnz = s.nnz;
    m.rowptr.alloc(m.n + 1);
    m.col.alloc(std::max<int64_t>(m.nnz, 1));
    m.val.alloc(std::max<int64_t>(m.nnz, 1));
    m.val32.alloc(std::max<int64_t>(m.nnz, 1));
    m.jai.alloc(std::max<int64_t>(m.nnz, 1));
    SBDG_CHECK(hipMemcpyAsync(m.rowptr.p, s.rowptr.p, sizeof(int64_t) * (m.n + 1), hipMemcpyDeviceToDevice, st));
    if (m.nnz) {
      SBDG_CHECK(hipMemcpyAsync(m.col.p, s.col.p, sizeof(int32_t) * m.nnz, hipMemcpyDeviceToDevice, st));
      SBDG_CHECK(hipMemcpyAsync(m.val.p, s.val.p, sizeof(double) * m.nnz, hipMemcpyDeviceToDevice, st));
      SBDG_CHECK(hipMemcpyAsync(m.val32.p, s.val32.p, sizeof(float) * m.nnz, hipMemcpyDeviceToDevice, st));
      hipLaunchKernelGGL(k_jai, dim3(blocks_for(m.nnz, 256)), dim3(256), 0, st, m.nnz, L, s.exc.p, m.jai.p);
      SBDG_CHECK(hipGetLastError());
    }
    SBDG_CHECK(hipStreamSynchronize(st));
    return;
  }
  m.n = s.n;
  m.nnz = s.nnz + d.nnz;
  m.rowptr.alloc(m.n + 1);
  m.col.alloc(std::max<int64_t>(m.nnz, 1));
  m.val.alloc(std::max<int64_t>(m.nnz, 1));
  m.val32.alloc(std::max<int64_t>(m.nnz, 1));
  m.jai.alloc(std::max<int64_t>(m.nnz, 1));
  if (m.n > 0) {
    hipLaunchKernelGGL(k_merge, dim3((unsigned)m.n), dim3(TB), 0, st, m.n, L, s.rowptr.p, s.col.p, s.val.p,
                       s.val32.p, s.exc.p, d.rowptr.p, d.col.p, d.val.p, d.val32.p, m.rowptr.p, m.col.p, m.val.p,
                       m.val32.p, m.jai.p);
    SBDG_CHECK(hipGetLastError());
  } else {
    SBDG_CHECK(hipMemsetAsync(m.rowptr.p, 0, sizeof(int64_t), st));
  }
}

void build_ab(const Table& sb, int L, int64_t chunk, AbTable& ab, hipStream_t st) {
  if (sb.n > (1 << 18)) throw Error(SBDG_ERR_LIMIT, "beta string set beyond 2^18 strings");
  ab.chunk = chunk;
  ab.nch = (int)((sb.n + chunk - 1) / chunk);
  if (ab.nch < 1) ab.nch = 1;
  ab.pack.alloc(std::max<int64_t>(sb.nnz, 1));
  ab.cptr.alloc((size_t)sb.n * (ab.nch + 1) + 1);
  if (sb.nnz > 0)
    hipLaunchKernelGGL(k_abpack, dim3(blocks_for(sb.nnz, 256)), dim3(256), 0, st, sb.nnz, L, sb.col.p,
                       sb.exc.p, ab.pack.p);
  const int64_t nt = sb.n * (ab.nch + 1);
  if (nt > 0)
    hipLaunchKernelGGL(k_abchunk, dim3(blocks_for(nt, 256)), dim3(256), 0, st, sb.n, ab.nch, chunk,
                       sb.rowptr.p, sb.col.p, ab.cptr.p);
  SBDG_CHECK(hipGetLastError());
}

// Alpha-beta launch geometry: nt threads, m columns per thread, so that one workgroup
// covers a whole row (the ket rows and integral slices are then staged once per bra row,
// not once per column tile).  The FP64 accumulators bound m: 1024 threads (four waves per
// SIMD, 128 VGPRs) hold up to 24 columns each; beyond 24,576 columns FP64 takes 768 threads
// (three waves per SIMD, 168 VGPRs) with up to 53 columns each.  FP32 holds up to 48
// columns at 1024 threads.  Wider rows take several column tiles at the largest m.
void ab_geometry(int64_t nb, int elem, int& nt, int& m) {
  nt = (int)std::min<int64_t>(1024, std::max<int64_t>(64, (nb + 63) / 64 * 64));
  static const int m1024_64[] = {1, 2, 4, 8, 12, 16, 20, 24};
  static const int m768_64[] = {32, 40, 48, 53};
  static const int m1024_32[] = {1, 2, 4, 8, 12, 16, 20, 24, 40, 48};
  const int* ms = elem == 4 ? m1024_32 : m1024_64;
  int nms = elem == 4 ? 10 : 8;
  int64_t need = (nb + nt - 1) / nt;
  if (elem != 4 && need > 24) {
    nt = 768;
    need = (nb + nt - 1) / nt;
    ms = m768_64;
    nms = 4;
  }
  m = ms[nms - 1];
  for (int i = 0; i < nms; ++i)
    if (need <= ms[i]) {
      m = ms[i];
      break;
    }
}

// Order of the entries of the lanes of one (group, chunk) block, step by step, so that the
// lanes of each aligned pass group (P lanes) read distinct LDS slots (S slots of the
// element size) in both the ket-row array and the integral array; equal addresses
// broadcast and cost nothing.  Padding (the zero slot, pair 0) may be placed at any step
// while a lane has slack.  Greedy: at each step the lanes with least slack choose first,
// each taking the entry that adds the fewest new addresses to slots already in use.
// lists[l] holds lane l's entries (ket byte address | pair << 18 | sign << 31); out gets
// K*64.  Slots count 16-byte LDS vectors: the pair's slice vector is the pair index, the
// ket's is its byte address / 16.
static void schedule_block(std::vector<std::vector<uint32_t>>& lists, int K, int P, int S, uint32_t pad,
                           uint32_t* out) {
  struct Use {
    uint32_t addr[8];
    int n;
  };
  std::vector<int> rem(64), order(P);
  for (int l = 0; l < 64; ++l) rem[l] = (int)lists[l].size();
  auto cost1 = [](const Use& u, uint32_t a) {
    for (int i = 0; i < u.n; ++i)
      if (u.addr[i] == a) return 0;
    return u.n;
  };
  auto add1 = [](Use& u, uint32_t a) {
    for (int i = 0; i < u.n; ++i)
      if (u.addr[i] == a) return;
    if (u.n < 8) u.addr[u.n++] = a;
  };
  std::vector<Use> uc(S), uv(S);
  for (int p0 = 0; p0 < 64; p0 += P) {
    for (int e = 0; e < K; ++e) {
      for (int x = 0; x < S; ++x) uc[x].n = uv[x].n = 0;
      for (int i = 0; i < P; ++i) order[i] = p0 + i;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rem[a] > rem[b]; });
      for (int l : order) {
        const int slack = (K - e) - rem[l];
        auto& ls = lists[l];
        int best = -1, bc = 1 << 30;
        for (int j = 0; j < rem[l]; ++j) {
          const uint32_t w = ls[j];
          const uint32_t ca = (w & 0x3FFFFu) >> 4;
          const uint32_t va = (w >> 18) & 0xFFFu;
          const int cst = cost1(uc[ca % S], ca) + cost1(uv[va % S], va);
          if (cst < bc) {
            bc = cst;
            best = j;
            if (cst == 0) break;
          }
        }
        uint32_t w = pad;
        if (slack > 0) {
          const uint32_t ca = (pad & 0x3FFFFu) >> 4, va = (pad >> 18) & 0xFFFu;
          const int cst = cost1(uc[ca % S], ca) + cost1(uv[va % S], va);
          if (best < 0 || cst < bc) best = -1;
        }
        if (best >= 0) {
          w = ls[best];
          ls[best] = ls[rem[l] - 1];  // order of the rest does not matter
          --rem[l];
        }
        const uint32_t ca = (w & 0x3FFFFu) >> 4, va = (w >> 18) & 0xFFFu;
        add1(uc[ca % S], ca);
        add1(uv[va % S], va);
        out[(size_t)e * 64 + l] = w;
      }
    }
  }
}

void build_ell(const Table& sb, int L, int64_t chunk, int prec_bytes, EllTable& e, hipStream_t st) {
  const int64_t nb = sb.n;
  if (L > 64) throw Error(SBDG_ERR_LIMIT, "norb beyond 64");
  e.kbase = ab_slice_span(L);
  if (16 * (e.kbase + chunk + 1) > (1 << 18)) throw Error(SBDG_ERR_LIMIT, "alpha-beta LDS beyond 256 KiB");
  e.chunk = chunk;
  e.nch = (int)std::max<int64_t>(1, (nb + chunk - 1) / chunk);
  ab_geometry(nb, prec_bytes, e.nt, e.m);
  const int64_t tw = (int64_t)e.nt * e.m;
  e.ntiles = std::max<int64_t>(1, (nb + tw - 1) / tw);
  const int64_t npos = e.ntiles * tw;
  e.groups = npos / 64;
  const int nch = e.nch;
  const int64_t ng = e.groups * nch;
  // host copies of the beta singles
  SBDG_CHECK(hipStreamSynchronize(st));
  std::vector<int64_t> rp(nb + 1);
  SBDG_CHECK(hipMemcpy(rp.data(), sb.rowptr.p, sizeof(int64_t) * (nb + 1), hipMemcpyDeviceToHost));
  const int64_t nnz = rp[nb];
  std::vector<int32_t> col(nnz);
  std::vector<uint32_t> exc(nnz);
  if (nnz) {
    SBDG_CHECK(hipMemcpy(col.data(), sb.col.p, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    SBDG_CHECK(hipMemcpy(exc.data(), sb.exc.p, sizeof(uint32_t) * nnz, hipMemcpyDeviceToHost));
  }
  // entries of column ib in chunk k: its singles with ket in the chunk
  std::vector<int32_t> cnt((size_t)nb * nch, 0);
  for (int64_t ib = 0; ib < nb; ++ib)
    for (int64_t q = rp[ib]; q < rp[ib + 1]; ++q) ++cnt[(size_t)ib * nch + col[q] / chunk];
  // column tiles: columns by decreasing total count, cut into tiles of tw lane positions;
  // inside a tile, each chunk orders its columns by decreasing count in that chunk, so the
  // 64 lanes of a group carry similar lists (little padding).  A column stays in its tile
  // for every chunk: one workgroup owns it.
  std::vector<int32_t> order(nb);
  for (int64_t i = 0; i < nb; ++i) order[i] = (int32_t)i;
  std::vector<int32_t> tot(nb, 0);
  for (int64_t ib = 0; ib < nb; ++ib)
    for (int k = 0; k < nch; ++k) tot[ib] += cnt[(size_t)ib * nch + k];
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return tot[x] > tot[y]; });
  std::vector<int32_t> perm((size_t)std::max<int64_t>(npos * nch, 1), -1);
  for (int k = 0; k < nch; ++k) {
    for (int64_t t = 0; t < e.ntiles; ++t) {
      const int64_t lo = t * tw, hi = std::min<int64_t>(nb, lo + tw);
      if (lo >= hi) break;
      std::vector<int32_t> cols(order.begin() + lo, order.begin() + hi);
      std::stable_sort(cols.begin(), cols.end(), [&](int32_t x, int32_t y) {
        return cnt[(size_t)x * nch + k] > cnt[(size_t)y * nch + k];
      });
      // groups of 64 columns in count order; each goes to the wave with the fewest rows so
      // far that still has a free slot (the kernel walks slot j of wave w as group w + j*W)
      const int Wv = e.nt / 64;
      std::vector<int> load(Wv, 0), used(Wv, 0);
      for (int j = 0; j < Wv * e.m; ++j) {
        const size_t i0 = (size_t)j * 64;
        const int mx = i0 < cols.size() ? cnt[(size_t)cols[i0] * nch + k] : 0;  // first lane holds the most
        int best = -1;
        for (int w = 0; w < Wv; ++w)
          if (used[w] < e.m && (best < 0 || load[w] < load[best])) best = w;
        const int64_t slot = (int64_t)best + (int64_t)used[best] * Wv;
        ++used[best];
        load[best] += (mx + 3) & ~3;
        for (size_t l = 0; l < 64 && i0 + l < cols.size(); ++l) perm[(size_t)k * npos + lo + slot * 64 + l] = cols[i0 + l];
      }
    }
  }
  // rows of each (group, chunk) block: the longest list of its lanes, rounded up to four
  std::vector<int32_t> K((size_t)std::max<int64_t>(ng, 1), 0);
  for (int64_t g = 0; g < e.groups; ++g)
    for (int k = 0; k < nch; ++k) {
      int mx = 0;
      for (int l = 0; l < 64; ++l) {
        const int32_t ib = perm[(size_t)k * npos + g * 64 + l];
        if (ib >= 0) mx = std::max(mx, cnt[(size_t)ib * nch + k]);
      }
      K[g * nch + k] = (mx + 3) & ~3;
    }
  // one stream per (tile, wave, chunk): the wave's groups j = 0..m-1 one after another
  const int W = e.nt / 64;
  const int64_t nstream = e.ntiles * W * nch;
  std::vector<int64_t> soff((size_t)nstream + 1, 0), off((size_t)std::max<int64_t>(ng, 1), 0);
  std::vector<int32_t> nq((size_t)std::max<int64_t>(nstream * e.m, 1), 0);
  for (int64_t t = 0; t < e.ntiles; ++t)
    for (int w = 0; w < W; ++w)
      for (int k = 0; k < nch; ++k) {
        const int64_t s = (t * W + w) * nch + k;
        int64_t o = soff[s];
        for (int j = 0; j < e.m; ++j) {
          const int64_t g = t * W * e.m + w + (int64_t)j * W;
          off[g * nch + k] = o;
          nq[s * e.m + j] = K[g * nch + k] / 4;
          o += (int64_t)K[g * nch + k] * 64;
        }
        soff[s + 1] = o;
      }
  const int64_t total = soff[nstream];
  const uint32_t pad = ((uint32_t)e.kbase + (uint32_t)chunk) << 4;  // zero slot, pair 0
  std::vector<uint32_t> ent((size_t)std::max<int64_t>(total, 1), pad);
  // pass groups of the kernel's 16-byte LDS reads, as measured on gfx950 (the alpha-beta
  // kernel at norb=36, 20,000^2 with four candidate models; this one was fastest): 32 lanes
  // over 16 slots of 16 bytes (256 bytes of banks)
  const int P = 32, S = 16;
  const uint32_t kbase = (uint32_t)e.kbase;  // ket vectors follow the slice vectors
  std::vector<std::vector<uint32_t>> lists(64);
  for (int64_t g = 0; g < e.groups; ++g)
    for (int k = 0; k < nch; ++k) {
      const int64_t j0 = (int64_t)k * chunk;
      const int64_t base = off[g * nch + k];
      for (int l = 0; l < 64; ++l) {
        auto& lst = lists[l];
        lst.clear();
        const int32_t ib = perm[(size_t)k * npos + g * 64 + l];
        if (ib < 0) continue;
        for (int64_t q = rp[ib]; q < rp[ib + 1]; ++q) {
          if (col[q] < j0 || col[q] >= j0 + chunk) continue;
          const uint32_t x = exc[q];
          const uint32_t j = x & 63u, b = (x >> 6) & 63u, neg = (x >> 24) & 1u;
          lst.push_back(((kbase + (uint32_t)(col[q] - j0)) << 4) | (ab_pair(b, j) << 18) | (neg << 31));
        }
      }
      schedule_block(lists, K[g * nch + k], P, S, pad, ent.data() + base);
    }
  e.colperm.alloc(perm.size());
  SBDG_CHECK(hipMemcpy(e.colperm.p, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice));
  e.nq.alloc(nq.size());
  SBDG_CHECK(hipMemcpy(e.nq.p, nq.data(), sizeof(int32_t) * nq.size(), hipMemcpyHostToDevice));
  e.soff.alloc(soff.size());
  SBDG_CHECK(hipMemcpy(e.soff.p, soff.data(), sizeof(int64_t) * soff.size(), hipMemcpyHostToDevice));
  e.ent.alloc(ent.size());
  SBDG_CHECK(hipMemcpy(e.ent.p, ent.data(), sizeof(uint32_t) * ent.size(), hipMemcpyHostToDevice));
}

void build_cores(const uint64_t* d_str, int64_t n, int L, int nel, const double* h_eri, CoreTable& ct,
                 Table& direct, hipStream_t st) {
  ct.ncore = ct.nmem = ct.nblk = 0;
  ct.maxm = 0;
  const int64_t npp = (int64_t)nel * (nel - 1) / 2;
  const int64_t total = n * npp;
  std::vector<int32_t> start(1, 0), row, pair, blk, blk_x0, grp, slot_ptr(n + 1, 0), slot;
  // direct pairs of the small cores: (row, column, value), value eps_x eps_y W[P_x][P_y]
  struct Direct {
    int32_t r, c;
    double v;
  };
  std::vector<Direct> dp;
  auto unpair = [](int P, int& lo, int& hi) {
    int b = 1;
    while ((b + 1) * b / 2 <= P) ++b;
    hi = b;
    lo = P - b * (b - 1) / 2;
  };
  auto eri4h = [&](int p, int q, int r, int t) { return h_eri[(((int64_t)p * L + q) * L + r) * L + t]; };
  if (total > 0) {
    if (total >= (int64_t)INT32_MAX) throw Error(SBDG_ERR_LIMIT, "too many core memberships");
    DBuf<uint64_t> key, val, key2, val2;
    key.alloc(total);
    val.alloc(total);
    key2.alloc(total);
    val2.alloc(total);
    hipLaunchKernelGGL(k_core_enum, dim3(blocks_for(total, 256)), dim3(256), 0, st, d_str, n, nel, key.p, val.p);
    SBDG_CHECK(hipGetLastError());
    size_t tmp = 0;
    SBDG_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, key.p, key2.p, val.p, val2.p, (int)total, 0,
                                                  std::max(1, L), st));
    DBuf<unsigned char> tb;
    tb.alloc(tmp + 1);
    SBDG_CHECK(hipcub::DeviceRadixSort::SortPairs(tb.p, tmp, key.p, key2.p, val.p, val2.p, (int)total, 0,
                                                  std::max(1, L), st));
    std::vector<uint64_t> hk(total), hv(total);
    SBDG_CHECK(hipMemcpyAsync(hk.data(), key2.p, sizeof(uint64_t) * total, hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipMemcpyAsync(hv.data(), val2.p, sizeof(uint64_t) * total, hipMemcpyDeviceToHost, st));
    SBDG_CHECK(hipStreamSynchronize(st));
    // keep, core by core, the members whose pair is disjoint from some other member's pair:
    // any other member is a single excitation away (or the same string), contributes
    // nothing and receives nothing.  Radix sort is stable: members ascend by string.
    auto pair_mask = [](uint64_t v) {
      const int P = (int)((v >> 32) & 0x7FFFFFFFull);
      int b = 1;
      while ((b + 1) * b / 2 <= P) ++b;
      const int a = P - b * (b - 1) / 2;
      return (1ull << a) | (1ull << b);
    };
    std::vector<uint64_t> masks;
    std::vector<int64_t> keep;
    for (int64_t e = 0; e < total;) {
      int64_t f = e + 1;
      while (f < total && hk[f] == hk[e]) ++f;
      masks.clear();
      keep.clear();
      for (int64_t q = e; q < f; ++q) masks.push_back(pair_mask(hv[q]));
      for (int64_t x = 0; x < f - e; ++x)
        for (int64_t y = 0; y < f - e; ++y)
          if ((masks[x] & masks[y]) == 0) {
            keep.push_back(e + x);
            break;
          }
      const int64_t m = (int64_t)keep.size();
      if (m >= 2 && m < CORE_KMIN) {
        for (int64_t x = 0; x < m; ++x)
          for (int64_t y = 0; y < m; ++y) {
            const uint64_t vx = hv[keep[x]], vy = hv[keep[y]];
            if (x == y || (pair_mask(vx) & pair_mask(vy))) continue;  // W = 0
            int A, B, I, J;
            unpair((int)((vx >> 32) & 0x7FFFFFFFull), A, B);
            unpair((int)((vy >> 32) & 0x7FFFFFFFull), I, J);
            double v = eri4h(A, I, B, J) - eri4h(A, J, B, I);
            if ((vx >> 63) != (vy >> 63)) v = -v;
            dp.push_back({(int32_t)(vx & 0xFFFFFFFFull), (int32_t)(vy & 0xFFFFFFFFull), v});
          }
      } else if (m >= 2) {
        const int32_t c = (int32_t)(start.size() - 1);
        for (int64_t x0 = 0; x0 < m; x0 += CORE_XB) {
          blk.push_back(c);
          blk_x0.push_back((int32_t)x0);
        }
        for (int64_t q : keep) {
          row.push_back((int32_t)(hv[q] & 0xFFFFFFFFull));
          pair.push_back((int32_t)((hv[q] >> 32) & 0x7FFFFFFFull) | (int32_t)((hv[q] >> 63) << 31));
        }
        start.push_back((int32_t)row.size());
        ct.maxm = std::max<int>(ct.maxm, (int)m);
      }
      e = f;
    }
    // workgroups: consecutive product blocks until about CORE_WORK member rows
    grp.push_back(0);
    int64_t work = 0;
    for (size_t q = 0; q < blk.size(); ++q) {
      work += start[blk[q] + 1] - start[blk[q]];
      if (work >= CORE_WORK || q + 1 == blk.size()) {
        grp.push_back((int32_t)(q + 1));
        work = 0;
      }
    }
    // per-string slots (counting sort by row)
    for (int32_t r : row) ++slot_ptr[r + 1];
    for (int64_t r = 0; r < n; ++r) slot_ptr[r + 1] += slot_ptr[r];
    slot.resize(row.size());
    std::vector<int32_t> fillp(slot_ptr.begin(), slot_ptr.end() - 1);
    for (size_t q = 0; q < row.size(); ++q) slot[fillp[row[q]]++] = (int32_t)q;
  }
  // the direct pairs as a CSR table (rows ascending; any order inside a row)
  {
    std::vector<int64_t> rp(n + 1, 0);
    for (const Direct& d : dp) ++rp[d.r + 1];
    for (int64_t r = 0; r < n; ++r) rp[r + 1] += rp[r];
    std::vector<int32_t> dc(dp.size());
    std::vector<double> dv(dp.size());
    std::vector<float> dv32(dp.size());
    std::vector<int64_t> fill(rp.begin(), rp.end() - 1);
    for (const Direct& d : dp) {
      const int64_t q = fill[d.r]++;
      dc[q] = d.c;
      dv[q] = d.v;
      dv32[q] = (float)d.v;
    }
    direct.n = n;
    direct.nnz = (int64_t)dp.size();
    direct.rowptr.alloc(n + 1);
    direct.col.alloc(std::max<size_t>(dp.size(), 1));
    direct.exc.alloc(1);
    direct.val.alloc(std::max<size_t>(dp.size(), 1));
    direct.val32.alloc(std::max<size_t>(dp.size(), 1));
    SBDG_CHECK(hipMemcpy(direct.rowptr.p, rp.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
    if (!dp.empty()) {
      SBDG_CHECK(hipMemcpy(direct.col.p, dc.data(), sizeof(int32_t) * dp.size(), hipMemcpyHostToDevice));
      SBDG_CHECK(hipMemcpy(direct.val.p, dv.data(), sizeof(double) * dp.size(), hipMemcpyHostToDevice));
      SBDG_CHECK(hipMemcpy(direct.val32.p, dv32.data(), sizeof(float) * dp.size(), hipMemcpyHostToDevice));
    }
  }
  ct.ncore = (int64_t)start.size() - 1;
  ct.nmem = (int64_t)row.size();
  ct.nblk = (int64_t)blk.size();
  ct.ngrp = grp.empty() ? 0 : (int64_t)grp.size() - 1;
  auto up = [&](DBuf<int32_t>& d, const std::vector<int32_t>& h) {
    d.alloc(std::max<size_t>(h.size(), 1));
    if (!h.empty()) SBDG_CHECK(hipMemcpy(d.p, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
  };
  up(ct.start, start);
  up(ct.row, row);
  up(ct.pair, pair);
  up(ct.blk, blk);
  up(ct.blk_x0, blk_x0);
  up(ct.slot_ptr, slot_ptr);
  up(ct.slot, slot);
  up(ct.grp, grp);
}

void build_occ_table(const uint64_t* d_str, int64_t n, int L, const double* d_eri, double* J, float* J32,
                     hipStream_t st) {
  const int64_t total = ((int64_t)L * L + 1) * n;
  if (total == 0) return;
  hipLaunchKernelGGL(k_occ_table, dim3(blocks_for(total, 256)), dim3(256), 0, st, d_str, n, L, d_eri, J, J32);
  SBDG_CHECK(hipGetLastError());
}

void build_pair_integrals(const double* d_eri, int L, double* W, float* W32, hipStream_t st) {
  const int64_t npair = (int64_t)L * (L - 1) / 2;
  if (npair == 0) return;
  hipLaunchKernelGGL(k_pair_int, dim3(blocks_for(npair * npair, 256)), dim3(256), 0, st, d_eri, L, npair, W, W32);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipStreamSynchronize(st));
}

void compute_hii(const uint64_t* d_a, int64_t na, const uint64_t* d_b, int64_t nb, int L, double core,
                 const double* d_h1, const double* d_eri, double* d_hii, hipStream_t st) {
  DBuf<double> Ea, Eb;
  Ea.alloc(std::max<int64_t>(na, 1));
  Eb.alloc(std::max<int64_t>(nb, 1));
  if (na > 0) hipLaunchKernelGGL(k_strdiag, dim3(blocks_for(na, 256)), dim3(256), 0, st, d_a, na, L, d_h1, d_eri, Ea.p);
  if (nb > 0) hipLaunchKernelGGL(k_strdiag, dim3(blocks_for(nb, 256)), dim3(256), 0, st, d_b, nb, L, d_h1, d_eri, Eb.p);
  if (na > 0 && nb > 0)
    hipLaunchKernelGGL(k_hii, dim3((unsigned)na), dim3(TB), 0, st, d_a, d_b, nb, L, core, Ea.p, Eb.p, d_eri, d_hii);
  SBDG_CHECK(hipGetLastError());
  SBDG_CHECK(hipStreamSynchronize(st));
}

}  // namespace sbdg
