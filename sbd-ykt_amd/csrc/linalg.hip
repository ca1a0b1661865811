// Vector kernels of the device-resident Davidson (tpb/davidson.h:661-924 with the
// BLAS-1 of framework/dm_vector.h:20-84) and small helpers.
//
// Reductions are deterministic: each block reduces a fixed strided set of elements in a
// fixed tree order into a partial, and one block sums the partials in index order.
// Basis vectors live in one allocation with a leading dimension, V_k = V + k*ld.
#include <algorithm>

#include "sbdg_internal.h"

namespace sbdg {

namespace {

constexpr int RT = 256;   // reduction block
constexpr int RB = 1024;  // reduction blocks (upper bound; the partial-sum buffer holds RB rows)

inline unsigned nblocks(int64_t n) {
  int64_t b = (n + RT * 4 - 1) / (RT * 4);
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(RB, b));
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block tree sum of v (all threads receive it); sh has RT/64 doubles
__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  return t;
}

// VEC consecutive elements at p as doubles (VEC == 2 or 4: p is VEC-element aligned;
// 16-byte loads for FP64, 8- or 16-byte for FP32), and the store back
template <int VEC, class T>
__device__ inline void ld_vec(const T* __restrict__ p, double (&o)[VEC]) {
  if constexpr (VEC == 4) {
    if constexpr (sizeof(T) == 8) {
      const double2 a = *reinterpret_cast<const double2*>(p);
      const double2 b = *reinterpret_cast<const double2*>(p + 2);
      o[0] = a.x;
      o[1] = a.y;
      o[2] = b.x;
      o[3] = b.y;
    } else {
      const float4 v = *reinterpret_cast<const float4*>(p);
      o[0] = (double)v.x;
      o[1] = (double)v.y;
      o[2] = (double)v.z;
      o[3] = (double)v.w;
    }
  } else if constexpr (VEC == 2) {
    if constexpr (sizeof(T) == 8) {
      const double2 v = *reinterpret_cast<const double2*>(p);
      o[0] = v.x;
      o[1] = v.y;
    } else {
      const float2 v = *reinterpret_cast<const float2*>(p);
      o[0] = (double)v.x;
      o[1] = (double)v.y;
    }
  } else {
    o[0] = (double)p[0];
  }
}
template <int VEC, class T>
__device__ inline void st_vec(T* __restrict__ p, const double (&v)[VEC]) {
  if constexpr (VEC == 4) {
    if constexpr (sizeof(T) == 8) {
      *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
      *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
    } else {
      *reinterpret_cast<float4*>(p) = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
    }
  } else if constexpr (VEC == 2) {
    if constexpr (sizeof(T) == 8) *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    else *reinterpret_cast<float2*>(p) = make_float2((float)v[0], (float)v[1]);
  } else {
    p[0] = (T)v[0];
  }
}

// partial[blk*m + k] -> out[k], one block, fixed order
__global__ void __launch_bounds__(RT) k_final(const double* __restrict__ part, int nblk, int m,
                                              double* __restrict__ out) {
  __shared__ double sh[RT / 64];
  for (int k = 0; k < m; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += RT) v += part[(int64_t)b * m + k];
    v = block_sum(v, sh);
    if (threadIdx.x == 0) out[k] = v;
  }
}

// in may equal out (each element is read, then written, by one thread)
__global__ void k_sign(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, int64_t nb,
                       int64_t N, const double* in, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const int64_t ia = t / nb, ib = t - ia * nb;
  uint64_t A = a[ia];
  const uint64_t B = b[ib];
  int par = 0;
  while (A) {
    const int p = __builtin_ctzll(A);
    A &= A - 1;
    par += __popcll(B & ((1ull << p) - 1ull));
  }
  out[t] = (par & 1) ? -in[t] : in[t];
}

// deterministic pseudo-random values in [-1, 1) (splitmix64 of the index)
__global__ void k_fill(double* __restrict__ x, int64_t n, uint64_t seed) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = (uint64_t)t + seed * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[t] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
  }
}

__global__ void k_d2f(const double* __restrict__ in, float* __restrict__ out, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (float)in[t];
}
__global__ void k_f2d(const float* __restrict__ in, double* __restrict__ out, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    out[t] = (double)in[t];
}

template <class TY, int MM, int VEC>
__global__ void __launch_bounds__(RT) k_mdot(const double* __restrict__ X, int64_t ldx, int m,
                                             const TY* __restrict__ y, int64_t n,
                                             double* __restrict__ part) {
  __shared__ double sh[RT / 64];
  double acc[MM];
#pragma unroll
  for (int k = 0; k < MM; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t u = (int64_t)blockIdx.x * RT + threadIdx.x; u < n / VEC; u += stride) {
    const int64_t i = u * VEC;
    double yv[VEC];
    ld_vec<VEC>(y + i, yv);
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double xv[VEC];
        ld_vec<VEC>(X + k * ldx + i, xv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[k] += xv[e] * yv[e];
      }
  }
#pragma unroll
  for (int k = 0; k < MM; ++k) {
    if (k < m) {
      double v = block_sum(acc[k], sh);
      if (threadIdx.x == 0) part[(int64_t)blockIdx.x * m + k] = v;
    }
  }
}

// out partials of <Y_k, x> (k < m, Y FP32) and, with Z, of <z, x>: x is read once
template <int MM, int VEC, bool Z>
__global__ void __launch_bounds__(RT) k_mdot_rows(const float* __restrict__ Y, int64_t ldy, int m,
                                                  const double* __restrict__ x, const double* __restrict__ z,
                                                  int64_t n, double* __restrict__ part) {
  __shared__ double sh[RT / 64];
  double acc[MM], az = 0.0;
#pragma unroll
  for (int k = 0; k < MM; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t u = (int64_t)blockIdx.x * RT + threadIdx.x; u < n / VEC; u += stride) {
    const int64_t i = u * VEC;
    double xv[VEC];
    ld_vec<VEC>(x + i, xv);
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double yv[VEC];
        ld_vec<VEC>(Y + k * ldy + i, yv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) acc[k] += yv[e] * xv[e];
      }
    if constexpr (Z) {
      double zv[VEC];
      ld_vec<VEC>(z + i, zv);
#pragma unroll
      for (int e = 0; e < VEC; ++e) az += zv[e] * xv[e];
    }
  }
  const int w = m + (Z ? 1 : 0);
#pragma unroll
  for (int k = 0; k < MM; ++k) {
    if (k < m) {
      const double v = block_sum(acc[k], sh);
      if (threadIdx.x == 0) part[(int64_t)blockIdx.x * w + k] = v;
    }
  }
  if constexpr (Z) {
    const double v = block_sum(az, sh);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * w + m] = v;
  }
}

template <class TH, int MM, int VEC>
__global__ void __launch_bounds__(RT) k_ritz(const double* __restrict__ V, int64_t ldv,
                                             const TH* __restrict__ HV, int64_t ldh, int m,
                                             const double* __restrict__ u,
                                             const double* __restrict__ theta,
                                             double* __restrict__ W, double* __restrict__ R, int64_t n,
                                             double* __restrict__ part) {
  __shared__ double sh[RT / 64];
  double uk[MM];
#pragma unroll
  for (int k = 0; k < MM; ++k) uk[k] = (k < m) ? u[k] : 0.0;
  const double th = theta[0];
  double sw = 0.0, sr = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t q = (int64_t)blockIdx.x * RT + threadIdx.x; q < n / VEC; q += stride) {
    const int64_t i = q * VEC;
    double w[VEC], h[VEC], r[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) w[e] = h[e] = 0.0;
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double v[VEC], hv[VEC];
        ld_vec<VEC>(V + k * ldv + i, v);
        ld_vec<VEC>(HV + k * ldh + i, hv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          w[e] += uk[k] * v[e];
          h[e] += uk[k] * hv[e];
        }
      }
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      r[e] = h[e] - th * w[e];
      sw += w[e] * w[e];
      sr += r[e] * r[e];
    }
    st_vec<VEC>(W + i, w);
    st_vec<VEC>(R + i, r);
  }
  sw = block_sum(sw, sh);
  sr = block_sum(sr, sh);
  if (threadIdx.x == 0) {
    part[(int64_t)blockIdx.x * 2] = sw;
    part[(int64_t)blockIdx.x * 2 + 1] = sr;
  }
}

// One pass over the basis for a step that continues without restart: with the Ritz pair
// (theta, u) of the projected matrix, r = sum_k u_k HV_k - theta sum_k u_k V_k,
// t = r / (theta - hii) (regularised as k_precond; the missing 1/|r| is a common factor
// that the normalisation after Gram-Schmidt removes) written to T, and the partials of
// V_k . t (k < m), |r|^2 and |t|^2.  The Ritz vector itself is not formed.
template <class TH, int MM, int VEC>
__global__ void __launch_bounds__(RT) k_ritz_expand(const double* __restrict__ V, int64_t ldv,
                                                    const TH* __restrict__ HV, int64_t ldh, int m,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ theta,
                                                    const double* __restrict__ hii, double* __restrict__ T,
                                                    int64_t n, double* __restrict__ part) {
  __shared__ double sh[RT / 64];
  double uk[MM], d[MM];
#pragma unroll
  for (int k = 0; k < MM; ++k) {
    uk[k] = (k < m) ? u[k] : 0.0;
    d[k] = 0.0;
  }
  const double th = theta[0];
  const double eps_reg = 1.0e-12;
  double sr = 0.0, st = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t q = (int64_t)blockIdx.x * RT + threadIdx.x; q < n / VEC; q += stride) {
    const int64_t i = q * VEC;
    double w[VEC], h[VEC], t[VEC], hd[VEC];
#pragma unroll
    for (int e = 0; e < VEC; ++e) w[e] = h[e] = 0.0;
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double v[VEC], hv[VEC];
        ld_vec<VEC>(V + k * ldv + i, v);
        ld_vec<VEC>(HV + k * ldh + i, hv);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          w[e] += uk[k] * v[e];
          h[e] += uk[k] * hv[e];
        }
      }
    ld_vec<VEC>(hii + i, hd);
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
      const double r = h[e] - th * w[e];
      const double den = th - hd[e];
      t[e] = (fabs(den) > eps_reg) ? r / den : r / (den - eps_reg);
      sr += r * r;
      st += t[e] * t[e];
    }
    st_vec<VEC>(T + i, t);
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double v[VEC];
        ld_vec<VEC>(V + k * ldv + i, v);  // the same elements as above: L2, usually registers
#pragma unroll
        for (int e = 0; e < VEC; ++e) d[k] += v[e] * t[e];
      }
  }
#pragma unroll
  for (int k = 0; k < MM; ++k) {
    if (k < m) {
      const double v = block_sum(d[k], sh);
      if (threadIdx.x == 0) part[(int64_t)blockIdx.x * (m + 2) + k] = v;
    }
  }
  sr = block_sum(sr, sh);
  st = block_sum(st, sh);
  if (threadIdx.x == 0) {
    part[(int64_t)blockIdx.x * (m + 2) + m] = sr;
    part[(int64_t)blockIdx.x * (m + 2) + m + 1] = st;
  }
}

template <int VEC>
__global__ void k_scale_cast(double* __restrict__ x, const double* __restrict__ n2, float* __restrict__ y,
                             int64_t n) {
  const double f = 1.0 / sqrt(n2[0]);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n / VEC; q += (int64_t)gridDim.x * blockDim.x) {
    double v[VEC];
    ld_vec<VEC>(x + q * VEC, v);
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] *= f;
    st_vec<VEC>(x + q * VEC, v);
    st_vec<VEC>(y + q * VEC, v);
  }
}

template <int VEC>
__global__ void k_scale(double* __restrict__ x, const double* __restrict__ n2, int64_t n) {
  const double f = 1.0 / sqrt(n2[0]);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n / VEC; q += (int64_t)gridDim.x * blockDim.x) {
    double v[VEC];
    ld_vec<VEC>(x + q * VEC, v);
#pragma unroll
    for (int e = 0; e < VEC; ++e) v[e] *= f;
    st_vec<VEC>(x + q * VEC, v);
  }
}

__global__ void k_precond(const double* __restrict__ R, const double* __restrict__ rn2,
                          const double* __restrict__ hii, const double* __restrict__ theta,
                          double* __restrict__ t, int64_t n) {
  const double f = 1.0 / sqrt(rn2[0]);
  const double th = theta[0];
  const double eps_reg = 1.0e-12;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double den = th - hii[i];
    const double r = R[i] * f;
    t[i] = (fabs(den) > eps_reg) ? r / den : r / (den - eps_reg);
  }
}

template <int MM, int VEC>
__global__ void __launch_bounds__(RT) k_gs(const double* __restrict__ V, int64_t ldv, int m,
                                           const double* __restrict__ s, double* __restrict__ t, int64_t n,
                                           double* __restrict__ part) {
  __shared__ double sh[RT / 64];
  double sk[MM];
#pragma unroll
  for (int k = 0; k < MM; ++k) sk[k] = (k < m) ? s[k] : 0.0;
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * RT;
  for (int64_t q = (int64_t)blockIdx.x * RT + threadIdx.x; q < n / VEC; q += stride) {
    const int64_t i = q * VEC;
    double v[VEC];
    ld_vec<VEC>(t + i, v);
#pragma unroll
    for (int k = 0; k < MM; ++k)
      if (k < m) {
        double x[VEC];
        ld_vec<VEC>(V + k * ldv + i, x);
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] -= sk[k] * x[e];
      }
    st_vec<VEC>(t + i, v);
#pragma unroll
    for (int e = 0; e < VEC; ++e) acc += v[e] * v[e];
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

template <class T, int MM, int VEC>
__global__ void k_rotate(T* __restrict__ X, int64_t ldx, int m, const double* __restrict__ U, int ld, int mk,
                         int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n / VEC; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q * VEC;
    double x[MM][VEC];
#pragma unroll
    for (int k = 0; k < MM; ++k) {
      if (k < m) ld_vec<VEC>(X + k * ldx + i, x[k]);
      else
#pragma unroll
        for (int e = 0; e < VEC; ++e) x[k][e] = 0.0;
    }
    for (int j = 0; j < mk; ++j) {
      double y[VEC];
#pragma unroll
      for (int e = 0; e < VEC; ++e) y[e] = 0.0;
#pragma unroll
      for (int k = 0; k < MM; ++k)
        if (k < m)
#pragma unroll
          for (int e = 0; e < VEC; ++e) y[e] += U[k + ld * j] * x[k][e];
      st_vec<VEC>(X + j * ldx + i, y);
    }
  }
}

// cyclic Jacobi on one wave; A, V in LDS (m <= 64)
__global__ void __launch_bounds__(64) k_jacobi(const double* __restrict__ H, int m, int ld,
                                               double* __restrict__ E, double* __restrict__ U) {
  __shared__ double A[64 * 64];
  __shared__ double Vv[64 * 64];
  __shared__ int ord[64];
  const int k = threadIdx.x;
  for (int j = 0; j < m; ++j)
    if (k < m) {
      const int r = k <= j ? k : j, c = k <= j ? j : k;  // upper triangle is authoritative
      A[k + 64 * j] = H[r + ld * c];
      Vv[k + 64 * j] = (k == j) ? 1.0 : 0.0;
    }
  __syncthreads();
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    if (k < m)
      for (int j = 0; j < m; ++j) {
        const double v = A[k + 64 * j];
        tot += v * v;
        if (j != k) off += v * v;
      }
    off = wave_sum(off);
    tot = wave_sum(tot);
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[p + 64 * q];
        if (fabs(apq) < 1e-300) continue;
        const double app = A[p + 64 * p], aqq = A[q + 64 * q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        __syncthreads();
        if (k < m) {
          const double akp = A[k + 64 * p], akq = A[k + 64 * q];
          A[k + 64 * p] = c * akp - s * akq;
          A[k + 64 * q] = s * akp + c * akq;
        }
        __syncthreads();
        if (k < m) {
          const double apk = A[p + 64 * k], aqk = A[q + 64 * k];
          A[p + 64 * k] = c * apk - s * aqk;
          A[q + 64 * k] = s * apk + c * aqk;
          const double vkp = Vv[k + 64 * p], vkq = Vv[k + 64 * q];
          Vv[k + 64 * p] = c * vkp - s * vkq;
          Vv[k + 64 * q] = s * vkp + c * vkq;
        }
        __syncthreads();
      }
  }
  if (k == 0) {
    for (int i = 0; i < m; ++i) ord[i] = i;
    for (int i = 0; i < m; ++i) {
      int mi = i;
      for (int j = i + 1; j < m; ++j)
        if (A[ord[j] * 65] < A[ord[mi] * 65]) mi = j;
      const int t = ord[i];
      ord[i] = ord[mi];
      ord[mi] = t;
    }
  }
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    if (k == 0) E[j] = A[ord[j] * 65];
    if (k < m) U[k + ld * j] = Vv[k + 64 * ord[j]];
  }
}

// row sums of W^2 and the per-orbital occupations -> density[2i+s]
__global__ void __launch_bounds__(RT) k_rowsq(const double* __restrict__ W, int64_t nb, double* __restrict__ rs) {
  __shared__ double sh[RT / 64];
  const int64_t ia = blockIdx.x;
  double v = 0.0;
  for (int64_t ib = threadIdx.x; ib < nb; ib += RT) {
    const double w = W[ia * nb + ib];
    v += w * w;
  }
  v = block_sum(v, sh);
  if (threadIdx.x == 0) rs[ia] = v;
}

__global__ void k_colsq(const double* __restrict__ W, int64_t na, int64_t nb, double* __restrict__ cs) {
  const int64_t ib = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nb) return;
  double v = 0.0;
  for (int64_t ia = 0; ia < na; ++ia) {
    const double w = W[ia * nb + ib];
    v += w * w;
  }
  cs[ib] = v;
}

__global__ void __launch_bounds__(RT) k_occ(const double* __restrict__ rs, const uint64_t* __restrict__ a, int64_t na,
                                            const double* __restrict__ cs, const uint64_t* __restrict__ b,
                                            int64_t nb, int L, double* __restrict__ dens) {
  __shared__ double sh[RT / 64];
  for (int o = 0; o < L; ++o) {
    double va = 0.0, vb = 0.0;
    for (int64_t i = threadIdx.x; i < na; i += RT)
      if ((a[i] >> o) & 1ull) va += rs[i];
    for (int64_t i = threadIdx.x; i < nb; i += RT)
      if ((b[i] >> o) & 1ull) vb += cs[i];
    va = block_sum(va, sh);
    vb = block_sum(vb, sh);
    if (threadIdx.x == 0) {
      dens[2 * o] = va;
      dens[2 * o + 1] = vb;
    }
  }
}

template <int MM, class F>
void dispatch_m(int m, F&& f) {
  if (m <= 8) f(std::integral_constant<int, 8>{});
  else if (m <= 16) f(std::integral_constant<int, 16>{});
  else if (m <= 32) f(std::integral_constant<int, 32>{});
  else if (m <= 64) f(std::integral_constant<int, 64>{});
  else throw Error(SBDG_ERR_LIMIT, "more than 64 subspace vectors");
}

// two or four consecutive elements per thread (16-byte loads, longer runs per stream)
// when every vector starts 2- or 4-element aligned
inline bool pairs(int64_t n, int64_t ld1, int64_t ld2 = 0) { return n % 2 == 0 && ld1 % 2 == 0 && ld2 % 2 == 0; }
inline bool quads(int64_t n, int64_t ld1, int64_t ld2 = 0) { return n % 4 == 0 && ld1 % 4 == 0 && ld2 % 4 == 0; }

}  // namespace

void sign_transform(const uint64_t* a, int64_t na, const uint64_t* b, int64_t nb, const double* in, double* out,
                    hipStream_t st) {
  const int64_t N = na * nb;
  if (N == 0) return;
  hipLaunchKernelGGL(k_sign, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, a, b, nb, N, in, out);
  SBDG_CHECK(hipGetLastError());
}

void fill_test_vector(double* x, int64_t n, uint64_t seed, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n, seed);
  SBDG_CHECK(hipGetLastError());
}

void cast_d2f(const double* in, float* out, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_d2f, dim3(nblocks(n) * 4), dim3(256), 0, st, in, out, n);
}
void cast_f2d(const float* in, double* out, int64_t n, hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_f2d, dim3(nblocks(n) * 4), dim3(256), 0, st, in, out, n);
}

void multi_dot_impl(const double* X, int64_t ldx, int m, const double* y, int64_t n, double* out, double* part,
                    hipStream_t st) {
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    if (quads(n, ldx))
      hipLaunchKernelGGL((k_mdot<double, decltype(MM)::value, 4>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
    else if (pairs(n, ldx))
      hipLaunchKernelGGL((k_mdot<double, decltype(MM)::value, 2>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
    else
      hipLaunchKernelGGL((k_mdot<double, decltype(MM)::value, 1>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, m, out);
  SBDG_CHECK(hipGetLastError());
}

void multi_dot_impl(const double* X, int64_t ldx, int m, const float* y, int64_t n, double* out, double* part,
                    hipStream_t st) {
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    if (quads(n, ldx))
      hipLaunchKernelGGL((k_mdot<float, decltype(MM)::value, 4>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
    else if (pairs(n, ldx))
      hipLaunchKernelGGL((k_mdot<float, decltype(MM)::value, 2>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
    else
      hipLaunchKernelGGL((k_mdot<float, decltype(MM)::value, 1>), dim3(g), dim3(RT), 0, st, X, ldx, m, y, n, part);
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, m, out);
  SBDG_CHECK(hipGetLastError());
}

template <class TH>
void ritz_combine_impl(const double* V, int64_t ldv, const TH* HV, int64_t ldh, int m, const double* u,
                       const double* theta, double* W, double* R, int64_t n, double* out, double* part,
                       hipStream_t st) {
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    if (quads(n, ldv, ldh))
      hipLaunchKernelGGL((k_ritz<TH, decltype(MM)::value, 4>), dim3(g), dim3(RT), 0, st, V, ldv, HV, ldh, m, u,
                         theta, W, R, n, part);
    else if (pairs(n, ldv, ldh))
      hipLaunchKernelGGL((k_ritz<TH, decltype(MM)::value, 2>), dim3(g), dim3(RT), 0, st, V, ldv, HV, ldh, m, u,
                         theta, W, R, n, part);
    else
      hipLaunchKernelGGL((k_ritz<TH, decltype(MM)::value, 1>), dim3(g), dim3(RT), 0, st, V, ldv, HV, ldh, m, u,
                         theta, W, R, n, part);
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, 2, out);
  SBDG_CHECK(hipGetLastError());
}
template void ritz_combine_impl<double>(const double*, int64_t, const double*, int64_t, int, const double*,
                                        const double*, double*, double*, int64_t, double*, double*, hipStream_t);
template void ritz_combine_impl<float>(const double*, int64_t, const float*, int64_t, int, const double*,
                                       const double*, double*, double*, int64_t, double*, double*, hipStream_t);

template <class TH>
void ritz_expand_impl(const double* V, int64_t ldv, const TH* HV, int64_t ldh, int m, const double* u,
                      const double* theta, const double* hii, double* T, int64_t n, double* out, double* part,
                      hipStream_t st) {
  if (m > 32) throw Error(SBDG_ERR_LIMIT, "fused Ritz step beyond 32 vectors");
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    if constexpr (decltype(MM)::value <= 32) {
      if (decltype(MM)::value <= 16 && quads(n, ldv, ldh))
        hipLaunchKernelGGL((k_ritz_expand<TH, decltype(MM)::value, (decltype(MM)::value <= 16 ? 4 : 2)>), dim3(g),
                           dim3(RT), 0, st, V, ldv, HV, ldh, m, u, theta, hii, T, n, part);
      else if (pairs(n, ldv, ldh))
        hipLaunchKernelGGL((k_ritz_expand<TH, decltype(MM)::value, 2>), dim3(g), dim3(RT), 0, st, V, ldv, HV, ldh,
                           m, u, theta, hii, T, n, part);
      else
        hipLaunchKernelGGL((k_ritz_expand<TH, decltype(MM)::value, 1>), dim3(g), dim3(RT), 0, st, V, ldv, HV, ldh,
                           m, u, theta, hii, T, n, part);
    }
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, m + 2, out);
  SBDG_CHECK(hipGetLastError());
}
template void ritz_expand_impl<double>(const double*, int64_t, const double*, int64_t, int, const double*,
                                       const double*, const double*, double*, int64_t, double*, double*,
                                       hipStream_t);
template void ritz_expand_impl<float>(const double*, int64_t, const float*, int64_t, int, const double*,
                                      const double*, const double*, double*, int64_t, double*, double*,
                                      hipStream_t);

void scale_cast_impl(double* x, const double* n2, float* y, int64_t n, hipStream_t st) {
  if (n == 0) return;
  if (quads(n, 0)) hipLaunchKernelGGL(k_scale_cast<4>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, y, n);
  else if (pairs(n, 0)) hipLaunchKernelGGL(k_scale_cast<2>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, y, n);
  else hipLaunchKernelGGL(k_scale_cast<1>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, y, n);
}

void multi_dot_rows_impl(const float* Y, int64_t ldy, int m, const double* x, const double* z, int64_t n,
                         double* out, double* part, hipStream_t st) {
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    constexpr int mm = decltype(MM)::value;
    const bool v2 = pairs(n, ldy), v4 = quads(n, ldy);
    if (z) {
      if (v4) hipLaunchKernelGGL((k_mdot_rows<mm, 4, true>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
      else if (v2) hipLaunchKernelGGL((k_mdot_rows<mm, 2, true>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
      else hipLaunchKernelGGL((k_mdot_rows<mm, 1, true>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
    } else {
      if (v4) hipLaunchKernelGGL((k_mdot_rows<mm, 4, false>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
      else if (v2) hipLaunchKernelGGL((k_mdot_rows<mm, 2, false>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
      else hipLaunchKernelGGL((k_mdot_rows<mm, 1, false>), dim3(g), dim3(RT), 0, st, Y, ldy, m, x, z, n, part);
    }
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, m + (z ? 1 : 0), out);
  SBDG_CHECK(hipGetLastError());
}

void scale_impl(double* x, const double* n2, int64_t n, hipStream_t st) {
  if (n == 0) return;
  if (quads(n, 0)) hipLaunchKernelGGL(k_scale<4>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, n);
  else if (pairs(n, 0)) hipLaunchKernelGGL(k_scale<2>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, n);
  else hipLaunchKernelGGL(k_scale<1>, dim3(nblocks(n) * 4), dim3(256), 0, st, x, n2, n);
}

void precond_impl(const double* R, const double* rn2, const double* hii, const double* theta, double* t, int64_t n,
                  hipStream_t st) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_precond, dim3(nblocks(n) * 4), dim3(256), 0, st, R, rn2, hii, theta, t, n);
}

void gs_impl(const double* V, int64_t ldv, int m, const double* s, double* t, int64_t n, double* out, double* part,
             hipStream_t st) {
  const unsigned g = nblocks(n);
  dispatch_m<0>(m, [&](auto MM) {
    if (quads(n, ldv))
      hipLaunchKernelGGL((k_gs<decltype(MM)::value, 4>), dim3(g), dim3(RT), 0, st, V, ldv, m, s, t, n, part);
    else if (pairs(n, ldv))
      hipLaunchKernelGGL((k_gs<decltype(MM)::value, 2>), dim3(g), dim3(RT), 0, st, V, ldv, m, s, t, n, part);
    else
      hipLaunchKernelGGL((k_gs<decltype(MM)::value, 1>), dim3(g), dim3(RT), 0, st, V, ldv, m, s, t, n, part);
  });
  hipLaunchKernelGGL(k_final, dim3(1), dim3(RT), 0, st, part, (int)g, 1, out);
  SBDG_CHECK(hipGetLastError());
}

template <class T>
void rotate_impl(T* X, int64_t ldx, int m, const double* U, int ld, int mk, int64_t n, hipStream_t st) {
  if (n == 0) return;
  dispatch_m<0>(m, [&](auto MM) {
    constexpr int mm = decltype(MM)::value;
    if (mm <= 16 && pairs(n, ldx))  // x[MM][2] in registers
      hipLaunchKernelGGL((k_rotate<T, mm, (mm <= 16 ? 2 : 1)>), dim3(nblocks(n) * 4), dim3(256), 0, st, X, ldx, m, U,
                         ld, mk, n);
    else
      hipLaunchKernelGGL((k_rotate<T, mm, 1>), dim3(nblocks(n) * 4), dim3(256), 0, st, X, ldx, m, U, ld, mk, n);
  });
  SBDG_CHECK(hipGetLastError());
}
template void rotate_impl<double>(double*, int64_t, int, const double*, int, int, int64_t, hipStream_t);
template void rotate_impl<float>(float*, int64_t, int, const double*, int, int, int64_t, hipStream_t);

void syev_device(const double* d_H, int m, int ld, double* d_E, double* d_U, hipStream_t st) {
  if (m > 64) throw Error(SBDG_ERR_LIMIT, "subspace beyond 64");
  hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(64), 0, st, d_H, m, ld, d_E, d_U);
  SBDG_CHECK(hipGetLastError());
}

void row_sumsq(const double* W, int64_t na, int64_t nb, double* d_rs, hipStream_t st) {
  if (na > 0) hipLaunchKernelGGL(k_rowsq, dim3((unsigned)na), dim3(RT), 0, st, W, nb, d_rs);
  SBDG_CHECK(hipGetLastError());
}

void density(const double* W, const uint64_t* a, int64_t na, const uint64_t* b, int64_t nb, int L, double* d_dens,
             double* d_part, hipStream_t st) {
  double* rs = d_part;
  double* cs = d_part + na;
  if (na > 0) hipLaunchKernelGGL(k_rowsq, dim3((unsigned)na), dim3(RT), 0, st, W, nb, rs);
  if (nb > 0) hipLaunchKernelGGL(k_colsq, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, W, na, nb, cs);
  hipLaunchKernelGGL(k_occ, dim3(1), dim3(RT), 0, st, rs, a, na, cs, b, nb, L, d_dens);
  SBDG_CHECK(hipGetLastError());
}

}  // namespace sbdg
