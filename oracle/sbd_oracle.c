/*
 * sbd_oracle.c -- CPU restatement of the selected-basis diagonalization hot path of
 * walkup/sbd-ykt (tensor-product basis, method 0).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.  Nothing in
 * the product links it, calls it or falls back to it.
 *
 * It follows the reference literally, in the reference's own conventions:
 *   - a determinant interleaves alpha orbital i at spin-orbital 2i and beta orbital i at
 *     2i+1 (include/sbd/chemistry/basic/determinants.h:96-142, non-TRADMODE branch);
 *   - the fermionic sign of an excitation counts the occupied spin-orbitals of a
 *     determinant strictly between the two indices (determinants.h:228-279);
 *   - matrix elements are ZeroExcite / OneExcite / TwoExcite selected by the bit
 *     difference between bra and ket (determinants.h:397-587);
 *   - integrals are the spin-orbital one-body matrix and the 8-fold packed two-body
 *     store (chemistry/basic/integrals.h:10-67, makeintegrals.h:10-55);
 *   - excitation tables list, per string, the ket strings at Hamming distance 2 and 4 in
 *     ascending index order (chemistry/tpb/helper.h:132-197);
 *   - sigma adds the diagonal, then the alpha-alpha task, the beta-beta task and the
 *     alpha-beta task, per bra row in table order (chemistry/tpb/mult.h:376-632 with the
 *     one-rank task schedule of helper.h:463-515);
 *   - Davidson, restart to one vector, is davidson.h:661-924 with dm_vector.h:20-84;
 *   - RDMs are chemistry/tpb/correlation.h:15-444.
 *
 * Strings are uint64 (one word, orbitals 0..63); determinants are 128-bit.
 * Parity status of each function is recorded in the tests that pin it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

#ifndef PLANT
#define PLANT 0 /* test hook: >0 compiles a deliberately wrong variant (see the oracle tests) */
#endif

typedef struct {
  int L;
  double I0;
  double *I1;   /* (2L)^2 spin-orbital one-body, I1[i*2L+j]  (integrals.h:10-22) */
  double *I2;   /* packed (ij|kl), size B(B+1)/2, B=L(L+1)/2   (integrals.h:24-67) */
  double *Dir;  /* Direct(i,j)   = (ii|jj) at i+L*j  (makeintegrals.h:47-53) */
  double *Exc;  /* Exchange(i,j) = (ij|ji) at i+L*j */
  int64_t na, nb;
  uint64_t *a, *b;
  int64_t *rpSa, *rpDa, *rpSb, *rpDb; /* CSR row pointers (n+1) */
  int64_t *Sa, *Da, *Sb, *Db;         /* ket indices, ascending */
} orc_problem;

/* ------------------------------------------------------------------ bits */

static inline int pc128(u128 x) {
  return __builtin_popcountll((uint64_t)x) + __builtin_popcountll((uint64_t)(x >> 64));
}

static inline int ctz128(u128 x) {
  uint64_t lo = (uint64_t)x;
  if (lo) return __builtin_ctzll(lo);
  return 64 + __builtin_ctzll((uint64_t)(x >> 64));
}

/* DetFromAlphaBeta, determinants.h:96-142: alpha bit i -> 2i, beta bit i -> 2i+1. */
static inline u128 interleave(uint64_t A, uint64_t B, int L) {
  u128 D = 0;
  for (int i = 0; i < L; ++i) {
    if ((A >> i) & 1ull) D |= ((u128)1) << (2 * i + (PLANT == 4));
    if ((B >> i) & 1ull) D |= ((u128)1) << (2 * i + (PLANT != 4));
  }
  return D;
}

/* parity(dets,bit_length,start,end,sgn), determinants.h:228-279: flip for every
 * occupied bit in [start,end), and once more when bit start itself is occupied. */
static inline double parity(u128 det, int start, int end, double sgn) {
  u128 mask = ((((u128)1) << end) - 1) ^ ((((u128)1) << start) - 1);
  if (pc128(det & mask) & 1) sgn = -sgn;
  if (PLANT != 2 && ((det >> start) & 1)) sgn = -sgn;
  return sgn;
}

/* ------------------------------------------------------------- integrals */

static inline int64_t pair_index(int I, int J) {
  return I > J ? (int64_t)I * (I + 1) / 2 + J : (int64_t)J * (J + 1) / 2 + I;
}

static inline int64_t packed_index(int I, int J, int K, int L_) {
  int64_t ij = pair_index(I, J), kl = pair_index(K, L_);
  int64_t a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
  return a * (a + 1) / 2 + b;
}

/* twoInt::Value, integrals.h:50-59: spin-orbital indices, zero unless the spins pair. */
static inline double I2v(const orc_problem *p, int i, int j, int k, int l) {
  if (!(((i & 1) == (j & 1)) && ((k & 1) == (l & 1)))) return 0.0;
  return p->I2[packed_index(i >> 1, j >> 1, k >> 1, l >> 1)];
}

static inline double I1v(const orc_problem *p, int i, int j) { return p->I1[i * 2 * p->L + j]; }

int64_t orc_packed_size(int L) {
  int64_t B = (int64_t)L * (L + 1) / 2;
  return B * (B + 1) / 2;
}

/* SetupIntegrals, makeintegrals.h:10-55, from FCIDUMP tuples (1-based indices; the
 * "0 0 0 0" line is the core energy; "i j 0 0" a one-body entry).  Later tuples overwrite
 * earlier ones exactly as the reference's stores do.  h1 is L*L row-major (spatial). */
int orc_setup_from_fcidump(int L, int64_t n, const double *val, const int32_t *idx, double *I0,
                           double *h1, double *packed) {
  int64_t np = orc_packed_size(L);
  memset(h1, 0, sizeof(double) * L * L);
  memset(packed, 0, sizeof(double) * np);
  *I0 = 0.0;
  for (int64_t t = 0; t < n; ++t) {
    int i = idx[4 * t], j = idx[4 * t + 1], k = idx[4 * t + 2], l = idx[4 * t + 3];
    double v = val[t];
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      *I0 = v;
    } else if (k == 0 && l == 0) {
      if (i < 1 || j < 1 || i > L || j > L) return -1;
      h1[(i - 1) * L + (j - 1)] = v;
      h1[(j - 1) * L + (i - 1)] = v;
    } else {
      if (i < 1 || j < 1 || k < 1 || l < 1 || i > L || j > L || k > L || l > L) return -1;
      packed[packed_index(i - 1, j - 1, k - 1, l - 1)] = v;
    }
  }
  return 0;
}

/* Pack a full (pq|rs) tensor (row-major L^4) into the reference store: each canonical
 * slot takes the value at [p,q,r,s] with p>=q, r>=s, pq>=rs. */
void orc_pack_eri(int L, const double *eri, double *packed) {
  for (int p = 0; p < L; ++p)
    for (int q = 0; q <= p; ++q)
      for (int r = 0; r < L; ++r)
        for (int s = 0; s <= r; ++s) {
          int64_t pq = pair_index(p, q), rs = pair_index(r, s);
          if (pq < rs) continue;
          packed[pq * (pq + 1) / 2 + rs] = eri[(((int64_t)p * L + q) * L + r) * L + s];
        }
}

/* ---------------------------------------------------------------- problem */

static int64_t *alloc_i64(int64_t n) { return (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1)); }

/* GenerateExcitation, helper.h:132-197: kets at XOR popcount 2 (singles) and 4
 * (doubles), ascending ket index. */
static void build_tables(int64_t n, const uint64_t *s, int64_t **rpS, int64_t **S, int64_t **rpD,
                         int64_t **D) {
  int64_t *cs = alloc_i64(n + 1), *cd = alloc_i64(n + 1);
  cs[0] = cd[0] = 0;
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    int64_t ns = 0, nd = 0;
    for (int64_t j = 0; j < n; ++j) {
      int d = __builtin_popcountll(s[i] ^ s[j]);
      ns += (d == 2);
      nd += (d == 4);
    }
    cs[i + 1] = ns;
    cd[i + 1] = nd;
  }
  for (int64_t i = 0; i < n; ++i) {
    cs[i + 1] += cs[i];
    cd[i + 1] += cd[i];
  }
  int64_t *ls = alloc_i64(cs[n]), *ld = alloc_i64(cd[n]);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    int64_t ps = cs[i], pd = cd[i];
    for (int64_t j = 0; j < n; ++j) {
      int d = __builtin_popcountll(s[i] ^ s[j]);
      if (d == 2) ls[ps++] = j;
      else if (d == 4) ld[pd++] = j;
    }
  }
  *rpS = cs;
  *S = ls;
  *rpD = cd;
  *D = ld;
}

orc_problem *orc_create(int L, double I0, const double *h1, const double *packed, int64_t na,
                        const uint64_t *a, int64_t nb, const uint64_t *b) {
  if (L < 1 || L > 64) return NULL;
  orc_problem *p = (orc_problem *)calloc(1, sizeof(orc_problem));
  p->L = L;
  p->I0 = I0;
  int n2 = 2 * L;
  p->I1 = (double *)calloc((size_t)n2 * n2, sizeof(double));
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) {
      p->I1[(2 * i) * n2 + 2 * j] = h1[i * L + j];
      p->I1[(2 * i + 1) * n2 + 2 * j + 1] = h1[i * L + j];
    }
  int64_t np = orc_packed_size(L);
  p->I2 = (double *)malloc(sizeof(double) * np);
  memcpy(p->I2, packed, sizeof(double) * np);
  p->Dir = (double *)malloc(sizeof(double) * L * L);
  p->Exc = (double *)malloc(sizeof(double) * L * L);
  for (int i = 0; i < L; ++i)
    for (int j = 0; j < L; ++j) {
      p->Dir[i + L * j] = p->I2[packed_index(i, i, j, j)];
      p->Exc[i + L * j] = p->I2[packed_index(i, j, j, i)];
    }
  p->na = na;
  p->nb = nb;
  p->a = (uint64_t *)malloc(sizeof(uint64_t) * (na > 0 ? na : 1));
  p->b = (uint64_t *)malloc(sizeof(uint64_t) * (nb > 0 ? nb : 1));
  memcpy(p->a, a, sizeof(uint64_t) * na);
  memcpy(p->b, b, sizeof(uint64_t) * nb);
  build_tables(na, p->a, &p->rpSa, &p->Sa, &p->rpDa, &p->Da);
  build_tables(nb, p->b, &p->rpSb, &p->Sb, &p->rpDb, &p->Db);
  return p;
}

void orc_free(orc_problem *p) {
  if (!p) return;
  free(p->I1); free(p->I2); free(p->Dir); free(p->Exc); free(p->a); free(p->b);
  free(p->rpSa); free(p->rpDa); free(p->rpSb); free(p->rpDb);
  free(p->Sa); free(p->Da); free(p->Sb); free(p->Db);
  free(p);
}

/* Table access for tests: which = 0 Sa, 1 Da, 2 Sb, 3 Db. */
int64_t orc_table_nnz(const orc_problem *p, int which) {
  switch (which) {
    case 0: return p->rpSa[p->na];
    case 1: return p->rpDa[p->na];
    case 2: return p->rpSb[p->nb];
    case 3: return p->rpDb[p->nb];
  }
  return -1;
}

void orc_table_get(const orc_problem *p, int which, int64_t *rowptr, int64_t *cols) {
  int64_t n = (which < 2) ? p->na : p->nb;
  const int64_t *rp = which == 0 ? p->rpSa : which == 1 ? p->rpDa : which == 2 ? p->rpSb : p->rpDb;
  const int64_t *cl = which == 0 ? p->Sa : which == 1 ? p->Da : which == 2 ? p->Sb : p->Db;
  memcpy(rowptr, rp, sizeof(int64_t) * (n + 1));
  memcpy(cols, cl, sizeof(int64_t) * rp[n]);
}

/* ------------------------------------------------------ matrix elements */

/* ZeroExcite, determinants.h:397-421. */
static double zero_excite(const orc_problem *p, u128 det) {
  double e = 0.0;
  int occ[128], n = 0;
  u128 t = det;
  while (t) {
    int k = ctz128(t);
    occ[n++] = k;
    t &= t - 1;
  }
  for (int i = 0; i < n; ++i) {
    int I = occ[i];
    e += I1v(p, I, I);
    for (int j = i + 1; j < n; ++j) {
      int J = occ[j];
      e += p->Dir[(I / 2) + p->L * (J / 2)];
      if ((I % 2) == (J % 2)) e -= p->Exc[(I / 2) + p->L * (J / 2)];
    }
  }
  return e + p->I0;
}

/* OneExcite(ket, i, a), determinants.h:423-445. */
static double one_excite(const orc_problem *p, u128 ket, int i, int a) {
  double sgn = parity(ket, i < a ? i : a, i < a ? a : i, 1.0);
  double e = I1v(p, a, i);
  u128 t = ket;
  while (t) {
    int j = ctz128(t);
    e += (I2v(p, a, i, j, j) - (PLANT == 1 ? 0.0 : I2v(p, a, j, j, i)));
    t &= t - 1;
  }
  return e * sgn;
}

/* TwoExcite(ket, i, j, a, b), determinants.h:447-465. */
static double two_excite(const orc_problem *p, u128 ket, int i, int j, int a, int b) {
  int I = i < j ? i : j, J = i < j ? j : i, A = a < b ? a : b, B = a < b ? b : a;
  double sgn = 1.0;
  sgn = parity(ket, I < A ? I : A, I < A ? A : I, sgn);
  sgn = parity(ket, J < B ? J : B, J < B ? B : J, sgn);
  if (PLANT != 3 && (A > J || B < I)) sgn = -sgn;
  return sgn * (I2v(p, A, I, B, J) - I2v(p, A, J, B, I));
}

/* Hij(DetA=bra, DetB=ket), determinants.h:467-526. */
static double hij(const orc_problem *p, u128 bra, u128 ket) {
  u128 cdiff = bra & ~ket, ddiff = ket & ~bra;
  int nc = pc128(cdiff);
  if (nc == 0) return zero_excite(p, ket);
  int c0 = ctz128(cdiff), d0 = ctz128(ddiff);
  if (nc == 1) return one_excite(p, ket, d0, c0);
  if (nc == 2) {
    u128 c1m = cdiff & (cdiff - 1), d1m = ddiff & (ddiff - 1);
    int c1 = ctz128(c1m), d1 = ctz128(d1m);
    return two_excite(p, ket, d0, d1, c0, c1);
  }
  return 0.0;
}

/* Matrix elements for explicit (bra, ket) string pairs, in the reference convention. */
void orc_hij_batch(const orc_problem *p, int64_t n, const uint64_t *abra, const uint64_t *bbra,
                   const uint64_t *aket, const uint64_t *bket, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < n; ++t)
    out[t] = hij(p, interleave(abra[t], bbra[t], p->L), interleave(aket[t], bket[t], p->L));
}

/* makeQChamDiagTerms, tpb/qcham.h:314-366. */
void orc_hii(const orc_problem *p, double *hii) {
  int64_t nb = p->nb;
#pragma omp parallel for schedule(static)
  for (int64_t ia = 0; ia < p->na; ++ia)
    for (int64_t ib = 0; ib < nb; ++ib)
      hii[ia * nb + ib] = zero_excite(p, interleave(p->a[ia], p->b[ib], p->L));
}

/* mult (direct), tpb/mult.h:376-632, one rank: Wb += hii*T, then the alpha task
 * (type 2), the beta task (type 1), the alpha-beta task (type 0). */
void orc_sigma(const orc_problem *p, const double *hii, const double *T, double *Wb) {
  const int64_t na = p->na, nb = p->nb;
  const int L = p->L;
  for (int64_t k = 0; k < na * nb; ++k) Wb[k] += hii[k] * T[k];
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t ia = 0; ia < na; ++ia) {
    for (int64_t ib = 0; ib < nb; ++ib) {
      u128 I = interleave(p->a[ia], p->b[ib], L);
      double acc = Wb[ia * nb + ib];
      for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
        int64_t ja = p->Sa[q];
        acc += hij(p, I, interleave(p->a[ja], p->b[ib], L)) * T[ja * nb + ib];
      }
      for (int64_t q = p->rpDa[ia]; q < p->rpDa[ia + 1]; ++q) {
        int64_t ja = p->Da[q];
        acc += hij(p, I, interleave(p->a[ja], p->b[ib], L)) * T[ja * nb + ib];
      }
      Wb[ia * nb + ib] = acc;
    }
    for (int64_t ib = 0; ib < nb; ++ib) {
      u128 I = interleave(p->a[ia], p->b[ib], L);
      double acc = Wb[ia * nb + ib];
      for (int64_t q = p->rpSb[ib]; q < p->rpSb[ib + 1]; ++q) {
        int64_t jb = p->Sb[q];
        acc += hij(p, I, interleave(p->a[ia], p->b[jb], L)) * T[ia * nb + jb];
      }
      for (int64_t q = p->rpDb[ib]; q < p->rpDb[ib + 1]; ++q) {
        int64_t jb = p->Db[q];
        acc += hij(p, I, interleave(p->a[ia], p->b[jb], L)) * T[ia * nb + jb];
      }
      Wb[ia * nb + ib] = acc;
    }
    for (int64_t ib = 0; ib < nb; ++ib) {
      u128 I = interleave(p->a[ia], p->b[ib], L);
      double acc = Wb[ia * nb + ib];
      for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
        int64_t ja = p->Sa[q];
        for (int64_t r = p->rpSb[ib]; r < p->rpSb[ib + 1]; ++r) {
          int64_t jb = p->Sb[r];
          acc += hij(p, I, interleave(p->a[ja], p->b[jb], L)) * T[ja * nb + jb];
        }
      }
      Wb[ia * nb + ib] = acc;
    }
  }
}

/* sigma restricted to bra rows [ia0, ia1): the bounded CPU sample used as a timing
 * baseline.  Writes Wb rows ia0..ia1-1 (Wb indexed like the full vector); hii may be
 * NULL, then the diagonal is evaluated in place.  Every (bra row, bra column) element is
 * independent, so the threads share the elements in runs of 64 columns. */
void orc_sigma_rows(const orc_problem *p, const double *hii, const double *T, double *Wb,
                    int64_t ia0, int64_t ia1) {
  const int64_t nb = p->nb;
  const int L = p->L;
#pragma omp parallel for collapse(2) schedule(dynamic, 64)
  for (int64_t ia = ia0; ia < ia1; ++ia) {
    for (int64_t ib = 0; ib < nb; ++ib) {
      u128 I = interleave(p->a[ia], p->b[ib], L);
      double acc = (hii ? hii[ia * nb + ib] : zero_excite(p, I)) * T[ia * nb + ib];
      for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
        int64_t ja = p->Sa[q];
        acc += hij(p, I, interleave(p->a[ja], p->b[ib], L)) * T[ja * nb + ib];
      }
      for (int64_t q = p->rpDa[ia]; q < p->rpDa[ia + 1]; ++q) {
        int64_t ja = p->Da[q];
        acc += hij(p, I, interleave(p->a[ja], p->b[ib], L)) * T[ja * nb + ib];
      }
      for (int64_t q = p->rpSb[ib]; q < p->rpSb[ib + 1]; ++q) {
        int64_t jb = p->Sb[q];
        acc += hij(p, I, interleave(p->a[ia], p->b[jb], L)) * T[ia * nb + jb];
      }
      for (int64_t q = p->rpDb[ib]; q < p->rpDb[ib + 1]; ++q) {
        int64_t jb = p->Db[q];
        acc += hij(p, I, interleave(p->a[ia], p->b[jb], L)) * T[ia * nb + jb];
      }
      for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
        int64_t ja = p->Sa[q];
        for (int64_t r = p->rpSb[ib]; r < p->rpSb[ib + 1]; ++r) {
          int64_t jb = p->Sb[r];
          acc += hij(p, I, interleave(p->a[ja], p->b[jb], L)) * T[ja * nb + jb];
        }
      }
      Wb[ia * nb + ib] = acc;
    }
  }
}

/* Number of structurally connected (bra, ket) pairs per sigma, diagonal included. */
int64_t orc_term_count(const orc_problem *p) {
  int64_t na = p->na, nb = p->nb;
  int64_t sa = p->rpSa[na], da = p->rpDa[na], sb = p->rpSb[nb], db = p->rpDb[nb];
  return na * nb + (sa + da) * nb + (sb + db) * na + sa * sb;
}

/* ---------------------------------------------------------- Davidson */

/* InnerProduct / Normalize, dm_vector.h:20-84: compensated (Kahan) sums. */
static double kahan_dot(const double *x, const double *y, int64_t n) {
  double sum = 0.0, eps = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    double val = x[i] * y[i] - eps;
    double tmp = sum + val;
    eps = (tmp - sum) - val;
    sum = tmp;
  }
  return sum;
}

static double normalize(double *x, int64_t n) {
  double r = sqrt(kahan_dot(x, x, n));
  double f = 1.0 / r;
  for (int64_t i = 0; i < n; ++i) x[i] *= f;
  return r;
}

/* Symmetric eigensolver for the n x n leading block of the column-major matrix U
 * (leading dimension lda), standing in for LAPACK dsyev (hp_numeric.h:497-516):
 * eigenvalues ascending in E, orthonormal eigenvectors in the columns of U.  Cyclic
 * Jacobi rotations, then a selection sort on the eigenvalues. */
void orc_syev(int n, double *U, int lda, double *E) {
  double *A = (double *)malloc(sizeof(double) * n * n);
  double *V = (double *)calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) A[i + n * j] = U[(i <= j ? i : j) + lda * (i <= j ? j : i)];
  for (int i = 0; i < n; ++i) V[i + n * i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += A[i + n * j] * A[i + n * j];
        if (i != j) off += A[i + n * j] * A[i + n * j];
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int pp = 0; pp < n - 1; ++pp)
      for (int q = pp + 1; q < n; ++q) {
        double apq = A[pp + n * q];
        if (fabs(apq) < 1e-300) continue;
        double app = A[pp + n * pp], aqq = A[q + n * q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k + n * pp], akq = A[k + n * q];
          A[k + n * pp] = c * akp - s * akq;
          A[k + n * q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[pp + n * k], aqk = A[q + n * k];
          A[pp + n * k] = c * apk - s * aqk;
          A[q + n * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = V[k + n * pp], vkq = V[k + n * q];
          V[k + n * pp] = c * vkp - s * vkq;
          V[k + n * q] = s * vkp + c * vkq;
        }
      }
  }
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int k = i + 1; k < n; ++k)
      if (A[ord[k] + n * ord[k]] < A[ord[m] + n * ord[m]]) m = k;
    int t = ord[i];
    ord[i] = ord[m];
    ord[m] = t;
  }
  for (int j = 0; j < n; ++j) {
    E[j] = A[ord[j] + n * ord[j]];
    for (int i = 0; i < n; ++i) U[i + lda * j] = V[i + n * ord[j]];
  }
  free(A);
  free(V);
  free(ord);
}

/* Davidson(...,max_time), tpb/davidson.h:661-924, lowest root, restart to one vector.
 * W is the initial vector on entry (BasisInitVector) and the Ritz vector on exit.
 * Returns the number of sigma builds; *E0 the last Ritz value, *rnorm the last
 * residual norm.  energies (optional, length max_it*nb) receives E0 per step. */
int orc_davidson(const orc_problem *p, const double *hii, double *W, int max_it, int nb,
                 double eps, double *E0, double *rnorm, double *energies) {
  const int64_t n = p->na * p->nb;
  const double eps_reg = 1.0e-12;
  double **C = (double **)malloc(sizeof(double *) * nb);
  double **HC = (double **)malloc(sizeof(double *) * nb);
  for (int k = 0; k < nb; ++k) {
    C[k] = (double *)calloc(n, sizeof(double));
    HC[k] = (double *)calloc(n, sizeof(double));
  }
  double *R = (double *)calloc(n, sizeof(double));
  double *H = (double *)calloc((size_t)nb * nb, sizeof(double));
  double *U = (double *)calloc((size_t)nb * nb, sizeof(double));
  double *E = (double *)calloc(nb, sizeof(double));
  int nsig = 0, done = 0, step = 0;
  double last_r = 0.0;
  for (int it = 0; it < max_it && !done; ++it) {
    memcpy(C[0], W, sizeof(double) * n);
    for (int ib = 0; ib < nb; ++ib) {
      memset(HC[ib], 0, sizeof(double) * n);
      orc_sigma(p, hii, C[ib], HC[ib]);
      ++nsig;
      for (int jb = 0; jb <= ib; ++jb) {
        H[jb + nb * ib] = kahan_dot(C[jb], HC[ib], n);
        H[ib + nb * jb] = H[jb + nb * ib];
      }
      for (int jb = 0; jb <= ib; ++jb)
        for (int kb = 0; kb <= ib; ++kb) U[jb + nb * kb] = H[jb + nb * kb];
      orc_syev(ib + 1, U, nb, E);
      for (int64_t s = 0; s < n; ++s) W[s] = C[0][s] * U[0];
      for (int64_t s = 0; s < n; ++s) R[s] = HC[0][s] * (-U[0]);
      for (int kb = 1; kb <= ib; ++kb) {
        double x = U[kb];
        for (int64_t s = 0; s < n; ++s) W[s] += C[kb][s] * x;
        x = -U[kb];
        for (int64_t s = 0; s < n; ++s) R[s] += HC[kb][s] * x;
      }
      for (int64_t s = 0; s < n; ++s) R[s] += E[0] * W[s];
      normalize(W, n);
      double nr = normalize(R, n);
      last_r = nr;
      if (energies) energies[step] = E[0];
      ++step;
      if (nr < eps) {
        done = 1;
        break;
      }
      if (ib < nb - 1) {
        for (int64_t s = 0; s < n; ++s) {
          double den = E[0] - hii[s];
          C[ib + 1][s] = (fabs(den) > eps_reg) ? R[s] / den : R[s] / (den - eps_reg);
        }
        for (int kb = 0; kb < ib + 1; ++kb) {
          double olap = -kahan_dot(C[kb], C[ib + 1], n);
          for (int64_t s = 0; s < n; ++s) C[ib + 1][s] += C[kb][s] * olap;
        }
        normalize(C[ib + 1], n);
      }
    }
  }
  *E0 = E[0];
  *rnorm = last_r;
  for (int k = 0; k < nb; ++k) {
    free(C[k]);
    free(HC[k]);
  }
  free(C); free(HC); free(R); free(H); free(U); free(E);
  return nsig;
}

/* Energy of a normalized vector as the reference reports it, sbdiag.h:250-273:
 * one more sigma, then <W|HW>. */
double orc_energy(const orc_problem *p, const double *hii, const double *W) {
  int64_t n = p->na * p->nb;
  double *C = (double *)calloc(n, sizeof(double));
  orc_sigma(p, hii, W, C);
  double e = kahan_dot(W, C, n);
  free(C);
  return e;
}

/* ------------------------------------------------------------- RDMs */

#define OB(s, i) one[(size_t)(s) * L2 + (i)]
#define TB(s, i) two[(size_t)(s) * L4 + (i)]

/* ZeroDiffCorrelation, correlation.h:15-42. */
static void zero_diff(u128 det, double wi, int L, double *one, double *two) {
  const int64_t L2 = (int64_t)L * L, L3 = L2 * L, L4 = L3 * L;
  int occ[128], n = 0;
  u128 t = det;
  while (t) {
    occ[n++] = ctz128(t);
    t &= t - 1;
  }
  double w2 = wi * wi;
  for (int i = 0; i < n; ++i) {
    int oi = occ[i] / 2, si = occ[i] % 2;
    OB(si, oi + L * oi) += w2;
    for (int j = i + 1; j < n; ++j) {
      int oj = occ[j] / 2, sj = occ[j] % 2;
      TB(si + 2 * sj, oi + L * oj + L2 * oi + L3 * oj) += w2;
      TB(sj + 2 * si, oj + L * oi + L2 * oj + L3 * oi) += w2;
      if (si == sj) {
        TB(si + 2 * sj, oi + L * oj + L2 * oj + L3 * oi) += -w2;
        TB(sj + 2 * si, oj + L * oi + L2 * oi + L3 * oj) += -w2;
      }
    }
  }
}

/* OneDiffCorrelation, correlation.h:47-83 (i: bra-only, a: ket-only). */
static void one_diff(u128 detI, double wi, double wj, int L, int i, int a, double *one,
                     double *two) {
  const int64_t L2 = (int64_t)L * L, L3 = L2 * L, L4 = L3 * L;
  double sgn = parity(detI, i < a ? i : a, i < a ? a : i, 1.0);
  int oi = i / 2, si = i % 2, oa = a / 2;
  double w = wi * wj * sgn;
  OB(si, oi + L * oa) += w;
  u128 t = detI;
  while (t) {
    int soj = ctz128(t);
    t &= t - 1;
    int oj = soj / 2, sj = soj % 2;
    TB(si + 2 * sj, oa + oj * L + oi * L2 + oj * L3) += w;
    TB(sj + 2 * si, oj + oa * L + oj * L2 + oi * L3) += w;
    if (si == sj) {
      TB(si + 2 * sj, oa + oj * L + oj * L2 + oi * L3) += -w;
      TB(sj + 2 * si, oj + oa * L + oi * L2 + oj * L3) += -w;
    }
  }
}

/* TwoDiffCorrelation, correlation.h:88-130 (i,j: bra-only, a,b: ket-only). */
static void two_diff(u128 detI, double wi, double wj, int L, int i, int j, int a, int b,
                     double *one, double *two) {
  (void)one;
  const int64_t L2 = (int64_t)L * L, L3 = L2 * L, L4 = L3 * L;
  int I = i < j ? i : j, J = i < j ? j : i, A = a < b ? a : b, B = a < b ? b : a;
  double sgn = 1.0;
  sgn = parity(detI, I < A ? I : A, I < A ? A : I, sgn);
  sgn = parity(detI, J < B ? J : B, J < B ? B : J, sgn);
  if (A > J || B < I) sgn = -sgn;
  int oi = I / 2, si = I % 2, oa = A / 2, sa = A % 2, oj = J / 2, sj = J % 2, ob = B / 2,
      sb = B % 2;
  double w = sgn * wi * wj;
  if (si == sa) {
    TB(si + 2 * sj, oa + L * ob + L2 * (oi + L * oj)) += w;
    TB(sj + 2 * si, ob + L * oa + L2 * (oj + L * oi)) += w;
  }
  if (si == sb) {
    TB(si + 2 * sj, oa + L * ob + L2 * (oj + L * oi)) += -w;
    TB(sj + 2 * si, ob + L * oa + L2 * (oi + L * oj)) += -w;
  }
}

/* CorrelationTermAddition, correlation.h:135-189. */
static void term_add(u128 detI, u128 detJ, double wi, double wj, int L, double *one, double *two) {
  u128 cdiff = detI & ~detJ, ddiff = detJ & ~detI;
  int nc = pc128(cdiff);
  if (nc == 0) {
    zero_diff(detI, wi, L, one, two);
    return;
  }
  int c0 = ctz128(cdiff), d0 = ctz128(ddiff);
  if (nc == 1) {
    one_diff(detI, wi, wj, L, c0, d0, one, two);
  } else if (nc == 2) {
    int c1 = ctz128(cdiff & (cdiff - 1)), d1 = ctz128(ddiff & (ddiff - 1));
    two_diff(detI, wi, wj, L, c0, c1, d0, d1, one, two);
  }
}

/* Correlation, correlation.h:194-444: one[2][L^2] and two[4][L^4] from W, per-thread
 * partial arrays summed in thread order. */
void orc_rdm(const orc_problem *p, const double *W, double *one, double *two) {
  const int L = p->L;
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  const int64_t na = p->na, nb = p->nb;
  memset(one, 0, sizeof(double) * 2 * L2);
  memset(two, 0, sizeof(double) * 4 * L4);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  double *one_t = (double *)calloc((size_t)nth * 2 * L2, sizeof(double));
  double *two_t = (double *)calloc((size_t)nth * 4 * L4, sizeof(double));
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *o = one_t + (size_t)tid * 2 * L2, *t = two_t + (size_t)tid * 4 * L4;
#pragma omp for schedule(dynamic, 1)
    for (int64_t ia = 0; ia < na; ++ia)
      for (int64_t ib = 0; ib < nb; ++ib)
        zero_diff(interleave(p->a[ia], p->b[ib], L), W[ia * nb + ib], L, o, t);
#pragma omp for schedule(dynamic, 1)
    for (int64_t ia = 0; ia < na; ++ia)
      for (int64_t ib = 0; ib < nb; ++ib) {
        u128 I = interleave(p->a[ia], p->b[ib], L);
        double wi = W[ia * nb + ib];
        for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
          int64_t ja = p->Sa[q];
          term_add(I, interleave(p->a[ja], p->b[ib], L), wi, W[ja * nb + ib], L, o, t);
        }
        for (int64_t q = p->rpDa[ia]; q < p->rpDa[ia + 1]; ++q) {
          int64_t ja = p->Da[q];
          term_add(I, interleave(p->a[ja], p->b[ib], L), wi, W[ja * nb + ib], L, o, t);
        }
      }
#pragma omp for schedule(dynamic, 1)
    for (int64_t ia = 0; ia < na; ++ia)
      for (int64_t ib = 0; ib < nb; ++ib) {
        u128 I = interleave(p->a[ia], p->b[ib], L);
        double wi = W[ia * nb + ib];
        for (int64_t q = p->rpSb[ib]; q < p->rpSb[ib + 1]; ++q) {
          int64_t jb = p->Sb[q];
          term_add(I, interleave(p->a[ia], p->b[jb], L), wi, W[ia * nb + jb], L, o, t);
        }
        for (int64_t q = p->rpDb[ib]; q < p->rpDb[ib + 1]; ++q) {
          int64_t jb = p->Db[q];
          term_add(I, interleave(p->a[ia], p->b[jb], L), wi, W[ia * nb + jb], L, o, t);
        }
      }
#pragma omp for schedule(dynamic, 1)
    for (int64_t ia = 0; ia < na; ++ia)
      for (int64_t ib = 0; ib < nb; ++ib) {
        u128 I = interleave(p->a[ia], p->b[ib], L);
        double wi = W[ia * nb + ib];
        for (int64_t q = p->rpSa[ia]; q < p->rpSa[ia + 1]; ++q) {
          int64_t ja = p->Sa[q];
          for (int64_t r = p->rpSb[ib]; r < p->rpSb[ib + 1]; ++r) {
            int64_t jb = p->Sb[r];
            term_add(I, interleave(p->a[ja], p->b[jb], L), wi, W[ja * nb + jb], L, o, t);
          }
        }
      }
  }
  for (int k = 0; k < nth; ++k) {
    for (int64_t i = 0; i < 2 * L2; ++i) one[i] += one_t[(size_t)k * 2 * L2 + i];
    for (int64_t i = 0; i < 4 * L4; ++i) two[i] += two_t[(size_t)k * 4 * L4 + i];
  }
  free(one_t);
  free(two_t);
}

/* OccupationDensity, tpb/occupation.h:10-95: density[2i+s] = <n_{i,s}>. */
void orc_density(const orc_problem *p, const double *W, double *density) {
  const int L = p->L;
  const int64_t na = p->na, nb = p->nb;
  for (int i = 0; i < 2 * L; ++i) density[i] = 0.0;
  for (int64_t ia = 0; ia < na; ++ia)
    for (int64_t ib = 0; ib < nb; ++ib) {
      double w2 = W[ia * nb + ib] * W[ia * nb + ib];
      for (int o = 0; o < L; ++o) {
        if ((p->a[ia] >> o) & 1ull) density[2 * o] += w2;
        if ((p->b[ib] >> o) & 1ull) density[2 * o + 1] += w2;
      }
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
