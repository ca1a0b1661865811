// Internal declarations of the MI355X selected-basis diagonalization library.
//
// Conventions used on the device (the C-ABI converts at its boundary):
//  * alpha strings index rows, beta strings columns: a wavefunction is row-major
//    [ia][ib], beta fastest, exactly the reference layout (tpb/qcham.h:359, mult.h:487);
//  * determinants are phased in the separated order (all alpha creators, ascending, then
//    all beta creators, ascending).  The reference phases them in interleaved order
//    (alpha p -> spin-orbital 2p, beta q -> 2q+1; determinants.h:96-142).  The two differ
//    by the diagonal sign s(A,B) = (-1)^(sum over p in A of #{q in B : q < p}), so
//    H_ref = S H_sep S, sigma_ref(c) = S sigma_sep(S c), and every matrix element keeps
//    its magnitude.  In the separated order an alpha-alpha element depends on the beta
//    string only through the beta occupation (singles) or not at all (doubles).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <vector>

#include "../../include/sbdg.h"

#define SBDG_CHECK(expr)                                                         \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) throw sbdg::Error(SBDG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

namespace sbdg {

struct Error {
  int code;
  std::string msg;
  Error(int c, std::string m) : code(c), msg(std::move(m)) {}
};

// Device bytes held by the library's buffers (all devices of the process) and their
// high-water mark since the last reset; sbdg_stats reports them.
struct MemCount {
  std::atomic<int64_t> now{0}, peak{0};
  void add(int64_t b) {
    const int64_t v = now.fetch_add(b) + b;
    int64_t p = peak.load();
    while (v > p && !peak.compare_exchange_weak(p, v)) {
    }
  }
  void reset_peak() { peak.store(now.load()); }
};
inline MemCount& mem_count() {
  static MemCount m;
  return m;
}

// Device buffer owned by the context.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    if (count == n && p) return;
    free();
    if (count) {
      SBDG_CHECK(hipMalloc(&p, sizeof(T) * count));
      n = count;
      mem_count().add((int64_t)(sizeof(T) * count));
    }
  }
  void free() {
    if (p) {
      (void)hipFree(p);
      mem_count().add(-(int64_t)(sizeof(T) * n));
    }
    p = nullptr;
    n = 0;
  }
  ~DBuf() { free(); }
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
};

// Excitation table of one string set, CSR with ket indices ascending per bra row
// (membership and order of chemistry/tpb/helper.h:132-197).
//  exc word (singles): bits 0-5 ket-only orbital i, 6-11 bra-only orbital a,
//                      bit 24 set when the separated-order sign is negative.
//  exc word (doubles): bits 0-5 I, 6-11 J (ket-only, I<J), 12-17 A, 18-23 B (bra-only,
//                      A<B), bit 24 sign.
//  val: the spin-string-only part of the element in the separated order:
//       singles  sign * (h_ai + sum_{k in occ(ket)\{i}} [(ai|kk) - (ak|ki)])
//       doubles  sign * ((AI|BJ) - (AJ|BI))
struct Table {
  int64_t n = 0, nnz = 0;
  DBuf<int64_t> rowptr;  // n+1
  DBuf<int32_t> col;     // nnz
  DBuf<uint32_t> exc;    // nnz
  DBuf<double> val;      // nnz
  DBuf<float> val32;     // nnz
};

// Row-merged (singles then doubles) table feeding the row-streaming product.
struct SpmmTable {
  int64_t n = 0, nnz = 0;
  DBuf<int64_t> rowptr;
  DBuf<int32_t> col;
  DBuf<double> val;
  DBuf<float> val32;
  // other-spin occupation term of a single: +-(p + 1) with p = a*L + i its orbital pair
  // and the sign of the single; 0 for a double.  The product adds sign * J[p + 1][c] per
  // column (J: occupation table of the column strings, row 0 zero).
  DBuf<int32_t> jai;
};

// Packed beta singles for the alpha-beta RDM kernel: bits 0-17 ket index jb, 18-29
// orbital pair b*L+j, bit 30 sign.  cptr[ib*(nch+1)+k] is the first entry of row ib whose
// ket lies in column chunk k (entries ascend in jb, so chunks are contiguous runs).
struct AbTable {
  int nch = 0;
  int64_t chunk = 0;
  DBuf<uint32_t> pack;
  DBuf<int32_t> cptr;
};

// Beta singles in an ELL layout for the alpha-beta sigma kernel.  Columns are cut into
// tiles of nt*m lane positions (one workgroup each); in ket chunk k, colperm[k*npos + pos]
// is the column at lane position pos (-1: none).  Per chunk, a tile's columns are sorted by
// their entry count in that chunk and cut into 64-lane groups (lists of similar length);
// wave w of tile t walks the groups at slots g = t*W*m + w + j*W, j < m (W = nt/64), and
// the groups are dealt to the slots longest first, each to the wave with the fewest rows
// so far, so the waves of a workgroup reach its barriers together.  Per chunk k a wave's
// groups form one stream starting at soff[(t*W + w)*nch + k], group after group, each a
// multiple of four rows of 64 lane-interleaved entries (nq[stream*m + j] batches of four
// rows) padded with a zero-weight entry: the wave reads its stream coalesced, four rows
// at a time, with the next four rows in flight across group boundaries.
// Entry: bits 0-17 LDS byte address of the ket vector (16 * (kbase + column offset inside
// the chunk); offset `chunk` is a zero vector: padding), bits 18-29 the slice vector of
// orbital pair (b, j) (ab_pair(b, j)), bit 31 sign.  Within a lane the entries are ordered
// so that lanes read distinct LDS slots where possible.
struct EllTable {
  int nch = 0;
  int nt = 0, m = 0;        // alpha-beta launch: threads per workgroup, columns per thread
  int64_t chunk = 0, groups = 0, ntiles = 0;
  int64_t kbase = 0;        // first ket vector (16-byte units): after the slice vectors
  DBuf<int64_t> soff;       // ntiles*(nt/64)*nch + 1: start of each wave's entry stream per chunk
  DBuf<int32_t> nq;         // ntiles*(nt/64)*nch*m: four-row batches of each group of the stream
  DBuf<uint32_t> ent;
  DBuf<int32_t> colperm;    // nch * groups*64: column of each lane position in each chunk (-1: none)
  void release() {
    soff.free();
    nq.free();
    ent.free();
    colperm.free();
    nch = nt = m = 0;
  }
};

// Double excitations of one spin grouped by their (n-2)-electron core: a double pair
// (x, y) shares exactly one core k = x & y (two electrons fewer than x), and its element
// in the separated order is eps_x * eps_y * W[P_x][P_y], with P_x = (a < b) the pair added
// to k to make x, eps_x = <x| c+_a c+_b |k> and W[(A,B)][(I,J)] = (AI|BJ) - (AJ|BI)
// (zero when the pairs overlap).  Members are stored core by core (only cores with at
// least CORE_KMIN members); row slots list, per string, the member entries it owns.  The
// pairs of the smaller cores (two members and up) are cheaper read directly: they go to a
// CSR table of their own that the row-streaming product merges with the singles.
struct CoreTable {
  int64_t ncore = 0, nmem = 0, nblk = 0;
  int maxm = 0;
  DBuf<int32_t> start;    // ncore + 1
  DBuf<int32_t> row;      // nmem
  DBuf<int32_t> pair;     // nmem: P(a,b) = b(b-1)/2 + a, bit 31 set when eps is negative
  DBuf<int32_t> blk;      // nblk: core of each product block
  DBuf<int32_t> blk_x0;   // nblk: first member of the block
  DBuf<int32_t> slot_ptr; // n + 1
  DBuf<int32_t> slot;     // nmem
  int64_t ngrp = 0;
  DBuf<int32_t> grp;      // ngrp + 1: product blocks of each workgroup
};
constexpr int CORE_XB = 16;     // output members per product block
constexpr int CORE_COLS = 512;  // column granularity of the core-product passes
constexpr int CORE_WORK = 256;  // member rows streamed per core-product workgroup (target)
// Smallest core taken through the workspace: a core of k members costs about 3k row
// transfers there (member rows in, Z out, Z gathered) and at most k(k-1) read directly,
// fewer where members' pairs overlap, and the direct reads of one core's member rows hit
// in cache.  FP64 sigma at norb=36, 20,000^2 strings with the cut at 2 (all through the
// workspace) / 4 / 6 / 8 / 10 / 12 / 17 members: 204.2 / 198.9 / 194.8 / 189.9 / 190.5 /
// 191.9 / 193.0 ms.
constexpr int CORE_KMIN = 10;

// kernels (tables.hip)
void build_table(const uint64_t* d_str, int64_t n, int L, const double* d_h1, const double* d_eri,
                 Table& t_single, Table& t_double, hipStream_t st);
void build_spmm(const Table& s, const Table& d, SpmmTable& m, int L, hipStream_t st, bool with_doubles = true);
// h_eri: the integrals on the host (full L^4), for the values of the direct pairs
void build_cores(const uint64_t* d_str, int64_t n, int L, int nel, const double* h_eri, CoreTable& ct,
                 Table& direct, hipStream_t st);
// J[(p+1)*n + r] = sum_{o in occ(s[r])} eri[p*L^2 + o*L + o] for p < L^2 (row 0 zero):
// the occupation part of the other spin's singles, (ai|oo) summed over the string's orbitals
void build_occ_table(const uint64_t* d_str, int64_t n, int L, const double* d_eri, double* J, float* J32,
                     hipStream_t st);
// W[P*npair + Q] (FP64 and FP32), npair = L(L-1)/2
void build_pair_integrals(const double* d_eri, int L, double* W, float* W32, hipStream_t st);
void build_ab(const Table& sb, int L, int64_t chunk, AbTable& ab, hipStream_t st);
// prec_bytes: 8 (FP64) or 4 (FP32) -- the launch geometry depends on the accumulator size
void build_ell(const Table& sb, int L, int64_t chunk, int prec_bytes, EllTable& e, hipStream_t st);
void ab_geometry(int64_t nb, int prec_bytes, int& nt, int& m);
void compute_hii(const uint64_t* d_a, int64_t na, const uint64_t* d_b, int64_t nb, int L, double core,
                 const double* d_h1, const double* d_eri, double* d_hii, hipStream_t st);

// sigma (sigma.hip)
struct SigmaPlan {
  int64_t na = 0, nb = 0;
  int L = 0;
  const uint64_t* a = nullptr;
  const uint64_t* b = nullptr;
  const SpmmTable* ma = nullptr;  // alpha merged
  const SpmmTable* mb = nullptr;  // beta merged
  const Table* sa = nullptr;      // alpha singles (exc, col)
  const EllTable* ell = nullptr;
  const double* eri = nullptr;
  const float* eri32 = nullptr;
  const double* hii = nullptr;
  const float* hii32 = nullptr;
  const double* Jb = nullptr;  // beta-occupation table (columns of c), alpha singles
  const float* Jb32 = nullptr;
  const double* Ja = nullptr;  // alpha-occupation table (columns of c^T), beta singles
  const float* Ja32 = nullptr;
  const CoreTable* ca = nullptr;  // alpha doubles by core (null: doubles in ma/mb)
  const CoreTable* cb = nullptr;
  const double* W = nullptr;
  const float* W32 = nullptr;
  int64_t npair = 0;
  void* Z = nullptr;    // core-product workspace of zbytes
  size_t zbytes = 0;
  hipStream_t st2 = nullptr;  // when set: the alpha-beta terms run there, beside the spin phases
  hipEvent_t fork = nullptr, join = nullptr;
};
// sigma = H c in the separated convention.  Scratch: three vectors of na*nb elements of T
// (c transposed, the beta product on it, the alpha-beta terms when they run on the second
// stream), any device memory the caller does not use while sigma runs; none may overlap
// c or sigma.
template <class T>
struct SigmaScratch {
  T* cT = nullptr;
  T* sT = nullptr;
  T* sab = nullptr;
};
template <class T>
void sigma_apply(const SigmaPlan& P, const T* c, T* sigma, const SigmaScratch<T>& scr, hipStream_t st,
                 hipEvent_t* ev = nullptr);
size_t ab_lds_bytes(int64_t kbase, int64_t chunk);

// The alpha-beta kernel's LDS in 16-byte vectors: one integral slice vector per orbital pair
// b > j at ab_pair(b, j) = b(b-1)/2 + j (a single moves an electron between two different
// orbitals, and (ai|bj) = (ai|jb) for the 8-fold symmetric integrals the library takes), at
// least one vector; then the ket vectors of one chunk.
inline int64_t ab_pairs(int L) {
  return std::max<int64_t>(1, (int64_t)L * (L - 1) / 2);
}
inline uint32_t ab_pair(uint32_t b, uint32_t j) {
  const uint32_t h = b > j ? b : j, l = b > j ? j : b;
  return h * (h - 1) / 2 + l;
}
// Ket chunk of the RDM's alpha-beta kernel: the chunk's FP64 values and the L*L
// accumulators of one orbital pair fill a CU's 160 KiB of LDS.
inline int64_t rdm_chunk(int64_t nb, int L) {
  const int64_t ch = std::max<int64_t>(256, (int64_t)(160 * 1024 / 8) - (int64_t)L * L - 8);
  return std::min<int64_t>(std::max<int64_t>(nb, 1), ch);
}
inline int64_t ab_slice_span(int L) {
  return ab_pairs(L);
}

// vector kernels (linalg.hip)
void sign_transform(const uint64_t* a, int64_t na, const uint64_t* b, int64_t nb, const double* in,
                    double* out, hipStream_t st);
void cast_d2f(const double* in, float* out, int64_t n, hipStream_t st);
void fill_test_vector(double* x, int64_t n, uint64_t seed, hipStream_t st);
void cast_f2d(const float* in, double* out, int64_t n, hipStream_t st);
// Basis vectors are contiguous: X_k = X + k*ldx.
// out[k] = <X_k, y> for k < m (deterministic two-stage reduction); y FP64 or FP32.
void multi_dot_impl(const double* X, int64_t ldx, int m, const double* y, int64_t n, double* out, double* part,
                    hipStream_t st);
void multi_dot_rows_impl(const float* Y, int64_t ldy, int m, const double* x, const double* z, int64_t n,
                         double* out, double* part, hipStream_t st);
void multi_dot_impl(const double* X, int64_t ldx, int m, const float* y, int64_t n, double* out, double* part,
                    hipStream_t st);
// W = sum_k u[k] V_k ; R = sum_k u[k] HV_k - theta W ; out[0] = |W|^2, out[1] = |R|^2
template <class TH>
void ritz_combine_impl(const double* V, int64_t ldv, const TH* HV, int64_t ldh, int m, const double* u,
                       const double* theta, double* W, double* R, int64_t n, double* out, double* part,
                       hipStream_t st);
// fused Ritz step (m <= 32): out[0..m) = V_k . t, out[m] = |r|^2, out[m+1] = |t|^2, t -> T
template <class TH>
void ritz_expand_impl(const double* V, int64_t ldv, const TH* HV, int64_t ldh, int m, const double* u,
                      const double* theta, const double* hii, double* T, int64_t n, double* out, double* part,
                      hipStream_t st);
void scale_impl(double* x, const double* n2, int64_t n, hipStream_t st);  // x /= sqrt(n2[0])
// x /= sqrt(n2[0]) and y = (float)x in the same pass
void scale_cast_impl(double* x, const double* n2, float* y, int64_t n, hipStream_t st);
// out[k] = <Y_k, x> for k < m (Y FP32, Y_k = Y + k*ldy), and out[m] = <z, x> when z is set:
// one pass that reads x once
// t = (R/|R|) / (theta - d), with the 1e-12 regularisation of davidson.h:855-862
void precond_impl(const double* R, const double* rn2, const double* hii, const double* theta, double* t, int64_t n,
                  hipStream_t st);
// t -= sum_k s[k] V_k; out[0] = |t|^2
void gs_impl(const double* V, int64_t ldv, int m, const double* s, double* t, int64_t n, double* out, double* part,
             hipStream_t st);
// in place: X_j = sum_k U[k + ld*j] X_k for j < mk (thick-restart rotation)
template <class T>
void rotate_impl(T* X, int64_t ldx, int m, const double* U, int ld, int mk, int64_t n, hipStream_t st);
// symmetric eigensolve of the m x m column-major d_H (ld) on the device: ascending
// eigenvalues in d_E, eigenvectors in the columns of d_U (ld).
void syev_device(const double* d_H, int m, int ld, double* d_E, double* d_U, hipStream_t st);
void density(const double* W, const uint64_t* a, int64_t na, const uint64_t* b, int64_t nb, int L,
             double* d_dens, double* d_part, hipStream_t st);
// d_rs[ia] = sum_ib W[ia,ib]^2 (fixed-order block sums)
void row_sumsq(const double* W, int64_t na, int64_t nb, double* d_rs, hipStream_t st);

// RDMs (rdm.hip)
struct RdmPlan {
  int64_t na = 0, nb = 0;
  int L = 0;
  const uint64_t* a = nullptr;
  const uint64_t* b = nullptr;
  const Table* sa = nullptr;
  const Table* da = nullptr;
  const Table* sb = nullptr;
  const Table* db = nullptr;
  const AbTable* ab = nullptr;
  const CoreTable* ca = nullptr;  // larger cores of the same-spin doubles (row-pair sums by block), or null
  const CoreTable* cb = nullptr;
};
// one[2][L^2], two[4][L^4] in the reference layout (correlation.h:194-444) from the
// separated-convention vector W; WT: scratch of na*nb doubles (W transposed).  Bitwise
// reproducible: every sum is taken in a fixed order or as exact integer (fixed-point) adds.
void rdm_compute(const RdmPlan& P, const double* W, double* WT, double* d_one, double* d_two, hipStream_t st);
// Hellmann-Feynman gradient: dh[p*L+q] = g*(one_a+one_b)[p+L*q];
// deri[((p*L+q)*L+r)*L+s] = g/2 * sum_blk two[blk][p + L*r + L^2*q + L^3*s]
void rdm_backward(const double* d_one, const double* d_two, int L, double g, double* d_dh,
                  double* d_deri, hipStream_t st);

}  // namespace sbdg
