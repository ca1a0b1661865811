// C-ABI of libsbdg and the device-resident Davidson driver.
//
// The driver follows Davidson(...,max_time) (tpb/davidson.h:661-924): lowest root,
// diagonal preconditioner on the normalised residual with the 1e-12 regularisation,
// Gram-Schmidt against the basis, convergence on |HW - E W| < eps, the max_time guard.
// Differences, all numerically neutral: Gram-Schmidt is classical (each pass one
// multi-dot and one update; a second pass only when the first removed more than half of
// the vector) instead of one modified pass; at a restart
// the kept Ritz vectors carry their sigma (HV rotated with the same coefficients), so
// the sigma of the restarted vector is not recomputed; restart_keep > 1 keeps that many
// Ritz vectors (thick restart).  Every vector stays in HBM; the host sees only the
// subspace scalars.
//
// Mixed precision (options.precision = 1): basis vectors, Gram-Schmidt, the subspace
// eigensolve and all residual bookkeeping stay FP64; the sigma of each new basis vector
// is built in FP32 (FP32 vectors, integrals and per-row accumulation).  Every refresh_k
// steps the basis is rotated onto the Ritz vectors of the subspace (no vector dropped;
// a full basis restarts to restart_keep as in the FP64 path), so the current Ritz
// vector x is basis vector 0, and its sigma is rebuilt in FP64.  That gives the exact
// Rayleigh quotient and residual of x (the only quantities used for convergence and
// reported), an exact first row/column of the projected matrix, and the exact residual
// drives the next correction vector.  With refresh_k <= 0 the refreshes follow the
// residual instead: one when the FP32 estimate has fallen tenfold since the last exact
// residual, or is below eps; a full basis then restarts on its FP32 sigma alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <random>
#include <string>
#include <unordered_map>

#include "sbdg_internal.h"

using namespace sbdg;


namespace {
constexpr int MAXSUB = 48;
constexpr size_t LDS_BUDGET = 160 * 1024;  // alpha-beta kernel: all of a CU's LDS, one 1024-thread workgroup

// Scratch of sbdg_sigma_device, the sigma entry point that allocates nothing: one array
// per device, held from the library's load until the process ends.  On the host the
// array's name stands for no device address: sbdg_create resolves each device's copy once
// (hipGetSymbolAddress, with that device current) and every use goes through the resolved
// address.  FP64 sigma uses 4 vectors of na*nb doubles of it, FP32 5 vectors of floats.
constexpr size_t SCRATCH_BYTES = (size_t)256 << 20;
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

struct ScratchTable {
  std::mutex mu;
  void* ptr[SCRATCH_DEVICES] = {};
  hipError_t rc[SCRATCH_DEVICES] = {};
};
ScratchTable& scratch_table() {
  static ScratchTable t;
  return t;
}

// The device address of `scratch` on `device` (which must be current), resolved on the
// first call for that device; nullptr when the runtime cannot resolve it.
void* scratch_address(int device) {
  if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
  ScratchTable& t = scratch_table();
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.ptr[device] && t.rc[device] == hipSuccess) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) t.ptr[device] = p;
    else t.rc[device] = e == hipSuccess ? hipErrorInvalidSymbol : e;
  }
  return t.ptr[device];
}
}  // namespace

struct sbdg_ctx {
  int device = 0;
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;  // alpha-beta kernel, concurrent with the spin phases
  hipEvent_t fork = nullptr, join = nullptr;
  std::string err;
  sbdg_options opt{};
  bool have_ints = false, have_strs = false, built = false, solved = false, have_rdm = false,
       have_dens = false, have_user_vec = false;
  int L = 0;
  double core = 0.0;
  std::vector<double> h1, eri;  // eri full L^4
  std::vector<uint64_t> a, b;
  std::vector<double> user_vec;
  DBuf<double> d_h1, d_eri, d_hii;
  DBuf<float> d_eri32, d_hii32;
  DBuf<uint64_t> d_a, d_b;
  Table Sa, Da, Sb, Db;
  SpmmTable Ma, Mb;
  CoreTable Ca, Cb;       // same-spin doubles by core (used when use_ca / use_cb)
  Table Xa, Xb;           // the pairs of the small cores, read directly (with Ca / Cb)
  bool use_ca = false, use_cb = false;
  DBuf<double> Ja, Jb;    // occupation tables of the alpha / beta strings
  DBuf<float> Ja32, Jb32;
  DBuf<double> Wp;        // antisymmetrised pair integrals
  DBuf<float> Wp32;
  DBuf<double> Zbuf;      // core-product workspace
  AbTable ab64;     // beta singles by chunk (RDMs)
  EllTable ell;      // beta singles, ELL (alpha-beta sigma, FP64; FP32 too when the geometries agree)
  EllTable ell32;    // the FP32 ELL when its launch geometry differs (wide rows)
  bool own_ell32 = false;
  // solver state
  DBuf<double> V, HV, W, R, T, HX, scr64, small;
  DBuf<float> HV32, c32, s32, scr32;
  DBuf<double> d_one, d_two, d_dens;
  DBuf<double> bx64, by64;  // benchmark vectors (filled once on the device)
  DBuf<float> bx32, by32;
  sbdg_stats stats{};
  std::vector<double> traj_t, traj_e;
  double energy = 0.0;
  void* scratch = nullptr;  // this device's copy of the static scratch (sbdg_sigma_device)
};

namespace {

int fail(sbdg_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

template <class F>
int guard(sbdg_ctx* c, F&& f) {
  if (!c) return SBDG_ERR_ARG;
  try {
    c->err.clear();
    SBDG_CHECK(hipSetDevice(c->device));
    return f();
  } catch (const Error& e) {
    return fail(c, e.code, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(c, SBDG_ERR_HIP, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(c, SBDG_ERR_ARG, e.what());
  }
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// column chunk of the alpha-beta kernel: ket-row chunks and the integral slices of 16-byte
// row vectors (2 FP64 or 4 FP32 rows) in LDS
int64_t choose_chunk(int64_t nb, int L) {
  const int64_t budget = (int64_t)(LDS_BUDGET / 16) - ab_slice_span(L) - 1;
  int64_t ch = std::max<int64_t>(256, budget);
  return std::min<int64_t>(std::max<int64_t>(nb, 1), ch);
}

SigmaPlan plan(sbdg_ctx* c, bool fp32) {
  SigmaPlan P;
  P.na = (int64_t)c->a.size();
  P.nb = (int64_t)c->b.size();
  P.L = c->L;
  P.a = c->d_a.p;
  P.b = c->d_b.p;
  P.ma = &c->Ma;
  P.mb = &c->Mb;
  P.sa = &c->Sa;
  P.ell = (fp32 && c->own_ell32) ? &c->ell32 : &c->ell;
  P.eri = c->d_eri.p;
  P.eri32 = c->d_eri32.p;
  P.hii = c->d_hii.p;
  P.hii32 = c->d_hii32.p;
  P.Ja = c->Ja.p;
  P.Ja32 = c->Ja32.p;
  P.Jb = c->Jb.p;
  P.Jb32 = c->Jb32.p;
  P.ca = c->use_ca ? &c->Ca : nullptr;
  P.cb = c->use_cb ? &c->Cb : nullptr;
  P.W = c->Wp.p;
  P.W32 = c->Wp32.p;
  P.npair = (int64_t)c->L * (c->L - 1) / 2;
  P.Z = c->Zbuf.p;
  P.zbytes = c->Zbuf.n * sizeof(double);
  P.st2 = c->opt.overlap ? c->st2 : nullptr;
  P.fork = c->fork;
  P.join = c->join;
  return P;
}

void build(sbdg_ctx* c) {
  if (!c->have_ints) throw Error(SBDG_ERR_STATE, "integrals not set");
  if (!c->have_strs) throw Error(SBDG_ERR_STATE, "strings not set");
  const double t0 = now();
  const int L = c->L;
  const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
  const int64_t L4 = (int64_t)L * L * L * L;
  c->d_h1.alloc(L * L);
  c->d_eri.alloc(L4);
  c->d_eri32.alloc(L4);
  SBDG_CHECK(hipMemcpyAsync(c->d_h1.p, c->h1.data(), sizeof(double) * L * L, hipMemcpyHostToDevice, c->st));
  SBDG_CHECK(hipMemcpyAsync(c->d_eri.p, c->eri.data(), sizeof(double) * L4, hipMemcpyHostToDevice, c->st));
  cast_d2f(c->d_eri.p, c->d_eri32.p, L4, c->st);
  c->d_a.alloc(std::max<int64_t>(na, 1));
  c->d_b.alloc(std::max<int64_t>(nb, 1));
  SBDG_CHECK(hipMemcpyAsync(c->d_a.p, c->a.data(), sizeof(uint64_t) * na, hipMemcpyHostToDevice, c->st));
  SBDG_CHECK(hipMemcpyAsync(c->d_b.p, c->b.data(), sizeof(uint64_t) * nb, hipMemcpyHostToDevice, c->st));
  build_table(c->d_a.p, na, L, c->d_h1.p, c->d_eri.p, c->Sa, c->Da, c->st);
  build_table(c->d_b.p, nb, L, c->d_h1.p, c->d_eri.p, c->Sb, c->Db, c->st);
  // same-spin doubles: excitation table (as mult) or grouped by core.  Auto picks cores
  // when the table would stream more than four times the member rows a core pass moves.
  const int nel_a = na ? __builtin_popcountll(c->a[0]) : 0, nel_b = nb ? __builtin_popcountll(c->b[0]) : 0;
  c->use_ca = c->use_cb = false;
  if (c->opt.doubles != 1) {
    build_cores(c->d_a.p, na, L, nel_a, c->eri.data(), c->Ca, c->Xa, c->st);
    build_cores(c->d_b.p, nb, L, nel_b, c->eri.data(), c->Cb, c->Xb, c->st);
    // auto: the core form (its workspace for the larger cores, the pairs of the smaller
    // ones read directly) where it moves fewer rows and one 512-column pass of its
    // workspace fits core_gib; otherwise the excitation-table form
    const size_t cap = (size_t)(c->opt.core_gib * 1073741824.0);
    auto fits = [&](const CoreTable& ct) { return (size_t)ct.nmem * CORE_COLS * sizeof(double) <= cap; };
    c->use_ca = c->opt.doubles == 2 || (c->Da.nnz > 4 * c->Ca.nmem + c->Xa.nnz && fits(c->Ca));
    c->use_cb = c->opt.doubles == 2 || (c->Db.nnz > 4 * c->Cb.nmem + c->Xb.nnz && fits(c->Cb));
  }
  // row products: the singles with the doubles of the excitation table, or with the direct
  // pairs of the small cores when the doubles go by core
  build_spmm(c->Sa, c->use_ca ? c->Xa : c->Da, c->Ma, L, c->st, !c->use_ca || c->Xa.nnz > 0);
  build_spmm(c->Sb, c->use_cb ? c->Xb : c->Db, c->Mb, L, c->st, !c->use_cb || c->Xb.nnz > 0);
  // occupation parts of the singles: beta occupations for the alpha product (columns of
  // c), alpha occupations for the beta product (columns of c^T)
  const int64_t LL1 = (int64_t)L * L + 1;
  c->Jb.alloc(std::max<int64_t>(LL1 * nb, 1));
  c->Jb32.alloc(std::max<int64_t>(LL1 * nb, 1));
  c->Ja.alloc(std::max<int64_t>(LL1 * na, 1));
  c->Ja32.alloc(std::max<int64_t>(LL1 * na, 1));
  build_occ_table(c->d_b.p, nb, L, c->d_eri.p, c->Jb.p, c->Jb32.p, c->st);
  build_occ_table(c->d_a.p, na, L, c->d_eri.p, c->Ja.p, c->Ja32.p, c->st);
  if (c->use_ca || c->use_cb) {
    const int64_t npair = (int64_t)L * (L - 1) / 2;
    c->Wp.alloc(std::max<int64_t>(npair * npair, 1));
    c->Wp32.alloc(std::max<int64_t>(npair * npair, 1));
    build_pair_integrals(c->d_eri.p, L, c->Wp.p, c->Wp32.p, c->st);
    auto need = [](const CoreTable& ct, int64_t ncols) {
      return (size_t)ct.nmem * (size_t)((ncols + CORE_COLS - 1) / CORE_COLS * CORE_COLS) * sizeof(double);
    };
    auto least = [](const CoreTable& ct) { return (size_t)ct.nmem * CORE_COLS * sizeof(double); };
    size_t want = 0, low = 0;
    if (c->use_ca) want = std::max(want, need(c->Ca, nb)), low = std::max(low, least(c->Ca));
    if (c->use_cb) want = std::max(want, need(c->Cb, na)), low = std::max(low, least(c->Cb));
    const size_t cap = (size_t)(c->opt.core_gib * 1073741824.0);
    if (low > cap) throw Error(SBDG_ERR_LIMIT, "core_gib below one 512-column pass of the core products");
    c->Zbuf.alloc(std::max<size_t>(std::min(want, cap) / sizeof(double), 1));
  } else {
    c->Zbuf.free();
  }
  build_ab(c->Sb, L, rdm_chunk(nb, L), c->ab64, c->st);
  build_ell(c->Sb, L, choose_chunk(nb, L), 8, c->ell, c->st);
  {
    int nt64, m64, nt32, m32;
    ab_geometry(nb, 8, nt64, m64);
    ab_geometry(nb, 4, nt32, m32);
    c->own_ell32 = nt64 != nt32 || m64 != m32;
    if (c->own_ell32) build_ell(c->Sb, L, choose_chunk(nb, L), 4, c->ell32, c->st);
    else c->ell32.release();
  }
  c->d_hii.alloc(std::max<int64_t>(N, 1));
  c->d_hii32.alloc(std::max<int64_t>(N, 1));
  compute_hii(c->d_a.p, na, c->d_b.p, nb, L, c->core, c->d_h1.p, c->d_eri.p, c->d_hii.p, c->st);
  cast_d2f(c->d_hii.p, c->d_hii32.p, N, c->st);
  SBDG_CHECK(hipStreamSynchronize(c->st));
  c->built = true;
  c->solved = c->have_rdm = c->have_dens = false;
  c->stats = sbdg_stats{};
  c->stats.t_tables = now() - t0;
  c->stats.nnz[0] = c->Sa.nnz;
  c->stats.nnz[1] = c->Da.nnz;
  c->stats.nnz[2] = c->Sb.nnz;
  c->stats.nnz[3] = c->Db.nnz;
  c->stats.core_members[0] = c->use_ca ? c->Ca.nmem : 0;
  c->stats.core_members[1] = c->use_cb ? c->Cb.nmem : 0;
  c->stats.core_direct[0] = c->use_ca ? c->Xa.nnz : 0;
  c->stats.core_direct[1] = c->use_cb ? c->Xb.nnz : 0;
  c->stats.terms = N + (c->Sa.nnz + c->Da.nnz) * nb + (c->Sb.nnz + c->Db.nnz) * na + c->Sa.nnz * c->Sb.nnz;
}

void ensure_built(sbdg_ctx* c) {
  if (!c->built) build(c);
}

// sigma with explicit scratch vectors (the Davidson loop lends it vectors that are idle
// while sigma runs), or with the context's own scratch (standalone sigma and benchmarks)
void sigma64(sbdg_ctx* c, const double* x, double* y, const SigmaScratch<double>& scr, hipEvent_t* ev = nullptr) {
  sigma_apply<double>(plan(c, false), x, y, scr, c->st, ev);
}

void sigma32(sbdg_ctx* c, const float* x, float* y, const SigmaScratch<float>& scr, hipEvent_t* ev = nullptr) {
  sigma_apply<float>(plan(c, true), x, y, scr, c->st, ev);
}

SigmaScratch<double> own_scratch64(sbdg_ctx* c) {
  const int64_t N = (int64_t)c->a.size() * c->b.size();
  c->scr64.alloc(std::max<int64_t>(3 * N, 1));
  return {c->scr64.p, c->scr64.p + N, c->scr64.p + 2 * N};
}

SigmaScratch<float> own_scratch32(sbdg_ctx* c) {
  const int64_t N = (int64_t)c->a.size() * c->b.size();
  c->scr32.alloc(std::max<int64_t>(3 * N, 1));
  return {c->scr32.p, c->scr32.p + N, c->scr32.p + 2 * N};
}

// Device vectors of the Davidson loop for N determinants (elements of 8 bytes unless
// stated): the basis V (nbmax), the sigma of every basis vector (HV, nbmax; in the mixed
// path HV32, nbmax FP32), W, R, T, and in the mixed path the anchor's FP64 sigma HX and the
// FP32 input of the next sigma c32 (FP32).  sigma's three scratch vectors are W, R and T
// (free while sigma runs), so they add nothing.
int64_t davidson_bytes(const sbdg_options& o, int64_t N) {
  const int64_t nbmax = std::max(2, std::min(o.max_nb, MAXSUB));
  if (o.precision == 1) return N * (nbmax * 8 + nbmax * 4 + 3 * 8 + 8 + 4);
  return N * (nbmax * 8 * 2 + 3 * 8);
}

// Resident device bytes after the build (an upper bound where a table's padding depends on
// the strings): integrals, strings, excitation tables, row-product tables, core tables and
// workspace, occupation tables, the alpha-beta tables, the diagonal.  nnz: alpha singles,
// alpha doubles, beta singles, beta doubles; mem/dir: core memberships and directly read
// pairs per spin (0 members: that spin's doubles go by excitation table).
int64_t build_bytes(const sbdg_options& o, int L, int64_t na, int64_t nb, const int64_t* nnz, const int64_t* mem,
                    const int64_t* dir) {
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2, N = na * nb, npair = (int64_t)L * (L - 1) / 2;
  int64_t b = 8 * L2 + 12 * L4 + 8 * (na + nb) + 12 * N;  // h1, eri FP64+FP32, strings, hii FP64+FP32
  bool any_core = false;
  int64_t zwant = 0;
  const int64_t ns[2] = {na, nb}, nc[2] = {nb, na};
  for (int sp = 0; sp < 2; ++sp) {
    const int64_t n = ns[sp], S = nnz[2 * sp], D = nnz[2 * sp + 1];
    const bool core = mem[sp] > 0;
    b += 2 * 8 * (n + 1) + 20 * (S + D);                          // singles and doubles tables
    b += 8 * (n + 1) + 20 * (S + (core ? dir[sp] : D));          // row-product table
    b += 12 * (L2 + 1) * n;                                        // occupation tables
    if (core) {
      any_core = true;
      b += 4 * (n + 1) + 28 * mem[sp] + 8 * (n + 1) + 16 * dir[sp];  // members, blocks, slots; direct pairs
      zwant = std::max(zwant, mem[sp] * ((nc[sp] + CORE_COLS - 1) / CORE_COLS * CORE_COLS) * 8);
    }
  }
  if (any_core) b += 12 * npair * npair + std::min<int64_t>(zwant, (int64_t)(o.core_gib * 1073741824.0));
  // alpha-beta tables: RDM pack and chunk pointers; ELL entries (a group pads to its longest
  // list, at most the group's whole sum, plus three rows), lane maps, stream offsets
  const int64_t Sb = nnz[2];
  const int64_t ach = rdm_chunk(nb, L);
  b += 4 * Sb + 4 * (nb * ((nb + ach - 1) / ach + 1) + 1);
  const int64_t nch = (nb + choose_chunk(nb, L) - 1) / choose_chunk(nb, L);
  const int64_t groups = (nb + 63) / 64 + 1024;  // tiles round the lane positions up
  int nt64, m64, nt32, m32;
  ab_geometry(nb, 8, nt64, m64);
  ab_geometry(nb, 4, nt32, m32);
  const int nell = (nt64 != nt32 || m64 != m32) ? 2 : 1;  // an FP32 ELL of its own (wide rows)
  b += nell * (4 * (64 * Sb + 192 * groups * nch) + 4 * nch * groups * 64 + 16 * groups * nch);
  return b;
}

// Transient bytes of the RDMs beyond the resident build and the vectors (which the RDMs
// borrow): the largest of the per-table passes, plus the outputs 2 L^2 + 4 L^4 doubles.
int64_t rdm_bytes(int L, const int64_t* nnz) {
  const int64_t L2 = (int64_t)L * L, L4 = L2 * L2;
  int64_t peak = 0;
  for (int sp = 0; sp < 2; ++sp) {
    const int64_t S = nnz[2 * sp], D = nnz[2 * sp + 1];
    peak = std::max(peak, 8 * S + 8 * S * L + 16 * S + 16 * S + 4 * (L2 + 1));  // gram, keys, sort
    peak = std::max(peak, 8 * D + 32 * D + 16 * D);                            // gram, keys, scan, sort
  }
  const int64_t Sa = nnz[0];
  peak = std::max(peak, 8 * Sa + 16 * Sa + 16 * Sa + 8 * (Sa / 32 + L2 + 1) * L2);  // alpha-beta partials
  return peak + 8 * (2 * L2 + 4 * L4);
}

struct Small {  // device scalars carved from one buffer
  double *H, *U, *E, *dots, *norms, *part;
};

Small carve(sbdg_ctx* c) {
  const size_t ms = MAXSUB * MAXSUB;
  c->small.alloc(4 * ms + 4 * MAXSUB + 64 + 1024 * (MAXSUB + 2));  // partials of up to 1024 reduction blocks
  Small s;
  s.H = c->small.p;
  s.U = s.H + ms;
  s.E = s.U + ms;
  s.dots = s.E + MAXSUB;
  s.norms = s.dots + MAXSUB;
  s.part = s.norms + 64;
  return s;
}

// The start vector, formed on the device in T (reference phases), normalised, then
// converted to the separated phases into W.  init 0: the first determinant (BasisInitVector,
// tpb/davidson.h:130-133); 1: values in [-1, 1) from a counter-based hash (splitmix64 of
// seed and index); 2: the caller's vector.
void initial_vector(sbdg_ctx* c, Small& s, double* W) {
  const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
  if (c->opt.init == 2) {
    if (!c->have_user_vec) throw Error(SBDG_ERR_STATE, "init=2 but no vector set");
    SBDG_CHECK(hipMemcpyAsync(c->T.p, c->user_vec.data(), sizeof(double) * N, hipMemcpyHostToDevice, c->st));
  } else if (c->opt.init == 1) {
    fill_test_vector(c->T.p, N, c->opt.seed, c->st);
  } else {
    const double one = 1.0;
    SBDG_CHECK(hipMemsetAsync(c->T.p, 0, sizeof(double) * N, c->st));
    SBDG_CHECK(hipMemcpyAsync(c->T.p, &one, sizeof(double), hipMemcpyHostToDevice, c->st));
  }
  multi_dot_impl(c->T.p, N, 1, c->T.p, N, s.dots, s.part, c->st);
  double nn = 0.0;
  SBDG_CHECK(hipMemcpyAsync(&nn, s.dots, sizeof(double), hipMemcpyDeviceToHost, c->st));
  SBDG_CHECK(hipStreamSynchronize(c->st));
  if (!(nn > 0.0) || !std::isfinite(nn)) throw Error(SBDG_ERR_ARG, "initial vector has zero or non-finite norm");
  scale_impl(c->T.p, s.dots, N, c->st);
  sign_transform(c->d_a.p, na, c->d_b.p, nb, c->T.p, W, c->st);  // reference phases -> separated
  SBDG_CHECK(hipStreamSynchronize(c->st));
}

// t = precondition(R) then classical Gram-Schmidt twice against V[0..m), normalised
// into V[m].  Returns |t| before the final normalisation.
// Gram-Schmidt of t = V[m] against V[0..m) given the first-pass dots in s.dots[0..m) and
// |t|^2 before projection at n0_dev; normalises t and returns |t| after projection (0 when
// nothing is left).  A second pass runs only when the first removed most of t
// (|t1| < |t0| / sqrt(2)): "twice is enough".
// y32 set: the normalised vector is also written there in FP32 (the next FP32 sigma's input)
double finish_gs(sbdg_ctx* c, Small& s, int m, int64_t N, const double* n0_dev, float* y32 = nullptr) {
  double* vm = c->V.p + (int64_t)m * N;
  double nt = 0.0, n0 = 0.0;
  gs_impl(c->V.p, N, m, s.dots, vm, N, s.norms + 8, s.part, c->st);
  SBDG_CHECK(hipMemcpyAsync(&n0, n0_dev, sizeof(double), hipMemcpyDeviceToHost, c->st));
  SBDG_CHECK(hipMemcpyAsync(&nt, s.norms + 8, sizeof(double), hipMemcpyDeviceToHost, c->st));
  SBDG_CHECK(hipStreamSynchronize(c->st));
  if (!(nt >= 0.5 * n0)) {
    multi_dot_impl(c->V.p, N, m, vm, N, s.dots, s.part, c->st);
    gs_impl(c->V.p, N, m, s.dots, vm, N, s.norms + 8, s.part, c->st);
    SBDG_CHECK(hipMemcpyAsync(&nt, s.norms + 8, sizeof(double), hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
  }
  if (!std::isfinite(nt)) throw Error(SBDG_ERR_NUMERIC, "non-finite correction vector");
  if (nt <= 0.0) return 0.0;
  if (y32) scale_cast_impl(vm, s.norms + 8, y32, N, c->st);
  else scale_impl(vm, s.norms + 8, N, c->st);
  return std::sqrt(nt);
}

// correction vector from the stored residual R: preconditioned, then Gram-Schmidt (the new
// vector sits at V[m], so one more column of the dot pass gives |t|^2 before projection)
double expand(sbdg_ctx* c, Small& s, int m, int64_t N, const double* theta_dev, const double* rn2_dev,
              float* y32 = nullptr) {
  double* vm = c->V.p + (int64_t)m * N;
  precond_impl(c->R.p, rn2_dev, c->d_hii.p, theta_dev, vm, N, c->st);
  multi_dot_impl(c->V.p, N, m + 1, vm, N, s.dots, s.part, c->st);
  return finish_gs(c, s, m, N, s.dots + m, y32);
}

void davidson(sbdg_ctx* c) {
  const sbdg_options& o = c->opt;
  const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
  const bool mixed = o.precision == 1;
  const int nbmax = std::max(2, std::min(o.max_nb, MAXSUB));
  const int keep = std::max(1, std::min(o.restart_keep, nbmax - 1));
  const int max_steps = std::max(1, o.max_it) * nbmax;
  // only the vectors this precision uses stay allocated (davidson_bytes); the standalone
  // sigma's scratch and the benchmark vectors are released first
  c->scr64.free();
  c->scr32.free();
  c->bx64.free();
  c->by64.free();
  c->bx32.free();
  c->by32.free();
  if (mixed) {
    c->HV.free();
  } else {
    c->HV32.free();
    c->HX.free();
    c->c32.free();
  }
  c->V.alloc((size_t)nbmax * N);
  c->W.alloc(N);
  c->R.alloc(N);
  c->T.alloc(N);
  if (mixed) {
    c->HV32.alloc((size_t)nbmax * N);
    c->HX.alloc(N);
    c->c32.alloc(N);
  } else {
    c->HV.alloc((size_t)nbmax * N);
  }
  // sigma's scratch: W, R and T hold nothing that is needed while a sigma runs (W is formed
  // again by the next Ritz step, R was consumed by the last expansion, T is a staging
  // vector); an FP32 sigma fits its three FP32 vectors into T and R
  const SigmaScratch<double> scr64{c->T.p, c->R.p, c->W.p};
  const SigmaScratch<float> scr32{reinterpret_cast<float*>(c->T.p), reinterpret_cast<float*>(c->T.p) + N,
                                  reinterpret_cast<float*>(c->R.p)};
  Small s = carve(c);
  // device scalar slots: norms[0] |W|^2, [1] |R|^2 (combine), [2] theta and [3] |R|^2 used
  // by the next expansion, [6] the constant 1, [8] |t|^2
  double* th_exp = s.norms + 2;
  double* rr_exp = s.norms + 3;
  double* one_dev = s.norms + 6;
  std::vector<double> Hs((size_t)MAXSUB * MAXSUB, 0.0), E(MAXSUB), dots(MAXSUB);
  hipEvent_t e0, e1;
  SBDG_CHECK(hipEventCreate(&e0));
  SBDG_CHECK(hipEventCreate(&e1));
  auto timed = [&](auto&& f) {
    SBDG_CHECK(hipEventRecord(e0, c->st));
    f();
    SBDG_CHECK(hipEventRecord(e1, c->st));
    SBDG_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    SBDG_CHECK(hipEventElapsedTime(&ms, e0, e1));
    c->stats.t_sigma += 1e-3 * ms;
  };
  auto h2d = [&](double* d, double x) {
    SBDG_CHECK(hipMemcpyAsync(d, &x, sizeof(double), hipMemcpyHostToDevice, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
  };

  {
    const sbdg_stats keep_st = c->stats;
    c->stats = sbdg_stats{};
    c->stats.t_tables = keep_st.t_tables;
    c->stats.terms = keep_st.terms;
    for (int k = 0; k < 4; ++k) c->stats.nnz[k] = keep_st.nnz[k];
    for (int k = 0; k < 2; ++k) c->stats.core_members[k] = keep_st.core_members[k];
    for (int k = 0; k < 2; ++k) c->stats.core_direct[k] = keep_st.core_direct[k];
  }
  initial_vector(c, s, c->V.p);
  h2d(one_dev, 1.0);
  const double tstart = now();
  c->traj_t.clear();
  c->traj_e.clear();
  int h = 0, v = 1, steps = 0, cycle = 0, since = 0;
  double theta = 0.0, rnorm = 0.0;
  bool converged = false, stop = false;
  std::vector<double> step_times;

  auto log_step = [&](bool exact) {
    const double el = now() - tstart;
    c->stats.iterations = ++steps;
    c->stats.ritz = theta;
    c->stats.resid = rnorm;
    if (exact) {
      c->traj_t.push_back(el);
      c->traj_e.push_back(theta);
    }
    if (o.verbose)
      printf(" Davidson step %d (basis %d, %s): E = %.15f |r| = %.3e  t = %.3f s\n", steps, h,
             exact ? "fp64" : "fp32 sigma", theta, rnorm, el);
    fflush(stdout);
    return el;
  };
  // projected-matrix eigenpair and Ritz vector/residual over the first m vectors
  bool w_valid = false;  // W holds the Ritz vector of the current projected matrix
  auto ritz = [&](int m) {
    SBDG_CHECK(hipMemcpyAsync(s.H, Hs.data(), sizeof(double) * MAXSUB * MAXSUB, hipMemcpyHostToDevice, c->st));
    syev_device(s.H, m, MAXSUB, s.E, s.U, c->st);
    if (mixed)
      ritz_combine_impl<float>(c->V.p, N, c->HV32.p, N, m, s.U, s.E, c->W.p, c->R.p, N, s.norms, s.part, c->st);
    else
      ritz_combine_impl<double>(c->V.p, N, c->HV.p, N, m, s.U, s.E, c->W.p, c->R.p, N, s.norms, s.part, c->st);
    double nr[2];
    SBDG_CHECK(hipMemcpyAsync(nr, s.norms, sizeof(double) * 2, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipMemcpyAsync(E.data(), s.E, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    if (!std::isfinite(nr[0]) || !std::isfinite(nr[1]) || !std::isfinite(E[0]))
      throw Error(SBDG_ERR_NUMERIC, "non-finite Ritz value or residual");
    scale_impl(c->W.p, s.norms, N, c->st);  // Normalize(W) (davidson.h:818-819)
    SBDG_CHECK(hipMemcpyAsync(th_exp, s.E, sizeof(double), hipMemcpyDeviceToDevice, c->st));
    SBDG_CHECK(hipMemcpyAsync(rr_exp, s.norms + 1, sizeof(double), hipMemcpyDeviceToDevice, c->st));
    theta = E[0];
    rnorm = std::sqrt(nr[1]);
    w_valid = true;
  };
  // a step that goes on to expand without restart: Ritz pair, residual norm and the
  // preconditioned correction with its first Gram-Schmidt dots in one pass (t -> V[m]);
  // the Ritz vector W is formed later only if needed (w_valid)
  auto ritz_fused = [&](int m) {
    SBDG_CHECK(hipMemcpyAsync(s.H, Hs.data(), sizeof(double) * MAXSUB * MAXSUB, hipMemcpyHostToDevice, c->st));
    syev_device(s.H, m, MAXSUB, s.E, s.U, c->st);
    double* tm = c->V.p + (int64_t)m * N;
    if (mixed)
      ritz_expand_impl<float>(c->V.p, N, c->HV32.p, N, m, s.U, s.E, c->d_hii.p, tm, N, s.dots, s.part, c->st);
    else
      ritz_expand_impl<double>(c->V.p, N, c->HV.p, N, m, s.U, s.E, c->d_hii.p, tm, N, s.dots, s.part, c->st);
    double rr = 0.0;
    SBDG_CHECK(hipMemcpyAsync(&rr, s.dots + m, sizeof(double), hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipMemcpyAsync(E.data(), s.E, sizeof(double) * m, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    if (!std::isfinite(rr) || !std::isfinite(E[0])) throw Error(SBDG_ERR_NUMERIC, "non-finite Ritz value or residual");
    theta = E[0];
    rnorm = std::sqrt(rr);
    w_valid = false;
  };
  // mixed: c32 holds basis vector c32_of in FP32 when the normalisation wrote it there
  int c32_of = -1;
  auto finish_fused = [&]() {
    const double nt = finish_gs(c, s, v, N, s.dots + v + 1, mixed ? c->c32.p : nullptr);
    c32_of = mixed ? v : -1;
    if (nt == 0.0) {
      stop = true;  // the residual lies in the span of the basis
      return;
    }
    ++v;
  };
  // mixed: FP64 sigma of basis vector 0 (= current Ritz vector): exact row/column 0 of
  // the projected matrix, exact Rayleigh quotient and residual
  auto refresh = [&]() {
    // W holds a copy of V_0 here (restored below), so it may serve as scratch
    timed([&] { sigma64(c, c->V.p, c->HX.p, scr64); });
    c->stats.sigma_fp64++;
    multi_dot_impl(c->V.p, N, v, c->HX.p, N, s.dots, s.part, c->st);
    SBDG_CHECK(hipMemcpyAsync(dots.data(), s.dots, sizeof(double) * v, hipMemcpyDeviceToHost, c->st));
    cast_d2f(c->HX.p, c->HV32.p, N, c->st);
    SBDG_CHECK(hipStreamSynchronize(c->st));
    for (int k = 0; k < v; ++k) Hs[k] = Hs[MAXSUB * k] = dots[k];
    theta = dots[0];
    h2d(th_exp, theta);
    ritz_combine_impl<double>(c->V.p, N, c->HX.p, N, 1, one_dev, th_exp, c->T.p, c->R.p, N, s.norms, s.part, c->st);
    double nr[2];
    SBDG_CHECK(hipMemcpyAsync(nr, s.norms, sizeof(double) * 2, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipMemcpyAsync(rr_exp, s.norms + 1, sizeof(double), hipMemcpyDeviceToDevice, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    if (!std::isfinite(nr[1]) || !std::isfinite(theta)) throw Error(SBDG_ERR_NUMERIC, "non-finite residual");
    rnorm = std::sqrt(nr[1]);
    SBDG_CHECK(hipMemcpyAsync(c->W.p, c->V.p, sizeof(double) * N, hipMemcpyDeviceToDevice, c->st));
    since = 0;
  };
  auto check_time = [&](double el, double ts) {
    step_times.push_back(now() - ts);
    double avg = 0.0;
    for (double x : step_times) avg += x;
    avg /= step_times.size();
    return el + avg > o.max_time;  // davidson.h:880-903
  };
  auto expand_or_stop = [&]() {
    if (v >= nbmax) return;
    const double nt = expand(c, s, v, N, th_exp, rr_exp, mixed ? c->c32.p : nullptr);
    c32_of = mixed ? v : -1;
    if (nt == 0.0) {
      stop = true;  // the residual lies in the span of the basis
      return;
    }
    ++v;
  };

  // mixed: the first FP64 sigma comes at the first refresh; until then the projected
  // matrix is all FP32-sigma entries
  bool have_anchor = false;
  double prev_exact = NAN;
  // mixed with refresh_k <= 0: refresh when the FP32 residual estimate has fallen by
  // ADAPT_DROP since the last exact residual (the first estimate before any), when it
  // is below eps (to confirm convergence), or when the FP32 Ritz value moved less than
  // ADAPT_DE in one step (the energy has settled: an exact one is worth its FP64 sigma);
  // a full basis restarts without a refresh.  FP32 Ritz values sit within ~1e-7 Ha of the
  // exact Rayleigh quotients on the benchmark workloads, so ADAPT_DE = 1e-6 Ha triggers only
  // once the energy has converged to about the accuracy asked of the mixed path.
  const bool adaptive = mixed && o.refresh_k <= 0;
  constexpr double ADAPT_DROP = 0.1;
  constexpr double ADAPT_DE = 1e-6;
  double r_ref = -1.0, theta32_prev = NAN;
  while (!converged && !stop && steps < max_steps) {
    const double ts = now();
    // sigma of the newest basis vector and column h of the projected matrix
    if (mixed) {
      timed([&] {
        if (c32_of != h) cast_d2f(c->V.p + (int64_t)h * N, c->c32.p, N, c->st);
        sigma32(c, c->c32.p, c->HV32.p + (int64_t)h * N, scr32);
      });
      c32_of = -1;
      c->stats.sigma_fp32++;
      // column h of the projected matrix from the rows of the (symmetric) operator,
      // <H V_k, V_h> with the FP32 sigma of every basis vector (V_h read once, the FP32
      // vectors at half the bytes), and the exact <V_h, H x> for the anchor row
      multi_dot_rows_impl(c->HV32.p, N, h + 1, c->V.p + (int64_t)h * N, have_anchor ? c->HX.p : nullptr, N,
                          s.dots, s.part, c->st);
      SBDG_CHECK(hipMemcpyAsync(dots.data(), s.dots, sizeof(double) * (h + 1), hipMemcpyDeviceToHost, c->st));
      if (have_anchor)
        SBDG_CHECK(hipMemcpyAsync(&dots[0], s.dots + h + 1, sizeof(double), hipMemcpyDeviceToHost, c->st));
    } else {
      timed([&] { sigma64(c, c->V.p + (int64_t)h * N, c->HV.p + (int64_t)h * N, scr64); });
      c->stats.sigma_fp64++;
      multi_dot_impl(c->V.p, N, h + 1, c->HV.p + (int64_t)h * N, N, s.dots, s.part, c->st);
      SBDG_CHECK(hipMemcpyAsync(dots.data(), s.dots, sizeof(double) * (h + 1), hipMemcpyDeviceToHost, c->st));
    }
    SBDG_CHECK(hipStreamSynchronize(c->st));
    for (int k = 0; k <= h; ++k) Hs[k + MAXSUB * h] = Hs[h + MAXSUB * k] = dots[k];
    ++h;
    ++since;
    bool do_refresh = mixed && !adaptive && (since >= o.refresh_k || v == nbmax);
    const bool fuse = v < nbmax && !do_refresh && h <= 32;  // expansion follows directly
    if (fuse) ritz_fused(h);
    else ritz(h);
    if (adaptive) {
      if (r_ref < 0.0) r_ref = rnorm;
      do_refresh = rnorm < o.eps || (since >= 2 && rnorm < ADAPT_DROP * r_ref) ||
                   (since >= 2 && std::fabs(theta - theta32_prev) < ADAPT_DE);
      theta32_prev = do_refresh ? NAN : theta;
      if (do_refresh && !w_valid) ritz(h);  // the fused step formed no Ritz vector
    }
    if (!mixed) {
      const double el = log_step(true);
      if (rnorm < o.eps) {
        converged = true;
        break;
      }
      if (check_time(el, ts)) break;
      if (v == nbmax) {
        // restart (davidson.h:912-916) keeping `keep` Ritz vectors with their sigma
        if (cycle + 1 >= std::max(1, o.max_it)) break;
        ++cycle;
        rotate_impl<double>(c->V.p, N, h, s.U, MAXSUB, keep, N, c->st);
        rotate_impl<double>(c->HV.p, N, h, s.U, MAXSUB, keep, N, c->st);
        std::fill(Hs.begin(), Hs.end(), 0.0);
        for (int k = 0; k < keep; ++k) Hs[k + MAXSUB * k] = E[k];
        v = h = keep;
      }
      if (fuse) finish_fused();
      else expand_or_stop();
    } else {
      if (!do_refresh) {
        const double el = log_step(false);
        if (check_time(el, ts)) break;
        if (v == nbmax) {
          // adaptive schedule, full basis: restart onto `keep` Ritz vectors with their
          // FP32 sigma, no FP64 sigma; basis vector 0 changes, so no exact anchor row
          ++cycle;
          rotate_impl<double>(c->V.p, N, h, s.U, MAXSUB, keep, N, c->st);
          rotate_impl<float>(c->HV32.p, N, h, s.U, MAXSUB, keep, N, c->st);
          std::fill(Hs.begin(), Hs.end(), 0.0);
          for (int k = 0; k < keep; ++k) Hs[k + MAXSUB * k] = E[k];
          v = h = keep;
          have_anchor = false;
          expand_or_stop();
          continue;
        }
        if (fuse) finish_fused();
        else expand_or_stop();
        continue;
      }
      // the basis becomes the Ritz vectors of the subspace (all of them, or `keep` when
      // the basis is full: thick restart), so the Ritz vector x is basis vector 0
      const int kk = (v == nbmax) ? keep : h;
      rotate_impl<double>(c->V.p, N, h, s.U, MAXSUB, kk, N, c->st);
      rotate_impl<float>(c->HV32.p, N, h, s.U, MAXSUB, kk, N, c->st);
      SBDG_CHECK(hipMemcpyAsync(c->V.p, c->W.p, sizeof(double) * N, hipMemcpyDeviceToDevice, c->st));
      std::fill(Hs.begin(), Hs.end(), 0.0);
      for (int k = 1; k < kk; ++k) Hs[k + MAXSUB * k] = E[k];
      v = h = kk;
      ++cycle;
      refresh();
      have_anchor = true;
      r_ref = rnorm;
      const double el = log_step(true);
      if (rnorm < o.eps) {
        converged = true;
        break;
      }
      if (o.target_de > 0.0 && std::isfinite(prev_exact) && std::fabs(theta - prev_exact) < o.target_de) {
        converged = true;
        break;
      }
      prev_exact = theta;
      if (check_time(el, ts)) break;
      expand_or_stop();
    }
  }
  if (!w_valid) ritz(h);  // the last step was fused: form the Ritz vector
  c->stats.converged = converged ? 1 : 0;
  c->stats.t_davidson = now() - tstart;
  SBDG_CHECK(hipEventDestroy(e0));
  SBDG_CHECK(hipEventDestroy(e1));
  // final energy: one more FP64 sigma of the normalised W, then <W|HW> (sbdiag.h:250-273);
  // the sigma vectors of the basis are free now: the result goes to HX (mixed) or HV_0,
  // the third scratch vector to the FP32 sigma vectors (mixed) or HV_1
  const double tf = now();
  double* hw = mixed ? c->HX.p : c->HV.p;
  double* third = mixed ? reinterpret_cast<double*>(c->HV32.p) : c->HV.p + N;
  sigma64(c, c->W.p, hw, SigmaScratch<double>{c->T.p, c->R.p, third});
  c->stats.sigma_fp64++;
  multi_dot_impl(c->W.p, N, 1, hw, N, s.dots, s.part, c->st);
  double e = 0.0;
  SBDG_CHECK(hipMemcpyAsync(&e, s.dots, sizeof(double), hipMemcpyDeviceToHost, c->st));
  SBDG_CHECK(hipStreamSynchronize(c->st));
  c->energy = e;
  c->stats.energy = e;
  c->stats.t_final = now() - tf;
  c->solved = true;
}

void rdm_or_density(sbdg_ctx* c) {
  const double t0 = now();
  const int L = c->L;
  const int64_t na = c->a.size(), nb = c->b.size();
  if (c->opt.do_rdm) {
    c->d_one.alloc(2 * L * L);
    c->d_two.alloc((size_t)4 * L * L * L * L);
    RdmPlan P;
    P.na = na;
    P.nb = nb;
    P.L = L;
    P.a = c->d_a.p;
    P.b = c->d_b.p;
    P.sa = &c->Sa;
    P.da = &c->Da;
    P.sb = &c->Sb;
    P.db = &c->Db;
    P.ab = &c->ab64;
    if (c->opt.doubles != 1) {  // the core tables exist (built for sigma, used or not)
      P.ca = &c->Ca;
      P.cb = &c->Cb;
    }
    c->T.alloc(na * nb);  // the transposed vector
    rdm_compute(P, c->W.p, c->T.p, c->d_one.p, c->d_two.p, c->st);
    c->have_rdm = true;
  }
  c->d_dens.alloc(2 * L);
  c->small.alloc(std::max<size_t>(c->small.n, (size_t)(na + nb + 16)));
  DBuf<double> part;
  part.alloc(na + nb + 1);
  density(c->W.p, c->d_a.p, na, c->d_b.p, nb, L, c->d_dens.p, part.p, c->st);
  SBDG_CHECK(hipStreamSynchronize(c->st));
  c->have_dens = true;
  c->stats.t_rdm = now() - t0;
}

}  // namespace

namespace {

// A string as the reference's word vector: orbital i in word i / bit_length, bit
// i % bit_length (from_string, bit_manipulation.h:158-172); words = ceil(norb/bit_length).
std::vector<uint64_t> to_words(uint64_t s, int L, int bl) {
  std::vector<uint64_t> w((size_t)((L + bl - 1) / bl), 0);
  for (int i = 0; i < L; ++i)
    if ((s >> i) & 1ull) w[(size_t)(i / bl)] |= 1ull << (i % bl);
  return w;
}

// inverse of to_words for `words` words; false when a bit lies beyond orbital 63 or
// beyond a word's bit_length (such a string cannot be in a set of this context)
bool from_words(const uint64_t* w, size_t words, int bl, uint64_t& s) {
  s = 0;
  for (size_t k = 0; k < words; ++k) {
    uint64_t x = w[k];
    if (bl < 64 && (x >> bl)) return false;
    while (x) {
      const int p = __builtin_ctzll(x);
      x &= x - 1;
      const size_t orb = k * (size_t)bl + (size_t)p;
      if (orb >= 64) return false;
      s |= 1ull << orb;
    }
  }
  return true;
}

std::vector<double> host_vector(sbdg_ctx* c) {  // W with interleaved-order phases
  const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
  std::vector<double> w(N);
  sign_transform(c->d_a.p, na, c->d_b.p, nb, c->W.p, c->T.p, c->st);
  SBDG_CHECK(hipMemcpyAsync(w.data(), c->T.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
  SBDG_CHECK(hipStreamSynchronize(c->st));
  return w;
}

int check_bit_length(sbdg_ctx* c, int bl) {
  if (bl < 1 || bl > 64) return fail(c, SBDG_ERR_ARG, "bit_length must be in 1..64");
  return SBDG_OK;
}

}  // namespace

extern "C" {

sbdg_ctx* sbdg_create(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return nullptr;
  if (hipSetDevice(device) != hipSuccess) return nullptr;
  auto* c = new sbdg_ctx();
  c->device = device;
  if (hipStreamCreate(&c->st) != hipSuccess || hipStreamCreate(&c->st2) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess) {
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->st2) (void)hipStreamDestroy(c->st2);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return nullptr;
  }
  sbdg_default_options(&c->opt);
  c->scratch = scratch_address(device);  // once per device; checked where it is used
  return c;
}

void sbdg_destroy(sbdg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->st);
  (void)hipStreamSynchronize(c->st2);
  (void)hipEventDestroy(c->fork);
  (void)hipEventDestroy(c->join);
  (void)hipStreamDestroy(c->st2);
  (void)hipStreamDestroy(c->st);
  delete c;
}

const char* sbdg_last_error(const sbdg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sbdg_default_options(sbdg_options* o) {
  if (!o) return SBDG_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  o->max_it = 1;
  o->max_nb = 10;
  o->eps = 1.0e-4;
  o->max_time = 86400.0;
  o->init = 0;
  o->do_rdm = 0;
  o->ratio = 0.0;
  o->threshold = 0.01;
  o->restart_keep = 1;
  o->precision = 0;
  o->refresh_k = 4;
  o->target_de = 0.0;
  o->verbose = 0;
  o->seed = 12345;
  o->doubles = 0;
  o->core_gib = 16.0;
  o->overlap = 1;
  return SBDG_OK;
}

int sbdg_set_integrals(sbdg_ctx* c, int norb, double core, const double* h1, const double* eri, int layout) {
  return guard(c, [&]() {
    if (norb < 1 || norb > 64) return fail(c, SBDG_ERR_ARG, "norb must be in 1..64");
    if (!h1 || !eri) return fail(c, SBDG_ERR_ARG, "null integral pointer");
    const int L = norb;
    const int64_t L4 = (int64_t)L * L * L * L;
    c->L = L;
    c->core = core;
    c->h1.assign(h1, h1 + L * L);
    c->eri.assign(L4, 0.0);
    if (layout == 0) {
      std::memcpy(c->eri.data(), eri, sizeof(double) * L4);
    } else if (layout == 1) {
      // twoInt::store (integrals.h:24-67): slot of (ij|kl) = a(a+1)/2+b, a=max(ij,kl),
      // b=min(ij,kl), ij = max(i,j)(max(i,j)+1)/2 + min(i,j)
      auto pr = [](int64_t i, int64_t j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; };
      for (int p = 0; p < L; ++p)
        for (int q = 0; q < L; ++q)
          for (int r = 0; r < L; ++r)
            for (int s_ = 0; s_ < L; ++s_) {
              const int64_t ij = pr(p, q), kl = pr(r, s_);
              const int64_t aa = std::max(ij, kl), bb = std::min(ij, kl);
              c->eri[(((int64_t)p * L + q) * L + r) * L + s_] = eri[aa * (aa + 1) / 2 + bb];
            }
    } else {
      return fail(c, SBDG_ERR_ARG, "eri_layout must be 0 or 1");
    }
    for (double x : c->h1)
      if (!std::isfinite(x)) return fail(c, SBDG_ERR_ARG, "non-finite one-electron integral");
    for (double x : c->eri)
      if (!std::isfinite(x)) return fail(c, SBDG_ERR_ARG, "non-finite two-electron integral");
    if (layout == 0) {
      // the kernels read (pq|rs) through its images ((ai|bj) for b < j from (ai|jb) in the
      // alpha-beta kernel): the tensor must carry the 8-fold symmetry of real orbitals
      double mx = 0.0;
      for (double x : c->eri) mx = std::max(mx, std::fabs(x));
      const double tol = 1e-10 * mx;
      const double* e = c->eri.data();
      auto at = [&](int64_t p, int64_t q, int64_t r, int64_t s_) { return e[((p * L + q) * L + r) * L + s_]; };
      for (int64_t p = 0; p < L; ++p)
        for (int64_t q = 0; q < L; ++q)
          for (int64_t r = 0; r < L; ++r)
            for (int64_t s_ = 0; s_ < L; ++s_) {
              const double x = at(p, q, r, s_);
              if (std::fabs(x - at(q, p, r, s_)) > tol || std::fabs(x - at(p, q, s_, r)) > tol ||
                  std::fabs(x - at(r, s_, p, q)) > tol)
                return fail(c, SBDG_ERR_ARG, "two-electron integrals lack the 8-fold symmetry (pq|rs) = (qp|rs) = (pq|sr) = (rs|pq)");
            }
    }
    c->have_ints = true;
    c->built = c->solved = c->have_rdm = c->have_dens = false;
    return SBDG_OK;
  });
}

static std::string check_set(const uint64_t* s, int64_t n, int L, const char* name) {
  if (n < 1) return std::string(name) + " string set is empty";
  const int pc = __builtin_popcountll(s[0]);
  std::vector<uint64_t> v(s, s + n);
  for (int64_t i = 0; i < n; ++i) {
    if (L < 64 && (s[i] >> L) != 0) return std::string(name) + " string has orbitals beyond norb";
    if (__builtin_popcountll(s[i]) != pc) return std::string(name) + " strings differ in electron count";
  }
  std::sort(v.begin(), v.end());
  if (std::adjacent_find(v.begin(), v.end()) != v.end()) return std::string(name) + " strings are not unique";
  return "";
}

int sbdg_set_strings(sbdg_ctx* c, const uint64_t* a, int64_t na, const uint64_t* b, int64_t nb) {
  return guard(c, [&]() {
    if (!c->have_ints) return fail(c, SBDG_ERR_STATE, "set integrals before strings");
    if (!a || !b) return fail(c, SBDG_ERR_ARG, "null string pointer");
    std::string m = check_set(a, na, c->L, "alpha");
    if (m.empty()) m = check_set(b, nb, c->L, "beta");
    if (!m.empty()) return fail(c, SBDG_ERR_ARG, m);
    if (nb > (1 << 18)) return fail(c, SBDG_ERR_LIMIT, "beta string set beyond 262144 strings");
    if (na > (int64_t)INT32_MAX) return fail(c, SBDG_ERR_LIMIT, "alpha string set beyond 2^31 strings");
    c->a.assign(a, a + na);
    c->b.assign(b, b + nb);
    c->have_strs = true;
    c->built = c->solved = c->have_rdm = c->have_dens = c->have_user_vec = false;
    return SBDG_OK;
  });
}

int sbdg_set_options(sbdg_ctx* c, const sbdg_options* o) {
  return guard(c, [&]() {
    if (!o) return fail(c, SBDG_ERR_ARG, "null options");
    if (o->max_nb < 1 || o->max_nb > MAXSUB) return fail(c, SBDG_ERR_ARG, "max_nb must be in 1..48");
    if (o->max_it < 1) return fail(c, SBDG_ERR_ARG, "max_it must be >= 1");
    if (o->precision != 0 && o->precision != 1) return fail(c, SBDG_ERR_ARG, "precision must be 0 or 1");
    if (o->init < 0 || o->init > 2) return fail(c, SBDG_ERR_ARG, "init must be 0, 1 or 2");
    if (!(o->eps >= 0.0)) return fail(c, SBDG_ERR_ARG, "eps must be >= 0");
    if (o->doubles < 0 || o->doubles > 2) return fail(c, SBDG_ERR_ARG, "doubles must be 0, 1 or 2");
    if (!(o->core_gib > 0.0)) return fail(c, SBDG_ERR_ARG, "core_gib must be > 0");
    if (o->overlap != 0 && o->overlap != 1) return fail(c, SBDG_ERR_ARG, "overlap must be 0 or 1");
    if (c->built && (o->doubles != c->opt.doubles || o->core_gib != c->opt.core_gib)) c->built = false;
    c->opt = *o;
    return SBDG_OK;
  });
}

int sbdg_set_vector(sbdg_ctx* c, const double* w) {
  return guard(c, [&]() {
    if (!c->have_strs) return fail(c, SBDG_ERR_STATE, "strings not set");
    if (!w) return fail(c, SBDG_ERR_ARG, "null vector");
    const int64_t N = (int64_t)c->a.size() * c->b.size();
    c->user_vec.assign(w, w + N);
    c->have_user_vec = true;
    return SBDG_OK;
  });
}

int sbdg_build(sbdg_ctx* c) {
  return guard(c, [&]() {
    mem_count().reset_peak();
    build(c);
    c->stats.mem_now = mem_count().now.load();
    c->stats.mem_peak = mem_count().peak.load();
    return SBDG_OK;
  });
}

int sbdg_run(sbdg_ctx* c) {
  return guard(c, [&]() {
    mem_count().reset_peak();
    ensure_built(c);
    const int64_t N = (int64_t)c->a.size() * c->b.size();
    c->T.alloc(N);
    davidson(c);
    rdm_or_density(c);
    c->stats.mem_now = mem_count().now.load();
    c->stats.mem_peak = mem_count().peak.load();
    return SBDG_OK;
  });
}

int sbdg_memory_plan(const sbdg_options* o, int norb, int64_t na, int64_t nb, const int64_t* nnz,
                     const int64_t* mem, const int64_t* dir, int64_t* out) {
  if (!o || !nnz || !mem || !dir || !out || norb < 1 || norb > 64 || na < 0 || nb < 0) return SBDG_ERR_ARG;
  out[0] = build_bytes(*o, norb, na, nb, nnz, mem, dir);
  out[1] = davidson_bytes(*o, na * nb);
  out[2] = rdm_bytes(norb, nnz);
  return SBDG_OK;
}

int sbdg_get_stats(const sbdg_ctx* c, sbdg_stats* st) {
  if (!c || !st) return SBDG_ERR_ARG;
  *st = c->stats;
  st->mem_now = mem_count().now.load();
  return SBDG_OK;
}

int sbdg_get_energy(const sbdg_ctx* c, double* e, double* r, int* it) {
  if (!c) return SBDG_ERR_ARG;
  if (!c->solved) return SBDG_ERR_STATE;
  if (e) *e = c->energy;
  if (r) *r = c->stats.resid;
  if (it) *it = c->stats.iterations;
  return SBDG_OK;
}

int sbdg_get_trajectory(const sbdg_ctx* c, double* t, double* e, int cap, int* n) {
  if (!c || !n) return SBDG_ERR_ARG;
  const int m = (int)c->traj_t.size();
  *n = m;
  for (int k = 0; k < std::min(m, cap); ++k) {
    if (t) t[k] = c->traj_t[k];
    if (e) e[k] = c->traj_e[k];
  }
  return SBDG_OK;
}

int sbdg_get_vector(sbdg_ctx* c, double* w) {
  return guard(c, [&]() {
    if (!c->solved) return fail(c, SBDG_ERR_STATE, "no solution");
    if (!w) return fail(c, SBDG_ERR_ARG, "null output");
    const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
    sign_transform(c->d_a.p, na, c->d_b.p, nb, c->W.p, c->T.p, c->st);
    SBDG_CHECK(hipMemcpyAsync(w, c->T.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    return SBDG_OK;
  });
}

int sbdg_get_density(sbdg_ctx* c, double* d) {
  return guard(c, [&]() {
    if (!c->have_dens) return fail(c, SBDG_ERR_STATE, "no density");
    SBDG_CHECK(hipMemcpy(d, c->d_dens.p, sizeof(double) * 2 * c->L, hipMemcpyDeviceToHost));
    return SBDG_OK;
  });
}

int sbdg_compute_rdm(sbdg_ctx* c) {
  return guard(c, [&]() {
    if (!c->solved) return fail(c, SBDG_ERR_STATE, "no solution");
    const int save = c->opt.do_rdm;
    c->opt.do_rdm = 1;
    rdm_or_density(c);
    c->opt.do_rdm = save;
    return SBDG_OK;
  });
}

int sbdg_get_rdm(sbdg_ctx* c, double* one, double* two) {
  return guard(c, [&]() {
    if (!c->have_rdm) return fail(c, SBDG_ERR_STATE, "no RDMs (run with do_rdm=1 or call sbdg_compute_rdm)");
    const int64_t L = c->L;
    if (one) SBDG_CHECK(hipMemcpy(one, c->d_one.p, sizeof(double) * 2 * L * L, hipMemcpyDeviceToHost));
    if (two) SBDG_CHECK(hipMemcpy(two, c->d_two.p, sizeof(double) * 4 * L * L * L * L, hipMemcpyDeviceToHost));
    return SBDG_OK;
  });
}

int sbdg_rdm_backward(sbdg_ctx* c, double g, double* dh1, double* deri) {
  return guard(c, [&]() {
    if (!c->have_rdm) return fail(c, SBDG_ERR_STATE, "no RDMs");
    const int64_t L = c->L, L4 = L * L * L * L;
    DBuf<double> dh, de;
    dh.alloc(L * L);
    de.alloc(L4);
    rdm_backward(c->d_one.p, c->d_two.p, (int)L, g, dh.p, de.p, c->st);
    if (dh1) SBDG_CHECK(hipMemcpyAsync(dh1, dh.p, sizeof(double) * L * L, hipMemcpyDeviceToHost, c->st));
    if (deri) SBDG_CHECK(hipMemcpyAsync(deri, de.p, sizeof(double) * L4, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    return SBDG_OK;
  });
}

int sbdg_save_wavefunction(sbdg_ctx* c, const char* prefix, int bit_length) {
  return guard(c, [&]() {
    if (!c->solved) return fail(c, SBDG_ERR_STATE, "no solution");
    if (!prefix) return fail(c, SBDG_ERR_ARG, "null file name");
    if (check_bit_length(c, bit_length)) return SBDG_ERR_ARG;
    const int L = c->L;
    const std::vector<double> w = host_vector(c);
    // SaveWavefunction (restart.h:20-67) at one rank: the file of rank 0
    const std::string name = std::string(prefix) + "000000";
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return fail(c, SBDG_ERR_ARG, "cannot open " + name);
    const uint64_t na = c->a.size(), nb = c->b.size(), words = (uint64_t)((L + bit_length - 1) / bit_length);
    bool ok = std::fwrite(&na, 8, 1, f) == 1 && std::fwrite(&nb, 8, 1, f) == 1 && std::fwrite(&words, 8, 1, f) == 1;
    for (uint64_t x : c->a) ok = ok && std::fwrite(to_words(x, L, bit_length).data(), 8, words, f) == words;
    for (uint64_t x : c->b) ok = ok && std::fwrite(to_words(x, L, bit_length).data(), 8, words, f) == words;
    ok = ok && std::fwrite(w.data(), 8, w.size(), f) == w.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(c, SBDG_ERR_ARG, "write failed: " + name);
    return SBDG_OK;
  });
}

int sbdg_load_wavefunction(sbdg_ctx* c, const char* prefix, int bit_length) {
  return guard(c, [&]() {
    if (!c->have_strs) return fail(c, SBDG_ERR_STATE, "strings not set");
    if (!prefix) return fail(c, SBDG_ERR_ARG, "null file name");
    if (check_bit_length(c, bit_length)) return SBDG_ERR_ARG;
    // LoadWavefunction (restart.h:414-657) at one rank: reads the file of rank 0
    const std::string name = std::string(prefix) + "000000";
    FILE* f = std::fopen(name.c_str(), "rb");
    if (!f) return fail(c, SBDG_ERR_ARG, "failed to open input file " + name);
    uint64_t hd[3] = {0, 0, 0};
    bool ok = std::fread(hd, 8, 3, f) == 3;
    const uint64_t la = hd[0], lb = hd[1], words = hd[2];
    if (ok && (words == 0 || words > 64 || la > ((uint64_t)1 << 40) || lb > ((uint64_t)1 << 40))) ok = false;
    std::vector<uint64_t> wa, wb;
    std::vector<double> lw;
    if (ok) {
      wa.resize(la * words);
      wb.resize(lb * words);
      lw.resize(la * lb);
      ok = std::fread(wa.data(), 8, wa.size(), f) == wa.size() && std::fread(wb.data(), 8, wb.size(), f) == wb.size() &&
           std::fread(lw.data(), 8, lw.size(), f) == lw.size();
    }
    std::fclose(f);
    if (!ok) return fail(c, SBDG_ERR_ARG, "malformed restart file " + name);
    const int L = c->L;
    const uint64_t mine = (uint64_t)((L + bit_length - 1) / bit_length);
    const int64_t na = c->a.size(), nb = c->b.size();
    // position of each loaded string in the current sets (find on equal word vectors)
    auto place = [&](const std::vector<uint64_t>& wv, uint64_t n, const std::vector<uint64_t>& set) {
      std::vector<int64_t> pos(n, -1);
      if (words != mine) return pos;  // word vectors of another length never compare equal
      std::unordered_map<uint64_t, int64_t> at;
      for (size_t i = 0; i < set.size(); ++i) at.emplace(set[i], (int64_t)i);
      for (uint64_t i = 0; i < n; ++i) {
        uint64_t s;
        if (!from_words(&wv[i * words], words, bit_length, s)) continue;
        auto it = at.find(s);
        if (it != at.end()) pos[i] = it->second;
      }
      return pos;
    };
    const std::vector<int64_t> pa = place(wa, la, c->a), pb = place(wb, lb, c->b);
    std::vector<double> w((size_t)(na * nb), 0.0);
    for (uint64_t i = 0; i < la; ++i) {
      if (pa[i] < 0) continue;
      for (uint64_t j = 0; j < lb; ++j)
        if (pb[j] >= 0) w[(size_t)(pa[i] * nb + pb[j])] += lw[i * lb + j];
    }
    double nrm = 0.0;
    for (double x : w) nrm += x * x;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0)) return fail(c, SBDG_ERR_NUMERIC, "no determinant of " + name + " is in the current basis");
    for (double& x : w) x /= nrm;
    c->user_vec = std::move(w);
    c->have_user_vec = true;
    c->opt.init = 2;  // the loaded vector is the initial guess, as diag does with a loadname
    return SBDG_OK;
  });
}

int sbdg_save_matrix_form(sbdg_ctx* c, const char* filename, int bit_length) {
  return guard(c, [&]() {
    if (!c->solved) return fail(c, SBDG_ERR_STATE, "no solution");
    if (!filename) return fail(c, SBDG_ERR_ARG, "null file name");
    if (check_bit_length(c, bit_length)) return SBDG_ERR_ARG;
    // SaveMatrixFormWF (restart.h:662-780): by extension
    const std::string name(filename);
    const size_t dot = name.find_last_of('.');
    const std::string ext = dot == std::string::npos ? std::string() : name.substr(dot + 1);
    if (ext != "bin" && ext != "txt" && ext != "dat") return SBDG_OK;  // nothing written
    const std::vector<double> w = host_vector(c);
    const int64_t na = c->a.size(), nb = c->b.size();
    if (ext == "bin") {
      std::ofstream o(name, std::ios::binary);
      if (!o) return fail(c, SBDG_ERR_ARG, "cannot open " + name);
      o.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(sizeof(double) * w.size()));
      if (!o) return fail(c, SBDG_ERR_ARG, "write failed: " + name);
      return SBDG_OK;
    }
    const int L = c->L;
    auto str = [&](uint64_t s) {  // makestring (bit_manipulation.h:135-149): orbital L-1 first
      std::string t((size_t)L, '0');
      for (int i = 0; i < L; ++i)
        if ((s >> i) & 1ull) t[(size_t)(L - 1 - i)] = '1';
      return t;
    };
    std::ofstream o(name);  // default stream precision (six significant digits), as the reference
    if (!o) return fail(c, SBDG_ERR_ARG, "cannot open " + name);
    for (int64_t ia = 0; ia < na; ++ia) {
      const std::string sa = str(c->a[ia]);
      for (int64_t ib = 0; ib < nb; ++ib)
        o << w[(size_t)(ia * nb + ib)] << " # " << ia << ": " << sa << " " << ib << ": " << str(c->b[ib]) << "\n";
    }
    if (!o) return fail(c, SBDG_ERR_ARG, "write failed: " + name);
    return SBDG_OK;
  });
}

int sbdg_carryover(sbdg_ctx* c, double ratio, double threshold, uint64_t* out, int64_t cap, int64_t* count,
                   double* truncated) {
  return guard(c, [&]() {
    if (!c->solved) return fail(c, SBDG_ERR_STATE, "no solution");
    const int64_t na = c->a.size(), nb = c->b.size();
    // DiagonalReducedDensityMatrix (tpb/rdmat.h:11-58): row weights of |W|^2, on the device
    std::vector<double> D(na, 0.0);
    DBuf<double> rs;
    rs.alloc(std::max<int64_t>(na, 1));
    row_sumsq(c->W.p, na, nb, rs.p, c->st);
    SBDG_CHECK(hipMemcpyAsync(D.data(), rs.p, sizeof(double) * na, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    std::vector<int64_t> idx(na);
    for (int64_t i = 0; i < na; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t x, int64_t y) { return D[x] > D[y]; });
    int64_t kept;
    double tw = 0.0;
    if (ratio > 0.0) {
      kept = std::min<int64_t>(na, (int64_t)(ratio * na));
      double s_ = 0.0;
      for (int64_t i = 0; i < kept; ++i) s_ += D[idx[i]];
      tw = 1.0 - s_;
    } else {
      // CarryOverAlphaDet by threshold (tpb/rdmat.h:60-100): smallest prefix whose running
      // weight passes 1 - threshold (upper_bound + 1)
      const double target = 1.0 - threshold;
      double s_ = 0.0;
      kept = na;
      for (int64_t i = 0; i < na; ++i) {
        s_ += D[idx[i]];
        if (s_ > target) {  // first running sum above the target, kept inclusively
          kept = i + 1;
          break;
        }
      }
      kept = std::min<int64_t>(kept, na);
    }
    if (count) *count = kept;
    if (truncated) *truncated = tw;
    for (int64_t i = 0; i < std::min(kept, cap); ++i)
      if (out) out[i] = c->a[idx[i]];
    return SBDG_OK;
  });
}

int sbdg_table_nnz(const sbdg_ctx* c, int which, int64_t* nnz) {
  if (!c || !nnz || which < 0 || which > 3) return SBDG_ERR_ARG;
  if (!c->built) return SBDG_ERR_STATE;
  const Table* t[4] = {&c->Sa, &c->Da, &c->Sb, &c->Db};
  *nnz = t[which]->nnz;
  return SBDG_OK;
}

int sbdg_get_table(sbdg_ctx* c, int which, int64_t* rowptr, int32_t* col, uint32_t* exc, double* val) {
  return guard(c, [&]() {
    if (which < 0 || which > 3) return fail(c, SBDG_ERR_ARG, "which must be 0..3");
    ensure_built(c);
    const Table* t[4] = {&c->Sa, &c->Da, &c->Sb, &c->Db};
    const Table& tb = *t[which];
    if (rowptr) SBDG_CHECK(hipMemcpy(rowptr, tb.rowptr.p, sizeof(int64_t) * (tb.n + 1), hipMemcpyDeviceToHost));
    if (tb.nnz) {
      if (col) SBDG_CHECK(hipMemcpy(col, tb.col.p, sizeof(int32_t) * tb.nnz, hipMemcpyDeviceToHost));
      if (exc) SBDG_CHECK(hipMemcpy(exc, tb.exc.p, sizeof(uint32_t) * tb.nnz, hipMemcpyDeviceToHost));
      if (val) SBDG_CHECK(hipMemcpy(val, tb.val.p, sizeof(double) * tb.nnz, hipMemcpyDeviceToHost));
    }
    return SBDG_OK;
  });
}

int sbdg_get_hii(sbdg_ctx* c, double* hii) {
  return guard(c, [&]() {
    ensure_built(c);
    const int64_t N = (int64_t)c->a.size() * c->b.size();
    SBDG_CHECK(hipMemcpy(hii, c->d_hii.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    return SBDG_OK;
  });
}

int sbdg_sigma(sbdg_ctx* c, const double* x, double* y, int precision) {
  return guard(c, [&]() {
    if (!x || !y) return fail(c, SBDG_ERR_ARG, "null vector");
    if (precision != 0 && precision != 1) return fail(c, SBDG_ERR_ARG, "precision must be 0 or 1");
    ensure_built(c);
    const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
    DBuf<double> in, out;
    in.alloc(N);
    out.alloc(N);
    SBDG_CHECK(hipMemcpyAsync(in.p, x, sizeof(double) * N, hipMemcpyHostToDevice, c->st));
    sign_transform(c->d_a.p, na, c->d_b.p, nb, in.p, in.p, c->st);  // in place: elementwise
    if (precision == 0) {
      sigma64(c, in.p, out.p, own_scratch64(c));
    } else {
      DBuf<float> x32, y32;
      x32.alloc(N);
      y32.alloc(N);
      cast_d2f(in.p, x32.p, N, c->st);
      sigma32(c, x32.p, y32.p, own_scratch32(c));
      cast_f2d(y32.p, out.p, N, c->st);
      SBDG_CHECK(hipStreamSynchronize(c->st));
    }
    sign_transform(c->d_a.p, na, c->d_b.p, nb, out.p, out.p, c->st);
    SBDG_CHECK(hipMemcpyAsync(y, out.p, sizeof(double) * N, hipMemcpyDeviceToHost, c->st));
    SBDG_CHECK(hipStreamSynchronize(c->st));
    return SBDG_OK;
  });
}

int sbdg_sigma_device(sbdg_ctx* c, const double* x, double* y, int precision) {
  return guard(c, [&]() {
    if (!x || !y) return fail(c, SBDG_ERR_ARG, "null vector");
    if (x == y) return fail(c, SBDG_ERR_ARG, "input and output vectors must differ");
    if (precision != 0 && precision != 1) return fail(c, SBDG_ERR_ARG, "precision must be 0 or 1");
    if (!c->built) return fail(c, SBDG_ERR_STATE, "sbdg_sigma_device needs a built context (sbdg_build)");
    if (!c->scratch) return fail(c, SBDG_ERR_HIP, "the device scratch array has no address on this device");
    const int64_t na = c->a.size(), nb = c->b.size(), N = na * nb;
    const size_t need = precision == 0 ? (size_t)N * 4 * sizeof(double) : (size_t)N * 5 * sizeof(float);
    if (need > SCRATCH_BYTES)
      return fail(c, SBDG_ERR_LIMIT,
                  "sbdg_sigma_device: " + std::to_string(N) + " determinants need " + std::to_string(need) +
                      " scratch bytes, beyond its " + std::to_string(SCRATCH_BYTES) +
                      " (at most " + std::to_string(SCRATCH_BYTES / 32) + " determinants in FP64, " +
                      std::to_string(SCRATCH_BYTES / 20) + " in FP32)");
    if (precision == 0) {
      double* S = static_cast<double*>(c->scratch);
      sign_transform(c->d_a.p, na, c->d_b.p, nb, x, S, c->st);  // separated phases
      sigma64(c, S, y, SigmaScratch<double>{S + N, S + 2 * N, S + 3 * N});
    } else {
      float* F = static_cast<float*>(c->scratch);
      sign_transform(c->d_a.p, na, c->d_b.p, nb, x, y, c->st);  // y: staging of the FP64 input
      cast_d2f(y, F, N, c->st);
      sigma32(c, F, F + N, SigmaScratch<float>{F + 2 * N, F + 3 * N, F + 4 * N});
      cast_f2d(F + N, y, N, c->st);
    }
    sign_transform(c->d_a.p, na, c->d_b.p, nb, y, y, c->st);  // back to the reference phases
    SBDG_CHECK(hipStreamSynchronize(c->st));
    return SBDG_OK;
  });
}

int sbdg_bench_sigma(sbdg_ctx* c, int precision, int reps, double* ms) {
  return guard(c, [&]() {
    if (!ms || reps < 1 || (precision != 0 && precision != 1)) return fail(c, SBDG_ERR_ARG, "bad arguments");
    ensure_built(c);
    const int64_t N = (int64_t)c->a.size() * c->b.size();
    if (c->bx64.n != (size_t)N) {
      c->bx64.alloc(N);
      fill_test_vector(c->bx64.p, N, 7, c->st);
    }
    if (precision == 1 && c->bx32.n != (size_t)N) {
      c->bx32.alloc(N);
      cast_d2f(c->bx64.p, c->bx32.p, N, c->st);
    }
    if (precision == 0) c->by64.alloc(N);
    else c->by32.alloc(N);
    SBDG_CHECK(hipStreamSynchronize(c->st));
    hipEvent_t ev[6];
    for (auto& e : ev) SBDG_CHECK(hipEventCreate(&e));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int r = 0; r < reps; ++r) {
      if (precision == 1) sigma32(c, c->bx32.p, c->by32.p, own_scratch32(c), ev);
      else sigma64(c, c->bx64.p, c->by64.p, own_scratch64(c), ev);
      SBDG_CHECK(hipEventSynchronize(ev[5]));
      float t;
      SBDG_CHECK(hipEventElapsedTime(&t, ev[0], ev[5]));
      acc[0] += t;
      for (int k = 1; k < 6; ++k) {
        SBDG_CHECK(hipEventElapsedTime(&t, ev[k - 1], ev[k]));
        acc[k] += t;
      }
    }
    for (int k = 0; k < 6; ++k) ms[k] = acc[k] / reps;
    for (auto& e : ev) SBDG_CHECK(hipEventDestroy(e));
    return SBDG_OK;
  });
}

}  // extern "C"
