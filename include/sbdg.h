/*
 * sbdg.h -- C-ABI of the MI355X selected-basis diagonalization library (libsbdg.so).
 *
 * Drop-in boundary for the hot path of walkup/sbd-ykt's tensor-product-basis solver,
 * `sbd::tpb::diag` (include/sbd/chemistry/tpb/sbdiag.h:113-124), method 0:
 * excitation helpers, diagonal, Davidson with on-the-fly sigma, final energy, RDMs.
 * Plain pointers and sizes only; the caller owns every host buffer and the library
 * copies in and out; device memory belongs to the context.  Every function returns
 * SBDG_OK (0) or a negative status; sbdg_last_error() gives the message.  One context
 * per host thread; calls on one context are serialised on its HIP stream (sigma's
 * alpha-beta kernel runs on a second stream of the context, joined before sigma returns).
 *
 * Conventions at this boundary are the reference's:
 *  - a string is a uint64 whose bit i is orbital i (bit_manipulation.h:158-172 reads
 *    the rightmost character as orbital 0), at most 64 orbitals;
 *  - the wavefunction is row-major [ia][ib] over the given alpha and beta string order
 *    (tpb/qcham.h:359), determinants phased in the interleaved spin-orbital order
 *    (determinants.h:96-142);
 *  - one[s][i + L*j] = <c+_{i s} c_{j s}>, two[s + 2t][i + L j + L^2 k + L^3 l] =
 *    <c+_{i s} c+_{j t} c_{l t} c_{k s}> (docs/user-guide.md:138-139); the two mixed-spin
 *    blocks are the physical ones (their sum equals the reference's sum).
 */
#ifndef SBDG_H
#define SBDG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBDG_OK 0
#define SBDG_ERR_ARG -1      /* invalid argument: std::invalid_argument in the reference */
#define SBDG_ERR_STATE -2    /* call out of order (no problem set, no solution yet) */
#define SBDG_ERR_HIP -3      /* HIP runtime error (no device, out of memory, launch failure) */
#define SBDG_ERR_NUMERIC -4  /* NaN/Inf in the solver: std::runtime_error in the reference */
#define SBDG_ERR_LIMIT -5    /* size beyond a documented limit */

typedef struct sbdg_ctx sbdg_ctx;

/* Davidson options.  The first eight fields are sbd::tpb::SBD's (sbdiag.h:12-33) with the
 * same meaning and defaults; the rest are additions with reference-neutral defaults. */
typedef struct {
  int max_it;         /* SBD::max_it (--iteration): restart cycles, default 1 */
  int max_nb;         /* SBD::max_nb (--block): subspace size, default 10 */
  double eps;         /* SBD::eps (--tolerance): stop when |HW - E W| < eps, default 1e-4 */
  double max_time;    /* SBD::max_time: stop when the next step would end later (s) */
  int init;           /* SBD::init: 0 first determinant (HF), 1 random, 2 vector set by caller */
  int do_rdm;         /* SBD::do_rdm: 0 density only, else 1- and 2-RDMs */
  double ratio;       /* SBD::ratio: carry-over by ratio when > 0 */
  double threshold;   /* SBD::threshold: carry-over by weight, default 0.01 */
  int restart_keep;   /* Ritz vectors kept at a restart, default 1 (the reference's) */
  int precision;      /* 0: FP64 sigma; 1: mixed (FP32 sigma + FP64 refresh) */
  int refresh_k;      /* mixed: FP64 sigma of the Ritz vector every refresh_k steps (4);
                         <= 0: when the FP32 residual has fallen tenfold since the last
                         exact one or is below eps, or the FP32 Ritz value moved less than
                         1e-6 Ha in one step (a full basis then restarts without) */
  double target_de;   /* mixed: also stop when |E - E_prev refresh| < target_de (0: off) */
  int verbose;        /* 1: one line per Davidson step on stdout */
  uint64_t seed;      /* init==1: random start vector seed */
  int doubles;        /* same-spin doubles in sigma: 0 auto, 1 excitation tables (as
                         mult), 2 grouped by their (n-2)-electron core */
  double core_gib;    /* workspace cap in GiB for the per-core products (default 16); with
                         doubles==0 a spin whose 512-column pass does not fit uses the
                         excitation-table form, with doubles==2 it is an error */
  int overlap;        /* 1 (default): the alpha-beta sigma kernel runs on a second stream
                         beside the spin phases; 0: all sigma kernels in one stream, in
                         order (sbdg_bench_sigma's per-phase times need 0) */
} sbdg_options;

typedef struct {
  double energy;        /* <W|H|W> from a final FP64 sigma (sbdiag.h:250-273) */
  double ritz;          /* last Ritz value */
  double resid;         /* last |HW - E W| */
  int iterations;       /* Davidson steps */
  int sigma_fp64;       /* FP64 sigma builds, final one included */
  int sigma_fp32;       /* FP32 sigma builds */
  int converged;        /* 1 if resid < eps (or target_de met) */
  double t_tables;      /* seconds: excitation tables + diagonal */
  double t_davidson;    /* seconds: Davidson loop */
  double t_sigma;       /* seconds inside sigma builds (device events) */
  double t_final;       /* seconds: final energy sigma */
  double t_rdm;         /* seconds: RDMs or density */
  int64_t terms;        /* structurally connected (bra, ket) pairs per sigma, diagonal included */
  int64_t nnz[4];       /* alpha singles, alpha doubles, beta singles, beta doubles */
  int64_t core_members[2]; /* alpha, beta: memberships of the core-grouped doubles (0: the
                              excitation-table form is used for that spin) */
  int64_t core_direct[2];  /* alpha, beta: double pairs of cores with fewer than ten members,
                              read directly by the row product beside the core-grouped ones */
  int64_t mem_now;      /* device bytes the library holds when the stats are read (all
                           contexts of the process) */
  int64_t mem_peak;     /* their high-water mark since the start of the last build or run */
} sbdg_stats;

/* Context on a device.  Replaces the communicator set-up of TaskCommunicator
 * (tpb/helper.h:199-233): one GPU, one HBM-resident wavefunction. */
sbdg_ctx *sbdg_create(int device);
void sbdg_destroy(sbdg_ctx *ctx);
const char *sbdg_last_error(const sbdg_ctx *ctx);
int sbdg_default_options(sbdg_options *opt);

/* Integrals, replacing SetupIntegrals (makeintegrals.h:10-55) on a parsed FCIDump.
 * h1: L*L row-major; eri_layout 0: full (pq|rs) row-major L^4, 8-fold symmetric
 * ((pq|rs) = (qp|rs) = (pq|sr) = (rs|pq) within 1e-10 of the largest element, else
 * SBDG_ERR_ARG); eri_layout 1: the reference's packed store twoInt::store
 * (integrals.h:24-67), symmetric by construction. */
int sbdg_set_integrals(sbdg_ctx *ctx, int norb, double core, const double *h1, const double *eri,
                       int eri_layout);

/* Alpha and beta string sets (the adet/bdet arguments of diag, sbdiag.h:116-117).  Each
 * set must hold unique strings of one popcount below 2^norb. */
int sbdg_set_strings(sbdg_ctx *ctx, const uint64_t *astr, int64_t na, const uint64_t *bstr,
                     int64_t nb);

int sbdg_set_options(sbdg_ctx *ctx, const sbdg_options *opt);

/* Initial vector for init==2 (the loadname path of diag, sbdiag.h:197-203), na*nb. */
int sbdg_set_vector(sbdg_ctx *ctx, const double *w);

/* Excitation tables (MakeHelpers/GenerateExcitation, helper.h:132-197,447-589) and the
 * diagonal (makeQChamDiagTerms, tpb/qcham.h:314-366), built on the device. */
int sbdg_build(sbdg_ctx *ctx);

/* Davidson (tpb/davidson.h:661-924), final energy (sbdiag.h:250-273), then RDMs
 * (Correlation, tpb/correlation.h:194-444) or density (OccupationDensity,
 * tpb/occupation.h:10-95) as options.do_rdm asks.  Builds first if needed. */
int sbdg_run(sbdg_ctx *ctx);

int sbdg_get_stats(const sbdg_ctx *ctx, sbdg_stats *st);
int sbdg_get_energy(const sbdg_ctx *ctx, double *energy, double *resid, int *iterations);
/* Converged vector W (na*nb, interleaved-order phases, unit norm). */
int sbdg_get_vector(sbdg_ctx *ctx, double *w);
/* density[2i+s] (sbdiag.h:367-416). */
int sbdg_get_density(sbdg_ctx *ctx, double *density);
/* one: 2*L*L, two: 4*L^4 (Correlation's output arrays). */
int sbdg_get_rdm(sbdg_ctx *ctx, double *one, double *two);
/* Computes the RDMs of the current vector on demand (no Davidson re-run). */
int sbdg_compute_rdm(sbdg_ctx *ctx);

/* Hellmann-Feynman gradient of the energy: with grad_e = dL/dE,
 * dh1[p*L+q] = grad_e * (one[0]+one[1])[p + L q],
 * deri[((p*L+q)*L+r)*L+s] = grad_e/2 * sum_blk two[blk][p + L r + L^2 q + L^3 s]
 * (the energy identity of samples/selected_basis_diagonalization/main.cc:166-216).
 * Either output may be NULL. */
int sbdg_rdm_backward(sbdg_ctx *ctx, double grad_e, double *dh1, double *deri);

/* Wavefunction files of diag (restart.h).  bit_length is the reference's string word size
 * (sbd::tpb::SBD::bit_length, 20 by default): orbital i sits in word i/bit_length at bit
 * i%bit_length, ceil(norb/bit_length) words per string.
 *
 * sbdg_save_wavefunction replaces diag's savename (SaveWavefunction, restart.h:20-67) at
 * one rank: file <prefix>000000 = size_t na, nb, words; the na alpha then nb beta strings
 * as words; W (na*nb doubles, the vector of sbdg_get_vector).  Needs a solved context. */
int sbdg_save_wavefunction(sbdg_ctx *ctx, const char *prefix, int bit_length);
/* diag's loadname (LoadWavefunction, restart.h:414-657) at one rank: reads
 * <prefix>000000, adds each coefficient whose alpha and beta strings are both in the
 * current sets at their place, normalises, and makes it the initial vector (sets
 * options.init = 2).  SBDG_ERR_NUMERIC when no coefficient lands in the basis (the
 * reference would divide by a zero norm). */
int sbdg_load_wavefunction(sbdg_ctx *ctx, const char *prefix, int bit_length);
/* SaveMatrixFormWF (restart.h:662-780): extension .bin = na*nb doubles row-major; .txt or
 * .dat = one line "<value> # <ia>: <alpha string> <ib>: <beta string>" per determinant,
 * six significant digits, strings with orbital norb-1 first; other extensions write
 * nothing. */
int sbdg_save_matrix_form(sbdg_ctx *ctx, const char *filename, int bit_length);

/* Carry-over alpha strings (CarryOverAlphaDet, tpb/rdmat.h:60-132): ratio > 0 keeps
 * floor(ratio*na) strings, else the smallest set of heaviest rows whose weight exceeds
 * 1 - threshold.  Writes up to cap strings; *count receives the number selected. */
int sbdg_carryover(sbdg_ctx *ctx, double ratio, double threshold, uint64_t *out, int64_t cap,
                   int64_t *count, double *truncated_weight);

/* Time-to-solution record of the last run: elapsed seconds since the Davidson start and
 * the energy at every step whose energy is an FP64 Rayleigh quotient (every step of the
 * FP64 path; every refresh of the mixed path).  Writes min(cap, *n) entries. */
int sbdg_get_trajectory(const sbdg_ctx *ctx, double *t, double *e, int cap, int *n);

/* Planned device bytes of a run, computed without a device: out[0] resident after the
 * build (an upper bound where table padding depends on the strings), out[1] the Davidson
 * vectors for opt's precision and max_nb (the basis, its sigma vectors, W, R, T; mixed:
 * the anchor's FP64 sigma and the FP32 input vector; sigma's scratch borrows W, R and T),
 * out[2] the RDMs' transient peak and outputs.  nnz, core_members and core_direct as
 * sbdg_stats reports them after a build (core_members[s] == 0: that spin's doubles go by
 * excitation table).  No counterpart in the reference, which sizes its Davidson memory by
 * hand (data/h2o/README.md:1). */
int sbdg_memory_plan(const sbdg_options *opt, int norb, int64_t na, int64_t nb, const int64_t *nnz,
                     const int64_t *core_members, const int64_t *core_direct, int64_t *out);

/* ---- lower-level entry points (parity tests and benchmarks) ---- */

/* which: 0 alpha singles, 1 alpha doubles, 2 beta singles, 3 beta doubles. */
int sbdg_table_nnz(const sbdg_ctx *ctx, int which, int64_t *nnz);
/* rowptr: n+1 int64; col: nnz int32; exc: nnz packed orbitals/sign; val: nnz FP64
 * separated-order string-pair values (see the library's internal header). */
int sbdg_get_table(sbdg_ctx *ctx, int which, int64_t *rowptr, int32_t *col, uint32_t *exc,
                   double *val);
int sbdg_get_hii(sbdg_ctx *ctx, double *hii);
/* sigma = H c for a host vector in the reference convention (mult, tpb/mult.h:376-632).
 * precision 0: FP64, 1: FP32 vectors/integrals (result widened to FP64). */
int sbdg_sigma(sbdg_ctx *ctx, const double *c, double *sigma, int precision);
/* sbdg_sigma on device vectors, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  x and y are distinct device buffers
 * of na*nb doubles owned by the caller (reference phases, as sbdg_sigma); the context must
 * be built (sbdg_build), else SBDG_ERR_STATE.  Its scratch is a static device array of
 * 256 MiB per device, held until the process ends: at most 8,388,608 determinants in FP64
 * (4 vectors), 13,421,772 in FP32 (5 FP32 vectors); larger problems get SBDG_ERR_LIMIT.
 * Every context on a device shares that device's array: calls that overlap on one device
 * (from several host threads) race.  Returns after the result is in y. */
int sbdg_sigma_device(sbdg_ctx *ctx, const double *x, double *y, int precision);
/* Times `reps` sigma builds on resident device vectors with HIP events on the context
 * stream.  ms[0] total per build; ms[1..5] per kernel phase (transpose, beta product,
 * alpha product + diagonal, transpose-add, alpha-beta) when options.overlap is 0.  With
 * overlap, ms[1..3] are the spin phases as they run beside the alpha-beta kernel, ms[4]
 * the wait for it after them and ms[5] the joining transpose-add. */
int sbdg_bench_sigma(sbdg_ctx *ctx, int precision, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* SBDG_H */
