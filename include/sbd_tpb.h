// sbd_tpb.h -- header-only C++ mirror of the reference's selected-basis API over libsbdg.
//
// Keeps the names, argument meaning and error behaviour of
//   sbd::FCIDump, sbd::LoadFCIDump            (framework/fcidump.h:18-139)
//   sbd::tpb::SBD, sbd::tpb::generate_sbd_data (chemistry/tpb/sbdiag.h:12-96)
//   sbd::tpb::diag                             (chemistry/tpb/sbdiag.h:113-124)
//   sbd::LoadAlphaDets, sbd::makestring       (makedeterminants.h:160-185,
//                                               bit_manipulation.h:135-150)
//   sbd::ShuffleDet                            (determinants.h:589-594)
// so a caller of the reference switches by including this header instead of sbd/sbd.h
// and linking libsbdg.so.  Differences:
//   * the communicator argument is accepted (any type) and ignored: one process, one GPU;
//   * FCIDUMP integrals keep full precision (the reference's MPI broadcast rounds them to
//     six digits, fcidump.h:156-169); SBD::round_integrals = 1 reproduces that rounding;
//   * the mixed-spin 2-RDM blocks two_p_rdm[1] and [2] are the physical ones (the
//     reference's sum two_p_rdm[1] + two_p_rdm[2] is unchanged);
//   * wavefunction files as one rank of the reference writes and reads them: savename
//     writes <savename>000000, loadname reads <loadname>000000 (restart.h:20-67,
//     414-657); dump_matrix_form_wf writes .bin / .txt / .dat (restart.h:662-780).
// Errors from the library are rethrown as std::invalid_argument (bad input) or
// std::runtime_error (everything else), as the reference throws.
#ifndef SBD_TPB_MI355X_H
#define SBD_TPB_MI355X_H

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "sbdg.h"

namespace sbd {

struct FCIDump {
  std::map<std::string, std::string> header;
  std::vector<std::tuple<double, int, int, int, int>> integrals;
};

// framework/fcidump.h:27-139: skip empty and '!' lines, concatenate the header (without
// "&FCI") up to the "&END" line, split it on ',' into trimmed KEY=VALUE pairs, then read
// "value i j k l" per line.
inline FCIDump LoadFCIDump(const std::string& filename) {
  FCIDump d;
  std::ifstream in(filename);
  if (!in.is_open()) throw std::runtime_error("Failed to open FCIDUMP file.");
  std::string line, head;
  bool in_header = true;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '!') continue;
    if (in_header) {
      if (line.find("&END") != std::string::npos) {
        in_header = false;
        continue;
      }
      size_t pos = line.find("&FCI");
      if (pos != std::string::npos) line.erase(pos, 4);
      head += line;
    } else {
      std::stringstream ss(line);
      double v;
      int i, j, k, l;
      if (ss >> v >> i >> j >> k >> l) d.integrals.emplace_back(v, i, j, k, l);
    }
  }
  std::stringstream hs(head);
  std::string tok;
  auto trim = [](std::string s) {
    s.erase(0, s.find_first_not_of(" \t"));
    s.erase(s.find_last_not_of(" \t") + 1);
    return s;
  };
  while (std::getline(hs, tok, ',')) {
    size_t eq = tok.find('=');
    if (eq != std::string::npos) d.header[trim(tok.substr(0, eq))] = trim(tok.substr(eq + 1));
  }
  return d;
}

// bit i of the string is orbital i, held in words of bit_length bits
inline std::vector<size_t> from_string(const std::string& s, size_t bit_length, size_t L) {
  std::vector<size_t> r((L + bit_length - 1) / bit_length, 0);
  for (size_t i = 0; i < L; ++i)
    if (s[L - 1 - i] == '1') r[i / bit_length] |= size_t(1) << (i % bit_length);
  return r;
}

inline std::string makestring(const std::vector<size_t>& c, size_t bit_length, size_t L) {
  std::string s;
  for (size_t i = L; i > 0; --i) s += (c[(i - 1) / bit_length] >> ((i - 1) % bit_length)) & 1 ? '1' : '0';
  return s;
}

// text files only (makedeterminants.h:160-185); sorted by value, duplicates removed
inline void LoadAlphaDets(const std::string& file, std::vector<std::vector<size_t>>& adet, size_t bit_length,
                          size_t L) {
  std::ifstream in(file);
  if (!in.is_open()) throw std::runtime_error("Failed to open alpha det file.");
  std::string line;
  adet.clear();
  while (std::getline(in, line))
    if (!line.empty()) adet.push_back(from_string(line, bit_length, L));
  auto less = [](const std::vector<size_t>& a, const std::vector<size_t>& b) {
    for (size_t n = a.size(); n > 0; --n)
      if (a[n - 1] != b[n - 1]) return a[n - 1] < b[n - 1];
    return false;
  };
  std::sort(adet.begin(), adet.end(), less);
  adet.erase(std::unique(adet.begin(), adet.end()), adet.end());
}

// ShuffleDet (determinants.h:589-594): the strings after the first (the Hartree-Fock one
// stays in front) in the order std::shuffle gives with a std::mt19937 seeded by `seed`;
// the sample driver applies it with seeds 1729 (alpha) and 137 (beta) under --shuffle
// (samples/selected_basis_diagonalization/main.cc:123-128).
inline void ShuffleDet(std::vector<std::vector<size_t>>& det, unsigned int seed) {
  if (det.size() <= 1) return;
  std::mt19937 g(seed);
  std::shuffle(det.begin() + 1, det.end(), g);
}

namespace tpb {

struct SBD {
  int task_comm_size = 1;  // accepted, unused: one GPU
  int adet_comm_size = 1;
  int bdet_comm_size = 1;
  int h_comm_size = 1;
  int method = 0;          // 0 only (direct sigma)
  int max_it = 1;
  int max_nb = 10;
  double eps = 1.0e-4;
  double max_time = 86400.0;
  int init = 0;
  int do_shuffle = 0;
  int do_rdm = 0;
  double ratio = 0.0;
  double threshold = 0.01;
  size_t bit_length = 20;
  std::string dump_matrix_form_wf;
  // additions (reference-neutral defaults)
  int restart_keep = 1;
  int precision = 0;       // 1: FP32 sigma with FP64 refresh
  int refresh_k = 4;
  int round_integrals = 0; // 1: six-significant-digit integrals, as the reference's broadcast
  int device = 0;
  int verbose = 0;
};

inline SBD generate_sbd_data(int argc, char* argv[]) {
  SBD d;
  for (int i = 0; i + 1 < argc; ++i) {
    const std::string a = argv[i];
    const char* v = argv[i + 1];
    bool hit = true;
    if (a == "--adet_comm_size") d.adet_comm_size = std::atoi(v);
    else if (a == "--bdet_comm_size") d.bdet_comm_size = std::atoi(v);
    else if (a == "--task_comm_size") d.task_comm_size = std::atoi(v);
    else if (a == "--method") d.method = std::atoi(v);
    else if (a == "--iteration") d.max_it = std::atoi(v);
    else if (a == "--block") d.max_nb = std::atoi(v);
    else if (a == "--tolerance") d.eps = std::atof(v);
    else if (a == "--carryover_ratio") d.ratio = std::atof(v);
    else if (a == "--carryover_threshold") d.threshold = std::atof(v);
    else if (a == "--max_time") d.max_time = std::atof(v);
    else if (a == "--shuffle") d.do_shuffle = std::atoi(v);
    else if (a == "--rdm") d.do_rdm = std::atoi(v);
    else if (a == "--bit_length") d.bit_length = (size_t)std::atoi(v);
    else if (a == "--dump_matrix_form_wf") d.dump_matrix_form_wf = v;
    else if (a == "--restart_keep") d.restart_keep = std::atoi(v);
    else if (a == "--precision") d.precision = std::atoi(v);
    else if (a == "--refresh") d.refresh_k = std::atoi(v);
    else if (a == "--round_integrals") d.round_integrals = std::atoi(v);
    else if (a == "--device") d.device = std::atoi(v);
    else if (a == "--verbose") d.verbose = std::atoi(v);
    else hit = false;
    if (hit) ++i;
  }
  return d;
}

namespace detail {

inline void check(sbdg_ctx* c, int rc) {
  if (rc == SBDG_OK) return;
  std::string m = c ? sbdg_last_error(c) : "no context";
  if (rc == SBDG_ERR_ARG) throw std::invalid_argument(m);
  throw std::runtime_error(m);
}

inline uint64_t to_u64(const std::vector<size_t>& s, size_t bit_length, int L) {
  uint64_t x = 0;
  for (int i = 0; i < L; ++i)
    if ((s[i / bit_length] >> (i % bit_length)) & 1) x |= uint64_t(1) << i;
  return x;
}

inline double round6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.6g", v);
  return std::atof(buf);
}

struct Ctx {
  sbdg_ctx* c;
  explicit Ctx(int dev) : c(sbdg_create(dev)) {
    if (!c) throw std::runtime_error("no HIP device");
  }
  ~Ctx() { sbdg_destroy(c); }
};

}  // namespace detail

// sbdiag.h:113-124 without the communicator.
inline void diag(const SBD& o, const sbd::FCIDump& fcidump, const std::vector<std::vector<size_t>>& adet,
                 const std::vector<std::vector<size_t>>& bdet, const std::string& loadname,
                 const std::string& savename, double& energy, std::vector<double>& density,
                 std::vector<std::vector<size_t>>& carryover_bitstrings, std::vector<std::vector<double>>& one_p_rdm,
                 std::vector<std::vector<double>>& two_p_rdm) {
  if (o.method != 0) throw std::invalid_argument("only method 0 (direct sigma) is supported");
  int L = 0;
  for (const auto& kv : fcidump.header)
    if (kv.first == "NORB") L = std::atoi(kv.second.c_str());
  if (L < 1 || L > 64) throw std::invalid_argument("NORB must be in 1..64");
  // SetupIntegrals (makeintegrals.h:10-55) into h1 and the packed store
  const int64_t B = (int64_t)L * (L + 1) / 2;
  std::vector<double> h1((size_t)L * L, 0.0), packed((size_t)(B * (B + 1) / 2), 0.0);
  double I0 = 0.0;
  auto pr = [](int64_t i, int64_t j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; };
  for (const auto& t : fcidump.integrals) {
    double v = std::get<0>(t);
    if (o.round_integrals) v = detail::round6(v);
    const int i = std::get<1>(t), j = std::get<2>(t), k = std::get<3>(t), l = std::get<4>(t);
    if (i == 0 && j == 0 && k == 0 && l == 0) {
      I0 = v;
    } else if (k == 0 && l == 0) {
      if (i < 1 || j < 1 || i > L || j > L) throw std::invalid_argument("FCIDUMP index out of range");
      h1[(size_t)(i - 1) * L + (j - 1)] = v;
      h1[(size_t)(j - 1) * L + (i - 1)] = v;
    } else {
      if (i < 1 || j < 1 || k < 1 || l < 1 || i > L || j > L || k > L || l > L)
        throw std::invalid_argument("FCIDUMP index out of range");
      const int64_t ij = pr(i - 1, j - 1), kl = pr(k - 1, l - 1);
      const int64_t a = ij > kl ? ij : kl, b = ij > kl ? kl : ij;
      packed[(size_t)(a * (a + 1) / 2 + b)] = v;
    }
  }
  std::vector<uint64_t> a(adet.size()), b(bdet.size());
  for (size_t i = 0; i < adet.size(); ++i) a[i] = detail::to_u64(adet[i], o.bit_length, L);
  for (size_t i = 0; i < bdet.size(); ++i) b[i] = detail::to_u64(bdet[i], o.bit_length, L);

  detail::Ctx ctx(o.device);
  sbdg_ctx* c = ctx.c;
  detail::check(c, sbdg_set_integrals(c, L, I0, h1.data(), packed.data(), 1));
  detail::check(c, sbdg_set_strings(c, a.data(), (int64_t)a.size(), b.data(), (int64_t)b.size()));
  sbdg_options so;
  sbdg_default_options(&so);
  so.max_it = o.max_it;
  so.max_nb = o.max_nb;
  so.eps = o.eps;
  so.max_time = o.max_time;
  so.init = o.init;
  so.do_rdm = o.do_rdm;
  so.ratio = o.ratio;
  so.threshold = o.threshold;
  so.restart_keep = o.restart_keep;
  so.precision = o.precision;
  so.refresh_k = o.refresh_k;
  so.verbose = o.verbose;
  detail::check(c, sbdg_set_options(c, &so));
  if (!loadname.empty())  // sbdiag.h:197-203
    detail::check(c, sbdg_load_wavefunction(c, loadname.c_str(), (int)o.bit_length));
  detail::check(c, sbdg_run(c));
  detail::check(c, sbdg_get_energy(c, &energy, nullptr, nullptr));
  density.assign(2 * L, 0.0);
  if (o.do_rdm) {
    const size_t L2 = (size_t)L * L, L4 = L2 * L2;
    std::vector<double> one(2 * L2), two(4 * L4);
    detail::check(c, sbdg_get_rdm(c, one.data(), two.data()));
    one_p_rdm.assign(2, std::vector<double>(L2));
    two_p_rdm.assign(4, std::vector<double>(L4));
    for (int s = 0; s < 2; ++s) std::copy(one.begin() + s * L2, one.begin() + (s + 1) * L2, one_p_rdm[s].begin());
    for (int s = 0; s < 4; ++s) std::copy(two.begin() + s * L4, two.begin() + (s + 1) * L4, two_p_rdm[s].begin());
    for (int i = 0; i < L; ++i) {  // sbdiag.h:412-416
      density[2 * i] = one_p_rdm[0][i + (size_t)L * i];
      density[2 * i + 1] = one_p_rdm[1][i + (size_t)L * i];
    }
  } else {
    detail::check(c, sbdg_get_density(c, density.data()));
  }
  std::vector<uint64_t> co(a.size());
  int64_t cnt = 0;
  double tw = 0.0;
  detail::check(c, sbdg_carryover(c, o.ratio, o.threshold, co.data(), (int64_t)co.size(), &cnt, &tw));
  carryover_bitstrings.clear();
  const size_t words = (L + o.bit_length - 1) / o.bit_length;
  for (int64_t k = 0; k < cnt && k < (int64_t)co.size(); ++k) {
    std::vector<size_t> w(words, 0);
    for (int i = 0; i < L; ++i)
      if ((co[k] >> i) & 1) w[i / o.bit_length] |= size_t(1) << (i % o.bit_length);
    carryover_bitstrings.push_back(w);
  }
  if (!savename.empty())  // sbdiag.h:459-465
    detail::check(c, sbdg_save_wavefunction(c, savename.c_str(), (int)o.bit_length));
  if (!o.dump_matrix_form_wf.empty())  // sbdiag.h:468-474
    detail::check(c, sbdg_save_matrix_form(c, o.dump_matrix_form_wf.c_str(), (int)o.bit_length));
}

// The reference signature with its communicator first (any type; ignored).
template <class Comm>
inline void diag(const Comm&, const SBD& o, const sbd::FCIDump& fcidump,
                 const std::vector<std::vector<size_t>>& adet, const std::vector<std::vector<size_t>>& bdet,
                 const std::string& loadname, const std::string& savename, double& energy,
                 std::vector<double>& density, std::vector<std::vector<size_t>>& carryover_bitstrings,
                 std::vector<std::vector<double>>& one_p_rdm, std::vector<std::vector<double>>& two_p_rdm) {
  diag(o, fcidump, adet, bdet, loadname, savename, energy, density, carryover_bitstrings, one_p_rdm, two_p_rdm);
}

}  // namespace tpb
}  // namespace sbd

#endif
