// ref_driver.cc -- runs the reference's own selected-basis code (walkup/sbd-ykt headers,
// compiled from where they lie) to produce golden vectors for the test suite.
//
// Used only in the container that holds the reference sources; its outputs are committed
// as fixtures under tests/golden/.  Inputs are full-precision files (no MPI broadcast of
// the FCIDUMP, which would round every integral to six digits: fcidump.h:156-169).
//
//   sbd_ref diag  --fcidump F --adet A --bdet B --block nb --iteration it --tolerance eps
//                 --rdm r --out PREFIX [--load NAME] [--save NAME] [--dump_matrix_form_wf FILE]
//       sbd::tpb::diag (sbdiag.h:113-477) -> PREFIX.energy.txt, PREFIX.wf.bin (unless
//       another matrix-form file is named), PREFIX.rdm1.bin (2*L^2), PREFIX.rdm2.bin
//       (4*L^4), PREFIX.density.bin (2*L), PREFIX.carryover.bin (carryover_bitstrings,
//       sbdiag.h:436-453: int64 count, int64 words per string, then the words);
//       --load/--save are diag's loadname/savename (restart.h:20-67,414-657)
//   sbd_ref sigma --fcidump F --adet A --bdet B --vec V --out PREFIX [--corr 1]
//       helpers (helper.h:447-589) + diagonal (tpb/qcham.h:314-366) + one direct mult
//       (tpb/mult.h:376-632) -> PREFIX.hii.bin, PREFIX.sigma.bin, PREFIX.tables.bin,
//       PREFIX.sigma_time.txt; with --corr 1 also Correlation (tpb/correlation.h:194-444)
//       of the vector scaled to unit norm -> PREFIX.rdm1.bin, PREFIX.rdm2.bin,
//       PREFIX.corr_time.txt
//   sbd_ref shuffle --fcidump F --adet A --bdet A --out PREFIX
//       ShuffleDet (determinants.h:589-594) of the alpha strings with the sample's seeds 1729
//       and 137 (main.cc:123-128) -> PREFIX.shuffle.bin (int64 count, words, then the
//       strings in the seed-1729 order, then in the seed-137 order)
#include <cassert>
#include <chrono>
#include <cmath>
#include <deque>
#include <fstream>
#include <iostream>
#include <numeric>
#include <random>
#include <unistd.h>
#include <vector>
using std::endl;  // tpb/davidson.h:901 uses an unqualified endl inside namespace sbd
#include "mpi.h"
#include "sbd/sbd.h"

static std::string arg(int argc, char** argv, const std::string& key, const std::string& def) {
  for (int i = 0; i + 1 < argc; ++i)
    if (key == argv[i]) return argv[i + 1];
  return def;
}

template <class T>
static void write_bin(const std::string& name, const std::vector<T>& v) {
  std::ofstream o(name, std::ios::binary);
  o.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
}

int main(int argc, char** argv) {
  int provided;
  MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
  MPI_Comm comm = MPI_COMM_WORLD;
  std::string mode = argc > 1 ? argv[1] : "diag";
  std::string out = arg(argc, argv, "--out", "ref");
  auto sbd_data = sbd::tpb::generate_sbd_data(argc, argv);
  sbd::FCIDump fcidump = sbd::LoadFCIDump(arg(argc, argv, "--fcidump", "fcidump.txt"));
  int L = 0, N = 0;
  for (const auto& [key, value] : fcidump.header)
    if (key == "NORB") L = std::atoi(value.c_str());
  std::vector<std::vector<size_t>> adet, bdet;
  sbd::LoadAlphaDets(arg(argc, argv, "--adet", "adet.txt"), adet, sbd_data.bit_length, L);
  sbd::LoadAlphaDets(arg(argc, argv, "--bdet", "bdet.txt"), bdet, sbd_data.bit_length, L);
  std::cout.precision(17);

  if (mode == "diag") {
    double energy;
    std::vector<double> density;
    std::vector<std::vector<size_t>> cobits;
    std::vector<std::vector<double>> one, two;
    if (sbd_data.dump_matrix_form_wf.empty()) sbd_data.dump_matrix_form_wf = out + ".wf.bin";
    const std::string loadname = arg(argc, argv, "--load", ""), savename = arg(argc, argv, "--save", "");
    auto t0 = std::chrono::high_resolution_clock::now();
    sbd::tpb::diag(comm, sbd_data, fcidump, adet, bdet, loadname, savename, energy, density, cobits, one, two);
    auto t1 = std::chrono::high_resolution_clock::now();
    std::ofstream e(out + ".energy.txt");
    e.precision(17);
    e << energy << "\n" << std::chrono::duration<double>(t1 - t0).count() << "\n";
    if (!one.empty()) {
      std::vector<double> o, t;
      for (auto& v : one) o.insert(o.end(), v.begin(), v.end());
      for (auto& v : two) t.insert(t.end(), v.begin(), v.end());
      write_bin(out + ".rdm1.bin", o);
      write_bin(out + ".rdm2.bin", t);
    }
    write_bin(out + ".density.bin", density);
    {
      std::vector<int64_t> co;
      co.push_back((int64_t)cobits.size());
      co.push_back(cobits.empty() ? 0 : (int64_t)cobits[0].size());
      for (auto& w : cobits)
        for (size_t x : w) co.push_back((int64_t)x);
      write_bin(out + ".carryover.bin", co);
    }
    std::cout << " ref energy " << energy << std::endl;
  } else if (mode == "sigma") {
    double I0 = 0.0;
    sbd::oneInt<double> I1;
    sbd::twoInt<double> I2;
    sbd::SetupIntegrals(fcidump, L, N, I0, I1, I2);
    MPI_Comm h_comm, b_comm, t_comm;
    sbd::TaskCommunicator(comm, 1, 1, 1, 1, h_comm, b_comm, t_comm);
    std::vector<sbd::TaskHelpers> helper;
    std::vector<std::vector<size_t>> shm;
    sbd::MakeHelpers(adet, bdet, sbd_data.bit_length, L, helper, shm, h_comm, b_comm, t_comm, 1, 1);
    std::vector<double> hii;
    sbd::makeQChamDiagTerms(adet, bdet, sbd_data.bit_length, L, helper, I0, I1, I2, hii, h_comm,
                            b_comm, t_comm);
    size_t n = adet.size() * bdet.size();
    std::vector<double> vec(n);
    {
      std::ifstream in(arg(argc, argv, "--vec", "vec.bin"), std::ios::binary);
      in.read(reinterpret_cast<char*>(vec.data()), sizeof(double) * n);
    }
    std::vector<double> sig(n, 0.0);
    auto t0 = std::chrono::high_resolution_clock::now();
    sbd::mult(hii, vec, sig, adet, bdet, sbd_data.bit_length, (size_t)L, 1, 1, helper, I0, I1, I2,
              h_comm, b_comm, t_comm);
    auto t1 = std::chrono::high_resolution_clock::now();
    write_bin(out + ".hii.bin", hii);
    write_bin(out + ".sigma.bin", sig);
    // tables: for each of Sa, Da, Sb, Db: the per-row lengths then the entries (int64).
    std::vector<int64_t> tab;
    const auto& h = helper[0];
    auto dump = [&](size_t rows, size_t* len, const std::vector<size_t*>& sm) {
      for (size_t i = 0; i < rows; ++i) tab.push_back((int64_t)len[i]);
      for (size_t i = 0; i < rows; ++i)
        for (size_t j = 0; j < len[i]; ++j) tab.push_back((int64_t)sm[i][j]);
    };
    dump(adet.size(), h.SinglesFromAlphaLen, h.SinglesFromAlphaSM);
    dump(adet.size(), h.DoublesFromAlphaLen, h.DoublesFromAlphaSM);
    dump(bdet.size(), h.SinglesFromBetaLen, h.SinglesFromBetaSM);
    dump(bdet.size(), h.DoublesFromBetaLen, h.DoublesFromBetaSM);
    write_bin(out + ".tables.bin", tab);
    std::ofstream e(out + ".sigma_time.txt");
    e.precision(17);
    e << std::chrono::duration<double>(t1 - t0).count() << "\n";
    if (arg(argc, argv, "--corr", "0") == "1") {
      double nn = 0.0;
      for (double x : vec) nn += x * x;
      for (double& x : vec) x /= std::sqrt(nn);
      std::vector<std::vector<double>> one, two;
      auto c0 = std::chrono::high_resolution_clock::now();
      sbd::Correlation(vec, adet, bdet, sbd_data.bit_length, (size_t)L, 1, 1, helper, h_comm, b_comm, t_comm,
                       one, two);
      auto c1 = std::chrono::high_resolution_clock::now();
      std::vector<double> o, t;
      for (auto& v : one) o.insert(o.end(), v.begin(), v.end());
      for (auto& v : two) t.insert(t.end(), v.begin(), v.end());
      write_bin(out + ".rdm1.bin", o);
      write_bin(out + ".rdm2.bin", t);
      std::ofstream ec(out + ".corr_time.txt");
      ec.precision(17);
      ec << std::chrono::duration<double>(c1 - c0).count() << "\n";
    }
  }
  if (mode == "shuffle") {
    std::vector<int64_t> outw;
    outw.push_back((int64_t)adet.size());
    outw.push_back(adet.empty() ? 0 : (int64_t)adet[0].size());
    for (unsigned int seed : {1729u, 137u}) {
      auto d = adet;
      sbd::ShuffleDet(d, seed);
      for (auto& w : d)
        for (size_t x : w) outw.push_back((int64_t)x);
    }
    write_bin(out + ".shuffle.bin", outw);
  }
  MPI_Finalize();
  return 0;
}
