// sbd_diag -- command-line driver with the reference sample's interface
// (samples/selected_basis_diagonalization/main.cc:17-226): the SBD options of
// generate_sbd_data plus --fcidump, --adetfile (beta strings = alpha strings, as the
// sample does), --bdetfile (separate beta strings), --carryoverfile,
// --loadname / --savename (wavefunction restart files), --shuffle 1 (the strings after the
// first in a seeded random order, main.cc:123-128).  With --rdm 1 it writes 1pRDM.txt and
// 2pRDM.txt (spin-summed, main.cc:172-205) in the working directory.
// Prints the energy, the density, the carry-over strings and, with --rdm 1, the
// one-body + two-body energy identity.
#include <cstdio>
#include <iostream>

#include "sbd_tpb.h"

static std::string arg(int argc, char** argv, const char* key, const char* def) {
  for (int i = 0; i + 1 < argc; ++i)
    if (std::string(argv[i]) == key) return argv[i + 1];
  return def;
}

int main(int argc, char** argv) {
  try {
    auto o = sbd::tpb::generate_sbd_data(argc, argv);
    const std::string fcifile = arg(argc, argv, "--fcidump", "fcidump.txt");
    const std::string afile = arg(argc, argv, "--adetfile", "alphadets.txt");
    const std::string bfile = arg(argc, argv, "--bdetfile", "");
    const std::string cofile = arg(argc, argv, "--carryoverfile", "");
    const std::string loadname = arg(argc, argv, "--loadname", "");  // main.cc:44-50
    const std::string savename = arg(argc, argv, "--savename", "");
    auto fcidump = sbd::LoadFCIDump(fcifile);
    int L = 0;
    for (const auto& kv : fcidump.header)
      if (kv.first == "NORB") L = std::atoi(kv.second.c_str());
    std::vector<std::vector<size_t>> adet, bdet;
    sbd::LoadAlphaDets(afile, adet, o.bit_length, L);
    if (bfile.empty()) bdet = adet;
    else sbd::LoadAlphaDets(bfile, bdet, o.bit_length, L);
    if (o.do_shuffle != 0) {  // main.cc:123-128
      sbd::ShuffleDet(adet, 1729);
      sbd::ShuffleDet(bdet, 137);
    }
    double energy = 0.0;
    std::vector<double> density;
    std::vector<std::vector<size_t>> cobits;
    std::vector<std::vector<double>> one, two;
    sbd::tpb::diag(o, fcidump, adet, bdet, loadname, savename, energy, density, cobits, one, two);
    std::cout.precision(16);
    std::cout << " Sample-based diagonalization: Energy = " << energy << std::endl;
    std::cout << " Sample-based diagonalization: density = ";
    for (size_t i = 0; i < density.size() / 2; ++i)
      std::cout << (i == 0 ? "[" : ",") << density[2 * i] + density[2 * i + 1];
    std::cout << "]" << std::endl;
    std::cout << " Sample-based diagonalization: carryover bitstrings = " << cobits.size() << std::endl;
    if (!cofile.empty()) {
      std::ofstream f(cofile);
      for (auto& c : cobits) f << sbd::makestring(c, o.bit_length, L) << "\n";
    }
    if (!one.empty()) {
      // energy identity of main.cc:166-216 from the RDMs and the integrals
      const size_t L2 = (size_t)L * L;
      std::vector<double> h(L2, 0.0), eri(L2 * L2, 0.0);
      double I0 = 0.0;
      for (auto& t : fcidump.integrals) {
        int i = std::get<1>(t) - 1, j = std::get<2>(t) - 1, k = std::get<3>(t) - 1, l = std::get<4>(t) - 1;
        double v = std::get<0>(t);
        if (o.round_integrals) v = sbd::tpb::detail::round6(v);
        if (i < 0 && j < 0 && k < 0 && l < 0) {
          I0 = v;
        } else if (k < 0 && l < 0) {
          h[i * L + j] = h[j * L + i] = v;
        } else {
          const int q[8][4] = {{i, j, k, l}, {j, i, k, l}, {i, j, l, k}, {j, i, l, k},
                               {k, l, i, j}, {l, k, i, j}, {k, l, j, i}, {l, k, j, i}};
          for (auto& x : q) eri[((size_t)(x[0] * L + x[1]) * L + x[2]) * L + x[3]] = v;
        }
      }
      double e1 = 0.0, e2 = 0.0;
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j) e1 += h[i * L + j] * (one[0][i + L * j] + one[1][i + L * j]);
      for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
          for (int k = 0; k < L; ++k)
            for (int l = 0; l < L; ++l) {
              const size_t x = i + (size_t)L * j + L2 * k + L2 * L * l;
              e2 += 0.5 * eri[((size_t)(i * L + k) * L + j) * L + l] *
                    (two[0][x] + two[1][x] + two[2][x] + two[3][x]);
            }
      // the spin-summed RDM dumps of main.cc:172-205: "i j value" and "i j k l value"
      {
        std::ofstream f1("1pRDM.txt");
        f1.precision(16);
        for (int i = 0; i < L; ++i)
          for (int j = 0; j < L; ++j) f1 << i << " " << j << " " << one[0][i + L * j] + one[1][i + L * j] << "\n";
        std::ofstream f2("2pRDM.txt");
        f2.precision(16);
        for (int i = 0; i < L; ++i)
          for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
              for (int l = 0; l < L; ++l) {
                const size_t x = i + (size_t)L * j + L2 * (k + (size_t)L * l);
                f2 << i << " " << j << " " << k << " " << l << " " << two[0][x] + two[1][x] + two[2][x] + two[3][x]
                   << "\n";
              }
      }
      std::cout << " One-Body energy = " << e1 << std::endl;
      std::cout << " Two-Body energy = " << e2 << std::endl;
      std::cout << " One-Body + Two-Body energy = " << e1 + e2 << " (+ core " << I0 << ")" << std::endl;
    }
  } catch (const std::exception& e) {
    std::cerr << "sbd_diag: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
