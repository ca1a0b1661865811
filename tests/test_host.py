"""Host-side checks that need no GPU: the C-ABI library loads and exports every symbol its
header declares, the FCIDUMP / string loaders keep the reference's semantics, the synthetic
workload generator is deterministic and follows its specification, and the benchmark
refuses more than one GPU."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
import sbdg
from conftest import ROOT
from helpers import dataset
from sbdg import fcidump as fio
from sbdg import synthetic as syn


def header_symbols():
    text = open(os.path.join(ROOT, "include", "sbdg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sbdg_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    l = C.CDLL(sbdg.LIBPATH)
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(l, s), s
    # the Python binding declares a signature for every one of them, and nothing else
    assert sorted(sbdg.SIGNATURES) == syms


def test_binding_loads_and_rejects_missing_device():
    l = sbdg.lib()
    o = sbdg.default_options()
    assert (o.max_it, o.max_nb, o.eps, o.init, o.do_rdm, o.threshold) == (1, 10, 1e-4, 0, 0, 0.01)
    assert o.max_time == 86400.0 and o.restart_keep == 1 and o.precision == 0
    # no GPU in a CPU-only run: creating a context on a missing device yields NULL
    import torch
    if not torch.cuda.is_available():
        assert not l.sbdg_create(0)
    assert not l.sbdg_create(1 << 20)


def test_fcidump_loader_matches_reference_setup():
    header, vals, idx, strings = dataset("n2")
    assert header["NORB"] == "18" and header["NELEC"] == "14"
    L, core, h1, eri = fio.integrals_from_fcidump(header, vals, idx)
    I0, h1o, packed = oracle.setup_from_fcidump(L, vals, idx)
    assert core == I0 == 25.92968333508
    assert np.array_equal(h1, h1o)
    assert np.array_equal(oracle.pack_eri(eri), packed)
    # 8-fold symmetry of the expansion
    assert np.array_equal(eri, eri.transpose(1, 0, 2, 3))
    assert np.array_equal(eri, eri.transpose(2, 3, 0, 1))
    assert len(strings) == 239 and strings[0] == (1 << 7) - 1
    assert np.all(np.diff(strings.astype(np.int64)) > 0)


def test_fcidump_roundtrip_and_rounding(tmp_path):
    rng = np.random.default_rng(4)
    h1, eri = syn.synthetic_integrals(5, rng)
    f = tmp_path / "fcidump.txt"
    fio.write_fcidump(str(f), 5, 4, h1, eri, core=1.25)
    header, vals, idx = fio.load_fcidump(str(f))
    L, core, h1b, erib = fio.integrals_from_fcidump(header, vals, idx)
    assert L == 5 and core == 1.25
    assert np.array_equal(h1, h1b) and np.array_equal(eri, erib)
    r = fio.round_like_broadcast(np.array([2.323747945954686, -7.283619520573128e-05, 25.92968333508]))
    assert list(r) == [2.32375, -7.28362e-05, 25.9297]


def test_string_file_io(tmp_path):
    s = np.array([0b0111, 0b1011, 0b0111, 0b1101], dtype=np.uint64)
    f = tmp_path / "a.txt"
    fio.write_strings(str(f), s, 6)
    assert open(f).read().split() == ["000111", "001011", "000111", "001101"]
    got = fio.load_alpha_strings(str(f))
    assert list(got) == [0b0111, 0b1011, 0b1101]


@pytest.mark.parametrize("name", list(syn.CONFIGS))
def test_config_shapes(name):
    cfg = syn.CONFIGS[name]
    if cfg.get("na", 0) and cfg["na"] > 5000:
        pytest.skip("large sampled set: covered by the generator property test")
    wl = syn.config_workload(name)
    L, na_e = cfg["norb"], cfg["nalpha"]
    if "na" in cfg:
        assert len(wl.astr) == cfg["na"] and len(wl.bstr) == cfg["nb"]
    else:
        from math import comb
        assert len(wl.astr) == comb(L, na_e)
    for s in (wl.astr, wl.bstr):
        assert s[0] == (1 << na_e) - 1
        assert np.all(np.diff(s.astype(np.float64)) > 0)
        assert all(bin(int(x)).count("1") == na_e for x in s)
    assert np.allclose(np.diag(wl.h1), np.linspace(-2.0, 1.0, L))
    assert np.array_equal(wl.h1, wl.h1.T)
    e = wl.eri
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        assert np.array_equal(e, e.transpose(perm))
    M = e.reshape(L * L, L * L)
    assert np.linalg.eigvalsh(M).min() > -1e-10  # positive semidefinite pair matrix


def test_generator_deterministic_and_levels():
    a1 = syn.sample_strings(26, 5, 600, np.random.default_rng(9))
    a2 = syn.sample_strings(26, 5, 600, np.random.default_rng(9))
    assert np.array_equal(a1, a2)
    hf = (1 << 5) - 1
    lv = np.array([bin(int(x) & ~hf).count("1") for x in a1])
    # level weights 0.5^k: level 1 is exhausted (5*21 = 105 strings) and level 2 dominates
    assert (lv == 0).sum() == 1 and (lv == 1).sum() <= 105 and (lv == 2).sum() > (lv == 4).sum()
    with pytest.raises(ValueError):
        syn.sample_strings(6, 3, 21, np.random.default_rng(0))


def test_bench_measures_one_gpu_only():
    """The benchmark has no multi-GPU path: --gpus above 1 exits at once with a message,
    before anything touches a device."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


HBM_BYTES = 288e9  # MI355X


def test_c4_mixed_davidson_fits_hbm():
    """The largest workload (norb=58, 12+12 electrons, 40,000 x 40,000 strings, 1.6e9
    determinants) with the mixed Davidson at its settings (subspace 8, restart to 3): the
    library's planned device bytes (tables from the oracle's excitation counts; every
    string's C(12,2) electron pairs taken as core memberships and every double as a directly
    read pair, both upper bounds) stay below the 288 GB of HBM.  The FP64 basis alone is
    8 x 12.8 GB; the plan also shows what the FP64 Davidson would need."""
    wl = syn.config_workload("c4_L58_12e12e_40000")
    p = oracle.Problem(wl.norb, wl.core, wl.h1, wl.astr, wl.bstr, eri=wl.eri)
    nnz = [oracle.lib().orc_table_nnz(p._p, w) for w in range(4)]
    na, nb, N = len(wl.astr), len(wl.bstr), wl.ndet
    mem = [na * 66, nb * 66]
    plan = sbdg.memory_plan(wl.norb, na, nb, nnz, mem, [nnz[1], nnz[3]], precision=1, max_nb=8, restart_keep=3)
    # the Davidson vectors: 8 FP64 basis + 8 FP32 sigma vectors, W, R, T, HX (FP64), c32 (FP32)
    assert plan["davidson"] == N * (8 * 8 + 8 * 4 + 4 * 8 + 4)
    assert plan["total"] < HBM_BYTES, plan
    # the build: at least the diagonal (FP64 + FP32) and the 16 GiB core workspace
    assert plan["build"] > 12 * N + 16 * 2 ** 30
    fp64 = sbdg.memory_plan(wl.norb, na, nb, nnz, mem, [nnz[1], nnz[3]], precision=0, max_nb=8)
    assert fp64["davidson"] == N * (8 * 8 * 2 + 3 * 8)
    print(f"c4 plan: mixed {plan['total'] / 1e9:.1f} GB (build {plan['build'] / 1e9:.1f}, Davidson "
          f"{plan['davidson'] / 1e9:.1f}, RDM {plan['rdm'] / 1e9:.1f}); FP64 {fp64['total'] / 1e9:.1f} GB")


SHUFFLE_PROG = r"""
#include <cstdio>
#include "sbd_tpb.h"
int main(int argc, char** argv) {
  std::vector<std::vector<size_t>> det;
  sbd::LoadAlphaDets(argv[1], det, 20, 18);
  for (unsigned seed : {1729u, 137u}) {
    auto d = det;
    sbd::ShuffleDet(d, seed);
    for (auto& w : d) {
      unsigned long long x = 0;
      for (size_t k = 0; k < w.size(); ++k) x |= (unsigned long long)w[k] << (20 * k);
      std::printf("%llu\n", x);
    }
  }
  return 0;
}
"""


def test_shuffle_matches_reference(tmp_path):
    """--shuffle: sbd::ShuffleDet of the drop-in header puts the N2 threshold-1e-3 strings in
    the orders the reference's ShuffleDet (determinants.h:589-594) gave them with the
    sample's seeds 1729 and 137 (golden ref_shuffle.npz)."""
    src = tmp_path / "shuffle.cc"
    src.write_text(SHUFFLE_PROG)
    exe = tmp_path / "shuffle"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe), os.path.join(ROOT, "tests", "golden", "n2_1em3-alpha.txt")], text=True)
    got = np.array([int(x) for x in out.split()], dtype=np.uint64).reshape(2, -1)
    g = np.load(os.path.join(ROOT, "tests", "golden", "ref_shuffle.npz"))
    assert np.array_equal(got[0], g["seed1729"]) and np.array_equal(got[1], g["seed137"])
    assert got[0][0] == got[0].min()  # the Hartree-Fock string stays first
