"""Writes the golden fixtures of this directory by running the reference's own code.

Needs the reference sources (compiled by ``make -C oracle ref`` into oracle/_ref/sbd_ref)
and therefore runs only where they are present.  Every fixture stores its inputs (or the
seeded recipe plus a checksum of them) next to the reference's outputs:

* ``ref_<case>.npz`` -- reference tables, diagonal and one sigma on a seeded vector
  (tpb/helper.h:132-197, tpb/qcham.h:314-366, tpb/mult.h:376-632), and a full
  ``sbd::tpb::diag`` (energy, 1-/2-RDM, density; sbdiag.h:113-477) at tol 1e-10.
* ``ref_c0.npz`` -- the norb=12, 4+4 full-space workload (seed 0): diag at tol 1e-8 with
  RDMs, plus checksums of one sigma on a seeded vector (the vector itself is 2 MB).
* ``ref_restart.npz`` -- the wavefunction files of diag (restart.h): the per-rank restart
  file written by ``savename`` (bytes), the matrix-form dump in text and binary form, and
  the vector obtained by loading that restart file into a different basis (``loadname``
  with a tolerance so loose that Davidson stops at the loaded, normalised vector).
* ``ref_n2_1em3.npz`` / ``ref_h2o_1em3.npz`` -- the reference datasets (data/n2, data/h2o)
  with the threshold-1e-3 string files: energy and 1-RDM from full-precision integrals,
  and the energy from integrals rounded to six digits (the published-table path).

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "sbd-ykt_amd"))
from sbdg import fcidump as fio  # noqa: E402
from sbdg import synthetic as syn  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "sbd_ref")
REFDATA = "/root/reference/data"


def run_ref(args, cwd):
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"),
               MKL_THREADING_LAYER="SEQUENTIAL")
    subprocess.check_call([REF] + args, cwd=cwd, env=env, stdout=subprocess.DEVNULL)


def rd(path, dtype=np.float64):
    return np.fromfile(path, dtype=dtype)


def checksum(*arrays):
    h = 0.0
    for a in arrays:
        a = np.asarray(a, dtype=np.float64).ravel()
        w = np.cos(np.arange(a.size) * 0.618034)
        h += float(a @ w)
    return h


def ref_case(name, wl, seed_vec, tol=1e-10, block=10, iteration=200, with_sigma=True,
             store_inputs=True, rdm=True):
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        L = wl.norb
        fio.write_fcidump(os.path.join(tmp, "fcidump.txt"), L, wl.nalpha + wl.nbeta, wl.h1, wl.eri,
                          core=wl.core, ms2=wl.nalpha - wl.nbeta)
        fio.write_strings(os.path.join(tmp, "a.txt"), wl.astr, L)
        fio.write_strings(os.path.join(tmp, "b.txt"), wl.bstr, L)
        out = {}
        common = ["--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "b.txt"]
        if with_sigma:
            vec = np.random.default_rng(seed_vec).normal(size=wl.ndet)
            vec.tofile(os.path.join(tmp, "vec.bin"))
            run_ref(["sigma"] + common + ["--vec", "vec.bin", "--out", "s"], tmp)
            out["vec_seed"] = seed_vec
            out["hii"] = rd(os.path.join(tmp, "s.hii.bin"))
            out["sigma"] = rd(os.path.join(tmp, "s.sigma.bin"))
            out["tables"] = rd(os.path.join(tmp, "s.tables.bin"), np.int64)
        run_ref(["diag"] + common + ["--block", str(block), "--iteration", str(iteration),
                                     "--tolerance", str(tol), "--rdm", "1" if rdm else "0",
                                     "--out", "d"], tmp)
        with open(os.path.join(tmp, "d.energy.txt")) as f:
            out["energy"] = float(f.readline())
        out["wf"] = rd(os.path.join(tmp, "d.wf.bin"))
        if rdm:
            out["rdm1"] = rd(os.path.join(tmp, "d.rdm1.bin")).reshape(2, L * L)
            out["rdm2"] = rd(os.path.join(tmp, "d.rdm2.bin")).reshape(4, L ** 4)
        out["density"] = rd(os.path.join(tmp, "d.density.bin"))
        if store_inputs:
            out.update(norb=L, nalpha=wl.nalpha, nbeta=wl.nbeta, core=wl.core, h1=wl.h1, eri=wl.eri,
                       astr=wl.astr, bstr=wl.bstr)
        out["tol"] = tol
        return out
    finally:
        shutil.rmtree(tmp)


def case_small():
    cases = {
        # alpha set != beta set, different electron counts
        "asym_L8": syn.Workload(*_sampled(8, 3, 2, 30, 20, 11)),
        "asym_L14": syn.Workload(*_sampled(14, 4, 3, 150, 120, 12)),
        # alpha set == beta set, full space
        "sym_L10": syn.make_workload(10, 3, 3, 13),
    }
    for name, wl in cases.items():
        out = ref_case(name, wl, seed_vec=100)
        np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), **out)
        print(name, wl.ndet, out["energy"])


def _sampled(L, na_e, nb_e, na, nb, seed):
    rng = np.random.default_rng(seed)
    h1, eri = syn.synthetic_integrals(L, rng)
    a = syn.sample_strings(L, na_e, na, rng)
    b = syn.sample_strings(L, nb_e, nb, rng)
    return (L, na_e, nb_e, 0.0, h1, eri, a, b)


def case_c0():
    wl = syn.config_workload("c0_L12_4e4e_full")
    out = ref_case("c0", wl, seed_vec=0, tol=1e-8, block=10, iteration=200, with_sigma=True,
                   store_inputs=False)
    sig = out.pop("sigma")
    vec = np.random.default_rng(0).normal(size=wl.ndet)
    hii = out.pop("hii")
    tables = out.pop("tables")
    wf = out.pop("wf")
    res = dict(energy=out["energy"], rdm1=out["rdm1"], rdm2=out["rdm2"], density=out["density"],
               tol=out["tol"],
               input_checksum=checksum(wl.h1, wl.eri, wl.astr.astype(np.float64), wl.bstr.astype(np.float64)),
               sigma_checksum=checksum(sig), sigma_norm=float(np.linalg.norm(sig)),
               sigma_dot_vec=float(sig @ vec), sigma_head=sig[:4096].copy(),
               hii_checksum=checksum(hii), hii_head=hii[:4096].copy(),
               tables_len=len(tables), tables_checksum=float(np.sum(tables.astype(np.float64) *
                                                                   np.cos(np.arange(len(tables)) * 0.618034))),
               wf_abs_head=np.abs(wf[:4096]).copy(), wf_norm=float(np.linalg.norm(wf)))
    np.savez_compressed(os.path.join(HERE, "ref_c0.npz"), **res)
    print("c0", wl.ndet, res["energy"])


def restart_bases():
    """The asym_L8 workload (basis 1) and basis 2: five alpha and two beta strings
    dropped, four alpha and three beta strings of the full space added."""
    import itertools
    wl = syn.Workload(*_sampled(8, 3, 2, 30, 20, 11))
    L = wl.norb

    def full(n):
        return [sum(1 << o for o in c) for c in itertools.combinations(range(L), n)]

    a1, b1 = [int(x) for x in wl.astr], [int(x) for x in wl.bstr]
    a2 = [x for i, x in enumerate(a1) if i not in (3, 7, 11, 19, 25)]
    a2 += [x for x in full(wl.nalpha) if x not in a1][:4]
    b2 = [x for i, x in enumerate(b1) if i not in (2, 9)]
    b2 += [x for x in full(wl.nbeta) if x not in b1][-3:]
    return wl, np.array(sorted(a2), dtype=np.uint64), np.array(sorted(b2), dtype=np.uint64)


def case_restart():
    wl, a2, b2 = restart_bases()
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        L = wl.norb
        fio.write_fcidump(os.path.join(tmp, "fcidump.txt"), L, wl.nalpha + wl.nbeta, wl.h1, wl.eri,
                          core=wl.core, ms2=wl.nalpha - wl.nbeta)
        fio.write_strings(os.path.join(tmp, "a.txt"), wl.astr, L)
        fio.write_strings(os.path.join(tmp, "b.txt"), wl.bstr, L)
        fio.write_strings(os.path.join(tmp, "a2.txt"), a2, L)
        fio.write_strings(os.path.join(tmp, "b2.txt"), b2, L)
        run_ref(["diag", "--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "b.txt", "--block", "10",
                 "--iteration", "200", "--tolerance", "1e-10", "--rdm", "0", "--out", "d",
                 "--save", "rs", "--dump_matrix_form_wf", "mf.txt"], tmp)
        with open(os.path.join(tmp, "d.energy.txt")) as f:
            e1 = float(f.readline())
        rs = np.fromfile(os.path.join(tmp, "rs000000"), dtype=np.uint8)
        with open(os.path.join(tmp, "mf.txt")) as f:
            mf_txt = f.read()
        run_ref(["diag", "--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "b.txt", "--block", "10",
                 "--iteration", "200", "--tolerance", "1e-10", "--rdm", "0", "--out", "d"], tmp)
        wf1 = rd(os.path.join(tmp, "d.wf.bin"))
        run_ref(["diag", "--fcidump", "fcidump.txt", "--adet", "a2.txt", "--bdet", "b2.txt", "--block", "2",
                 "--iteration", "1", "--tolerance", "1e10", "--rdm", "0", "--out", "l", "--load", "rs"], tmp)
        with open(os.path.join(tmp, "l.energy.txt")) as f:
            e_loaded = float(f.readline())
        wf_loaded = rd(os.path.join(tmp, "l.wf.bin"))
        res = dict(energy=e1, restart_bytes=rs, matrix_form_txt=np.array(mf_txt), wf=wf1, astr2=a2, bstr2=b2,
                   wf_loaded=wf_loaded, energy_loaded=e_loaded, bit_length=20)
        np.savez_compressed(os.path.join(HERE, "ref_restart.npz"), **res)
        print("restart", e1, len(rs), e_loaded, wf_loaded.size)
    finally:
        shutil.rmtree(tmp)


def case_dataset(tag, fcidump, adetfile, rounded_too=True):
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        shutil.copy(fcidump, os.path.join(tmp, "fcidump.txt"))
        shutil.copy(adetfile, os.path.join(tmp, "a.txt"))
        header, vals, idx = fio.load_fcidump(fcidump)
        L = int(header["NORB"])
        common = ["--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "a.txt"]
        run_ref(["diag"] + common + ["--block", "10", "--iteration", "100", "--tolerance", "1e-10",
                                     "--rdm", "1", "--out", "d"], tmp)
        with open(os.path.join(tmp, "d.energy.txt")) as f:
            e_full = float(f.readline())
        rdm1 = rd(os.path.join(tmp, "d.rdm1.bin")).reshape(2, L * L)
        rdm2 = rd(os.path.join(tmp, "d.rdm2.bin")).reshape(4, L ** 4)
        res = dict(energy=e_full, rdm1=rdm1, rdm2_sum_trace=np.einsum("bijij->", rdm2.reshape(4, L, L, L, L)),
                   rdm2_checksum=checksum(rdm2.sum(0)))
        if rounded_too:
            rv = fio.round_like_broadcast(vals)
            with open(os.path.join(tmp, "fcidump_r.txt"), "w") as f:
                with open(fcidump) as g:
                    for line in g:
                        f.write(line)
                        if "&END" in line:
                            break
                for v, (i, j, k, l) in zip(rv, idx):
                    f.write(f"{v:.17e} {i} {j} {k} {l}\n")
            run_ref(["diag", "--fcidump", "fcidump_r.txt", "--adet", "a.txt", "--bdet", "a.txt",
                     "--block", "10", "--iteration", "100", "--tolerance", "1e-10", "--rdm", "0",
                     "--out", "r"], tmp)
            with open(os.path.join(tmp, "r.energy.txt")) as f:
                res["energy_rounded"] = float(f.readline())
        np.savez_compressed(os.path.join(HERE, f"ref_{tag}.npz"), **res)
        print(tag, res)
    finally:
        shutil.rmtree(tmp)


def copy_datasets():
    for src, dst in ((f"{REFDATA}/n2/fcidump.txt", "n2_fcidump.txt.gz"),
                     (f"{REFDATA}/h2o/fcidump.txt", "h2o_fcidump.txt.gz")):
        with open(src, "rb") as f, gzip.open(os.path.join(HERE, dst), "wb", compresslevel=9) as g:
            g.write(f.read())
    shutil.copy(f"{REFDATA}/n2/1em3-alpha.txt", os.path.join(HERE, "n2_1em3-alpha.txt"))
    shutil.copy(f"{REFDATA}/h2o/h2o-1em3-alpha.txt", os.path.join(HERE, "h2o_1em3-alpha.txt"))


if __name__ == "__main__":
    want = sys.argv[1:] or ["small", "datasets", "c0", "restart"]
    if "small" in want:
        case_small()
    if "datasets" in want:
        copy_datasets()
        case_dataset("n2_1em3", f"{REFDATA}/n2/fcidump.txt", f"{REFDATA}/n2/1em3-alpha.txt")
        case_dataset("h2o_1em3", f"{REFDATA}/h2o/fcidump.txt", f"{REFDATA}/h2o/h2o-1em3-alpha.txt")
    if "c0" in want:
        case_c0()
    if "restart" in want:
        case_restart()
