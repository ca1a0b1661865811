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
* ``ref_c3_rdm.npz`` -- the norb=16, 5+5 full-space workload (seed 3, 19.1M determinants):
  one sigma of a seeded vector (numpy default_rng(33).normal), with checksums, and the
  1-/2-RDMs (Correlation, tpb/correlation.h:194-444) of that vector scaled to unit norm.
* ``ref_carryover.npz`` -- carry-over strings of diag (CarryOverAlphaDet, tpb/rdmat.h:60-132)
  by threshold and by ratio on two small cases.
* ``ref_shuffle.npz`` -- the N2 threshold-1e-3 alpha strings in the orders ShuffleDet
  (determinants.h:589-594) gives them with the sample's seeds 1729 and 137 (main.cc:123-128).
* ``ref_cpu_timings.json`` -- the reference's CPU path timed here (cores stated): the full
  Davidson of the norb=12 workload, one sigma of the norb=26 4,000 x 4,000 workload, and the
  sigma and Correlation of the norb=16 workload (from the ref_c3_rdm run).

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import gzip
import json
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


def ref_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "8"))


def write_inputs(tmp, wl):
    L = wl.norb
    fio.write_fcidump(os.path.join(tmp, "fcidump.txt"), L, wl.nalpha + wl.nbeta, wl.h1, wl.eri,
                      core=wl.core, ms2=wl.nalpha - wl.nbeta)
    fio.write_strings(os.path.join(tmp, "a.txt"), wl.astr, L)
    fio.write_strings(os.path.join(tmp, "b.txt"), wl.bstr, L)
    return ["--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "b.txt"]


def update_timings(**entries):
    path = os.path.join(HERE, "ref_cpu_timings.json")
    t = {}
    if os.path.exists(path):
        with open(path) as f:
            t = json.load(f)
    t.update(entries)
    t["host"] = ("this container: %d-core CPU (%s), 1 MPI rank x OMP_NUM_THREADS threads, "
                 "MKL sequential" % (os.cpu_count(), _cpu_name()))
    with open(path, "w") as f:
        json.dump(t, f, indent=1, sort_keys=True)
        f.write("\n")


def _cpu_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def case_c3_rdm():
    """One sigma and Correlation of the reference on a seeded unit vector over the norb=16,
    5+5 full space (19.1M determinants); about two hours on 8 cores."""
    wl = syn.config_workload("c3_L16_5e5e_full")
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        common = write_inputs(tmp, wl)
        vec = np.random.default_rng(33).normal(size=wl.ndet)
        vec.tofile(os.path.join(tmp, "vec.bin"))
        run_ref(["sigma"] + common + ["--vec", "vec.bin", "--out", "s", "--corr", "1"], tmp)
        sig = rd(os.path.join(tmp, "s.sigma.bin"))
        L = wl.norb
        with open(os.path.join(tmp, "s.sigma_time.txt")) as f:
            t_sigma = float(f.readline())
        with open(os.path.join(tmp, "s.corr_time.txt")) as f:
            t_corr = float(f.readline())
        res = dict(vec_seed=33, rdm1=rd(os.path.join(tmp, "s.rdm1.bin")).reshape(2, L * L),
                   rdm2=rd(os.path.join(tmp, "s.rdm2.bin")).reshape(4, L ** 4),
                   input_checksum=checksum(wl.h1, wl.eri, wl.astr.astype(np.float64), wl.bstr.astype(np.float64)),
                   sigma_checksum=checksum(sig), sigma_norm=float(np.linalg.norm(sig)),
                   sigma_dot_vec=float(sig @ vec), sigma_head=sig[:4096].copy())
        np.savez_compressed(os.path.join(HERE, "ref_c3_rdm.npz"), **res)
        update_timings(c3_L16_5e5e_full=dict(what="one sigma (mult) and one Correlation of a unit vector",
                                             determinants=int(wl.ndet), sigma_s=t_sigma, correlation_s=t_corr,
                                             threads=ref_threads()))
        print("c3", wl.ndet, t_sigma, t_corr)
    finally:
        shutil.rmtree(tmp)


def case_timing_c0():
    """The reference's full Davidson on the norb=12, 4+4 full space at tol 1e-8 (timed)."""
    wl = syn.config_workload("c0_L12_4e4e_full")
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        common = write_inputs(tmp, wl)
        run_ref(["diag"] + common + ["--block", "10", "--iteration", "200", "--tolerance", "1e-8",
                                     "--rdm", "0", "--out", "d"], tmp)
        with open(os.path.join(tmp, "d.energy.txt")) as f:
            e, t = float(f.readline()), float(f.readline())
        update_timings(c0_L12_4e4e_full=dict(what="diag: tables, Davidson to tol 1e-8 (block 10), final energy, "
                                                  "density", determinants=int(wl.ndet), diag_s=t, energy=e,
                                             threads=ref_threads()))
        print("c0", t, e)
    finally:
        shutil.rmtree(tmp)


def case_timing_c1():
    """One sigma of the reference on the norb=26, 5+5, 4,000 x 4,000 workload (timed)."""
    wl = syn.config_workload("c1_L26_5e5e_4000")
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        common = write_inputs(tmp, wl)
        np.random.default_rng(11).normal(size=wl.ndet).tofile(os.path.join(tmp, "vec.bin"))
        run_ref(["sigma"] + common + ["--vec", "vec.bin", "--out", "s"], tmp)
        with open(os.path.join(tmp, "s.sigma_time.txt")) as f:
            t = float(f.readline())
        update_timings(c1_L26_5e5e_4000=dict(what="one sigma (mult)", determinants=int(wl.ndet), sigma_s=t,
                                             threads=ref_threads()))
        print("c1", t)
    finally:
        shutil.rmtree(tmp)


def case_carryover():
    """Carry-over strings of diag by threshold (default 0.01 and 0.2) and by ratio (0.3) on
    the asym_L14 and sym_L10 cases."""
    cases = {"asym_L14": syn.Workload(*_sampled(14, 4, 3, 150, 120, 12)),
             "sym_L10": syn.make_workload(10, 3, 3, 13)}
    res = {}
    for name, wl in cases.items():
        tmp = tempfile.mkdtemp(prefix="golden_")
        try:
            common = write_inputs(tmp, wl)
            for tag, flag in (("thr001", ["--carryover_threshold", "0.01"]),
                              ("thr02", ["--carryover_threshold", "0.2"]),
                              ("ratio03", ["--carryover_ratio", "0.3"])):
                run_ref(["diag"] + common + ["--block", "10", "--iteration", "200", "--tolerance", "1e-10",
                                             "--rdm", "0", "--out", "d"] + flag, tmp)
                co = rd(os.path.join(tmp, "d.carryover.bin"), np.int64)
                n, words = int(co[0]), int(co[1])
                w = co[2:].reshape(n, words) if n else np.zeros((0, max(words, 1)), np.int64)
                strs = np.array([sum(int(w[i, k]) << (20 * k) for k in range(words)) for i in range(n)],
                                dtype=np.uint64)
                res[f"{name}_{tag}"] = strs
        finally:
            shutil.rmtree(tmp)
        print("carryover", name, {k: len(v) for k, v in res.items() if k.startswith(name)})
    np.savez_compressed(os.path.join(HERE, "ref_carryover.npz"), **res)


def case_shuffle():
    tmp = tempfile.mkdtemp(prefix="golden_")
    try:
        shutil.copy(os.path.join(HERE, "n2_1em3-alpha.txt"), os.path.join(tmp, "a.txt"))
        with gzip.open(os.path.join(HERE, "n2_fcidump.txt.gz"), "rb") as f, \
                open(os.path.join(tmp, "fcidump.txt"), "wb") as g:
            g.write(f.read())
        run_ref(["shuffle", "--fcidump", "fcidump.txt", "--adet", "a.txt", "--bdet", "a.txt", "--out", "s"], tmp)
        w = rd(os.path.join(tmp, "s.shuffle.bin"), np.int64)
        n, words = int(w[0]), int(w[1])
        orders = w[2:].reshape(2, n, words)
        strs = [np.array([sum(int(o[i, k]) << (20 * k) for k in range(words)) for i in range(n)], dtype=np.uint64)
                for o in orders]
        np.savez_compressed(os.path.join(HERE, "ref_shuffle.npz"), seed1729=strs[0], seed137=strs[1])
        print("shuffle", n)
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
    if "carryover" in want:
        case_carryover()
    if "shuffle" in want:
        case_shuffle()
    if "timing_c0" in want:
        case_timing_c0()
    if "timing_c1" in want:
        case_timing_c1()
    if "c3_rdm" in want:
        case_c3_rdm()
