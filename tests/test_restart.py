"""Wavefunction files of diag: the restart file of savename/loadname and the matrix-form
dump (restart.h:20-67, 414-657, 662-780), against files the reference itself wrote
(tests/golden/ref_restart.npz, made by tests/golden/make_golden.py).

CPU tests pin a restatement of the formats (below) on the reference's files; GPU tests
check the library's writer and loader against the same files.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_golden
from sbdg import fcidump as fio


# ---------------------------------------------------------------- format restatement

def words_of(s, L, bl):
    """from_string (bit_manipulation.h:158-172): orbital i in word i//bl, bit i%bl."""
    w = [0] * ((L + bl - 1) // bl)
    for i in range(L):
        if (int(s) >> i) & 1:
            w[i // bl] |= 1 << (i % bl)
    return w


def value_of(words, bl):
    v = 0
    for k, x in enumerate(words):
        for p in range(bl):
            if (int(x) >> p) & 1:
                v |= 1 << (k * bl + p)
    return v


def read_restart(raw):
    """(alpha word vectors, beta word vectors, W) of a SaveWavefunction file."""
    raw = bytes(raw)
    na, nb, dl = (int(x) for x in np.frombuffer(raw[:24], dtype=np.uint64))
    body = np.frombuffer(raw[24:24 + 8 * (na + nb) * dl], dtype=np.uint64).reshape(na + nb, dl)
    w = np.frombuffer(raw[24 + 8 * (na + nb) * dl:], dtype=np.float64)
    assert w.size == na * nb
    return body[:na], body[na:], w


def load_remap(raw, astr, bstr, L, bl):
    """LoadWavefunction at one rank: place every coefficient whose strings are both in the
    current sets (equal word vectors), then normalise."""
    la, lb, lw = read_restart(raw)
    pa = {tuple(words_of(s, L, bl)): i for i, s in enumerate(astr)}
    pb = {tuple(words_of(s, L, bl)): i for i, s in enumerate(bstr)}
    w = np.zeros(len(astr) * len(bstr))
    for i, wa in enumerate(la):
        ia = pa.get(tuple(int(x) for x in wa))
        if ia is None:
            continue
        for j, wb in enumerate(lb):
            ib = pb.get(tuple(int(x) for x in wb))
            if ib is not None:
                w[ia * len(bstr) + ib] += lw[i * len(lb) + j]
    return w / np.linalg.norm(w)


def makestring(s, L):
    return "".join("1" if (int(s) >> i) & 1 else "0" for i in range(L - 1, -1, -1))


def matrix_form_lines(w, astr, bstr, L):
    """SaveMatrixFormWF text form: default stream precision (six significant digits)."""
    return [f"{w[ia * len(bstr) + ib]:g} # {ia}: {makestring(a, L)} {ib}: {makestring(b, L)}"
            for ia, a in enumerate(astr) for ib, b in enumerate(bstr)]


def basis1():
    g = load_golden("ref_asym_L8.npz")
    return g, int(g["norb"])


# ---------------------------------------------------------------- CPU: the restatement

def test_restart_file_layout_matches_reference_file():
    r = load_golden("ref_restart.npz")
    g, L = basis1()
    bl = int(r["bit_length"])
    la, lb, w = read_restart(r["restart_bytes"])
    assert la.shape == (len(g["astr"]), (L + bl - 1) // bl)
    assert [value_of(x, bl) for x in la] == [int(s) for s in g["astr"]]
    assert [value_of(x, bl) for x in lb] == [int(s) for s in g["bstr"]]
    assert np.array_equal(w, r["wf"])  # the same W as the binary matrix form
    assert abs(np.linalg.norm(w) - 1.0) < 1e-12


def test_matrix_form_text_matches_reference_file():
    r = load_golden("ref_restart.npz")
    g, L = basis1()
    ours = matrix_form_lines(r["wf"], g["astr"], g["bstr"], L)
    assert ours == str(r["matrix_form_txt"]).splitlines()


def test_load_remap_matches_reference_loaded_vector():
    """The reference loaded its basis-1 file into basis 2 (five alpha and two beta strings
    dropped, others added) and stopped at the loaded vector."""
    r = load_golden("ref_restart.npz")
    _, L = basis1()
    w = load_remap(r["restart_bytes"], r["astr2"], r["bstr2"], L, int(r["bit_length"]))
    assert np.abs(w - r["wf_loaded"]).max() < 1e-15


def test_remap_restatement_catches_a_wrong_placement():
    """A transposed placement (alpha index for beta) does not reproduce the reference."""
    r = load_golden("ref_restart.npz")
    _, L = basis1()
    w = load_remap(r["restart_bytes"], r["astr2"], r["bstr2"], L, int(r["bit_length"]))
    na, nb = len(r["astr2"]), len(r["bstr2"])
    wrong = w.reshape(na, nb)[:, ::-1].ravel()
    assert np.abs(wrong - r["wf_loaded"]).max() > 1e-3


# ---------------------------------------------------------------- GPU: the library

def solver_basis1(**opt):
    import sbdg
    g, L = basis1()
    return sbdg.Solver(L, float(g["core"]), g["h1"], g["eri"], g["astr"], g["bstr"], **opt), g, L


@pytest.mark.gpu
def test_save_wavefunction_and_matrix_forms(tmp_path):
    r = load_golden("ref_restart.npz")
    s, g, L = solver_basis1(max_it=200, max_nb=10, eps=1e-10)
    e = s.run()
    assert abs(e - float(r["energy"])) < 1e-8
    w = s.vector()
    sign = 1.0 if w @ r["wf"] > 0 else -1.0
    prefix = str(tmp_path / "rs")
    s.save_wavefunction(prefix)
    raw = open(prefix + "000000", "rb").read()
    ref = bytes(r["restart_bytes"])
    nhead = 24 + 8 * (len(g["astr"]) + len(g["bstr"]))
    assert raw[:nhead] == ref[:nhead]  # sizes and strings byte-identical
    _, _, wf = read_restart(raw)
    assert np.array_equal(wf, w)
    assert np.abs(sign * wf - r["wf"]).max() < 1e-7
    s.save_matrix_form(str(tmp_path / "mf.bin"))
    assert np.array_equal(np.fromfile(tmp_path / "mf.bin"), w)
    s.save_matrix_form(str(tmp_path / "mf.txt"))
    ours = open(tmp_path / "mf.txt").read().splitlines()
    theirs = str(r["matrix_form_txt"]).splitlines()
    assert ours == matrix_form_lines(w, g["astr"], g["bstr"], L)
    assert [x.split(" # ")[1] for x in ours] == [x.split(" # ")[1] for x in theirs]
    a = np.array([float(x.split(" # ")[0]) for x in ours])
    b = np.array([float(x.split(" # ")[0]) for x in theirs])
    assert np.all(np.abs(sign * a - b) <= 1e-5 * np.abs(b) + 1e-6)
    s.save_matrix_form(str(tmp_path / "mf.other"))  # other extensions write nothing
    assert not os.path.exists(tmp_path / "mf.other")


@pytest.mark.gpu
def test_load_wavefunction_into_another_basis(tmp_path):
    import sbdg
    r = load_golden("ref_restart.npz")
    g, L = basis1()
    prefix = str(tmp_path / "ref")
    open(prefix + "000000", "wb").write(bytes(r["restart_bytes"]))
    s = sbdg.Solver(L, float(g["core"]), g["h1"], g["eri"], r["astr2"], r["bstr2"], max_it=1, max_nb=2, eps=1e10)
    s.load_wavefunction(prefix)
    e = s.run()
    assert np.abs(s.vector() - r["wf_loaded"]).max() < 1e-14
    assert abs(e - float(r["energy_loaded"])) < 1e-10
    # converging from the loaded vector reaches the basis-2 ground state
    s.set_options(max_it=200, max_nb=10, eps=1e-9)
    s.load_wavefunction(prefix)
    e2 = s.run()
    s0 = sbdg.Solver(L, float(g["core"]), g["h1"], g["eri"], r["astr2"], r["bstr2"], max_it=200, max_nb=10, eps=1e-9)
    assert abs(e2 - s0.run()) < 1e-9


@pytest.mark.gpu
def test_load_wavefunction_errors(tmp_path):
    import sbdg
    r = load_golden("ref_restart.npz")
    g, L = basis1()
    s = sbdg.Solver(L, float(g["core"]), g["h1"], g["eri"], g["astr"], g["bstr"])
    with pytest.raises(sbdg.SbdgError) as ei:
        s.load_wavefunction(str(tmp_path / "missing"))
    assert ei.value.code == sbdg.ERR_ARG
    # a file whose strings are all outside the basis
    other = np.array([1, 1, 1], dtype=np.uint64).tobytes()  # one 8-electron alpha string
    other += np.array([0b11111111, 0b11], dtype=np.uint64).tobytes() + np.array([1.0]).tobytes()
    open(str(tmp_path / "x000000"), "wb").write(other)
    with pytest.raises(sbdg.SbdgError) as ei:
        s.load_wavefunction(str(tmp_path / "x"))
    assert ei.value.code == sbdg.ERR_NUMERIC
    open(str(tmp_path / "t000000"), "wb").write(bytes(r["restart_bytes"])[:100])  # truncated
    with pytest.raises(sbdg.SbdgError) as ei:
        s.load_wavefunction(str(tmp_path / "t"))
    assert ei.value.code == sbdg.ERR_ARG


@pytest.mark.gpu
def test_cli_savename_and_loadname(tmp_path):
    """The sample-style driver writes the restart file and the matrix form, and warm-starts
    from a restart file (diag's loadname/savename through the C++ header)."""
    r = load_golden("ref_restart.npz")
    g, L = basis1()
    fio.write_fcidump(str(tmp_path / "fcidump.txt"), L, 5, g["h1"], g["eri"], core=float(g["core"]), ms2=1)
    fio.write_strings(str(tmp_path / "a.txt"), g["astr"], L)
    fio.write_strings(str(tmp_path / "b.txt"), g["bstr"], L)
    exe = os.path.join(ROOT, "sbd-ykt_amd", "sbdg", "sbd_diag")
    base = [exe, "--fcidump", "fcidump.txt", "--adetfile", "a.txt", "--bdetfile", "b.txt", "--block", "10",
            "--iteration", "200", "--tolerance", "1e-10"]
    subprocess.run(base + ["--savename", "rs", "--dump_matrix_form_wf", "mf.txt"], cwd=tmp_path, check=True,
                   capture_output=True, timeout=120)
    raw = open(tmp_path / "rs000000", "rb").read()
    nhead = 24 + 8 * (len(g["astr"]) + len(g["bstr"]))
    assert raw[:nhead] == bytes(r["restart_bytes"])[:nhead]
    assert len(open(tmp_path / "mf.txt").read().splitlines()) == len(g["astr"]) * len(g["bstr"])
    out = subprocess.run(base + ["--loadname", "rs"], cwd=tmp_path, check=True, capture_output=True, text=True,
                         timeout=120).stdout
    line = [x for x in out.splitlines() if "Energy =" in x][0]
    assert abs(float(line.split("=")[1]) - float(r["energy"])) < 1e-8


@pytest.mark.gpu
def test_cli_shuffle_and_rdm_dumps(tmp_path):
    """The sample driver with --shuffle 1 --rdm 1 on the N2 threshold-1e-3 strings (the
    strings after the first in ShuffleDet's order, main.cc:123-128): the energy is the
    reference's full-precision one, and 1pRDM.txt / 2pRDM.txt (main.cc:172-205) hold the
    spin-summed RDMs that give that energy through the energy identity."""
    import gzip
    import shutil

    g = load_golden("ref_n2_1em3.npz")
    with gzip.open(os.path.join(ROOT, "tests", "golden", "n2_fcidump.txt.gz"), "rb") as f, \
            open(tmp_path / "fcidump.txt", "wb") as o:
        o.write(f.read())
    shutil.copy(os.path.join(ROOT, "tests", "golden", "n2_1em3-alpha.txt"), tmp_path / "a.txt")
    exe = os.path.join(ROOT, "sbd-ykt_amd", "sbdg", "sbd_diag")
    out = subprocess.run([exe, "--fcidump", "fcidump.txt", "--adetfile", "a.txt", "--block", "10", "--iteration",
                          "100", "--tolerance", "1e-10", "--rdm", "1", "--shuffle", "1"], cwd=tmp_path, check=True,
                         capture_output=True, text=True, timeout=300).stdout
    e = float([x for x in out.splitlines() if "Energy =" in x][0].split("=")[1])
    assert abs(e - float(g["energy"])) < 1e-8
    L = 18
    one = np.loadtxt(tmp_path / "1pRDM.txt")
    two = np.loadtxt(tmp_path / "2pRDM.txt")
    assert one.shape == (L * L, 3) and two.shape == (L ** 4, 5)
    assert np.array_equal(one[:, 0], np.repeat(np.arange(L), L)) and np.array_equal(one[:, 1], np.tile(np.arange(L), L))
    g1 = np.zeros((L, L))
    g1[one[:, 0].astype(int), one[:, 1].astype(int)] = one[:, 2]       # g1[i, j] = gamma[i + L j]
    assert np.abs(g1 - (g["rdm1"][0] + g["rdm1"][1]).reshape(L, L, order="F")).max() < 1e-7
    i, j, k, l = (two[:, c].astype(int) for c in range(4))
    G = np.zeros((L, L, L, L))
    G[i, j, k, l] = two[:, 4]                                          # G[i,j,k,l] = Gamma[i+Lj+L^2k+L^3l]
    assert abs(np.einsum("ijij->", G) - float(g["rdm2_sum_trace"])) < 1e-6
    idline = [x for x in out.splitlines() if "One-Body + Two-Body" in x][0]
    e12 = float(idline.split("=")[1].split("(")[0])
    core = float(idline.split("core")[1].strip(" )"))
    assert abs(core + e12 - e) < 1e-8
