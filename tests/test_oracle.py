"""Pins the CPU restatement (the oracle) before anything is checked against it.

* against the reference's own outputs (golden fixtures written by running the reference:
  tables, diagonal, one sigma, Davidson energy, RDMs, density);
* against brute-force second quantisation on tiny bases (independent of the
  Slater-Condon formulas), against dense eigensolvers, against closed forms;
* against invariants: the energy/RDM identity, partial traces, Hermiticity;
* and with mistakes planted in a copy of the oracle's formulas, which these checks must
  catch.
"""
from math import comb

import numpy as np
import pytest

import bruteforce as bf
import oracle
from helpers import (SMALL_CASES, core_split, dataset, energy_from_rdm, rdm_close, small_case, tables_from_ref,
                     workload_small)
from sbdg import fcidump as fio
from sbdg import synthetic as syn


def problem(case):
    return oracle.Problem(case["norb"], case["core"], case["h1"], case["astr"], case["bstr"], eri=case["eri"])


# ------------------------------------------------------------------ reference goldens

@pytest.mark.parametrize("name", SMALL_CASES)
def test_tables_bit_exact_vs_reference(name):
    g, case = small_case(name)
    p = problem(case)
    ref = tables_from_ref(g["tables"], p.na, p.nb)
    for which in range(4):
        rp, cl = p.table(which)
        assert np.array_equal(rp, ref[which][0])
        assert np.array_equal(cl, ref[which][1])


@pytest.mark.parametrize("name", SMALL_CASES)
def test_hii_and_sigma_vs_reference(name):
    g, case = small_case(name)
    p = problem(case)
    hii = p.hii()
    assert np.abs(hii - g["hii"]).max() < 1e-13
    vec = np.random.default_rng(int(g["vec_seed"])).normal(size=p.ndet)
    sig = p.sigma(vec, hii=g["hii"])
    # same per-element summation order as mult.h: bit-identical up to reassociation-free FP
    assert np.abs(sig - g["sigma"]).max() <= 1e-12 * np.abs(g["sigma"]).max()


@pytest.mark.parametrize("name", SMALL_CASES)
def test_sigma_rows_vs_sigma_and_reference(name):
    """The row-restricted sigma (the checker of the full-size sigma tests and the CPU
    baseline's workload) against the whole-vector sigma and the reference's own sigma, on
    every row block of a split, with the diagonal passed in and evaluated in place."""
    g, case = small_case(name)
    p = problem(case)
    vec = np.random.default_rng(int(g["vec_seed"])).normal(size=p.ndet)
    full = p.sigma(vec, hii=g["hii"])
    nb = p.nb
    cuts = sorted({0, 1, p.na // 3, p.na - 1, p.na})
    for ia0, ia1 in zip(cuts[:-1], cuts[1:]):
        blk = slice(ia0 * nb, ia1 * nb)
        for hii in (g["hii"], None):
            rows = p.sigma_rows(vec, hii, ia0, ia1)
            assert np.abs(rows[blk] - full[blk]).max() <= 1e-13 * np.abs(full).max()
            assert np.abs(rows[blk] - g["sigma"][blk]).max() <= 1e-12 * np.abs(g["sigma"]).max()
            assert not rows[:ia0 * nb].any() and not rows[ia1 * nb:].any()  # nothing outside the rows


@pytest.mark.parametrize("name", SMALL_CASES)
def test_davidson_energy_rdm_vs_reference(name):
    g, case = small_case(name)
    p = problem(case)
    res = p.davidson(max_it=200, nb=10, eps=float(g["tol"]))
    W = res["W"]
    e = p.energy(W)
    assert abs(e - float(g["energy"])) < 1e-9
    one, two = p.rdm(W)
    # reference W has an arbitrary global sign; RDMs do not
    rdm_close(one, two, g["rdm1"], g["rdm2"], 1e-7)
    # with do_rdm the reference's density is the 1-RDM diagonal (sbdiag.h:412-416)
    assert np.abs(p.density(W) - g["density"]).max() < 1e-7


def test_dataset_n2_full_precision_and_rounded():
    header, vals, idx, s = dataset("n2")
    from conftest import load_golden
    g = load_golden("ref_n2_1em3.npz")
    L = int(header["NORB"])
    I0, h1, packed = oracle.setup_from_fcidump(L, vals, idx)
    p = oracle.Problem(L, I0, h1, s, s, packed=packed)
    res = p.davidson(max_it=100, nb=10, eps=1e-8)
    e = p.energy(res["W"])
    assert abs(e - float(g["energy"])) < 1e-8
    assert abs(e - (-109.0416210980518)) < 1e-8  # full-precision value measured on the reference
    # the published table (data/n2/README.md:8) comes from six-digit integrals
    I0r, h1r, packedr = oracle.setup_from_fcidump(L, fio.round_like_broadcast(vals), idx)
    pr = oracle.Problem(L, I0r, h1r, s, s, packed=packedr)
    er = pr.energy(pr.davidson(max_it=100, nb=10, eps=1e-8)["W"])
    assert abs(er - float(g["energy_rounded"])) < 1e-8
    assert round(er, 6) == -109.041511


# ------------------------------------------------------------------ brute force / maths

def tiny_cases():
    yield workload_small(6, 3, 2, 12, 9, 21)   # alpha set != beta set
    yield workload_small(6, 2, 2, None, None, 22)  # full space, alpha == beta
    yield workload_small(7, 3, 3, 20, 20, 23)


@pytest.mark.parametrize("case", list(tiny_cases()))
def test_matrix_elements_vs_bruteforce(case):
    p = problem(case)
    H = bf.hamiltonian(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])
    AB = [(x, y) for x in p.a for y in p.b]
    A = np.array([x for x, _ in AB], dtype=np.uint64)
    B = np.array([y for _, y in AB], dtype=np.uint64)
    n = len(AB)
    Ho = p.hij(np.repeat(A, n), np.repeat(B, n), np.tile(A, n), np.tile(B, n)).reshape(n, n)
    # Hij returns 0 beyond doubles; brute force has them 0 too
    assert np.abs(Ho - H).max() < 1e-13
    c = np.random.default_rng(1).normal(size=n)
    assert np.abs(p.sigma(c) - H @ c).max() < 1e-12
    w, v = np.linalg.eigh(H)
    res = p.davidson(max_it=50, nb=8, eps=1e-10)
    assert abs(res["energy"] - w[0]) < 1e-10
    one, two = p.rdm(v[:, 0])
    one_b, two_b = bf.rdms(case["norb"], p.a, p.b, v[:, 0])
    # the reference labels the two mixed-spin blocks of one double-excitation class
    # crosswise (correlation.h:125-128); only their sum is physical, for equal and for
    # different alpha/beta sets alike
    rdm_close(one, two, one_b, two_b, 1e-12, mixed_sum=True)
    assert abs(energy_from_rdm(case["core"], case["h1"], case["eri"], one, two) - w[0]) < 1e-10


@pytest.mark.parametrize("case", [workload_small(6, 3, 2, 12, 9, 21), workload_small(6, 2, 2, None, None, 22)])
def test_mixed_block_quirk_is_real(case):
    p = problem(case)
    H = bf.hamiltonian(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])
    v = np.linalg.eigh(H)[1][:, 0]
    _, two = p.rdm(v)
    _, two_b = bf.rdms(case["norb"], p.a, p.b, v)
    assert np.abs(two[1] - two_b[1]).max() > 1e-4  # individually off...
    assert np.abs(two[1] + two[2] - two_b[1] - two_b[2]).max() < 1e-12  # ...sum exact


@pytest.mark.parametrize("L,n", [(8, 3), (12, 4), (10, 5)])
def test_table_counts_closed_form(L, n):
    s = syn.full_space(L, n)
    assert len(s) == comb(L, n) and s[0] == (1 << n) - 1 and np.all(np.diff(s.astype(np.int64)) > 0)
    rng = np.random.default_rng(0)
    h1, eri = syn.synthetic_integrals(L, rng)
    p = oracle.Problem(L, 0.0, h1, s, s[:3], eri=eri)
    rpS, cS = p.table(0)
    rpD, cD = p.table(1)
    assert np.all(np.diff(rpS) == n * (L - n))
    assert np.all(np.diff(rpD) == comb(n, 2) * comb(L - n, 2))
    # brute-force membership and ascending order
    for i in (0, len(s) // 2, len(s) - 1):
        d = np.array([bin(int(s[i]) ^ int(x)).count("1") for x in s])
        assert np.array_equal(cS[rpS[i]:rpS[i + 1]], np.nonzero(d == 2)[0])
        assert np.array_equal(cD[rpD[i]:rpD[i + 1]], np.nonzero(d == 4)[0])


def test_core_split_holds_every_double_once():
    """Each same-spin double pair lies in exactly one (n-2)-electron core: with no core taken
    through the workspace, the brute-force core split's direct pairs are the oracle's
    doubles table, entry for entry in number (the split the GPU statistics are checked
    against)."""
    case = workload_small(26, 5, 5, 400, 300, 31)
    p = problem(case)
    for strs, which in ((case["astr"], 1), (case["bstr"], 3)):
        _, col = p.table(which)
        assert core_split(strs, kmin=1 << 30) == (0, len(col))
        nmem, direct = core_split(strs)
        assert 0 < direct < len(col) and nmem > 0


def test_hii_closed_form():
    case = workload_small(7, 3, 2, 15, 10, 5)
    p = problem(case)
    hii = p.hii()
    h1, eri, L = case["h1"], case["eri"], case["norb"]
    for k in (0, 7, 33, p.ndet - 1):
        A, B = int(p.a[k // p.nb]), int(p.b[k % p.nb])
        oa = [i for i in range(L) if (A >> i) & 1]
        ob = [i for i in range(L) if (B >> i) & 1]
        e = case["core"] + sum(h1[i, i] for i in oa + ob)
        for occ in (oa, ob):
            for x in range(len(occ)):
                for y in range(x + 1, len(occ)):
                    i, j = occ[x], occ[y]
                    e += eri[i, i, j, j] - eri[i, j, j, i]
        e += sum(eri[i, i, j, j] for i in oa for j in ob)
        assert abs(e - hii[k]) < 1e-12


def test_rdm_invariants():
    g, case = small_case("asym_L14")
    L = case["norb"]
    one, two = g["rdm1"], g["rdm2"]
    na_e = bin(int(case["astr"][0])).count("1")
    nb_e = bin(int(case["bstr"][0])).count("1")
    assert abs(one[0].reshape(L, L).trace() - na_e) < 1e-10
    assert abs(one[1].reshape(L, L).trace() - nb_e) < 1e-10
    for s in range(2):
        G = one[s].reshape(L, L)
        assert np.abs(G - G.T).max() < 1e-12
    # partial trace: sum_j Gamma^{st}[i j k j] = (N_t - delta_st) gamma^s[i k]
    for s in range(2):
        for t in range(2):
            G = two[s + 2 * t].reshape(L, L, L, L, order="F")
            Nt = (na_e, nb_e)[t] - (1 if s == t else 0)
            pt = np.einsum("ijkj->ik", G)
            if s == t:
                assert np.abs(pt - Nt * one[s].reshape(L, L, order="F")).max() < 1e-10
    pt12 = np.einsum("ijkj->ik", two[1].reshape(L, L, L, L, order="F")) + \
        np.einsum("ijkj->ik", two[2].reshape(L, L, L, L, order="F"))
    expect = nb_e * one[0].reshape(L, L, order="F") + na_e * one[1].reshape(L, L, order="F")
    assert np.abs(pt12 - expect).max() < 1e-10
    assert abs(energy_from_rdm(case["core"], case["h1"], case["eri"], one, two) - float(g["energy"])) < 1e-9


def test_syev_vs_numpy():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 24):
        A = rng.normal(size=(n, n))
        A = A + A.T
        E, U = oracle.syev(A)
        assert np.allclose(E, np.linalg.eigvalsh(A), atol=1e-12)
        assert np.allclose(U.T @ U, np.eye(n), atol=1e-12)
        assert np.allclose(A @ U, U * E, atol=1e-11)


def test_c0_reference_sigma_and_energy():
    from conftest import load_golden
    g = load_golden("ref_c0.npz")
    wl = syn.config_workload("c0_L12_4e4e_full")
    from golden.make_golden import checksum
    assert abs(checksum(wl.h1, wl.eri, wl.astr.astype(np.float64), wl.bstr.astype(np.float64))
               - float(g["input_checksum"])) < 1e-9
    p = oracle.Problem(wl.norb, wl.core, wl.h1, wl.astr, wl.bstr, eri=wl.eri)
    hii = p.hii()
    assert np.abs(hii[:4096] - g["hii_head"]).max() < 1e-13
    vec = np.random.default_rng(0).normal(size=p.ndet)
    sig = p.sigma(vec, hii=hii)
    assert np.abs(sig[:4096] - g["sigma_head"]).max() < 1e-11
    assert abs(float(sig @ vec) - float(g["sigma_dot_vec"])) < 1e-8 * abs(float(g["sigma_dot_vec"]))


# ------------------------------------------------------------------ planted mistakes

@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(plant, tmp_path):
    """Variants of the oracle compiled with a deliberate mistake (1: exchange term of the
    single excitation dropped, 2: parity without the start-bit flip, 3: double-excitation
    crossing sign dropped, 4: alpha/beta spin-orbital bits swapped) must fail the
    brute-force pin that the real oracle passes."""
    import ctypes as C
    import subprocess
    so = tmp_path / f"liboracle_plant{plant}.so"
    subprocess.check_call(["gcc", "-O1", "-fPIC", "-shared", "-fopenmp", f"-DPLANT={plant}",
                           "-o", str(so), oracle.HERE + "/sbd_oracle.c", "-lm"])
    saved = oracle._lib
    try:
        oracle._lib = None
        oracle.LIB, old = str(so), oracle.LIB
        case = workload_small(6, 3, 2, 12, 9, 21)
        H = bf.hamiltonian(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])
        p = problem(case)
        c = np.random.default_rng(1).normal(size=p.ndet)
        assert np.abs(p.sigma(c) - H @ c).max() > 1e-6
        del p
    finally:
        oracle._lib = saved
        oracle.LIB = old
