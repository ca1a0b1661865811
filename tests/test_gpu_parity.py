"""Parity of the HIP path (through the C-ABI) with the oracle and the reference goldens.

Bars: integer work (tables, orbitals, signs) bit-exact; FP64 sigma within 1e-12 of the
oracle relative to its largest element; FP64 energy within 1e-8 Ha and RDMs within
1e-7 of the reference; FP32 sigma within 2e-5 relative; mixed-precision energy within
1e-6 Ha of FP64.
"""
import numpy as np
import pytest

import bruteforce as bf
import oracle
import sbdg
from conftest import load_golden
from helpers import (SMALL_CASES, core_split, dataset, energy_from_rdm, rdm_close, rdm_invariants, small_case,
                     tables_from_ref, workload_small)
from sbdg import fcidump as fio
from sbdg import synthetic as syn

pytestmark = pytest.mark.gpu


def solver(case, **opt):
    return sbdg.Solver(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"], **opt)


def oproblem(case):
    return oracle.Problem(case["norb"], case["core"], case["h1"], case["astr"], case["bstr"], eri=case["eri"])


def sep_sign(a, b):
    """s(A,B) = (-1)^(sum over p in A of #{q in B: q < p}): interleaved vs separated order."""
    t = 0
    A, B = int(a), int(b)
    p = 0
    while A >> p:
        if (A >> p) & 1:
            t += bin(B & ((1 << p) - 1)).count("1")
        p += 1
    return -1.0 if t & 1 else 1.0


def medium_case():
    return workload_small(26, 5, 5, 400, 300, 31)


# ---------------------------------------------------------------- integer work

@pytest.mark.parametrize("name", SMALL_CASES)
def test_tables_bit_exact_vs_reference(name):
    g, case = small_case(name)
    s = solver(case)
    ref = tables_from_ref(g["tables"], s.na, s.nb)
    for which in range(4):
        rp, col, exc, val = s.table(which)
        assert np.array_equal(rp, ref[which][0]), which
        assert np.array_equal(col.astype(np.int64), ref[which][1]), which


def test_tables_bit_exact_vs_oracle_medium():
    case = medium_case()
    s, p = solver(case), oproblem(case)
    for which in range(4):
        rp, col, _, _ = s.table(which)
        rpo, colo = p.table(which)
        assert np.array_equal(rp, rpo) and np.array_equal(col.astype(np.int64), colo)


def test_pair_orbitals_and_signs():
    """Orbitals in every table entry, and the sign of every reconstructed matrix element
    (separated-order value times s(bra) s(ket)) equal the reference convention's."""
    case = workload_small(12, 4, 3, 200, 150, 41)
    s, p = solver(case), oproblem(case)
    rng = np.random.default_rng(0)
    A, B, L = p.a, p.b, case["norb"]
    for which in range(4):
        rp, col, exc, val = s.table(which)
        strs = A if which < 2 else B
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        for q in rng.choice(len(col), size=min(400, len(col)), replace=False):
            bra, ket = int(strs[rows[q]]), int(strs[col[q]])
            e = int(exc[q])
            if which % 2 == 0:
                i, a = e & 63, (e >> 6) & 63
                assert (ket & ~bra) == 1 << i and (bra & ~ket) == 1 << a
            else:
                I, J, Aa, Bb = e & 63, (e >> 6) & 63, (e >> 12) & 63, (e >> 18) & 63
                assert (ket & ~bra) == (1 << I) | (1 << J) and I < J
                assert (bra & ~ket) == (1 << Aa) | (1 << Bb) and Aa < Bb
            # the element in the reference convention for a sample of spectator strings
            for other in rng.choice(len(B if which < 2 else A), size=3, replace=False):
                if which < 2:
                    ab, bb, ak, bk = bra, int(B[other]), ket, int(B[other])
                else:
                    ab, bb, ak, bk = int(A[other]), bra, int(A[other]), ket
                href = p.hij(np.array([ab], np.uint64), np.array([bb], np.uint64),
                             np.array([ak], np.uint64), np.array([bk], np.uint64))[0]
                hsep = val[q]
                if which % 2 == 0:  # singles: add the other-spin occupation term
                    i, a = e & 63, (e >> 6) & 63
                    occ = bb if which == 0 else ab
                    neg = (e >> 24) & 1
                    extra = sum(case["eri"][a, i, k, k] for k in range(L) if (occ >> k) & 1)
                    hsep = hsep + (-extra if neg else extra)
                hgpu = sep_sign(ab, bb) * sep_sign(ak, bk) * hsep
                assert abs(hgpu - href) <= 1e-12 * max(1.0, abs(href))
                if abs(href) > 1e-10:
                    assert np.sign(hgpu) == np.sign(href)


# ---------------------------------------------------------------- sigma

@pytest.mark.parametrize("doubles", [1, 2])
@pytest.mark.parametrize("name", SMALL_CASES)
def test_hii_and_sigma_vs_reference(name, doubles):
    """doubles=1: same-spin doubles from the excitation tables; 2: grouped by core."""
    g, case = small_case(name)
    s = solver(case, doubles=doubles)
    assert np.abs(s.hii() - g["hii"]).max() < 1e-13
    vec = np.random.default_rng(int(g["vec_seed"])).normal(size=s.ndet)
    sig = s.sigma(vec)
    assert np.abs(sig - g["sigma"]).max() <= 1e-12 * np.abs(g["sigma"]).max()
    sig32 = s.sigma(vec, precision=1)
    assert np.abs(sig32 - g["sigma"]).max() <= 2e-5 * np.abs(g["sigma"]).max()


@pytest.mark.parametrize("doubles", [1, 2])
def test_sigma_columns_are_matrix_columns(doubles):
    """sigma(e_J) is column J of H: every element (sign included) of those columns.
    21 beta strings: row lengths that are not a multiple of the vector width."""
    case = workload_small(8, 3, 2, 30, 21, 51)
    s = solver(case, doubles=doubles)
    H = bf.hamiltonian(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])
    for J in (0, 1, 37, 199, s.ndet - 1):
        e = np.zeros(s.ndet)
        e[J] = 1.0
        col = s.sigma(e)
        assert np.abs(col - H[:, J]).max() < 1e-13
        assert np.array_equal(np.sign(np.round(col, 12)), np.sign(np.round(H[:, J], 12)))
        col32 = s.sigma(e, precision=1)
        assert np.abs(col32 - H[:, J]).max() < 2e-5 * np.abs(H[:, J]).max()


@pytest.mark.parametrize("L,na_e,nb_e,na,nb", [
    (6, 3, 2, 1, 15),  # one alpha string: no alpha excitation at all
    (6, 2, 3, 15, 1),  # one beta string: a single column, no beta excitation
    (5, 0, 2, 0, 0),   # no alpha electron (the empty string), beta full space
    (5, 2, 0, 0, 0),   # no beta electron
    (6, 1, 1, 0, 0),   # one electron per spin: no same-spin double anywhere
    (7, 7, 1, 0, 0),   # a full alpha shell: one alpha string holding every orbital
])
def test_edge_string_sets_vs_bruteforce(L, na_e, nb_e, na, nb):
    """Degenerate string sets: sigma (FP64 and FP32) and the ground-state energy against the
    dense brute-force Hamiltonian."""
    case = workload_small(L, na_e, nb_e, na, nb, 5)
    H = bf.hamiltonian(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"])
    s = solver(case, max_it=100, max_nb=8, eps=1e-10)
    x = np.random.default_rng(9).normal(size=s.ndet)
    hx = H @ x
    assert np.abs(s.sigma(x) - hx).max() <= 1e-12 * max(1.0, np.abs(hx).max())
    assert np.abs(s.sigma(x, precision=1) - hx).max() <= 2e-5 * max(1.0, np.abs(hx).max())
    e = s.run()
    assert abs(e - np.linalg.eigvalsh(H)[0]) < 1e-8


@pytest.mark.parametrize("doubles", [1, 2])
def test_sigma_medium_vs_oracle_and_symmetry(doubles):
    case = medium_case()
    s, p = solver(case, doubles=doubles), oproblem(case)
    if doubles == 2:  # both forms carry pairs here: cores of 10+ members and smaller ones
        s.build()
        st = s.stats()
        ea, eb = core_split(case["astr"]), core_split(case["bstr"])
        assert st["core_members"] == [ea[0], eb[0]] and st["core_direct"] == [ea[1], eb[1]]
        assert min(ea) > 0 and min(eb) > 0
        nd = st["nnz"]
        assert ea[1] < nd[1] and eb[1] < nd[3]  # the rest of the doubles go through the workspace
    rng = np.random.default_rng(2)
    x, y = rng.normal(size=s.ndet), rng.normal(size=s.ndet)
    sx, ox = s.sigma(x), p.sigma(x)
    assert np.abs(sx - ox).max() <= 1e-12 * np.abs(ox).max()
    sy = s.sigma(y)
    assert abs(x @ sy - y @ sx) < 1e-10 * abs(x @ sy)            # H symmetric
    sxy = s.sigma(2.0 * x - 3.0 * y)
    assert np.abs(sxy - (2.0 * sx - 3.0 * sy)).max() < 1e-11 * np.abs(sxy).max()  # linear


@pytest.mark.parametrize("doubles,core_gib", [(1, 8.0), (2, 8.0), (2, 0.2)])
def test_sigma_c1_shape_row_sample(doubles, core_gib):
    """Full c1 shape (norb=26, 5+5, 4000 x 4000 strings): GPU tables equal the oracle's,
    and sigma on a sample of rows equals the oracle restricted to those rows.  A 0.2 GiB
    core workspace splits the core products into several 512-column passes."""
    wl = syn.config_workload("c1_L26_5e5e_4000")
    case = dict(norb=wl.norb, core=wl.core, h1=wl.h1, eri=wl.eri, astr=wl.astr, bstr=wl.bstr)
    s, p = solver(case, doubles=doubles, core_gib=core_gib), oproblem(case)
    for which in range(4):
        rp, col, _, _ = s.table(which)
        rpo, colo = p.table(which)
        assert np.array_equal(rp, rpo) and np.array_equal(col.astype(np.int64), colo)
    x = np.random.default_rng(3).normal(size=s.ndet)
    sx = s.sigma(x)
    hii = p.hii()
    for r0 in (0, 1999):
        o = p.sigma_rows(x, hii, r0, r0 + 2)
        blk = slice(r0 * s.nb, (r0 + 2) * s.nb)
        assert np.abs(sx[blk] - o[blk]).max() <= 1e-12 * np.abs(o[blk]).max()
    # FP32 sigma on the same rows
    sx32 = s.sigma(x, precision=1)
    for r0 in (0, 3998):
        blk = slice(r0 * s.nb, (r0 + 2) * s.nb)
        assert np.abs(sx32[blk] - sx[blk]).max() <= 2e-5 * np.abs(sx[blk]).max()


def test_sigma_c2_full_size_rows_and_symmetry():
    """Full c2 shape (norb=36, 9+9 electrons, 20,000 x 20,000 strings, 400M determinants),
    the benchmark's workload: sigma on the first and the last alpha row equals the oracle
    restricted to those rows; <y, H x> = <H y, x> over the whole space; the alpha-beta
    kernel in sequence gives the same sigma as beside the spin phases (summation order
    only); FP32 sigma on the sampled rows within 2e-5."""
    wl = syn.config_workload("c2_L36_9e9e_20000")
    case = dict(norb=wl.norb, core=wl.core, h1=wl.h1, eri=wl.eri, astr=wl.astr, bstr=wl.bstr)
    s, p = solver(case), oproblem(case)
    nb = s.nb
    rng = np.random.default_rng(5)
    x = rng.normal(size=s.ndet)
    sx = s.sigma(x)
    for ia in (0, s.na - 1):
        o = p.sigma_rows(x, None, ia, ia + 1)
        blk = slice(ia * nb, (ia + 1) * nb)
        assert np.abs(sx[blk] - o[blk]).max() <= 1e-12 * np.abs(o[blk]).max()
        del o
    sx32 = s.sigma(x, precision=1)
    for ia in (0, s.na - 1):
        blk = slice(ia * nb, (ia + 1) * nb)
        assert np.abs(sx32[blk] - sx[blk]).max() <= 2e-5 * np.abs(sx[blk]).max()
    del sx32
    s.set_options(overlap=0)
    sq = s.sigma(x)
    assert np.abs(sq - sx).max() <= 1e-13 * np.abs(sx).max()
    del sq
    y = rng.normal(size=s.ndet)
    sy = s.sigma(y)
    a, b = float(np.dot(y, sx)), float(np.dot(sy, x))
    assert abs(a - b) <= 1e-12 * np.sqrt(float(np.dot(sx, sx)) * float(np.dot(y, y)))
    s.close()


@pytest.mark.parametrize("L", [45, 46, 48])
def test_sigma_slice_pairs_vs_oracle(L):
    """The alpha-beta kernel stages one integral slice vector per orbital pair b > j (990,
    1,035 and 1,128 pairs, decoded from the pair index on the GPU) and reads (ai|bj) for
    b < j from the pair (j, b): sigma against the oracle, which uses the full tensor."""
    case = workload_small(L, 4, 3, 300, 400, 61)
    s, p = solver(case), oproblem(case)
    x = np.random.default_rng(1).normal(size=s.ndet)
    sx, ox = s.sigma(x), p.sigma(x)
    assert np.abs(sx - ox).max() <= 1e-12 * np.abs(ox).max()
    assert np.abs(s.sigma(x, precision=1) - ox).max() <= 2e-5 * np.abs(ox).max()


@pytest.mark.parametrize("overlap", [1, 0])
def test_sigma_wide_beta_set_vs_oracle(overlap):
    """30,000 beta strings (norb=24, 5 electrons) against 12 alpha strings: one wide column
    tile per bra row (768 threads x 40 columns in FP64, 1024 x 40 in FP32), each chunk's
    sums added into sigma through LDS in column order.  overlap=1 runs the alpha-beta kernel
    on its own stream beside the spin phases, overlap=0 after them on one stream."""
    rng = np.random.default_rng(77)
    L = 24
    h1, eri = syn.synthetic_integrals(L, rng)
    astr = syn.sample_strings(L, 5, 12, rng)
    bstr = syn.sample_strings(L, 5, 30000, rng)
    case = dict(norb=L, core=0.5, h1=h1, eri=eri, astr=astr, bstr=bstr)
    s, p = solver(case, overlap=overlap), oproblem(case)
    x = rng.normal(size=s.ndet)
    sx, ox = s.sigma(x), p.sigma(x)
    assert np.abs(sx - ox).max() <= 1e-12 * np.abs(ox).max()
    sx32 = s.sigma(x, precision=1)
    assert np.abs(sx32 - ox).max() <= 2e-5 * np.abs(ox).max()
    ms = s.bench_sigma(0, 2)
    assert ms[0] > 0 and abs(ms[1:].sum() - ms[0]) <= 1e-3 * ms[0] + 1e-3
    assert np.array_equal(s.sigma(x), sx)  # repeatable


def test_sigma_beta_set_beyond_one_tile_vs_oracle():
    """53,130 beta strings (norb=25, 5 electrons, the full space) against 8 alpha strings:
    beyond the widest single tile (40,704 FP64, 49,152 FP32 columns), so the alpha-beta
    kernel runs two column tiles per bra row, each with its own per-chunk lane order, and
    adds each chunk's sums into sigma at the permuted columns directly.  FP64 and FP32
    against the oracle, the kernel on its own stream and in sequence."""
    rng = np.random.default_rng(78)
    L = 25
    h1, eri = syn.synthetic_integrals(L, rng)
    case = dict(norb=L, core=0.5, h1=h1, eri=eri, astr=syn.sample_strings(L, 5, 8, rng),
                bstr=syn.full_space(L, 5))
    p = oproblem(case)
    x = rng.normal(size=len(case["astr"]) * len(case["bstr"]))
    ox = p.sigma(x)
    for overlap in (1, 0):
        s = solver(case, overlap=overlap)
        sx = s.sigma(x)
        assert np.abs(sx - ox).max() <= 1e-12 * np.abs(ox).max(), overlap
        assert np.abs(s.sigma(x, precision=1) - ox).max() <= 2e-5 * np.abs(ox).max(), overlap
        assert np.array_equal(s.sigma(x), sx)  # repeatable
        s.close()


def test_rdm_beta_set_beyond_register_rows_vs_oracle():
    """The 1- and 2-RDMs of a seeded unit vector over 53,130 beta strings (norb=25, 5
    electrons, the full space) and 6 alpha strings against the oracle (the mixed-spin blocks
    as their sum: the oracle keeps the reference's block labels), each block's partial
    traces and Hermiticity.  At more than 40 bra values per thread (40,960 columns) the
    alpha-beta part leaves the register form for the kernel that re-reads the bra row per
    ket chunk; the same-spin doubles take the excitation-table form (doubles=1) and the
    core-block form (doubles=2)."""
    rng = np.random.default_rng(79)
    L = 25
    h1, eri = syn.synthetic_integrals(L, rng)
    case = dict(norb=L, core=0.5, h1=h1, eri=eri, astr=syn.sample_strings(L, 5, 6, rng),
                bstr=syn.full_space(L, 5))
    w = rng.normal(size=len(case["astr"]) * len(case["bstr"]))
    w /= np.linalg.norm(w)
    one_o, two_o = oproblem(case).rdm(w)
    for doubles in (1, 2):
        s = solver(case, doubles=doubles, max_it=1, max_nb=2, eps=1e10, init=2, do_rdm=1)
        s.set_vector(w)
        s.run()
        assert (s.stats()["core_members"][1] > 0) == (doubles == 2)
        one, two = s.rdm()
        rdm_close(one, two, one_o, two_o, 1e-11, mixed_sum=True)
        rdm_invariants(one, two, 5, 5, L, 1e-11)
        one2, two2 = s.rdm(compute=True)
        assert np.array_equal(one, one2) and np.array_equal(two, two2), doubles  # same bits
        s.close()


def test_core_workspace_cap():
    """A workspace cap below one 512-column pass: auto mode falls back to the
    excitation-table form (same sigma), forcing the core form is a LIMIT error."""
    case = medium_case()
    p = oproblem(case)
    x = np.random.default_rng(4).normal(size=len(case["astr"]) * len(case["bstr"]))
    ox = p.sigma(x)
    s = solver(case, doubles=0, core_gib=1e-6)
    s.build()
    assert s.stats()["core_members"] == [0, 0]
    assert np.abs(s.sigma(x) - ox).max() <= 1e-12 * np.abs(ox).max()
    s2 = solver(case, doubles=0)
    s2.build()
    assert min(s2.stats()["core_members"]) > 0  # the default cap takes the core form here
    s3 = solver(case, doubles=2, core_gib=1e-6)
    with pytest.raises(sbdg.SbdgError) as ei:
        s3.build()
    assert ei.value.code == sbdg.ERR_LIMIT


def test_core_doubles_energy_equals_table_doubles():
    """Davidson energies with either doubles formulation agree to round-off."""
    case = medium_case()
    e1 = solver(case, doubles=1, max_it=50, max_nb=10, eps=1e-9).run()
    e2 = solver(case, doubles=2, max_it=50, max_nb=10, eps=1e-9).run()
    assert abs(e1 - e2) < 1e-10


# ---------------------------------------------------------------- Davidson, RDMs

@pytest.mark.parametrize("name", SMALL_CASES)
def test_energy_rdm_density_vs_reference(name):
    g, case = small_case(name)
    s = solver(case, max_it=200, max_nb=10, eps=float(g["tol"]), do_rdm=1)
    e = s.run()
    assert abs(e - float(g["energy"])) < 1e-8
    one, two = s.rdm()
    rdm_close(one, two, g["rdm1"], g["rdm2"], 1e-7, mixed_sum=True)
    L = case["norb"]
    dens = np.empty(2 * L)
    dens[0::2] = one[0].reshape(L, L).diagonal()
    dens[1::2] = one[1].reshape(L, L).diagonal()
    assert np.abs(dens - g["density"]).max() < 1e-7
    assert np.abs(s.density() - g["density"]).max() < 1e-7
    assert abs(energy_from_rdm(case["core"], case["h1"], case["eri"], one, two) - e) < 1e-9
    w = s.vector()
    assert abs(np.linalg.norm(w) - 1.0) < 1e-12
    assert abs(abs(w @ g["wf"]) - 1.0) < 1e-7


def test_rdm_all_blocks_vs_bruteforce():
    """Every block, the mixed-spin ones included, against second quantisation."""
    for case in (workload_small(6, 3, 2, 12, 9, 21), workload_small(6, 2, 2, None, None, 22)):
        s = solver(case, max_it=50, max_nb=12, eps=1e-11, do_rdm=1)
        s.run()
        one, two = s.rdm()
        one_b, two_b = bf.rdms(case["norb"], s.astr, s.bstr, s.vector())
        rdm_close(one, two, one_b, two_b, 1e-9, mixed_sum=False)


def test_dataset_n2_energy():
    header, vals, idx, strings = dataset("n2")
    g = load_golden("ref_n2_1em3.npz")
    L, core, h1, eri = fio.integrals_from_fcidump(header, vals, idx)
    s = sbdg.Solver(L, core, h1, eri, strings, strings, max_it=100, max_nb=10, eps=1e-8, do_rdm=1)
    e = s.run()
    assert abs(e - float(g["energy"])) < 1e-8
    one, two = s.rdm()
    assert np.abs(one - g["rdm1"]).max() < 1e-7
    assert abs(np.einsum("bijij->", two.reshape(4, L, L, L, L)) - float(g["rdm2_sum_trace"])) < 1e-7
    # the published table value through the six-digit path (data/n2/README.md:8)
    Lr, core_r, h1r, erir = fio.integrals_from_fcidump(header, fio.round_like_broadcast(vals), idx)
    sr = sbdg.Solver(Lr, core_r, h1r, erir, strings, strings, max_it=100, max_nb=10, eps=1e-8)
    er = sr.run()
    assert abs(er - float(g["energy_rounded"])) < 1e-8 and round(er, 6) == -109.041511


def test_dataset_h2o_energy():
    header, vals, idx, strings = dataset("h2o")
    g = load_golden("ref_h2o_1em3.npz")
    L, core, h1, eri = fio.integrals_from_fcidump(header, vals, idx)
    s = sbdg.Solver(L, core, h1, eri, strings, strings, max_it=100, max_nb=10, eps=1e-8)
    e = s.run()
    assert abs(e - float(g["energy"])) < 1e-8
    Lr, core_r, h1r, erir = fio.integrals_from_fcidump(header, fio.round_like_broadcast(vals), idx)
    er = sbdg.Solver(Lr, core_r, h1r, erir, strings, strings, max_it=100, max_nb=10, eps=1e-8).run()
    assert abs(er - float(g["energy_rounded"])) < 1e-8 and round(er, 7) == -76.2359376


def test_c0_energy_and_rdm_vs_reference():
    g = load_golden("ref_c0.npz")
    wl = syn.config_workload("c0_L12_4e4e_full")
    s = sbdg.Solver(wl.norb, wl.core, wl.h1, wl.eri, wl.astr, wl.bstr, max_it=200, max_nb=10, eps=1e-8, do_rdm=1)
    e = s.run()
    assert abs(e - float(g["energy"])) < 1e-8
    one, two = s.rdm()
    rdm_close(one, two, g["rdm1"], g["rdm2"], 1e-7, mixed_sum=True)
    vec = np.random.default_rng(0).normal(size=s.ndet)
    sig = s.sigma(vec)
    assert np.abs(sig[:4096] - g["sigma_head"]).max() < 1e-11
    assert abs(float(sig @ vec) - float(g["sigma_dot_vec"])) < 1e-9 * abs(float(g["sigma_dot_vec"]))


def test_thick_restart_and_warm_start():
    case = medium_case()
    e1 = solver(case, max_it=100, max_nb=8, eps=1e-9).run()
    s2 = solver(case, max_it=100, max_nb=12, restart_keep=4, eps=1e-9)
    e2 = s2.run()
    assert abs(e1 - e2) < 1e-9
    w = s2.vector()
    s3 = solver(case, max_it=10, max_nb=8, eps=1e-8, init=2)
    s3.set_vector(w)
    e3 = s3.run()
    assert abs(e3 - e2) < 1e-9 and s3.stats()["iterations"] <= 2


@pytest.mark.parametrize("refresh_k", [4, 0])
def test_mixed_precision_energy(refresh_k):
    # refresh_k = 0: refreshes scheduled by the FP32 residual, restarts without refresh
    case = medium_case()
    e64 = solver(case, max_it=100, max_nb=12, restart_keep=4, eps=1e-8).run()
    s = solver(case, max_it=100, max_nb=12, restart_keep=4, eps=1e-6, precision=1, refresh_k=refresh_k)
    e = s.run()
    st = s.stats()
    assert st["sigma_fp32"] > 0 and st["converged"] == 1
    assert abs(e - e64) < 1e-6
    t, traj = s.trajectory()
    assert len(traj) >= 1 and abs(traj[-1] - e64) < 1e-6  # exact energies only


def test_hellmann_feynman_backward_vs_finite_differences():
    case = workload_small(8, 3, 3, 40, 40, 61)
    L = case["norb"]
    s = solver(case, max_it=100, max_nb=12, eps=1e-11, do_rdm=1)
    s.run()
    dh, de = s.rdm_backward(1.0)
    rng = np.random.default_rng(5)
    step = 1e-4

    def energy(h1, eri):
        return solver(dict(case, h1=h1, eri=eri), max_it=100, max_nb=12, eps=1e-11).run()

    for _ in range(4):
        p, q = rng.integers(0, L, size=2)
        dH = np.zeros((L, L))
        dH[p, q] += 1.0
        dH[q, p] = 1.0 if p != q else dH[p, q]
        fd = (energy(case["h1"] + step * dH, case["eri"]) - energy(case["h1"] - step * dH, case["eri"])) / (2 * step)
        an = np.sum(dh * dH)
        assert abs(fd - an) < 1e-6, (p, q, fd, an)
    for _ in range(4):
        p, q, r, t = rng.integers(0, L, size=4)
        dE = np.zeros((L, L, L, L))
        for a, b, c, d in ((p, q, r, t), (q, p, r, t), (p, q, t, r), (q, p, t, r),
                           (r, t, p, q), (t, r, p, q), (r, t, q, p), (t, r, q, p)):
            dE[a, b, c, d] = 1.0
        fd = (energy(case["h1"], case["eri"] + step * dE) - energy(case["h1"], case["eri"] - step * dE)) / (2 * step)
        an = np.sum(de * dE)
        assert abs(fd - an) < 1e-6, (p, q, r, t, fd, an)


def test_backward_vs_finite_differences_c3():
    """norb=16, 5+5 full space (19.1M determinants): the RDM contraction matches central
    differences (step 1e-4) of the FP64 Davidson energy to 1e-6 on 16 random symmetric
    one-electron and 16 random 8-fold-symmetric two-electron entries."""
    wl = syn.config_workload("c3_L16_5e5e_full")
    L = wl.norb
    s = sbdg.Solver(L, wl.core, wl.h1, wl.eri, wl.astr, wl.bstr, max_it=100, max_nb=12, restart_keep=4,
                    eps=1e-8, do_rdm=1)
    s.run()
    dh, de = s.rdm_backward(1.0)
    w0 = s.vector()
    s.set_options(do_rdm=0)
    step = 1e-4

    def energy(h1, eri):
        s.set_integrals(wl.core, h1, eri)
        s.set_vector(w0)  # warm start from the unperturbed ground state
        s.set_options(init=2)
        return s.run()

    rng = np.random.default_rng(32)
    errs = []
    for _ in range(16):
        p, q = rng.integers(0, L, size=2)
        dH = np.zeros((L, L))
        dH[p, q] = dH[q, p] = 1.0
        fd = (energy(wl.h1 + step * dH, wl.eri) - energy(wl.h1 - step * dH, wl.eri)) / (2 * step)
        errs.append(abs(fd - np.sum(dh * dH)))
    for _ in range(16):
        p, q, r, t = rng.integers(0, L, size=4)
        dE = np.zeros((L, L, L, L))
        for a, b, c, d in ((p, q, r, t), (q, p, r, t), (p, q, t, r), (q, p, t, r),
                           (r, t, p, q), (t, r, p, q), (r, t, q, p), (t, r, q, p)):
            dE[a, b, c, d] = 1.0
        fd = (energy(wl.h1, wl.eri + step * dE) - energy(wl.h1, wl.eri - step * dE)) / (2 * step)
        errs.append(abs(fd - np.sum(de * dE)))
    print(f"c3 finite differences: 32 entries, max |FD - backward| = {max(errs):.2e}", flush=True)
    assert max(errs) < 1e-6


def test_autograd_function():
    import torch
    case = workload_small(8, 3, 3, 40, 40, 61)
    s = solver(case, max_it=100, max_nb=12, eps=1e-11)
    h1 = torch.tensor(case["h1"], requires_grad=True)
    eri = torch.tensor(case["eri"], requires_grad=True)
    e = sbdg.ground_state_energy(h1, eri, s, case["core"])
    e.backward()
    one, two = s.rdm()
    L = case["norb"]
    assert torch.allclose(h1.grad, torch.tensor((one[0] + one[1]).reshape(L, L, order="F")))
    # energy identity through the gradients: E = core + <h, dE/dh> + <eri, dE/deri>
    ident = case["core"] + float((h1.grad * h1).sum()) + float((eri.grad * eri).sum())
    assert abs(ident - float(e)) < 1e-9


def test_autograd_two_forwards_one_solver():
    """Two energies from one solver before any backward: each backward uses the RDMs of its
    own forward (the second solve, at other integrals, does not leak into the first)."""
    import torch
    case = workload_small(8, 3, 3, 40, 40, 61)
    ref = solver(case, max_it=100, max_nb=12, eps=1e-11)
    h1 = torch.tensor(case["h1"], requires_grad=True)
    eri = torch.tensor(case["eri"], requires_grad=True)
    sbdg.ground_state_energy(h1, eri, ref, case["core"]).backward()
    g_h, g_e = h1.grad.clone(), eri.grad.clone()
    h1.grad = eri.grad = None
    s = solver(case, max_it=100, max_nb=12, eps=1e-11)
    e1 = sbdg.ground_state_energy(h1, eri, s, case["core"])
    e2 = sbdg.ground_state_energy(1.1 * h1, eri, s, case["core"])
    assert abs(float(e1) - float(e2)) > 1e-3
    e1.backward()
    assert torch.equal(h1.grad, g_h) and torch.equal(eri.grad, g_e)
    h1.grad = eri.grad = None
    (2.0 * e2).backward()
    h1b = torch.tensor(1.1 * case["h1"], requires_grad=True)
    sbdg.ground_state_energy(h1b, eri.detach(), solver(case, max_it=100, max_nb=12, eps=1e-11), case["core"]).backward()
    assert torch.allclose(h1.grad, 2.2 * h1b.grad, rtol=1e-12, atol=1e-14)  # the second point, chain rule included


def test_carryover_matches_row_weights():
    case = medium_case()
    s = solver(case, max_it=100, max_nb=10, eps=1e-8, threshold=0.01)
    s.run()
    w = s.vector().reshape(s.na, s.nb)
    D = (w * w).sum(1)
    order = np.argsort(-D, kind="stable")
    S = np.cumsum(D[order])
    kept = int(np.searchsorted(S, 1 - 0.01, side="right")) + 1
    co, _ = s.carryover()
    assert len(co) == kept and np.array_equal(co, s.astr[order[:kept]])
    co2, tw = s.carryover(ratio=0.1)
    assert len(co2) == int(0.1 * s.na) and abs(tw - (1 - D[order[:len(co2)]].sum())) < 1e-12


@pytest.mark.parametrize("name", ("asym_L14", "sym_L10"))
def test_carryover_vs_reference(name):
    """Carry-over strings against the ones the reference's diag wrote (CarryOverAlphaDet,
    tpb/rdmat.h:60-132) by threshold (0.01, 0.2) and by ratio (0.3): the same strings in the
    same order (heaviest rows first; the reference sorts with std::sort, so rows of equal
    weight may come in either order)."""
    g = load_golden("ref_carryover.npz")
    _, case = small_case(name)
    s = solver(case, max_it=200, max_nb=10, eps=1e-10)
    s.run()
    w = s.vector().reshape(s.na, s.nb)
    weight = dict(zip(s.astr.tolist(), (w * w).sum(1)))
    for tag, kw in (("thr001", dict(threshold=0.01, ratio=0.0)), ("thr02", dict(threshold=0.2, ratio=0.0)),
                    ("ratio03", dict(ratio=0.3))):
        want = g[f"{name}_{tag}"]
        got, _ = s.carryover(**kw)
        assert len(got) == len(want), tag
        assert set(got.tolist()) == set(want.tolist()), tag
        for x, y in zip(got.tolist(), want.tolist()):
            assert x == y or abs(weight[x] - weight[y]) < 1e-12, (tag, x, y)
    s.close()


def test_errors_from_the_host():
    case = workload_small(6, 3, 2, 12, 9, 21)
    with pytest.raises(sbdg.SbdgError) as ex:
        solver(dict(case, astr=np.concatenate([case["astr"], case["astr"][:1]])))
    assert ex.value.code == sbdg.ERR_ARG
    with pytest.raises(sbdg.SbdgError):
        solver(dict(case, bstr=np.array([0b111, 0b11], dtype=np.uint64)))
    with pytest.raises(sbdg.SbdgError):
        solver(case, max_nb=0)
    # a full tensor without the 8-fold symmetry ((01|23) != (10|23) ...) is refused
    bad = case["eri"].copy()
    bad[0, 1, 2, 3] += 1e-3
    with pytest.raises(sbdg.SbdgError) as ex:
        solver(dict(case, eri=bad))
    assert ex.value.code == sbdg.ERR_ARG and "symmetry" in str(ex.value)
    s = solver(case)
    with pytest.raises(sbdg.SbdgError) as ex:
        s.vector()
    assert ex.value.code == sbdg.ERR_STATE
    with pytest.raises(sbdg.SbdgError):
        s.set_options(init=2)
        s.run()
    # max_time ~ 0 stops after the first step with an unconverged vector
    s = solver(case, max_it=50, max_nb=8, eps=1e-12, max_time=1e-9)
    s.run()
    assert s.stats()["iterations"] == 1 and s.stats()["converged"] == 0
