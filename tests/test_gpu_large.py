"""The HIP path at the largest shapes and on the parity points the mixed path and the RDMs
need (through the C-ABI).

* norb=58, 12+12 electrons (the shape of the 1.6e9-determinant workload) on small sampled
  sets against the oracle: sigma, FP64 and mixed energies, RDMs;
* the full 40,000 x 40,000 workload: sigma rows against the oracle, symmetry, FP32 rows;
* norb=26, 5+5, 4,000 x 4,000 at its stated settings (tol 1e-8, subspace 24, restart to 4):
  the converged residual recomputed on oracle rows, and the mixed energy;
* norb=16, 5+5 full space (19.1M determinants): RDM invariants and bitwise reproducibility,
  and sigma and the RDMs of a seeded vector against the reference's own run;
* the mixed energy against the reference's own energies (golden fixtures), with the
  refresh every 4 steps, scheduled by the residual, with the large workload's subspace
  (8, restart to 3) and stopped by target_de;
* the device memory the library holds against its own plan.

Bars: FP64 sigma within 1e-12 of the oracle (relative to its largest element), FP32 within
2e-5; FP64 energies within 1e-8 Ha; mixed energies within 1e-6 Ha of the reference; RDMs
within 1e-7 (1e-10 when both sides use the same vector); invariants within 1e-10; RDMs
bit-identical between two evaluations.
"""
import time

import numpy as np
import pytest

import oracle
import sbdg
from conftest import load_golden
from helpers import SMALL_CASES, energy_from_rdm, rdm_close, rdm_invariants, small_case, workload_small
from sbdg import synthetic as syn

pytestmark = pytest.mark.gpu


def solver(case, **opt):
    return sbdg.Solver(case["norb"], case["core"], case["h1"], case["eri"], case["astr"], case["bstr"], **opt)


def oproblem(case):
    return oracle.Problem(case["norb"], case["core"], case["h1"], case["astr"], case["bstr"], eri=case["eri"])


def as_case(wl):
    return dict(norb=wl.norb, core=wl.core, h1=wl.h1, eri=wl.eri, astr=wl.astr, bstr=wl.bstr)


# ---------------------------------------------------------------- norb=58, 12+12

def test_norb58_12e_small_sets_vs_oracle():
    """The shape of the largest workload (58 orbitals, 12 electrons per spin: strings
    beyond 32 bits, integral slices of 58^2, the excitation fields at their widest) on
    300 x 400 sampled strings: tables, sigma, FP64 and mixed Davidson (subspace 8, restart
    to 3, refresh every 4), RDMs against the oracle."""
    case = workload_small(58, 12, 12, 300, 400, 71)
    s, p = solver(case), oproblem(case)
    for which in range(4):
        rp, col, _, _ = s.table(which)
        rpo, colo = p.table(which)
        assert np.array_equal(rp, rpo) and np.array_equal(col.astype(np.int64), colo)
    assert s.table(1)[0][-1] > 0 and s.table(3)[0][-1] > 0  # doubles present
    x = np.random.default_rng(1).normal(size=s.ndet)
    sx, ox = s.sigma(x), p.sigma(x)
    assert np.abs(sx - ox).max() <= 1e-12 * np.abs(ox).max()
    assert np.abs(s.sigma(x, precision=1) - ox).max() <= 2e-5 * np.abs(ox).max()
    ref = p.davidson(max_it=200, nb=10, eps=1e-9)
    s.set_options(max_it=200, max_nb=10, eps=1e-9, do_rdm=1)
    e = s.run()
    assert abs(e - ref["energy"]) < 1e-8
    one, two = s.rdm()
    one_o, two_o = p.rdm(ref["W"])
    rdm_close(one, two, one_o, two_o, 1e-7, mixed_sum=True)
    rdm_invariants(one, two, 12, 12, 58, 1e-10)
    assert abs(energy_from_rdm(case["core"], case["h1"], case["eri"], one, two) - e) < 1e-9
    s.set_options(max_it=200, max_nb=8, restart_keep=3, eps=1e-6, precision=1, refresh_k=4, do_rdm=0)
    em = s.run()
    assert s.stats()["sigma_fp32"] > 0 and s.stats()["converged"] == 1
    assert abs(em - ref["energy"]) < 1e-6
    s.close()


def test_sigma_c4_full_size_rows_and_symmetry():
    """Full norb=58, 12+12, 40,000 x 40,000 workload (1.6e9 determinants): sigma on the first
    and the last alpha row equals the oracle restricted to those rows; <y, H x> = <H y, x>
    over the whole space; FP32 sigma on the same rows within 2e-5."""
    t0 = time.time()
    wl = syn.config_workload("c4_L58_12e12e_40000")
    s, p = solver(as_case(wl)), oproblem(as_case(wl))
    s.build()
    st = s.stats()
    print(f"c4: workload + tables {time.time() - t0:.1f} s (GPU tables {st['t_tables']:.1f} s), "
          f"device bytes {st['mem_now'] / 1e9:.1f} GB", flush=True)
    nb = s.nb
    rng = np.random.default_rng(5)
    x = rng.standard_normal(s.ndet, dtype=np.float32).astype(np.float64)
    sx = s.sigma(x)
    for ia in (0, s.na - 1):
        o = p.sigma_rows(x, None, ia, ia + 1)
        blk = slice(ia * nb, (ia + 1) * nb)
        assert np.abs(sx[blk] - o[blk]).max() <= 1e-12 * np.abs(o[blk]).max()
        del o
    print(f"c4: sigma rows checked at {time.time() - t0:.1f} s", flush=True)
    sx32 = s.sigma(x, precision=1)
    for ia in (0, s.na - 1):
        blk = slice(ia * nb, (ia + 1) * nb)
        assert np.abs(sx32[blk] - sx[blk]).max() <= 2e-5 * np.abs(sx[blk]).max()
    del sx32
    y = rng.standard_normal(s.ndet, dtype=np.float32).astype(np.float64)
    sy = s.sigma(y)
    a, b = float(np.dot(y, sx)), float(np.dot(sy, x))
    assert abs(a - b) <= 1e-12 * np.sqrt(float(np.dot(sx, sx)) * float(np.dot(y, y)))
    print(f"c4: done at {time.time() - t0:.1f} s", flush=True)
    s.close()


# ---------------------------------------------------------------- stated settings

def test_c1_davidson_at_stated_settings():
    """norb=26, 5+5, 4,000 x 4,000 strings (16M determinants), FP64 Davidson at tol 1e-8 with
    subspace 24 and restart to 4: converged, and the residual H w - E w recomputed by the
    oracle on sampled rows is within the tolerance; the mixed Davidson (refresh every 4)
    reaches the same energy within 1e-6 Ha."""
    wl = syn.config_workload("c1_L26_5e5e_4000")
    case = as_case(wl)
    s, p = solver(case, max_it=50, max_nb=24, restart_keep=4, eps=1e-8), oproblem(case)
    e = s.run()
    st = s.stats()
    assert st["converged"] == 1 and st["resid"] < 1e-8
    w = s.vector()
    nb = s.nb
    hii = None
    for ia in (0, 1, 2000, s.na - 1):
        hw = p.sigma_rows(w, hii, ia, ia + 1)
        blk = slice(ia * nb, (ia + 1) * nb)
        r = hw[blk] - e * w[blk]
        assert np.linalg.norm(r) < 1e-8, (ia, np.linalg.norm(r))
    s.set_options(precision=1, refresh_k=4, eps=1e-6)
    em = s.run()
    assert s.stats()["converged"] == 1 and abs(em - e) < 1e-6
    print(f"c1: E = {e:.12f} in {st['iterations']} steps; mixed error {abs(em - e):.2e}", flush=True)
    s.close()


def test_c3_rdm_invariants_and_reproducibility():
    """norb=16, 5+5 full space (19.1M determinants): the RDMs of the converged vector satisfy
    tr gamma = N, Hermiticity, the partial traces and the energy identity against the
    Davidson energy; a second evaluation gives the same bits."""
    wl = syn.config_workload("c3_L16_5e5e_full")
    case = as_case(wl)
    s = solver(case, max_it=100, max_nb=12, restart_keep=4, eps=1e-8, do_rdm=1)
    e = s.run()
    one, two = s.rdm()
    rdm_invariants(one, two, 5, 5, 16, 1e-10)
    assert abs(energy_from_rdm(wl.core, wl.h1, wl.eri, one, two) - e) < 1e-9
    one2, two2 = s.rdm(compute=True)
    assert np.array_equal(one, one2) and np.array_equal(two, two2)
    s.close()


def weighted_sum(a):
    """The fixtures' checksum: sum of a[i] cos(0.618034 i)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    return float(a @ np.cos(np.arange(a.size) * 0.618034))


def test_c3_sigma_and_rdm_vs_reference():
    """norb=16, 5+5 full space (19.1M determinants) against the reference's own run (mult,
    tpb/mult.h, and Correlation, tpb/correlation.h:194-444, on one rank): sigma of the seeded
    vector numpy default_rng(33).normal, and the 1-/2-RDMs of that vector scaled to unit norm.
    The vector enters as the Davidson start with a tolerance so loose that Davidson stops at
    it, as the reference's loadname path does."""
    g = load_golden("ref_c3_rdm.npz")
    wl = syn.config_workload("c3_L16_5e5e_full")
    inputs = sum(weighted_sum(x) for x in (wl.h1, wl.eri, wl.astr.astype(np.float64), wl.bstr.astype(np.float64)))
    assert abs(inputs - float(g["input_checksum"])) < 1e-9 * abs(float(g["input_checksum"]))
    vec = np.random.default_rng(int(g["vec_seed"])).normal(size=wl.ndet)
    s = sbdg.Solver(wl.norb, wl.core, wl.h1, wl.eri, wl.astr, wl.bstr, max_it=1, max_nb=2, eps=1e10, init=2,
                    do_rdm=1)
    s.set_vector(vec)
    e = s.run()
    nn = float(vec @ vec)
    sig = s.sigma(vec)
    scale = np.abs(sig).max()
    assert np.abs(sig[:4096] - g["sigma_head"]).max() < 1e-12 * scale
    assert abs(float(sig @ vec) - float(g["sigma_dot_vec"])) < 1e-10 * abs(float(g["sigma_dot_vec"]))
    assert abs(float(np.linalg.norm(sig)) - float(g["sigma_norm"])) < 1e-10 * float(g["sigma_norm"])
    assert abs(weighted_sum(sig) - float(g["sigma_checksum"])) < 1e-9 * float(g["sigma_norm"])
    assert abs(e - float(g["sigma_dot_vec"]) / nn) < 1e-9  # the energy of the start vector
    one, two = s.rdm()
    rdm_close(one, two, g["rdm1"], g["rdm2"], 1e-10, mixed_sum=True)
    s.close()


# ---------------------------------------------------------------- mixed vs the reference

@pytest.mark.parametrize("name", SMALL_CASES + ("c0",))
def test_mixed_energy_vs_reference(name):
    """The mixed Davidson's energy against the reference's own FP64 energy (golden), with
    the refresh every 4 steps, scheduled by the residual, and with the largest workload's
    subspace (8, restart to 3)."""
    if name == "c0":
        g = load_golden("ref_c0.npz")
        case = as_case(syn.config_workload("c0_L12_4e4e_full"))
    else:
        g, case = small_case(name)
    e_ref = float(g["energy"])
    for opt in (dict(max_nb=10, refresh_k=4), dict(max_nb=12, restart_keep=4, refresh_k=0),
                dict(max_nb=8, restart_keep=3, refresh_k=4)):
        s = solver(case, max_it=200, eps=1e-6, precision=1, **opt)
        e = s.run()
        st = s.stats()
        assert st["converged"] == 1 and st["sigma_fp32"] > 0, opt
        assert abs(e - e_ref) < 1e-6, (opt, e, e_ref)
        s.close()


def test_mixed_target_de_stop():
    """target_de: the mixed run stops when two successive exact (FP64) energies agree to
    target_de, with the residual tolerance out of reach; the energy then lies within 1e-6 Ha
    of the reference's."""
    g = load_golden("ref_c0.npz")
    case = as_case(syn.config_workload("c0_L12_4e4e_full"))
    s = solver(case, max_it=400, max_nb=10, eps=0.0, precision=1, refresh_k=4, target_de=1e-9)
    e = s.run()
    st = s.stats()
    assert st["converged"] == 1 and st["resid"] > 0.0
    assert st["iterations"] < 400 * 10
    _, traj = s.trajectory()
    assert len(traj) >= 2 and abs(traj[-1] - traj[-2]) < 1e-9
    assert abs(e - float(g["energy"])) < 1e-6


# ---------------------------------------------------------------- memory and determinism

def test_device_memory_within_plan():
    """The device bytes the library holds: the Davidson vectors of each precision are
    exactly the plan's (sigma's scratch borrows W, R, T; HV is freed in the mixed path,
    HV32/HX/c32 in the FP64 path), and the run's high-water mark stays within the planned
    build + Davidson + RDM bytes."""
    case = workload_small(26, 5, 5, 400, 300, 31)
    L = case["norb"]
    s = solver(case, max_it=100, max_nb=8, restart_keep=3, eps=1e-8, do_rdm=1)
    s.build()
    st = s.stats()
    built = st["mem_now"]
    extra = 8 * (2 * L * L + 4 * L ** 4 + 2 * L)  # RDM and density outputs
    for prec in (1, 0, 1):
        s.set_options(precision=prec, refresh_k=4)
        s.run()
        st = s.stats()
        plan = sbdg.memory_plan(L, s.na, s.nb, st["nnz"], st["core_members"], st["core_direct"],
                                precision=prec, max_nb=8, restart_keep=3)
        held = st["mem_now"] - built
        assert abs(held - plan["davidson"] - extra) <= 1 << 20, (prec, held, plan)
        assert st["mem_peak"] - built <= plan["davidson"] + plan["rdm"] + (1 << 20), (prec, st["mem_peak"], plan)
    s.close()


def test_rdm_bitwise_reproducible():
    """The RDMs accumulate with exact integer adds where lanes share a slot: two evaluations
    of the same vector give identical bits (mixed-spin blocks included)."""
    case = workload_small(26, 5, 5, 400, 300, 31)
    s = solver(case, max_it=100, max_nb=10, eps=1e-9, do_rdm=1)
    s.run()
    one, two = s.rdm()
    for _ in range(2):
        one2, two2 = s.rdm(compute=True)
        assert np.array_equal(one, one2) and np.array_equal(two, two2)
    p = oproblem(case)
    one_o, two_o = p.rdm(s.vector())
    rdm_close(one, two, one_o, two_o, 1e-9, mixed_sum=True)
    s.close()


# ---------------------------------------------------------------- repeated calls, reused memory

def test_sigma_repeated_calls_several_ket_chunks():
    """Regression: 120 x 8,000 sampled strings at norb=36 (the alpha-beta kernel takes the
    ket columns in two chunks; a third of the alpha rows have no alpha single).  sigma on one
    context several times in a row, with the alpha-beta kernel beside the spin phases (its
    terms in a scratch vector that holds the previous call's values) and in sequence, each
    time equal to the oracle; then an FP64 Davidson and, on the same context (its vectors now
    holding FP64 data), the mixed one: the energies agree, and the FP64 energy equals the
    oracle's Rayleigh quotient of the returned vector."""
    case = workload_small(36, 9, 9, 120, 8000, 29)
    s, p = solver(case), oproblem(case)
    rp, _, _, _ = s.table(0)
    assert np.any(np.diff(rp) == 0)  # alpha rows without alpha singles
    rng = np.random.default_rng(3)
    xs = [rng.normal(size=s.ndet) for _ in range(2)]
    os_ = [p.sigma(x) for x in xs]
    for overlap in (1, 0, 1):
        s.set_options(overlap=overlap)
        for x, o in zip(xs, os_):
            assert np.abs(s.sigma(x) - o).max() <= 1e-12 * np.abs(o).max(), overlap
            assert np.abs(s.sigma(x, precision=1) - o).max() <= 2e-5 * np.abs(o).max(), overlap
    s.set_options(overlap=1, max_it=100, max_nb=10, restart_keep=3, eps=1e-7, precision=0)
    e64 = s.run()
    w = s.vector()
    e_rq = float(w @ p.sigma(w)) / float(w @ w)
    assert abs(e64 - e_rq) < 1e-10, (e64, e_rq)
    s.set_options(precision=1, refresh_k=4, eps=1e-6)
    em = s.run()
    assert s.stats()["sigma_fp32"] > 0 and abs(em - e64) < 1e-6, (em, e64)
    s.close()
