"""Shared test helpers: fixtures -> problems, RDM comparisons, energy identity."""
import gzip
import os

import numpy as np

from conftest import GOLDEN, load_golden
from sbdg import fcidump as fio
from sbdg import synthetic as syn

SMALL_CASES = ("asym_L8", "asym_L14", "sym_L10")


def small_case(name):
    g = load_golden(f"ref_{name}.npz")
    return g, dict(norb=int(g["norb"]), core=float(g["core"]), h1=g["h1"], eri=g["eri"],
                   astr=g["astr"], bstr=g["bstr"])


def tables_from_ref(flat, na, nb):
    """The reference's helper dump: per table the row lengths, then the entries."""
    out, pos = [], 0
    for n in (na, na, nb, nb):
        lens = flat[pos:pos + n]
        pos += n
        rp = np.zeros(n + 1, dtype=np.int64)
        rp[1:] = np.cumsum(lens)
        cols = flat[pos:pos + rp[-1]]
        pos += rp[-1]
        out.append((rp, cols))
    assert pos == len(flat)
    return out


def dataset(tag):
    """Reference dataset (data/n2 or data/h2o) with the threshold-1e-3 strings."""
    path = os.path.join(GOLDEN, f"{tag}_fcidump.txt.gz")
    tmp = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"sbdg_{tag}_{os.getpid()}.txt")
    with gzip.open(path, "rb") as f, open(tmp, "wb") as g:
        g.write(f.read())
    try:
        header, vals, idx = fio.load_fcidump(tmp)
    finally:
        os.remove(tmp)
    strings = fio.load_alpha_strings(os.path.join(GOLDEN, f"{tag}_1em3-alpha.txt"))
    return header, vals, idx, strings


def energy_from_rdm(core, h1, eri, one, two):
    """E = I0 + sum h_ij gamma[i+Lj] + 1/2 sum (ik|jl) Gamma[i+Lj+L^2k+L^3l]
    (samples/selected_basis_diagonalization/main.cc:166-216)."""
    L = h1.shape[0]
    g = (one[0] + one[1]).reshape(L, L, order="F")      # g[i,j] = gamma[i + L j]
    G = two.sum(0).reshape(L, L, L, L, order="F")       # G[i,j,k,l]
    e1 = np.einsum("ij,ij->", h1, g)
    e2 = 0.5 * np.einsum("ikjl,ijkl->", eri, G)
    return core + e1 + e2


def rdm_close(a1, a2, b1, b2, tol, mixed_sum=False):
    """1-RDM, same-spin blocks, and either every block or blocks 1+2 summed."""
    assert np.abs(a1 - b1).max() < tol, np.abs(a1 - b1).max()
    for k in (0, 3):
        assert np.abs(a2[k] - b2[k]).max() < tol, (k, np.abs(a2[k] - b2[k]).max())
    if mixed_sum:
        assert np.abs((a2[1] + a2[2]) - (b2[1] + b2[2])).max() < tol
    else:
        for k in (1, 2):
            assert np.abs(a2[k] - b2[k]).max() < tol, (k, np.abs(a2[k] - b2[k]).max())


def workload_small(L, na_e, nb_e, na, nb, seed):
    rng = np.random.default_rng(seed)
    h1, eri = syn.synthetic_integrals(L, rng)
    a = syn.sample_strings(L, na_e, na, rng) if na else syn.full_space(L, na_e)
    b = syn.sample_strings(L, nb_e, nb, rng) if nb else syn.full_space(L, nb_e)
    return dict(norb=L, core=0.25, h1=h1, eri=eri, astr=a, bstr=b)


def core_split(strs, kmin=10):
    """Brute-force split of the same-spin doubles by core: (memberships of the cores with at
    least kmin members, ordered pairs of the smaller cores).  A member counts when its pair
    is disjoint from another member's pair; a pair counts when the two pairs are disjoint."""
    members = {}
    for s in strs:
        s = int(s)
        occ = [i for i in range(64) if s >> i & 1]
        for i, a in enumerate(occ):
            for b in occ[i + 1:]:
                members.setdefault(s & ~(1 << a) & ~(1 << b), []).append((1 << a) | (1 << b))
    nmem = direct = 0
    for ms in members.values():
        keep = [m for j, m in enumerate(ms) if any((m & o) == 0 for t, o in enumerate(ms) if t != j)]
        if len(keep) < 2:
            continue
        if len(keep) >= kmin:
            nmem += len(keep)
        else:
            direct += sum(1 for j, x in enumerate(keep) for t, y in enumerate(keep) if t != j and (x & y) == 0)
    return nmem, direct


def rdm_invariants(one, two, na_e, nb_e, L, tol):
    """tr gamma^s = N_s; gamma Hermitian; sum_j Gamma^{st}[i j k j] = (N_t - delta_st)
    gamma^s[i k] for every block (the library's mixed blocks are the physical ones)."""
    assert abs(one[0].reshape(L, L).trace() - na_e) < tol
    assert abs(one[1].reshape(L, L).trace() - nb_e) < tol
    for s in range(2):
        G = one[s].reshape(L, L)
        assert np.abs(G - G.T).max() < tol
    for s in range(2):
        for t in range(2):
            G = two[s + 2 * t].reshape(L, L, L, L, order="F")
            Nt = (na_e, nb_e)[t] - (1 if s == t else 0)
            pt = np.einsum("ijkj->ik", G)
            assert np.abs(pt - Nt * one[s].reshape(L, L, order="F")).max() < tol, (s, t)
    # Hermiticity of the 2-RDM: Gamma[i j k l] = Gamma[k l i j] within each block
    for b in range(4):
        G = two[b].reshape(L, L, L, L, order="F")
        assert np.abs(G - G.transpose(2, 3, 0, 1)).max() < tol, b
