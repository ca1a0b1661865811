"""Seeded synthetic workloads: one-/two-electron integrals and sampled string sets.

The construction is the one the benchmark configurations name:

* ``h`` is symmetric; its diagonal is ``linspace(-2.0, 1.0, norb)`` (so the lowest
  ``n`` orbitals form the Hartree-Fock occupation) and each off-diagonal pair is drawn
  from N(0, 0.05), upper triangle in row-major order.
* ``(pq|rs) = sum_K B^K_pq B^K_rs`` over ``K = 1..2*norb``, each ``B^K`` symmetric with
  its upper triangle (diagonal included, row-major) drawn from N(0, 0.1).  The tensor is
  formed on the canonical pairs and expanded, so its 8-fold symmetry is exact and it is
  positive semidefinite as a pair matrix.  The core energy is 0.
* String sets either enumerate the full space (all ``C(norb, n)`` strings) or are
  sampled: the Hartree-Fock string is always present; further strings pick an
  excitation level ``k >= 1`` with probability proportional to ``0.5**k`` (``k`` up to
  ``min(n, norb - n)``), then ``k`` holes uniformly among the ``n`` lowest orbitals and
  ``k`` particles uniformly among the others (uniform within the level); duplicates are
  dropped until the count is reached, and the set is sorted by integer value.  Alpha is
  drawn before beta, from one ``numpy.random.default_rng(seed)`` stream shared with the
  integrals (integrals first).

Bit ``i`` of a string is orbital ``i`` (the reference's ``from_string`` reads the
rightmost character as orbital 0: bit_manipulation.h:158-172).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass

import numpy as np


@dataclass
class Workload:
    norb: int
    nalpha: int
    nbeta: int
    core: float
    h1: np.ndarray          # (norb, norb) float64
    eri: np.ndarray         # (norb,)*4 float64, chemists' (pq|rs)
    astr: np.ndarray        # uint64, sorted unique
    bstr: np.ndarray        # uint64, sorted unique

    @property
    def ndet(self) -> int:
        return int(len(self.astr)) * int(len(self.bstr))


def synthetic_integrals(norb: int, rng: np.random.Generator):
    L = norb
    h = np.zeros((L, L))
    iu = np.triu_indices(L, 1)
    off = rng.normal(0.0, 0.05, size=len(iu[0]))
    h[iu] = off
    h[(iu[1], iu[0])] = off
    h[np.diag_indices(L)] = np.linspace(-2.0, 1.0, L)
    # B^K on the canonical pairs p>=q (same draw order as the upper triangle row-major)
    iu0 = np.triu_indices(L)
    npair = len(iu0[0])
    Bu = np.empty((2 * L, npair))
    for K in range(2 * L):
        Bu[K] = rng.normal(0.0, 0.1, size=npair)
    M = Bu.T @ Bu                     # (pair, pair)
    M = np.tril(M) + np.tril(M, -1).T  # take the lower triangle as canonical: exact symmetry
    pidx = np.zeros((L, L), dtype=np.int64)
    pidx[iu0] = np.arange(npair)
    pidx[(iu0[1], iu0[0])] = np.arange(npair)
    eri = M[pidx.reshape(-1)][:, pidx.reshape(-1)].reshape(L, L, L, L)
    return h, np.ascontiguousarray(eri)


def full_space(norb: int, nelec: int) -> np.ndarray:
    out = []
    for occ in itertools.combinations(range(norb), nelec):
        s = 0
        for o in occ:
            s |= 1 << o
        out.append(s)
    return np.array(sorted(out), dtype=np.uint64)


def sample_strings(norb: int, nelec: int, count: int, rng: np.random.Generator) -> np.ndarray:
    from math import comb

    total = comb(norb, nelec)
    if count > total:
        raise ValueError(f"cannot sample {count} distinct strings from C({norb},{nelec})={total}")
    hf = (1 << nelec) - 1
    kmax = min(nelec, norb - nelec)
    ks = np.arange(1, kmax + 1)
    p = 0.5 ** ks
    p /= p.sum()
    occ = np.arange(nelec)
    vir = np.arange(nelec, norb)
    seen = {hf}
    while len(seen) < count:
        k = int(rng.choice(ks, p=p))
        holes = rng.choice(occ, size=k, replace=False)
        parts = rng.choice(vir, size=k, replace=False)
        s = hf
        for h in holes:
            s &= ~(1 << int(h))
        for q in parts:
            s |= 1 << int(q)
        seen.add(s)
    return np.array(sorted(seen), dtype=np.uint64)


def make_workload(norb: int, nalpha: int, nbeta: int, seed: int,
                  na: int | None = None, nb: int | None = None) -> Workload:
    """Full space when ``na``/``nb`` are None, else sampled sets of that size."""
    rng = np.random.default_rng(seed)
    h1, eri = synthetic_integrals(norb, rng)
    a = full_space(norb, nalpha) if na is None else sample_strings(norb, nalpha, na, rng)
    b = full_space(norb, nbeta) if nb is None else sample_strings(norb, nbeta, nb, rng)
    return Workload(norb, nalpha, nbeta, 0.0, h1, eri, a, b)


# The benchmark configurations, by their content.
CONFIGS = {
    "c0_L12_4e4e_full": dict(norb=12, nalpha=4, nbeta=4, seed=0),
    "c1_L26_5e5e_4000": dict(norb=26, nalpha=5, nbeta=5, seed=1, na=4000, nb=4000),
    "c2_L36_9e9e_20000": dict(norb=36, nalpha=9, nbeta=9, seed=2, na=20000, nb=20000),
    "c3_L16_5e5e_full": dict(norb=16, nalpha=5, nbeta=5, seed=3),
    "c4_L58_12e12e_40000": dict(norb=58, nalpha=12, nbeta=12, seed=4, na=40000, nb=40000),
}


def config_workload(name: str) -> Workload:
    return make_workload(**CONFIGS[name])
