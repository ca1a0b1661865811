"""Brute-force second quantisation for tiny product bases (test infrastructure).

Independent of the Slater-Condon formulas: every matrix element and RDM entry is formed
by applying creation/annihilation operators to 2L-bit occupation words.  Spin-orbital
``2p`` is alpha orbital ``p`` and ``2p+1`` is beta orbital ``p`` (the reference's
interleaved order, determinants.h:96-142); an operator on spin-orbital ``k`` picks up
``(-1)**(number of occupied spin-orbitals below k)``.
"""
from __future__ import annotations

import numpy as np


def interleave(a: int, b: int, L: int) -> int:
    d = 0
    for i in range(L):
        if (a >> i) & 1:
            d |= 1 << (2 * i)
        if (b >> i) & 1:
            d |= 1 << (2 * i + 1)
    return d


def _ann(d, k):
    if not (d >> k) & 1:
        return None, 0
    s = -1 if bin(d & ((1 << k) - 1)).count("1") & 1 else 1
    return d & ~(1 << k), s


def _cre(d, k):
    if (d >> k) & 1:
        return None, 0
    s = -1 if bin(d & ((1 << k) - 1)).count("1") & 1 else 1
    return d | (1 << k), s


def apply_ops(d, ops):
    """ops: list of (kind, k) applied right-to-left like an operator product."""
    sign = 1
    for kind, k in reversed(ops):
        d, s = (_cre if kind == "c" else _ann)(d, k)
        if d is None:
            return None, 0
        sign *= s
    return d, sign


def hamiltonian(L, core, h1, eri, astr, bstr):
    """Dense H over the product basis, row-major (ia, ib) with beta fastest."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    n = len(dets)
    H = np.zeros((n, n))
    for j, ket in enumerate(dets):
        H[j, j] += core
        occ = [k for k in range(2 * L) if (ket >> k) & 1]
        for q in occ:
            for p in range(2 * L):
                if p % 2 != q % 2:
                    continue
                d, s = apply_ops(ket, [("c", p), ("a", q)])
                if d is not None and d in index:
                    H[index[d], j] += s * h1[p // 2, q // 2]
        for q in occ:
            for s_ in occ:
                for p in range(2 * L):
                    if p % 2 != q % 2:
                        continue
                    for r in range(2 * L):
                        if r % 2 != s_ % 2:
                            continue
                        d, s = apply_ops(ket, [("c", p), ("c", r), ("a", s_), ("a", q)])
                        if d is not None and d in index:
                            H[index[d], j] += 0.5 * s * eri[p // 2, q // 2, r // 2, s_ // 2]
    return H


def rdms(L, astr, bstr, W):
    """gamma[s][i + L j] = <c+_{i s} c_{j s}>,
    Gamma[s + 2t][i + L j + L^2 k + L^3 l] = <c+_{i s} c+_{j t} c_{l t} c_{k s}>."""
    dets = [interleave(int(a), int(b), L) for a in astr for b in bstr]
    index = {d: i for i, d in enumerate(dets)}
    one = np.zeros((2, L * L))
    two = np.zeros((4, L ** 4))
    for jdx, ket in enumerate(dets):
        wk = W[jdx]
        if wk == 0.0:
            continue
        occ = [k for k in range(2 * L) if (ket >> k) & 1]
        for jj in occ:
            for ii in range(2 * L):
                if ii % 2 != jj % 2:
                    continue
                d, s = apply_ops(ket, [("c", ii), ("a", jj)])
                if d is not None and d in index:
                    one[ii % 2, ii // 2 + L * (jj // 2)] += W[index[d]] * s * wk
        for kk in occ:
            for ll in occ:
                for ii in range(2 * L):
                    if ii % 2 != kk % 2:
                        continue
                    for jj in range(2 * L):
                        if jj % 2 != ll % 2:
                            continue
                        d, s = apply_ops(ket, [("c", ii), ("c", jj), ("a", ll), ("a", kk)])
                        if d is not None and d in index:
                            blk = (ii % 2) + 2 * (jj % 2)
                            pos = ii // 2 + L * (jj // 2) + L * L * (kk // 2) + L ** 3 * (ll // 2)
                            two[blk, pos] += W[index[d]] * s * wk
    return one, two
