"""FCIDUMP and alpha-string file I/O with the reference's semantics.

* ``load_fcidump`` follows ``sbd::LoadFCIDump`` (framework/fcidump.h:27-139): lines that
  are empty or start with '!' are skipped; header lines are concatenated (with "&FCI"
  removed) until a line containing "&END"; the header is split on ',' into KEY=VALUE
  pairs (whitespace trimmed); every later line is ``value i j k l``.
* ``integrals_from_fcidump`` follows ``sbd::SetupIntegrals`` (makeintegrals.h:10-55):
  "0 0 0 0" is the core energy, "i j 0 0" sets h_ij and h_ji, anything else sets the
  8-fold-equivalent class of (ij|kl); later lines overwrite earlier ones.  Without a
  "0 0 0 0" line the reference leaves the core energy uninitialised; here it is 0.
* ``round_like_broadcast`` reproduces the six-significant-digit rounding that the
  reference's MPI broadcast applies to every integral (fcidump.h:156-169, 226-253) --
  only for comparisons against energies published from that path.
* ``load_alpha_strings`` follows ``sbd::LoadAlphaDets`` for text files
  (makedeterminants.h:160-185, bit_manipulation.h:158-172): one string per line, the
  rightmost character is orbital 0, then sort ascending and drop duplicates.
"""
from __future__ import annotations

import numpy as np


def load_fcidump(path):
    header = {}
    values, idx = [], []
    in_header = True
    head = ""
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line or line[0] == "!":
                continue
            if in_header:
                if "&END" in line:
                    in_header = False
                    continue
                pos = line.find("&FCI")
                if pos != -1:
                    line = line[:pos] + line[pos + 4:]
                head += line
            else:
                t = line.split()
                if len(t) < 5:
                    continue
                values.append(float(t[0]))
                idx.append((int(t[1]), int(t[2]), int(t[3]), int(t[4])))
    for tok in head.split(","):
        if "=" in tok:
            k, v = tok.split("=", 1)
            header[k.strip(" \t")] = v.strip(" \t")
    return header, np.array(values, dtype=np.float64), np.array(idx, dtype=np.int32).reshape(-1, 4)


def integrals_from_fcidump(header, values, idx):
    """Returns (norb, core, h1 (L,L), eri (L,L,L,L)) with exact 8-fold symmetry."""
    L = int(header["NORB"])
    core = 0.0
    h1 = np.zeros((L, L))
    eri = np.zeros((L, L, L, L))
    for v, (i, j, k, l) in zip(values, idx):
        if i == 0 and j == 0 and k == 0 and l == 0:
            core = float(v)
        elif k == 0 and l == 0:
            h1[i - 1, j - 1] = v
            h1[j - 1, i - 1] = v
        else:
            p, q, r, s = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                eri[a, b, c, d] = v
    return L, core, h1, eri


def round_like_broadcast(values):
    """Six significant digits, as a default-precision ostringstream writes them."""
    return np.array([float(f"{v:.6g}") for v in values])


def write_fcidump(path, norb, nelec, h1, eri, core=0.0, ms2=0, tol=0.0):
    """Full-precision FCIDUMP (17 significant digits), canonical 8-fold entries, with the
    "0 0 0 0" core line always written."""
    L = norb
    with open(path, "w") as f:
        f.write(f" &FCI NORB={L},NELEC={nelec},MS2={ms2},\n")
        f.write("  ORBSYM=" + "1," * L + "\n  ISYM=1,\n &END\n")
        for p in range(L):
            for q in range(p + 1):
                for r in range(L):
                    for s in range(r + 1):
                        if p * (p + 1) // 2 + q < r * (r + 1) // 2 + s:
                            continue
                        v = eri[p, q, r, s]
                        if abs(v) > tol or tol == 0.0:
                            f.write(f"{v:.17e} {p + 1} {q + 1} {r + 1} {s + 1}\n")
        for p in range(L):
            for q in range(p + 1):
                f.write(f"{h1[p, q]:.17e} {p + 1} {q + 1} 0 0\n")
        f.write(f"{core:.17e} 0 0 0 0\n")


def string_to_bits(s: str) -> int:
    L = len(s)
    x = 0
    for i in range(L):
        if s[L - 1 - i] == "1":
            x |= 1 << i
    return x


def bits_to_string(x: int, norb: int) -> str:
    return "".join("1" if (int(x) >> (norb - 1 - i)) & 1 else "0" for i in range(norb))


def load_alpha_strings(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line:
                out.append(string_to_bits(line))
    return np.array(sorted(set(out)), dtype=np.uint64)


def write_strings(path, strings, norb):
    with open(path, "w") as f:
        for x in strings:
            f.write(bits_to_string(int(x), norb) + "\n")
