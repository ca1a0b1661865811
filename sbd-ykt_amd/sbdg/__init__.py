"""Python binding of libsbdg (the MI355X selected-basis diagonalization library).

``Solver`` mirrors the reference's ``sbd::tpb::diag`` (chemistry/tpb/sbdiag.h:113-477):
integrals and alpha/beta strings in, ``SBD``-style options, energy / vector / density /
RDMs out.  ``GroundStateEnergy`` exposes the energy as a ``torch.autograd.Function`` over
the one- and two-electron integrals whose backward is the Hellmann-Feynman gradient
contracted on the device from the RDMs of the converged vector (no Davidson re-run).

The library is loaded from this directory (``libsbdg.so``, built by ``make -C
sbd-ykt_amd``).  There is no fallback: a missing library or a missing GPU is an error.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libsbdg.so")

OK, ERR_ARG, ERR_STATE, ERR_HIP, ERR_NUMERIC, ERR_LIMIT = 0, -1, -2, -3, -4, -5


class SbdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sbdg error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    """sbdg_options (include/sbdg.h); the first eight fields are sbd::tpb::SBD's."""
    _fields_ = [
        ("max_it", C.c_int), ("max_nb", C.c_int), ("eps", C.c_double), ("max_time", C.c_double),
        ("init", C.c_int), ("do_rdm", C.c_int), ("ratio", C.c_double), ("threshold", C.c_double),
        ("restart_keep", C.c_int), ("precision", C.c_int), ("refresh_k", C.c_int),
        ("target_de", C.c_double), ("verbose", C.c_int), ("seed", C.c_uint64),
        ("doubles", C.c_int), ("core_gib", C.c_double), ("overlap", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("energy", C.c_double), ("ritz", C.c_double), ("resid", C.c_double),
        ("iterations", C.c_int), ("sigma_fp64", C.c_int), ("sigma_fp32", C.c_int), ("converged", C.c_int),
        ("t_tables", C.c_double), ("t_davidson", C.c_double), ("t_sigma", C.c_double),
        ("t_final", C.c_double), ("t_rdm", C.c_double), ("terms", C.c_int64), ("nnz", C.c_int64 * 4),
        ("core_members", C.c_int64 * 2), ("core_direct", C.c_int64 * 2),
        ("mem_now", C.c_int64), ("mem_peak", C.c_int64),
    ]


_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int32)
_U32 = C.POINTER(C.c_uint32)
_U64 = C.POINTER(C.c_uint64)

# every entry point of include/sbdg.h with its ctypes signature (restype, argtypes)
SIGNATURES = {
    "sbdg_create": (_P, [C.c_int]),
    "sbdg_save_wavefunction": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "sbdg_load_wavefunction": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "sbdg_save_matrix_form": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "sbdg_destroy": (None, [_P]),
    "sbdg_last_error": (C.c_char_p, [_P]),
    "sbdg_default_options": (C.c_int, [C.POINTER(Options)]),
    "sbdg_set_integrals": (C.c_int, [_P, C.c_int, C.c_double, _D, _D, C.c_int]),
    "sbdg_set_strings": (C.c_int, [_P, _U64, C.c_int64, _U64, C.c_int64]),
    "sbdg_set_options": (C.c_int, [_P, C.POINTER(Options)]),
    "sbdg_set_vector": (C.c_int, [_P, _D]),
    "sbdg_build": (C.c_int, [_P]),
    "sbdg_run": (C.c_int, [_P]),
    "sbdg_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "sbdg_get_energy": (C.c_int, [_P, _D, _D, C.POINTER(C.c_int)]),
    "sbdg_get_vector": (C.c_int, [_P, _D]),
    "sbdg_get_density": (C.c_int, [_P, _D]),
    "sbdg_get_rdm": (C.c_int, [_P, _D, _D]),
    "sbdg_compute_rdm": (C.c_int, [_P]),
    "sbdg_rdm_backward": (C.c_int, [_P, C.c_double, _D, _D]),
    "sbdg_carryover": (C.c_int, [_P, C.c_double, C.c_double, _U64, C.c_int64, _I64, _D]),
    "sbdg_get_trajectory": (C.c_int, [_P, _D, _D, C.c_int, C.POINTER(C.c_int)]),
    "sbdg_table_nnz": (C.c_int, [_P, C.c_int, _I64]),
    "sbdg_get_table": (C.c_int, [_P, C.c_int, _I64, _I32, _U32, _D]),
    "sbdg_get_hii": (C.c_int, [_P, _D]),
    "sbdg_sigma": (C.c_int, [_P, _D, _D, C.c_int]),
    "sbdg_sigma_device": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_int]),
    "sbdg_bench_sigma": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "sbdg_memory_plan": (C.c_int, [C.POINTER(Options), C.c_int, C.c_int64, C.c_int64, _I64, _I64, _I64, _I64]),
}

_lib = None


def lib():
    """Load libsbdg.so (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built: run `make -C sbd-ykt_amd` (hipcc, gfx950)")
        l = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def device_memory():
    """(free, total) device bytes of the current device, from the HIP runtime the library
    loaded (torch ships another runtime of the same soname: it must not be initialised
    after this library's)."""
    lib()
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    rc = hip.hipMemGetInfo(C.byref(free), C.byref(total))
    if rc != 0:
        raise SbdgError(ERR_HIP, f"hipMemGetInfo failed ({rc})")
    return free.value, total.value


def _d(a):
    return a.ctypes.data_as(_D)


def default_options(**kw) -> Options:
    o = Options()
    lib().sbdg_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class Solver:
    """One GPU context holding a problem (integrals + strings) and its solution."""

    def __init__(self, norb, core, h1, eri, astr, bstr, device=0, eri_layout=0, **options):
        self._c = lib().sbdg_create(int(device))
        if not self._c:
            raise SbdgError(ERR_HIP, f"no HIP device {device}")
        self.norb = int(norb)
        self.na, self.nb = len(astr), len(bstr)
        self.astr = np.ascontiguousarray(astr, dtype=np.uint64)
        self.bstr = np.ascontiguousarray(bstr, dtype=np.uint64)
        self.set_integrals(core, h1, eri, eri_layout)
        self._chk(lib().sbdg_set_strings(self._c, self.astr.ctypes.data_as(_U64), self.na,
                                         self.bstr.ctypes.data_as(_U64), self.nb))
        self.options = default_options(**options)
        self._chk(lib().sbdg_set_options(self._c, C.byref(self.options)))

    def set_integrals(self, core, h1, eri, eri_layout=0):
        h1 = np.ascontiguousarray(h1, dtype=np.float64)
        eri = np.ascontiguousarray(eri, dtype=np.float64)
        self._chk(lib().sbdg_set_integrals(self._c, self.norb, float(core), _d(h1), _d(eri), int(eri_layout)))
        if hasattr(self, "astr"):
            self._chk(lib().sbdg_set_strings(self._c, self.astr.ctypes.data_as(_U64), self.na,
                                             self.bstr.ctypes.data_as(_U64), self.nb))

    def close(self):
        if getattr(self, "_c", None):
            lib().sbdg_destroy(self._c)
            self._c = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != OK:
            raise SbdgError(rc, lib().sbdg_last_error(self._c).decode())

    @property
    def ndet(self):
        return self.na * self.nb

    def set_options(self, **kw):
        for k, v in kw.items():
            if not hasattr(self.options, k):
                raise TypeError(f"unknown option {k}")
            setattr(self.options, k, v)
        self._chk(lib().sbdg_set_options(self._c, C.byref(self.options)))

    def build(self):
        self._chk(lib().sbdg_build(self._c))

    def set_vector(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        self._chk(lib().sbdg_set_vector(self._c, _d(w)))

    def run(self):
        self._chk(lib().sbdg_run(self._c))
        return self.energy()

    def energy(self):
        e = C.c_double()
        r = C.c_double()
        it = C.c_int()
        self._chk(lib().sbdg_get_energy(self._c, C.byref(e), C.byref(r), C.byref(it)))
        return e.value

    def stats(self) -> dict:
        s = Stats()
        self._chk(lib().sbdg_get_stats(self._c, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in Stats._fields_ if k not in ("nnz", "core_members", "core_direct")}
        d["nnz"] = list(s.nnz)
        d["core_members"] = list(s.core_members)
        d["core_direct"] = list(s.core_direct)
        return d

    def trajectory(self):
        n = C.c_int()
        self._chk(lib().sbdg_get_trajectory(self._c, None, None, 0, C.byref(n)))
        t = np.zeros(max(n.value, 1))
        e = np.zeros(max(n.value, 1))
        self._chk(lib().sbdg_get_trajectory(self._c, _d(t), _d(e), n.value, C.byref(n)))
        return t[: n.value], e[: n.value]

    def vector(self):
        w = np.zeros(self.ndet)
        self._chk(lib().sbdg_get_vector(self._c, _d(w)))
        return w

    def density(self):
        d = np.zeros(2 * self.norb)
        self._chk(lib().sbdg_get_density(self._c, _d(d)))
        return d

    def rdm(self, compute=False):
        if compute:
            self._chk(lib().sbdg_compute_rdm(self._c))
        L = self.norb
        one = np.zeros(2 * L * L)
        two = np.zeros(4 * L ** 4)
        self._chk(lib().sbdg_get_rdm(self._c, _d(one), _d(two)))
        return one.reshape(2, L * L), two.reshape(4, L ** 4)

    def rdm_backward(self, grad_e=1.0):
        L = self.norb
        dh = np.zeros((L, L))
        de = np.zeros((L, L, L, L))
        self._chk(lib().sbdg_rdm_backward(self._c, float(grad_e), _d(dh), _d(de)))
        return dh, de

    def carryover(self, ratio=None, threshold=None):
        ratio = self.options.ratio if ratio is None else ratio
        threshold = self.options.threshold if threshold is None else threshold
        out = np.zeros(self.na, dtype=np.uint64)
        cnt = C.c_int64()
        tw = C.c_double()
        self._chk(lib().sbdg_carryover(self._c, float(ratio), float(threshold), out.ctypes.data_as(_U64),
                                       self.na, C.byref(cnt), C.byref(tw)))
        return out[: min(cnt.value, self.na)], tw.value

    def table(self, which):
        """which: 0 alpha singles, 1 alpha doubles, 2 beta singles, 3 beta doubles."""
        nnz = C.c_int64()
        rc = lib().sbdg_table_nnz(self._c, which, C.byref(nnz))
        if rc == ERR_STATE:
            self.build()
            rc = lib().sbdg_table_nnz(self._c, which, C.byref(nnz))
        self._chk(rc)
        n = self.na if which < 2 else self.nb
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        exc = np.zeros(max(nnz.value, 1), dtype=np.uint32)
        val = np.zeros(max(nnz.value, 1))
        self._chk(lib().sbdg_get_table(self._c, which, rp.ctypes.data_as(_I64), col.ctypes.data_as(_I32),
                                       exc.ctypes.data_as(_U32), _d(val)))
        k = nnz.value
        return rp, col[:k], exc[:k], val[:k]

    def hii(self):
        out = np.zeros(self.ndet)
        self._chk(lib().sbdg_get_hii(self._c, _d(out)))
        return out

    def sigma(self, c, precision=0):
        c = np.ascontiguousarray(c, dtype=np.float64)
        out = np.zeros(self.ndet)
        self._chk(lib().sbdg_sigma(self._c, _d(c), _d(out), int(precision)))
        return out

    def sigma_device(self, x_ptr, y_ptr, precision=0):
        """sigma on caller-owned device buffers (addresses of na*nb doubles), allocating
        nothing: sbdg_sigma_device."""
        self._chk(lib().sbdg_sigma_device(self._c, C.c_void_p(int(x_ptr)), C.c_void_p(int(y_ptr)),
                                          int(precision)))

    def save_wavefunction(self, prefix, bit_length=20):
        """diag's savename: writes <prefix>000000 in the reference's restart layout."""
        self._chk(lib().sbdg_save_wavefunction(self._c, os.fsencode(prefix), int(bit_length)))

    def load_wavefunction(self, prefix, bit_length=20):
        """diag's loadname: <prefix>000000 remapped onto this basis becomes the initial vector."""
        self._chk(lib().sbdg_load_wavefunction(self._c, os.fsencode(prefix), int(bit_length)))
        self.options.init = 2

    def save_matrix_form(self, filename, bit_length=20):
        """dump_matrix_form_wf: .bin raw doubles, .txt/.dat one text line per determinant."""
        self._chk(lib().sbdg_save_matrix_form(self._c, os.fsencode(filename), int(bit_length)))

    def bench_sigma(self, precision=0, reps=3):
        ms = np.zeros(6)
        self._chk(lib().sbdg_bench_sigma(self._c, int(precision), int(reps), _d(ms)))
        return ms


def memory_plan(norb, na, nb, nnz, core_members, core_direct, **options):
    """Planned device bytes (no device needed): dict of the build's resident tables, the
    Davidson vectors for these options and the RDMs' transient peak (sbdg_memory_plan)."""
    o = default_options(**options)
    a = lambda v: (C.c_int64 * len(v))(*[int(x) for x in v])
    out = (C.c_int64 * 3)()
    rc = lib().sbdg_memory_plan(C.byref(o), int(norb), int(na), int(nb), a(nnz), a(core_members), a(core_direct),
                                out)
    if rc != OK:
        raise SbdgError(rc, "bad memory plan arguments")
    return dict(build=out[0], davidson=out[1], rdm=out[2], total=out[0] + out[1] + out[2])


@dataclass
class DiagResult:
    energy: float
    density: np.ndarray
    one_p_rdm: np.ndarray | None
    two_p_rdm: np.ndarray | None
    carryover: np.ndarray
    stats: dict


def diag(norb, core, h1, eri, adet, bdet, do_rdm=0, device=0, **options) -> DiagResult:
    """Function form of sbd::tpb::diag (sbdiag.h:113-124) without the communicator."""
    s = Solver(norb, core, h1, eri, adet, bdet, device=device, do_rdm=do_rdm, **options)
    try:
        e = s.run()
        one = two = None
        if do_rdm:
            one, two = s.rdm()
            dens = np.empty(2 * norb)
            dens[0::2] = one[0].reshape(norb, norb).diagonal()
            dens[1::2] = one[1].reshape(norb, norb).diagonal()
        else:
            dens = s.density()
        co, _ = s.carryover()
        return DiagResult(e, dens, one, two, co, s.stats())
    finally:
        s.close()


def _torch_function():
    import torch

    class GroundStateEnergy(torch.autograd.Function):
        """E(h1, eri) = lowest eigenvalue in the selected basis.

        forward: Davidson on the GPU (FP64), then the RDMs of the converged vector and their
                 contraction into dE/dh1[p,q] = gamma[p + L q] (spin-summed) and
                 dE/deri[p,q,r,s] = 1/2 sum_blk Gamma[blk][p + L r + L^2 q + L^3 s] on the device
                 (Hellmann-Feynman), kept with this node of the graph;
        backward: those derivatives times the incoming gradient (the Davidson is not re-run).
        The derivatives belong to this forward's solve: later solves on the same solver (a
        second energy in the graph, a line search) do not change them."""

        @staticmethod
        def forward(ctx, h1, eri, solver: Solver, core: float = 0.0):
            solver.set_integrals(core, h1.detach().cpu().double().numpy(), eri.detach().cpu().double().numpy())
            e = solver.run()
            solver.rdm(compute=True)
            dh, de = solver.rdm_backward(1.0)
            ctx.dh = torch.from_numpy(dh).to(h1.device)
            ctx.de = torch.from_numpy(de).to(h1.device)
            return torch.tensor(e, dtype=torch.float64, device=h1.device)

        @staticmethod
        def backward(ctx, g):
            return g * ctx.dh, g * ctx.de, None, None

    return GroundStateEnergy


def ground_state_energy(h1, eri, solver: Solver, core: float = 0.0):
    return _torch_function().apply(h1, eri, solver, core)
