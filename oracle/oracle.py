"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, never by the product package.  The oracle is
a restatement of the reference's algorithm (se2wave.c of
Nicolucas/se2dr_fork); see oracle/ref_cpu.c for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

STF_TYPES = {"ricker": 0, "gaussian": 1, "yoffe": 2}


class OrcScenario(C.Structure):
    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("p", C.c_int),
        ("scale", C.c_double * 2), ("shift", C.c_double * 2),
        ("lam", C.POINTER(C.c_double)), ("mu", C.POINTER(C.c_double)), ("rho", C.POINTER(C.c_double)),
        ("nsrc", C.c_int),
        ("src_xy", C.POINTER(C.c_double)), ("src_type", C.POINTER(C.c_int)), ("src_values", C.POINTER(C.c_double)),
        ("stf_type", C.POINTER(C.c_int)), ("stf_params", C.POINTER(C.c_double)),
        ("use_nearest_internal_gll", C.c_int),
        ("nrec", C.c_int), ("rec_xy", C.POINTER(C.c_double)),
        ("dt", C.c_double), ("nt", C.c_int), ("tmax", C.c_double), ("nthreads", C.c_int),
    ]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run make -C oracle")
        _lib = C.CDLL(path)
        _lib.orc_stf_eval.restype = C.c_double
        _lib.orc_stf_eval.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_double]
        _lib.orc_last_error.restype = C.c_char_p
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


class Problem:
    """Holds the ctypes scenario struct and keeps its arrays alive."""

    def __init__(self, sc, dt=None, nthreads=None, nt=None):
        mu = np.asarray(sc.vs) * np.asarray(sc.vs) * np.asarray(sc.rho)  # se2wave.c:1166-1167
        lam = np.asarray(sc.vp) * np.asarray(sc.vp) * np.asarray(sc.rho) - 2.0 * mu
        self._keep = []

        def keep(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a

        ns = len(sc.src_type)
        s = OrcScenario()
        s.mx, s.my, s.p = sc.mx, sc.my, sc.p
        s.scale[0], s.scale[1] = sc.scale
        s.shift[0], s.shift[1] = sc.shift
        s.lam, s.mu, s.rho = _dp(keep(lam)), _dp(keep(mu)), _dp(keep(sc.rho))
        s.nsrc = ns
        s.src_xy = _dp(keep(np.asarray(sc.src_xy).reshape(-1) if ns else np.zeros(2)))
        s.src_type = _ip(keep(np.asarray(sc.src_type if ns else [0]), np.int32))
        s.src_values = _dp(keep(np.asarray(sc.src_values).reshape(-1) if ns else np.zeros(4)))
        s.stf_type = _ip(keep(np.array([STF_TYPES[t[0]] for t in sc.stf] if ns else [0]), np.int32))
        prm = np.zeros((max(ns, 1), 4))
        for i, t in enumerate(sc.stf):
            prm[i, :3] = t[1:4]
        s.stf_params = _dp(keep(prm.reshape(-1)))
        s.use_nearest_internal_gll = int(bool(sc.use_nearest_internal_gll))
        s.nrec = len(sc.rec_xy)
        s.rec_xy = _dp(keep(np.asarray(sc.rec_xy).reshape(-1) if len(sc.rec_xy) else np.zeros(2)))
        s.dt = float(dt) if dt is not None else 0.0
        s.nt = int(nt if nt is not None else sc.nt)
        s.tmax = float(sc.tmax)
        s.nthreads = int(nthreads or os.cpu_count() or 1)
        self.s = s
        self.sc = sc
        self.nnodes = sc.nnodes
        self.npe = (sc.p + 1) ** 2

    def timestep(self) -> float:
        dt = C.c_double()
        _chk(lib().orc_compute_timestep(C.byref(self.s), C.byref(dt)))
        return dt.value

    def coords(self):
        c = np.zeros(2 * self.nnodes)
        gmin, gmax = np.zeros(2), np.zeros(2)
        _chk(lib().orc_mesh_coords(C.byref(self.s), _dp(c), _dp(gmin), _dp(gmax)))
        return c, gmin, gmax

    def mass(self):
        m = np.zeros(2 * self.nnodes)
        _chk(lib().orc_mass_vector(C.byref(self.s), _dp(m)))
        return m

    def assemble(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        f = np.zeros_like(u)
        _chk(lib().orc_assemble_forces(C.byref(self.s), _dp(u), _dp(f)))
        return f

    def source(self, s):
        e, g = C.c_int(), C.c_int()
        vals = np.zeros(2 * self.npe)
        _chk(lib().orc_source_values(C.byref(self.s), s, C.byref(e), C.byref(g), _dp(vals)))
        return e.value, g.value, vals

    def receivers(self):
        nr = len(self.sc.rec_xy)
        nid, eid, g = (np.zeros(max(nr, 1), dtype=np.int32) for _ in range(3))
        _chk(lib().orc_receiver_nodes(C.byref(self.s), _ip(nid), _ip(eid), _ip(g)))
        return nid[:nr], eid[:nr], g[:nr]

    def run(self, max_steps=0, u0=None, v0=None, a0=None, want_state=True):
        """Returns (times, trace[k][r][7], u, v, a)."""
        nt = self.s.nt if max_steps <= 0 else min(max_steps, self.s.nt)
        nr = len(self.sc.rec_xy)
        trace = np.zeros(max(nt * nr * 7, 1))
        times = np.zeros(max(nt, 1))
        done = C.c_int(0)
        n2 = 2 * self.nnodes
        u, v, a = (np.zeros(n2) for _ in range(3)) if want_state else (None, None, None)
        init = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (u0, v0, a0)]
        _chk(lib().orc_run(C.byref(self.s), max_steps, *[None if x is None else _dp(x) for x in init], _dp(trace),
                           _dp(times), C.byref(done), None if u is None else _dp(u), None if v is None else _dp(v),
                           None if a is None else _dp(a)))
        k = done.value
        return times[:k], trace[: k * nr * 7].reshape(k, nr, 7), u, v, a

    def write_receiver_files(self, directory, times, trace):
        t = np.ascontiguousarray(times, dtype=np.float64)
        tr = np.ascontiguousarray(trace, dtype=np.float64).reshape(-1)
        _chk(lib().orc_write_receiver_files(C.byref(self.s), directory.encode(), len(t), _dp(t), _dp(tr)))


def gll(p):
    xi, w = np.zeros(p + 1), np.zeros(p + 1)
    _chk(lib().orc_gll(p, _dp(xi), _dp(w)))
    return xi, w


def deriv_matrix(p):
    D = np.zeros((p + 1) ** 2)
    _chk(lib().orc_deriv_matrix(p, _dp(D)))
    return D.reshape(p + 1, p + 1)


def basis1d(p, x, derivative=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) * (p + 1))
    fn = lib().orc_deriv1d if derivative else lib().orc_basis1d
    _chk(fn(p, len(x), _dp(x), _dp(out)))
    return out.reshape(len(x), p + 1)


def stf(kind, params, t):
    prm = np.zeros(4)
    prm[: len(params)] = params
    return lib().orc_stf_eval(STF_TYPES[kind], _dp(prm), float(t))
