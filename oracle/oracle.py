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
        ("src_impl", C.POINTER(C.c_int)),
        ("nuv", C.c_int), ("uv_xy", C.POINTER(C.c_double)), ("uv_interp", C.POINTER(C.c_int)),
        ("rec_dir", C.c_char_p),
        ("src_h", C.POINTER(C.c_double)),
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

    def __init__(self, sc, dt=None, nthreads=None, nt=None, uv_xy=None, uv_interp=None, rec_dir=None):
        """uv_xy / uv_interp: RecordUV (0) or RecordUV_interp (1) stations;
        rec_dir: append each step's receiver row to the receiver file in that
        directory during run() (the reference's per-step file I/O)."""
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
        impl = getattr(sc, "src_impl", None)
        s.src_impl = _ip(keep(np.asarray(impl), np.int32)) if (impl is not None and ns) else None
        nuv = 0 if uv_xy is None else len(uv_xy)
        s.nuv = nuv
        s.uv_xy = _dp(keep(np.asarray(uv_xy).reshape(-1))) if nuv else None
        s.uv_interp = _ip(keep(np.asarray(uv_interp if uv_interp is not None else [0] * nuv), np.int32)) if nuv else None
        self._rec_dir = None if rec_dir is None else str(rec_dir).encode()
        s.rec_dir = self._rec_dir
        h = getattr(sc, "src_h", None)
        s.src_h = _dp(keep(np.asarray(h, dtype=np.float64))) if (h is not None and ns) else None
        self.nuv = nuv
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

    def assembled_source(self, s):
        """(F[2 nnodes], kernel width h, \\int W) of source s: the assembled
        nodal vector (h and \\int W are those of a cubic-spline source, else 0)."""
        F = np.zeros(2 * self.nnodes)
        h, iw = C.c_double(), C.c_double()
        _chk(lib().orc_source_assembled(C.byref(self.s), s, _dp(F), C.byref(h), C.byref(iw)))
        return F, h.value, iw.value

    def receivers(self):
        nr = len(self.sc.rec_xy)
        nid, eid, g = (np.zeros(max(nr, 1), dtype=np.int32) for _ in range(3))
        _chk(lib().orc_receiver_nodes(C.byref(self.s), _ip(nid), _ip(eid), _ip(g)))
        return nid[:nr], eid[:nr], g[:nr]

    def run(self, max_steps=0, u0=None, v0=None, a0=None, want_state=True, want_uv=False, want_trace=True):
        """Returns (times, trace[k][r][7], u, v, a), and uv_trace[k][r][4] of
        the single-station receivers as a sixth item with want_uv."""
        nt = self.s.nt if max_steps <= 0 else min(max_steps, self.s.nt)
        nr = len(self.sc.rec_xy)
        trace = np.zeros(max(nt * nr * 7, 1)) if want_trace else None
        times = np.zeros(max(nt, 1))
        uvt = np.zeros(max(nt * self.nuv * 4, 1)) if want_uv else None
        done = C.c_int(0)
        n2 = 2 * self.nnodes
        u, v, a = (np.zeros(n2) for _ in range(3)) if want_state else (None, None, None)
        init = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (u0, v0, a0)]
        _chk(lib().orc_run_ex(C.byref(self.s), max_steps, *[None if x is None else _dp(x) for x in init],
                              None if trace is None else _dp(trace), _dp(times), C.byref(done),
                              None if u is None else _dp(u), None if v is None else _dp(v),
                              None if a is None else _dp(a), None if uvt is None else _dp(uvt)))
        k = done.value
        out = (times[:k], None if trace is None else trace[: k * nr * 7].reshape(k, nr, 7), u, v, a)
        if want_uv:
            out = out + (uvt[: k * self.nuv * 4].reshape(k, self.nuv, 4),)
        return out

    def receiver_rows(self, u, v, a):
        """Receiver rows [nr][7] (ux uy vx vy ax ay curl) for the given state."""
        nr = len(self.sc.rec_xy)
        out = np.zeros(max(nr * 7, 1))
        arr = [np.ascontiguousarray(x, dtype=np.float64) for x in (u, v, a)]
        _chk(lib().orc_receiver_rows(C.byref(self.s), *[_dp(x) for x in arr], _dp(out)))
        return out[: nr * 7].reshape(nr, 7)

    def uv_rows(self, u=None, v=None):
        """Single-station rows [nuv][4] for the given state (None: only the
        stations), with the stations' elements and nodal weights."""
        out = np.zeros(max(self.nuv * 4, 1))
        w = np.zeros(max(self.nuv * self.npe, 1))
        el = np.zeros(max(self.nuv, 1), dtype=np.int32)
        arr = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (u, v)]
        _chk(lib().orc_uv_rows(C.byref(self.s), *[None if x is None else _dp(x) for x in arr], _dp(out), _dp(w),
                               _ip(el)))
        return (out[: self.nuv * 4].reshape(self.nuv, 4), el[: self.nuv], w[: self.nuv * self.npe].reshape(self.nuv, self.npe))

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
