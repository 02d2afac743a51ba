"""se2dr_fork_amd -- MI355X-native explicit spectral-element time step of se2wave.

Python mirror of the host C API (include/se2wave.h) and the GPU C ABI
(include/se2w.h), loaded from the in-tree libraries lib/libse2w_host.so and
lib/libse2w_hip.so.  The objects follow the reference's (se2wave.c of
Nicolucas/se2dr_fork): SpecFECtx -> SpecFE, SeismicSTF -> STF,
SeismicSource -> Source, RecordUVA_MultipleStations_NearestGLL -> Receivers,
the se2wave_demo time loop -> Solver.  There is no fallback: if the native
libraries are missing, importing the bindings raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
BIN_DIR = os.path.join(_HERE, "bin")

FP64, FP32, FP32_KAHAN = 0, 1, 2
KERNEL_LDS, KERNEL_REG, KERNEL_MFMA, KERNEL_LINE, KERNEL_LDS_WIDE, KERNEL_REG64 = 0, 1, 9, 10, 11, 12
SCHED_FUSED, SCHED_SPLIT = 0, 1
SOURCE_FORCE, SOURCE_MOMENT = 0, 1
IMPL_POINTWISE, IMPL_NEAREST_QPOINT, IMPL_SPLINE, IMPL_DEFAULT, IMPL_P0 = 0, 1, 2, 3, 4
PRECISIONS = {"fp64": FP64, "fp32": FP32, "fp32k": FP32_KAHAN}


class Se2wError(RuntimeError):
    pass


_hip: Optional[C.CDLL] = None
_host: Optional[C.CDLL] = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _ptr(a: np.ndarray, kind=_dp):
    return a.ctypes.data_as(kind)


def libraries():
    """Load (once) and return (libse2w_hip, libse2w_host).  Raises if missing."""
    global _hip, _host
    if _hip is None:
        hip_path = os.path.join(LIB_DIR, "libse2w_hip.so")
        host_path = os.path.join(LIB_DIR, "libse2w_host.so")
        for p in (hip_path, host_path):
            if not os.path.exists(p):
                raise Se2wError(f"native library {p} is missing: run __graft_entry__.build() or make -C se2dr_fork_amd")
        _hip = C.CDLL(hip_path, mode=C.RTLD_GLOBAL)
        _host = C.CDLL(host_path, mode=C.RTLD_GLOBAL)
        _declare(_hip, _host)
    return _hip, _host


def _declare(hip, host):
    v = C.c_void_p
    i32 = C.c_int32
    d = C.c_double
    hip.se2w_strerror.restype = C.c_char_p
    hip.se2w_last_error.restype = C.c_char_p
    hip.se2w_create.argtypes = [v, C.POINTER(v)]
    hip.se2w_set_tables.argtypes = [v, _dp, _dp, _dp, _dp]
    hip.se2w_set_cells.argtypes = [v, _dp, _dp]
    hip.se2w_set_mass.argtypes = [v, _dp]
    hip.se2w_set_sources.argtypes = [v, i32, _ip, _dp]
    hip.se2w_set_receivers.argtypes = [v, i32, _ip, _ip]
    hip.se2w_set_state.argtypes = [v, _dp, _dp, _dp]
    hip.se2w_get_state.argtypes = [v, _dp, _dp, _dp]
    hip.se2w_snapshot_begin.argtypes = [v]
    hip.se2w_snapshot_wait.argtypes = [v, C.POINTER(_dp), C.POINTER(_dp)]
    hip.se2w_run.argtypes = [v, i32, d, _dp]
    hip.se2w_reserve.argtypes = [v, i32]
    hip.se2w_get_reserved.argtypes = [v, C.POINTER(C.c_int64)]
    hip.se2w_get_counts.argtypes = [v, _ip]
    hip.se2w_get_receivers.argtypes = [v, _dp]
    hip.se2w_assemble_forces.argtypes = [v, _dp, _dp]
    hip.se2w_assemble_forces_noalloc.argtypes = [v, _dp, _dp]
    hip.se2w_norms.argtypes = [v, _dp]
    hip.se2w_set_profiling.argtypes = [v, i32]
    hip.se2w_get_timing.argtypes = [v, _dp, _dp]
    hip.se2w_get_tile.argtypes = [v, _ip, _ip]
    hip.se2w_set_kernel.argtypes = [v, i32]
    hip.se2w_get_kernel.argtypes = [v, _ip]
    hip.se2w_autotune.argtypes = [v, i32, i32, _ip, _dp]
    hip.se2w_destroy.argtypes = [v]
    hip.se2w_ensemble_create.argtypes = [i32, C.POINTER(v), C.POINTER(v)]
    hip.se2w_ensemble_run.argtypes = [v, i32, _dp, C.POINTER(_dp)]
    hip.se2w_ensemble_set_kernel.argtypes = [v, i32]
    hip.se2w_ensemble_get_kernel.argtypes = [v, _ip]
    hip.se2w_ensemble_autotune.argtypes = [v, i32, i32, _ip, _dp]
    hip.se2w_ensemble_set_graph.argtypes = [v, i32]
    hip.se2w_ensemble_get_timing.argtypes = [v, _dp, _dp]
    hip.se2w_ensemble_get_launch.argtypes = [v, _ip, _ip]
    hip.se2w_ensemble_get_member.argtypes = [v, i32, C.POINTER(v)]
    hip.se2w_ensemble_destroy.argtypes = [v]
    fp = C.POINTER(C.c_float)
    hip.se2w_mlp_create.argtypes = [i32, _ip, i32, C.c_uint64, i32, C.POINTER(v)]
    hip.se2w_mlp_param_count.argtypes = [v, C.POINTER(C.c_int64)]
    hip.se2w_mlp_get_params.argtypes = [v, fp]
    hip.se2w_mlp_set_params.argtypes = [v, fp]
    hip.se2w_mlp_forward.argtypes = [v, i32, fp, fp]
    hip.se2w_mlp_gradient.argtypes = [v, i32, fp, fp, fp, fp]
    hip.se2w_mlp_set_data.argtypes = [v, i32, fp, fp]
    hip.se2w_mlp_train.argtypes = [v, i32, C.c_float, C.c_uint64, fp]
    hip.se2w_mlp_adam_step.argtypes = [v, fp, C.c_float]
    hip.se2w_mlp_set_weight_decay.argtypes = [v, C.c_float]
    hip.se2w_mlp_set_lr_schedule.argtypes = [v, i32]
    hip.se2w_mlp_get_timing.argtypes = [v, _dp, _dp]
    hip.se2w_mlp_destroy.argtypes = [v]
    host.se2w_host_last_error.restype = C.c_char_p
    host.se2w_gll.argtypes = [C.c_int, _dp, _dp]
    host.se2w_basis1d.argtypes = [C.c_int, C.c_int, _dp, _dp]
    host.se2w_basis1d_deriv.argtypes = [C.c_int, C.c_int, _dp, _dp]
    host.se2w_deriv_matrix.argtypes = [C.c_int, _dp]
    host.se2w_spec_create.argtypes = [C.c_int, C.c_int, C.c_int, C.POINTER(v)]
    host.se2w_spec_destroy.argtypes = [C.POINTER(v)]
    host.se2w_spec_scale_coords.argtypes = [v, _dp, _dp]
    host.se2w_spec_set_constant_material.argtypes = [v, d, d, d]
    host.se2w_spec_set_constant_material_velocity.argtypes = [v, d, d, d]
    host.se2w_spec_set_cell_material.argtypes = [v, _dp, _dp, _dp]
    host.se2w_spec_set_cell_material_velocity.argtypes = [v, _dp, _dp, _dp]
    host.se2w_spec_set_perturbed_material_velocity.argtypes = [v, d, d, d, d, d, d]
    host.se2w_rand48_seed.argtypes = [C.POINTER(C.c_uint64), C.c_long]
    host.se2w_rand48_next.argtypes = [C.POINTER(C.c_uint64)]
    host.se2w_rand48_next.restype = d
    host.se2w_spec_get_info.argtypes = [v, v]
    host.se2w_spec_get_coords.argtypes = [v, _dp]
    host.se2w_spec_get_material.argtypes = [v, _dp, _dp, _dp]
    host.se2w_spec_get_tables.argtypes = [v, _dp, _dp, _dp, _dp]
    host.se2w_spec_get_element.argtypes = [v, _ip]
    host.se2w_spec_mass.argtypes = [v, _dp]
    host.se2w_spec_timestep.argtypes = [v, _dp]
    host.se2w_stf_create_ricker.argtypes = [d, d, d, C.POINTER(v)]
    host.se2w_stf_create_gaussian.argtypes = [d, d, d, C.POINTER(v)]
    host.se2w_stf_create_yoffe.argtypes = [d, d, d, C.POINTER(v)]
    host.se2w_stf_evaluate.argtypes = [v, d, _dp]
    host.se2w_stf_destroy.argtypes = [C.POINTER(v)]
    host.se2w_source_create.argtypes = [v, C.c_int, C.c_int, _dp, _dp, C.c_int, C.POINTER(v)]
    host.se2w_source_setup.argtypes = [v]
    host.se2w_source_destroy.argtypes = [C.POINTER(v)]
    host.se2w_source_get.argtypes = [v, _ip, _ip, _dp]
    host.se2w_source_set_kernel_width.argtypes = [v, d]
    host.se2w_source_get_kernel.argtypes = [v, _dp, _dp]
    host.se2w_source_get_elements.argtypes = [v, _ip, _ip, _dp]
    host.se2w_sources_evaluate.argtypes = [d, C.c_int, C.POINTER(v), C.POINTER(v), _dp]
    host.se2w_receivers_create.argtypes = [v, C.c_int, _dp, C.POINTER(v)]
    host.se2w_receivers_destroy.argtypes = [C.POINTER(v)]
    host.se2w_receivers_get.argtypes = [v, _ip, _ip, _ip]
    host.se2w_receivers_write_header.argtypes = [v, C.c_char_p]
    host.se2w_receivers_append_rows.argtypes = [v, C.c_int, _dp, _dp]
    host.se2w_rng_uniform.argtypes = [C.POINTER(C.c_uint64)]
    host.se2w_rng_uniform.restype = C.c_double
    host.se2w_material_random_cells.argtypes = [C.c_int, C.c_int, C.c_uint64, d, d, d, d, _dp, _dp, _dp]
    host.se2w_material_layered.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, d, d, d, d, _dp, _dp, _dp]
    host.se2w_solver_create.argtypes = [v, C.c_int, C.c_int, C.c_int, C.POINTER(v), C.POINTER(v), v, C.POINTER(v)]
    host.se2w_solver_destroy.argtypes = [C.POINTER(v)]
    host.se2w_solver_context.argtypes = [v]
    host.se2w_solver_context.restype = v
    host.se2w_solver_advance.argtypes = [v, C.c_int, d]
    host.se2w_solver_reserve.argtypes = [v, C.c_int]
    host.se2w_solver_time.argtypes = [v]
    host.se2w_solver_time.restype = d
    host.se2w_solver_get_trace.argtypes = [v, _dp, _dp]
    host.se2w_solvers_advance.argtypes = [C.c_int, C.POINTER(v), v, C.c_int, _dp]
    host.se2w_receiver_uv_create.argtypes = [v, _dp, C.c_int, C.c_int, C.POINTER(v)]
    host.se2w_receiver_uv_destroy.argtypes = [C.POINTER(v)]
    host.se2w_receiver_uv_get.argtypes = [v, _ip, _dp, _dp]
    host.se2w_receiver_uv_write_header.argtypes = [v, C.c_char_p]
    host.se2w_receiver_uv_append_rows.argtypes = [v, C.c_int, _dp, _dp]
    host.se2w_solver_set_uv_receivers.argtypes = [v, C.c_int, C.POINTER(v)]
    host.se2w_solver_get_uv_trace.argtypes = [v, _dp]
    hip.se2w_set_element_receivers.argtypes = [v, i32, _ip]
    hip.se2w_get_element_receivers.argtypes = [v, _dp, _dp]
    host.se2w_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]


def _hchk(code):
    if code != 0:
        _, host = libraries()
        raise Se2wError(f"se2wave host error {code}: {host.se2w_host_last_error().decode()}")


def _gchk(code):
    if code != 0:
        hip, _ = libraries()
        raise Se2wError(f"se2w error {code} ({hip.se2w_strerror(code).decode()}): {hip.se2w_last_error().decode()}")


# ---------------------------------------------------------------------------
# basis (se2wave.c:103-690)
# ---------------------------------------------------------------------------
def gll(p: int):
    """CreateGLLCoordsWeights: ascending nodes, weights in the unflipped order."""
    _, host = libraries()
    xi = np.zeros(p + 1)
    w = np.zeros(p + 1)
    _hchk(host.se2w_gll(p, _ptr(xi), _ptr(w)))
    return xi, w


def deriv_matrix(p: int) -> np.ndarray:
    _, host = libraries()
    D = np.zeros((p + 1) * (p + 1))
    _hchk(host.se2w_deriv_matrix(p, _ptr(D)))
    return D.reshape(p + 1, p + 1)


def basis1d(p: int, x, derivative=False) -> np.ndarray:
    _, host = libraries()
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x) * (p + 1))
    fn = host.se2w_basis1d_deriv if derivative else host.se2w_basis1d
    _hchk(fn(p, len(x), _ptr(x), _ptr(out)))
    return out.reshape(len(x), p + 1)


class _SpecInfo(C.Structure):
    _fields_ = [("mx", C.c_int32), ("my", C.c_int32), ("p", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("npe", C.c_int32), ("nnodes", C.c_int64), ("gmin", C.c_double * 2), ("gmax", C.c_double * 2)]


class SpecFE:
    """SpecFECtx: mesh of mx*my elements of degree p (DMDA node order)."""

    def __init__(self, mx: int, my: int, p: int):
        _, self._host = libraries()
        self.h = C.c_void_p()
        _hchk(self._host.se2w_spec_create(mx, my, p, C.byref(self.h)))
        self.mx, self.my, self.p = mx, my, p
        self.nx, self.ny = p * mx + 1, p * my + 1
        self.npe = (p + 1) ** 2
        self.nnodes = self.nx * self.ny

    def close(self):
        if self.h:
            self._host.se2w_spec_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scale_coords(self, scale, shift):
        s = np.asarray(scale, dtype=np.float64)
        t = np.asarray(shift, dtype=np.float64)
        _hchk(self._host.se2w_spec_scale_coords(self.h, _ptr(s), _ptr(t)))

    def set_constant_material_velocity(self, vp, vs, rho):
        _hchk(self._host.se2w_spec_set_constant_material_velocity(self.h, vp, vs, rho))

    def set_constant_material(self, lam, mu, rho):
        _hchk(self._host.se2w_spec_set_constant_material(self.h, lam, mu, rho))

    def set_cell_material_velocity(self, vp, vs, rho):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (vp, vs, rho)]
        _hchk(self._host.se2w_spec_set_cell_material_velocity(self.h, *[_ptr(x) for x in a]))

    def set_perturbed_material_velocity(self, vp0, dvp, vs0, dvs, rho0, drho):
        """The reference's RAND48-perturbed per-cell model (se2wave.c:1177-1236)."""
        _hchk(self._host.se2w_spec_set_perturbed_material_velocity(self.h, vp0, dvp, vs0, dvs, rho0, drho))

    def set_cell_material(self, lam, mu, rho):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (lam, mu, rho)]
        _hchk(self._host.se2w_spec_set_cell_material(self.h, *[_ptr(x) for x in a]))

    def info(self):
        inf = _SpecInfo()
        _hchk(self._host.se2w_spec_get_info(self.h, C.byref(inf)))
        return inf

    def coords(self) -> np.ndarray:
        c = np.zeros(2 * self.nnodes)
        _hchk(self._host.se2w_spec_get_coords(self.h, _ptr(c)))
        return c

    def material(self):
        ne = self.mx * self.my
        lam, mu, rho = np.zeros(ne), np.zeros(ne), np.zeros(ne)
        _hchk(self._host.se2w_spec_get_material(self.h, _ptr(lam), _ptr(mu), _ptr(rho)))
        return lam, mu, rho

    def tables(self):
        nb = self.p + 1
        D, w1d, dx, dy = np.zeros(nb * nb), np.zeros(nb), np.zeros(self.mx), np.zeros(self.my)
        _hchk(self._host.se2w_spec_get_tables(self.h, _ptr(D), _ptr(w1d), _ptr(dx), _ptr(dy)))
        return D, w1d, dx, dy

    def element_map(self) -> np.ndarray:
        el = np.zeros(self.mx * self.my * self.npe, dtype=np.int32)
        _hchk(self._host.se2w_spec_get_element(self.h, _ptr(el, _ip)))
        return el.reshape(self.mx * self.my, self.npe)

    def mass(self) -> np.ndarray:
        m = np.zeros(self.nnodes)
        _hchk(self._host.se2w_spec_mass(self.h, _ptr(m)))
        return m

    def timestep(self) -> float:
        dt = C.c_double()
        _hchk(self._host.se2w_spec_timestep(self.h, C.byref(dt)))
        return dt.value


class STF:
    """SeismicSTF object (function-pointer based; Ricker, Gaussian, Yoffe)."""

    def __init__(self, kind: str, *params):
        _, self._host = libraries()
        self.h = C.c_void_p()
        fn = {"ricker": self._host.se2w_stf_create_ricker, "gaussian": self._host.se2w_stf_create_gaussian,
              "yoffe": self._host.se2w_stf_create_yoffe}[kind]
        _hchk(fn(*[float(p) for p in params], C.byref(self.h)))
        self.kind, self.params = kind, tuple(params)

    def __call__(self, t: float) -> float:
        v = C.c_double()
        _hchk(self._host.se2w_stf_evaluate(self.h, t, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self._host.se2w_stf_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Source:
    """SeismicSource: created (located) and set up (nodal values).  impl is
    IMPL_NEAREST_QPOINT, IMPL_POINTWISE, IMPL_SPLINE (the cubic-spline moment
    source, kernel_h its width as -sm_h sets it, None for the default) or
    IMPL_P0 (the P0 moment source)."""

    def __init__(self, spec: SpecFE, kind: int, coor, values, impl=IMPL_NEAREST_QPOINT, use_nearest_internal_gll=True,
                 kernel_h=None):
        _, self._host = libraries()
        self.spec = spec
        self.impl = impl
        self.h = C.c_void_p()
        c = np.asarray(coor, dtype=np.float64)
        v = np.zeros(4)
        v[: len(values)] = values
        _hchk(self._host.se2w_source_create(spec.h, kind, impl, _ptr(c), _ptr(v), int(bool(use_nearest_internal_gll)),
                                            C.byref(self.h)))
        if kernel_h is not None:
            _hchk(self._host.se2w_source_set_kernel_width(self.h, float(kernel_h)))
        _hchk(self._host.se2w_source_setup(self.h))

    def get(self):
        """(element, closest GLL node, 2*npe nodal values) of a one-element source."""
        e, g = C.c_int32(), C.c_int32()
        vals = np.zeros(2 * self.spec.npe)
        _hchk(self._host.se2w_source_get(self.h, C.byref(e), C.byref(g), _ptr(vals)))
        return e.value, g.value, vals

    def elements(self):
        """(elem[n], vals[n, 2*npe]): the elements the source injects into and
        their nodal values (one element unless it is a cubic-spline source)."""
        n = C.c_int32()
        _hchk(self._host.se2w_source_get_elements(self.h, C.byref(n), None, None))
        el = np.zeros(max(n.value, 1), dtype=np.int32)
        vals = np.zeros((max(n.value, 1), 2 * self.spec.npe))
        _hchk(self._host.se2w_source_get_elements(self.h, C.byref(n), _ptr(el, _ip), _ptr(vals)))
        return el[: n.value], vals[: n.value]

    def kernel(self):
        """(h, int W) of a cubic-spline source: its kernel width and the GLL
        quadrature of the kernel over the mesh."""
        h, w = C.c_double(), C.c_double()
        _hchk(self._host.se2w_source_get_kernel(self.h, C.byref(h), C.byref(w)))
        return h.value, w.value

    def close(self):
        if self.h:
            self._host.se2w_source_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Receivers:
    """Nearest-GLL receivers and their output files."""

    def __init__(self, spec: SpecFE, xr):
        _, self._host = libraries()
        self.spec = spec
        self.xr = np.ascontiguousarray(np.asarray(xr, dtype=np.float64).reshape(-1))
        self.nr = len(self.xr) // 2
        self.h = C.c_void_p()
        _hchk(self._host.se2w_receivers_create(spec.h, self.nr, _ptr(self.xr), C.byref(self.h)))

    def nodes(self):
        nid, el, g = (np.zeros(max(self.nr, 1), dtype=np.int32) for _ in range(3))
        _hchk(self._host.se2w_receivers_get(self.h, _ptr(nid, _ip), _ptr(el, _ip), _ptr(g, _ip)))
        return nid[: self.nr], el[: self.nr], g[: self.nr]

    def write_header(self, directory: str):
        _hchk(self._host.se2w_receivers_write_header(self.h, directory.encode()))

    def append_rows(self, times, trace):
        t = np.ascontiguousarray(times, dtype=np.float64)
        tr = np.ascontiguousarray(trace, dtype=np.float64)
        _hchk(self._host.se2w_receivers_append_rows(self.h, len(t), _ptr(t), _ptr(tr)))

    def close(self):
        if self.h:
            self._host.se2w_receivers_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ReceiverUV:
    """Single-station receiver: RecordUV (interpolate=False, nearest node of
    the containing element) or RecordUV_interp (interpolate=True, Lagrange
    interpolation), se2wave.c:2351-2581."""

    def __init__(self, spec: SpecFE, xr, interpolate=False, source_implementation=1):
        _, self._host = libraries()
        self.spec = spec
        xy = np.ascontiguousarray(xr, dtype=np.float64)
        self.h = C.c_void_p()
        _hchk(self._host.se2w_receiver_uv_create(spec.h, _ptr(xy), 1 if interpolate else 0, source_implementation,
                                                 C.byref(self.h)))

    def get(self):
        """(element, weights[npe], mapped coordinate)."""
        e = C.c_int32()
        w = np.zeros((self.spec.p + 1) ** 2)
        m = np.zeros(2)
        _hchk(self._host.se2w_receiver_uv_get(self.h, C.byref(e), _ptr(w), _ptr(m)))
        return e.value, w, m

    def write_header(self, directory: str):
        _hchk(self._host.se2w_receiver_uv_write_header(self.h, directory.encode()))

    def append_rows(self, times, rows):
        t = np.ascontiguousarray(times, dtype=np.float64)
        r = np.ascontiguousarray(rows, dtype=np.float64)
        _hchk(self._host.se2w_receiver_uv_append_rows(self.h, len(t), _ptr(t), _ptr(r)))

    def close(self):
        if self.h:
            self._host.se2w_receiver_uv_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Solver:
    """se2wave_demo's time loop on the GPU (host-evaluated STFs, accumulated time)."""

    def __init__(self, spec: SpecFE, sources: Sequence[Source], stfs: Sequence[STF], receivers: Optional[Receivers],
                 precision=FP64, schedule=SCHED_FUSED):
        _, self._host = libraries()
        self.spec, self.sources, self.stfs, self.receivers = spec, list(sources), list(stfs), receivers
        ns = len(self.sources)
        srcs = (C.c_void_p * max(ns, 1))(*[s.h for s in self.sources])
        stfs_ = (C.c_void_p * max(ns, 1))(*[s.h for s in self.stfs])
        self.h = C.c_void_p()
        _hchk(self._host.se2w_solver_create(spec.h, precision, schedule, ns, srcs, stfs_,
                                            receivers.h if receivers else None, C.byref(self.h)))
        self.ctx = Context(_handle=self._host.se2w_solver_context(self.h), spec=spec, precision=precision)
        self.nr = receivers.nr if receivers else 0

    def reserve(self, nsteps: int):
        """Sizes the buffers of advances of up to nsteps steps ahead of them,
        so that such an advance allocates nothing."""
        _hchk(self._host.se2w_solver_reserve(self.h, nsteps))

    def advance(self, nsteps: int, dt: float):
        _hchk(self._host.se2w_solver_advance(self.h, nsteps, dt))
        times = np.zeros(nsteps)
        trace = np.zeros(nsteps * max(self.nr, 1) * 7)
        _hchk(self._host.se2w_solver_get_trace(self.h, _ptr(times), _ptr(trace)))
        return times, trace[: nsteps * self.nr * 7].reshape(nsteps, self.nr, 7)

    def set_uv_receivers(self, receivers: Sequence["ReceiverUV"]):
        self.uv_receivers = list(receivers)
        arr = (C.c_void_p * max(len(self.uv_receivers), 1))(*[r.h for r in self.uv_receivers])
        _hchk(self._host.se2w_solver_set_uv_receivers(self.h, len(self.uv_receivers), arr))

    def uv_trace(self, nsteps: int) -> np.ndarray:
        """[nsteps][n][4] (ux uy vx vy) of the last advance of nsteps steps."""
        n = len(getattr(self, "uv_receivers", []))
        out = np.zeros(max(nsteps * n * 4, 1))
        _hchk(self._host.se2w_solver_get_uv_trace(self.h, _ptr(out)))
        return out[: nsteps * n * 4].reshape(nsteps, n, 4)

    @property
    def time(self) -> float:
        return self._host.se2w_solver_time(self.h)

    def close(self):
        if self.h:
            self.ctx._handle = None
            self._host.se2w_solver_destroy(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Desc(C.Structure):
    _fields_ = [("mx", C.c_int32), ("my", C.c_int32), ("p", C.c_int32), ("precision", C.c_int32),
                ("schedule", C.c_int32), ("device", C.c_int32)]


class Context:
    """Direct handle on the GPU C ABI (include/se2w.h)."""

    def __init__(self, mx=None, my=None, p=None, precision=FP64, schedule=SCHED_FUSED, device=0, _handle=None,
                 spec: Optional[SpecFE] = None):
        self._hip, _ = libraries()
        self._owned = _handle is None
        if spec is not None:
            mx, my, p = spec.mx, spec.my, spec.p
        self.mx, self.my, self.p, self.precision = mx, my, p, precision
        self.nx, self.ny = p * mx + 1, p * my + 1
        self.nnodes = self.nx * self.ny
        self.npe = (p + 1) ** 2
        if _handle is None:
            h = C.c_void_p()
            desc = _Desc(mx, my, p, precision, schedule, device)
            _gchk(self._hip.se2w_create(C.byref(desc), C.byref(h)))
            self._handle = h.value
        else:
            self._handle = _handle

    def close(self):
        if self._owned and self._handle:
            self._hip.se2w_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sized(self, name, a, n, dtype=np.float64):
        """a as a contiguous array of exactly n values (ValueError otherwise):
        the library reads or writes that many through the pointer."""
        a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        if a.size != n:
            raise ValueError(f"{name}: expected {n} values, got {a.size}")
        return a

    def counts(self):
        """(sources, receivers, element receivers, steps of the last run)."""
        c = np.zeros(4, dtype=np.int32)
        _gchk(self._hip.se2w_get_counts(self._handle, _ptr(c, _ip)))
        return tuple(int(x) for x in c)

    def set_tables(self, D, w1d, elem_dx, elem_dy):
        nb = self.p + 1
        a = [self._sized("D", D, nb * nb), self._sized("w1d", w1d, nb), self._sized("elem_dx", elem_dx, self.mx),
             self._sized("elem_dy", elem_dy, self.my)]
        _gchk(self._hip.se2w_set_tables(self._handle, *[_ptr(x) for x in a]))

    def set_cells(self, lam, mu):
        ne = self.mx * self.my
        a = [self._sized("lambda", lam, ne), self._sized("mu", mu, ne)]
        _gchk(self._hip.se2w_set_cells(self._handle, _ptr(a[0]), _ptr(a[1])))

    def set_mass(self, m):
        m = self._sized("mass", m, self.nnodes)
        _gchk(self._hip.se2w_set_mass(self._handle, _ptr(m)))

    def set_sources(self, elem, vals):
        e = np.ascontiguousarray(elem, dtype=np.int32).reshape(-1)
        v = self._sized("vals", vals, len(e) * 2 * self.npe)
        _gchk(self._hip.se2w_set_sources(self._handle, len(e), _ptr(e, _ip), _ptr(v)))

    def set_receivers(self, elem, gll_idx):
        e = np.ascontiguousarray(elem, dtype=np.int32).reshape(-1)
        g = self._sized("gll", gll_idx, len(e), np.int32)
        _gchk(self._hip.se2w_set_receivers(self._handle, len(e), _ptr(e, _ip), _ptr(g, _ip)))

    def set_state(self, u=None, v=None, a=None):
        arrs = [None if x is None else self._sized(nm, x, 2 * self.nnodes) for nm, x in (("u", u), ("v", v), ("a", a))]
        _gchk(self._hip.se2w_set_state(self._handle, *[None if x is None else _ptr(x) for x in arrs]))

    def get_state(self):
        u, v, a = (np.zeros(2 * self.nnodes) for _ in range(3))
        _gchk(self._hip.se2w_get_state(self._handle, _ptr(u), _ptr(v), _ptr(a)))
        return u, v, a

    def snapshot_begin(self):
        """Enqueue a copy of u, v (FP64, DMDA order) to pinned host memory; returns at once."""
        _gchk(self._hip.se2w_snapshot_begin(self._handle))

    def snapshot_wait(self):
        """(u, v) of the last snapshot, copied out of the library's pinned buffer."""
        pu, pv = _dp(), _dp()
        _gchk(self._hip.se2w_snapshot_wait(self._handle, C.byref(pu), C.byref(pv)))
        n = 2 * self.nnodes
        return np.ctypeslib.as_array(pu, shape=(n,)).copy(), np.ctypeslib.as_array(pv, shape=(n,)).copy()

    def reserve(self, nsteps: int):
        """Sizes the run buffers for runs of up to nsteps steps (se2w_reserve)."""
        _gchk(self._hip.se2w_reserve(self._handle, nsteps))

    def reserved(self):
        """(psi values, receiver record values, element-receiver record values, pinned bytes)."""
        c = np.zeros(4, dtype=np.int64)
        _gchk(self._hip.se2w_get_reserved(self._handle, c.ctypes.data_as(C.POINTER(C.c_int64))))
        return tuple(int(x) for x in c)

    def run(self, nsteps: int, dt: float, psi=None):
        ns = self.counts()[0]
        p = None
        if ns > 0 or psi is not None:
            if psi is None:
                raise ValueError(f"psi: a table of {nsteps} x {ns} values is required (sources are set)")
            a = np.ascontiguousarray(psi, dtype=np.float64).reshape(-1)
            if a.size < nsteps * ns:
                raise ValueError(f"psi: expected {nsteps * ns} values, got {a.size}")
            p = _ptr(a)
        _gchk(self._hip.se2w_run(self._handle, nsteps, dt, p))

    def receivers(self, nsteps: Optional[int] = None, nr: Optional[int] = None) -> np.ndarray:
        """Receiver rows [steps][nr][7] of the last run, sized from the
        context's own receiver count and run length (the arguments, if given,
        must agree with them)."""
        _, n_r, _, steps = self.counts()
        if (nsteps is not None and nsteps != steps) or (nr is not None and nr != n_r):
            raise ValueError(f"the last run has {steps} steps and {n_r} receivers")
        out = np.zeros(max(steps * n_r * 7, 1))
        _gchk(self._hip.se2w_get_receivers(self._handle, _ptr(out)))
        return out[: steps * n_r * 7].reshape(steps, n_r, 7)

    def assemble_forces(self, u) -> np.ndarray:
        u = self._sized("u", u, 2 * self.nnodes)
        f = np.zeros_like(u)
        _gchk(self._hip.se2w_assemble_forces(self._handle, _ptr(u), _ptr(f)))
        return f

    def assemble_forces_noalloc(self, u, f=None) -> np.ndarray:
        """assemble_forces through the entry point that allocates no device
        memory (se2w_assemble_forces_noalloc; meshes up to its scratch limit)."""
        u = self._sized("u", u, 2 * self.nnodes)
        f = np.zeros_like(u) if f is None else self._sized("f", f, 2 * self.nnodes)
        _gchk(self._hip.se2w_assemble_forces_noalloc(self._handle, _ptr(u), _ptr(f)))
        return f

    def norms(self) -> np.ndarray:
        out = np.zeros(6)
        _gchk(self._hip.se2w_norms(self._handle, _ptr(out)))
        return out

    def set_profiling(self, on: bool):
        _gchk(self._hip.se2w_set_profiling(self._handle, int(on)))

    def timing(self):
        ms = C.c_double()
        k = np.zeros(2)
        _gchk(self._hip.se2w_get_timing(self._handle, C.byref(ms), _ptr(k)))
        return ms.value, k

    def set_kernel(self, variant: int):
        _gchk(self._hip.se2w_set_kernel(self._handle, variant))

    def autotune(self, steps=20, reps=3):
        best, ms = C.c_int32(), C.c_double()
        _gchk(self._hip.se2w_autotune(self._handle, steps, reps, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def kernel(self) -> int:
        v = C.c_int32()
        _gchk(self._hip.se2w_get_kernel(self._handle, C.byref(v)))
        return v.value

    def tile(self):
        a, b = C.c_int32(), C.c_int32()
        _gchk(self._hip.se2w_get_tile(self._handle, C.byref(a), C.byref(b)))
        return a.value, b.value


class Ensemble:
    """Solvers of one degree and precision advanced together: one batched GPU
    launch per step for all members (each with its own mesh, materials,
    sources, receivers and dt), the run's launches replayed from a HIP graph."""

    def __init__(self, solvers: Sequence[Solver]):
        self._hip, self._host = libraries()
        self.solvers = list(solvers)
        n = len(self.solvers)
        ctxs = (C.c_void_p * n)(*[sv.ctx._handle for sv in self.solvers])
        self.h = C.c_void_p()
        _gchk(self._hip.se2w_ensemble_create(n, ctxs, C.byref(self.h)))
        self._sv = (C.c_void_p * n)(*[sv.h for sv in self.solvers])

    def advance(self, nsteps: int, dt):
        """nsteps steps of every member (dt: one step size per member); returns
        per member (times, trace[nsteps][nr][7]) as Solver.advance does."""
        dts = np.ascontiguousarray(dt, dtype=np.float64)
        _hchk(self._host.se2w_solvers_advance(len(self.solvers), self._sv, self.h, nsteps, _ptr(dts)))
        out = []
        for sv in self.solvers:
            times = np.zeros(nsteps)
            trace = np.zeros(nsteps * max(sv.nr, 1) * 7)
            _hchk(self._host.se2w_solver_get_trace(sv.h, _ptr(times), _ptr(trace)))
            out.append((times, trace[: nsteps * sv.nr * 7].reshape(nsteps, sv.nr, 7)))
        return out

    def set_kernel(self, variant: int):
        _gchk(self._hip.se2w_ensemble_set_kernel(self.h, variant))

    def kernel(self) -> int:
        v = C.c_int32()
        _gchk(self._hip.se2w_ensemble_get_kernel(self.h, C.byref(v)))
        return v.value

    def autotune(self, steps=10, reps=3):
        best, ms = C.c_int32(), C.c_double()
        _gchk(self._hip.se2w_ensemble_autotune(self.h, steps, reps, C.byref(best), C.byref(ms)))
        return best.value, ms.value

    def set_graph(self, on: bool):
        _gchk(self._hip.se2w_ensemble_set_graph(self.h, 1 if on else 0))

    def timing(self):
        """(whole last run in ms, mean batched step in ms)."""
        a, b = C.c_double(), C.c_double()
        _gchk(self._hip.se2w_ensemble_get_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def launch(self):
        """(tiles of all members, workgroups of the batched launch: one per tile)."""
        a, b = C.c_int32(), C.c_int32()
        _gchk(self._hip.se2w_ensemble_get_launch(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            self._hip.se2w_ensemble_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a if shape is None else a.reshape(shape)


class Mlp:
    """The surrogate network on the GPU (include/se2w_mlp.h): BF16 MFMA
    products with FP32 accumulation, FP32 master weights, Adam."""

    def __init__(self, widths: Sequence[int], batch=256, seed=0, device=0):
        self._hip, _ = libraries()
        self.widths = [int(w) for w in widths]
        self.batch = batch
        w = np.array(self.widths, dtype=np.int32)
        self.h = C.c_void_p()
        _gchk(self._hip.se2w_mlp_create(len(w) - 1, _ptr(w, _ip), batch, C.c_uint64(seed), device, C.byref(self.h)))
        n = C.c_int64()
        _gchk(self._hip.se2w_mlp_param_count(self.h, C.byref(n)))
        self.nparams = n.value

    def params(self) -> np.ndarray:
        p = np.zeros(self.nparams, dtype=np.float32)
        _gchk(self._hip.se2w_mlp_get_params(self.h, _fptr(p)))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.nparams
        _gchk(self._hip.se2w_mlp_set_params(self.h, _fptr(p)))

    def forward(self, x) -> np.ndarray:
        x = _f32(x, (-1, self.widths[0]))
        y = np.zeros((x.shape[0], self.widths[-1]), dtype=np.float32)
        _gchk(self._hip.se2w_mlp_forward(self.h, x.shape[0], _fptr(x), _fptr(y)))
        return y

    def gradient(self, x, target):
        x = _f32(x, (-1, self.widths[0]))
        t = _f32(target, (-1, self.widths[-1]))
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        _gchk(self._hip.se2w_mlp_gradient(self.h, x.shape[0], _fptr(x), _fptr(t), C.byref(loss), _fptr(g)))
        return loss.value, g

    def set_lr_schedule(self, cosine: bool):
        _gchk(self._hip.se2w_mlp_set_lr_schedule(self.h, 1 if cosine else 0))

    def set_weight_decay(self, wd: float):
        _gchk(self._hip.se2w_mlp_set_weight_decay(self.h, wd))

    def adam_step(self, grad, lr=1e-3):
        g = _f32(grad)
        _gchk(self._hip.se2w_mlp_adam_step(self.h, _fptr(g), lr))

    def set_data(self, x, target):
        x = _f32(x, (-1, self.widths[0]))
        t = _f32(target, (-1, self.widths[-1]))
        _gchk(self._hip.se2w_mlp_set_data(self.h, x.shape[0], _fptr(x), _fptr(t)))

    def train(self, epochs: int, lr=1e-3, seed=0) -> np.ndarray:
        losses = np.zeros(epochs, dtype=np.float32)
        _gchk(self._hip.se2w_mlp_train(self.h, epochs, lr, C.c_uint64(seed), _fptr(losses)))
        return losses

    def timing(self):
        """(last training in ms, last forward in ms)."""
        a, b = C.c_double(), C.c_double()
        _gchk(self._hip.se2w_mlp_get_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if self.h:
            self._hip.se2w_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_program(args: Sequence[str]) -> int:
    """Runs the se2wave program (se2w_main) in-process with reference options."""
    _, host = libraries()
    argv = [b"se2wave"] + [a.encode() for a in args]
    arr = (C.c_char_p * len(argv))(*argv)
    return host.se2w_main(len(argv), arr)
