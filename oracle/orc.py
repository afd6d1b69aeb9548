"""ctypes binding of the CPU oracle (liborc_sph.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, the smoke check and the CPU-baseline leg of bench.py, as the checker and
the timed CPU baseline; never by the product library. Records cross in the reference's
AoS layouts (the same numpy dtypes the product binding uses).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc_sph.so")

SPLIT, MERGE, MC = 1, 2, 4
HASH, INDEX, START, END, SPLITFLAG, STATUS = 0, 1, 2, 3, 4, 5
DGRID, NORMAL, TRI = 10, 11, 12

_lib = None
_vp, _u32, _i32, _sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
_pu32 = ctypes.POINTER(_u32)
_PROTOS = {
    "orc_create": ([_i32, _vp, _u32, _i32], _vp),
    "orc_destroy": ([_vp], None),
    "orc_set_threads": ([_i32], None),
    "orc_get_threads": ([], _i32),
    "orc_set_params": ([_vp, _vp], _i32),
    "orc_set_particles": ([_vp, _vp, _u32], _i32),
    "orc_get_particles": ([_vp, _vp], _i32),
    "orc_num_particles": ([_vp], _u32),
    "orc_get_state_f64": ([_vp, _vp], _i32),
    "orc_calc_hash": ([_vp], _i32),
    "orc_sort": ([_vp], _i32),
    "orc_cell_ranges": ([_vp], _i32),
    "orc_density": ([_vp], _i32),
    "orc_force": ([_vp], _i32),
    "orc_integrate": ([_vp], _i32),
    "orc_energy": ([_vp], _i32),
    "orc_split": ([_vp, _pu32, _pu32], _i32),
    "orc_merge": ([_vp, _pu32], _i32),
    "orc_compact": ([_vp], _i32),
    "orc_density_grid": ([_vp], _i32),
    "orc_normals": ([_vp], _i32),
    "orc_marching_cubes": ([_vp, _pu32], _i32),
    "orc_step": ([_vp, _u32], _i32),
    "orc_get_u32": ([_vp, _i32, _vp, _sz], _i32),
    "orc_get_f32": ([_vp, _i32, _vp, _sz], _i32),
    "orc_count": ([_vp, _i32], _sz),
    "orc_set_field": ([_vp, _vp, _i32], _i32),
    "orc_set_field_f64": ([_vp, _vp, _i32], _i32),
    "orc_density_sample": ([_vp, _i32, _vp], _i32),
    "orc_set_field_tangent": ([_vp, _vp, _vp, _i32], _i32),
    "orc_set_state_tangent": ([_vp, _vp, _vp, _vp], _i32),
    "orc_density_sample_tangent": ([_vp, _i32, _vp, _vp], _i32),
    "orc_topology": ([_vp, _i32], _i32),
    "orc_topology_copy": ([_vp, _vp], _i32),
    "orc_topology_stats": ([_vp, _vp], _i32),
    "orc_expf": ([ctypes.c_float], ctypes.c_float),
    "orc_wavelet_f32": ([ctypes.c_float], ctypes.c_float),
    "orc_wavelet_f64": ([ctypes.c_double], ctypes.c_double),
}


def load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def set_threads(n):
    load().orc_set_threads(int(n))


def get_threads():
    return load().orc_get_threads()


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError("%s failed (%d)" % (what, rc))


class Oracle:
    """One scene on the CPU; fp64=True runs the double-precision drift oracle; fp64="dual"
    and fp64="dual32" the double- and single-precision builds with forward-mode float64
    derivatives (one direction of the field and the initial state)."""

    def __init__(self, variant, params, capacity, fp64=False):
        self.lib = load()
        self.variant = variant
        p = np.ascontiguousarray(params).reshape(1)
        self._pdtype = p.dtype
        self.pdt = None
        mode = {"dual": 2, "dual32": 3}.get(fp64, 1 if fp64 else 0)
        self.h = self.lib.orc_create(variant, p.ctypes.data, int(capacity), mode)
        if not self.h:
            raise OracleError("orc_create failed")
        self._rec = 56 if variant == 0 else 72

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = np.ascontiguousarray(params).reshape(1)
        _check(self.lib.orc_set_params(self.h, p.ctypes.data), "orc_set_params")

    def upload(self, particles):
        a = np.ascontiguousarray(particles)
        self._pdt = a.dtype
        _check(self.lib.orc_set_particles(self.h, a.ctypes.data, len(a)), "orc_set_particles")

    def download(self):
        n = self.num_particles()
        out = np.zeros(n, dtype=self._pdt)
        _check(self.lib.orc_get_particles(self.h, out.ctypes.data), "orc_get_particles")
        return out

    def state_f64(self):
        n = self.num_particles()
        out = np.zeros((n, 17), dtype=np.float64)
        _check(self.lib.orc_get_state_f64(self.h, out.ctypes.data), "orc_get_state_f64")
        return out

    def num_particles(self):
        return int(self.lib.orc_num_particles(self.h))

    def calc_hash(self):
        return self.lib.orc_calc_hash(self.h)

    def sort(self):
        _check(self.lib.orc_sort(self.h), "orc_sort")

    def cell_ranges(self):
        _check(self.lib.orc_cell_ranges(self.h), "orc_cell_ranges")

    def density(self):
        _check(self.lib.orc_density(self.h), "orc_density")

    def force(self):
        _check(self.lib.orc_force(self.h), "orc_force")

    def integrate(self):
        _check(self.lib.orc_integrate(self.h), "orc_integrate")

    def energy(self):
        _check(self.lib.orc_energy(self.h), "orc_energy")

    def split(self):
        a, b = _u32(), _u32()
        _check(self.lib.orc_split(self.h, ctypes.byref(a), ctypes.byref(b)), "orc_split")
        return a.value, b.value

    def merge(self):
        g = _u32()
        _check(self.lib.orc_merge(self.h, ctypes.byref(g)), "orc_merge")
        return g.value

    def compact(self):
        _check(self.lib.orc_compact(self.h), "orc_compact")
        return self.num_particles()

    def density_grid(self):
        _check(self.lib.orc_density_grid(self.h), "orc_density_grid")

    # differentiable-rollout model: learnable force field and density sample
    def set_field(self, theta, nf):
        """Node values: float32 arrays as float32; float64 arrays keep their precision in
        the fp64 build (finite differences of the float64 model)."""
        if np.asarray(theta).dtype == np.float64:
            t = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
            fn = self.lib.orc_set_field_f64
        else:
            t = np.ascontiguousarray(theta, dtype=np.float32).reshape(-1)
            fn = self.lib.orc_set_field
        assert nf == 0 or t.size == nf ** 3 * 3
        self._field = t
        _check(fn(self.h, t.ctypes.data, int(nf)), "orc_set_field")

    def density_sample(self, ns):
        out = np.zeros(ns ** 3, dtype=np.float64)
        _check(self.lib.orc_density_sample(self.h, int(ns), out.ctypes.data), "orc_density_sample")
        return out

    # derivative build (fp64="dual")
    def set_field_tangent(self, theta, dtheta, nf):
        t = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(dtheta, dtype=np.float64).reshape(-1)
        assert t.size == d.size == nf ** 3 * 3
        self._field, self._dfield = t, d
        _check(self.lib.orc_set_field_tangent(self.h, t.ctypes.data, d.ctypes.data, int(nf)),
               "orc_set_field_tangent")

    def set_state_tangent(self, dpos=None, dvel=None, dev=None):
        n = self.num_particles()
        arrs = [None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(n * 3)
                for x in (dpos, dvel, dev)]
        ptrs = [None if x is None else x.ctypes.data for x in arrs]
        _check(self.lib.orc_set_state_tangent(self.h, *ptrs), "orc_set_state_tangent")

    def density_sample_tangent(self, ns):
        v = np.zeros(ns ** 3, dtype=np.float64)
        d = np.zeros(ns ** 3, dtype=np.float64)
        _check(self.lib.orc_density_sample_tangent(self.h, int(ns), v.ctypes.data, d.ctypes.data),
               "orc_density_sample_tangent")
        return v, d

    # rollout topology: record each step's split parents and merge groups, or replay them
    def topology(self, mode):
        _check(self.lib.orc_topology(self.h, {"off": 0, "record": 1, "replay": 2}[mode]),
               "orc_topology")

    def topology_from(self, other):
        _check(self.lib.orc_topology_copy(self.h, other.h), "orc_topology_copy")

    def topology_stats(self):
        out = np.zeros(3, dtype=np.uint32)
        _check(self.lib.orc_topology_stats(self.h, out.ctypes.data), "orc_topology_stats")
        return {"steps": int(out[0]), "splits": int(out[1]), "merge_groups": int(out[2])}

    def normals(self):
        _check(self.lib.orc_normals(self.h), "orc_normals")

    def marching_cubes(self):
        t = _u32()
        _check(self.lib.orc_marching_cubes(self.h, ctypes.byref(t)), "orc_marching_cubes")
        return t.value

    def step(self, flags=0):
        return self.lib.orc_step(self.h, int(flags))

    def get_u32(self, which, n=None):
        n = self.lib.orc_count(self.h, which) if n is None else n
        out = np.zeros(n, dtype=np.uint32)
        _check(self.lib.orc_get_u32(self.h, which, out.ctypes.data, n), "orc_get_u32")
        return out

    def get_f32(self, which, n=None):
        n = self.lib.orc_count(self.h, which) if n is None else n
        out = np.zeros(n, dtype=np.float32)
        _check(self.lib.orc_get_f32(self.h, which, out.ctypes.data, n), "orc_get_f32")
        return out


def expf(t):
    return load().orc_expf(float(t))


def wavelet_f32(x):
    return load().orc_wavelet_f32(float(x))


def wavelet_f64(x):
    return load().orc_wavelet_f64(float(x))
