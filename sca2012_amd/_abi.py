"""ctypes binding of libsph_amd (include/sph_amd.h) and the reference's record layouts.

The structured dtypes below are byte-identical to the reference's ``Particle`` and
``SysParam`` classes (SPH-3D-2/sph_system.h:8-24, sph_param.h:8-87; SPH-2D/sph_system.h:7-22,
sph_param.h:6-72); the library static_asserts the same sizes. ``SPHContext`` is a thin
wrapper over the C ABI: every call returns the library's status, and a non-zero status
raises ``SphError``. There is no fallback: if the HIP library is missing or a call fails,
the error propagates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SPH_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("SPH_AMD_LIB") or os.path.join(_HERE, "libsph_amd.so")

V2D, V3D2, V3D4 = 0, 1, 2
STEP_SPLIT, STEP_MERGE, STEP_MC = 1, 2, 4
OK, EINVAL, EEMPTY, EDOMAIN, ECAP, EHIP, ETRI = 0, -1, -2, -3, -4, -5, -6
STATUS_NAMES = {0: "ok", -1: "EINVAL", -2: "EEMPTY", -3: "EDOMAIN", -4: "ECAP", -5: "EHIP", -6: "ETRI"}

Q_HASH, Q_INDEX, Q_START, Q_END, Q_SPLITFLAG, Q_STATUS, Q_COUNTERS = 0, 1, 2, 3, 4, 5, 6
Q_DGRID, Q_NORMAL, Q_TRI = 10, 11, 12

PHASES = ["hash", "sort", "cells", "reorder", "density", "force", "integrate", "energy",
          "split", "merge", "compact", "dgrid", "normals", "mc"]
COUNTER_NAMES = ["n", "m", "n_pre", "s_req", "s_done", "groups", "err", "ntri", "scan_total", "nmc"]


def _layout(fields, itemsize=None):
    """numpy dtype with C struct offsets. fields: (name, base, count, align)."""
    names, formats, offsets = [], [], []
    off = 0
    for name, base, count, align in fields:
        off = (off + align - 1) // align * align
        names.append(name)
        formats.append((base, count) if count > 1 else base)
        offsets.append(off)
        off += np.dtype(base).itemsize * count
    size = itemsize if itemsize is not None else off
    return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size})


SPHERE_DTYPE = _layout([("pos", "<f4", 3, 4), ("radius", "<f4", 1, 4), ("damping", "<f4", 1, 4)])

_P3 = [("num_particle", "<u4"), ("max_particle", "<u4"), ("world_width", "<f4"),
       ("world_height", "<f4"), ("world_length", "<f4"), ("large_radius", "<f4"),
       ("small_radius", "<f4"), ("large_kernel", "<f4"), ("small_kernel", "<f4"),
       ("large_mass", "<f4"), ("small_mass", "<f4"), ("cell_size", "<f4"), ("row_cell", "<u4"),
       ("col_cell", "<u4"), ("len_cell", "<u4"), ("total_cell", "<u4")]
_P3b = ["wall_damping", "rest_density", "gas_constant", "time_step", "viscosity",
        "surface_tension", "co_surface", "co_energy", "split_criteria", "merge_criteria",
        "large_poly6", "large_spiky", "large_visco", "large_grad_poly6", "large_lapc_poly6",
        "small_poly6", "small_spiky", "small_visco", "small_grad_poly6", "small_lapc_poly6",
        "large_kernel_2", "large_kernel_6", "small_kernel_2", "small_kernel_6",
        "large_poly6_radius", "small_poly6_radius"]
SYSPARAM3D_DTYPE = _layout(
    [(n, t, 1, 4) for n, t in _P3] + [("gravity", "<f4", 3, 4)] +
    [(n, "<f4", 1, 4) for n in _P3b] +
    [("dens_mul", "<u4", 1, 4), ("den_size", "<f4", 1, 4), ("row_dens", "<u4", 1, 4),
     ("col_dens", "<u4", 1, 4), ("len_dens", "<u4", 1, 4), ("tot_dens", "<u4", 1, 4)] +
    [("sphere%d" % k, SPHERE_DTYPE, 1, 4) for k in (1, 2, 3, 4)] +
    [("split_energy", "<f4", 1, 4), ("merge_energy", "<f4", 1, 4),
     ("flag0", "u1", 1, 1), ("flag1", "u1", 1, 1), ("flag2", "u1", 1, 1)], itemsize=296)

_P2a = ["world_width", "world_depth", "large_mass", "small_mass", "large_kernel",
        "small_kernel", "large_kernel2", "small_kernel2", "large_kernel6", "small_kernel6",
        "large_radius", "small_radius", "large_poly6", "small_poly6", "large_spiky",
        "small_spiky", "large_visco", "small_visco", "large_grad_poly6", "small_grad_poly6",
        "large_lapc_poly6", "small_lapc_poly6"]
_P2b = ["wall_damping", "rest_density", "gas_constant", "viscosity", "time_step",
        "surface_normal", "cell_size"]
SYSPARAM2D_DTYPE = _layout(
    [("num_particle", "<u4", 1, 4), ("max_particle", "<u4", 1, 4)] +
    [(n, "<f4", 1, 4) for n in _P2a] + [("gravity", "<f4", 2, 8)] +
    [(n, "<f4", 1, 4) for n in _P2b] +
    [("row_cell", "<u4", 1, 4), ("col_cell", "<u4", 1, 4), ("total_cell", "<u4", 1, 4),
     ("co_surface", "<f4", 1, 4), ("co_energy", "<f4", 1, 4), ("thresh_split", "<f4", 1, 4),
     ("thresh_merge", "<f4", 1, 4), ("force_waiting", "u1", 1, 1),
     ("extern_force", "<f4", 2, 8), ("extern_click", "<f4", 2, 8),
     ("extern_pos", "<u4", 2, 8), ("extern_len", "<f4", 1, 4), ("extern_coe", "<f4", 1, 4)],
    itemsize=200)

PARTICLE3D_DTYPE = _layout(
    [("pos", "<f4", 3, 4), ("vel", "<f4", 3, 4), ("ev", "<f4", 3, 4), ("acc", "<f4", 3, 4),
     ("dens", "<f4", 1, 4), ("press", "<f4", 1, 4), ("energy", "<f4", 1, 4),
     ("surface", "<f4", 1, 4), ("tension", "<f4", 1, 4), ("level", "<u4", 1, 4)], itemsize=72)

PARTICLE2D_DTYPE = _layout(
    [("pos", "<f4", 2, 8), ("vel", "<f4", 2, 8), ("ev", "<f4", 2, 8), ("acc", "<f4", 2, 8),
     ("dens", "<f4", 1, 4), ("pres", "<f4", 1, 4), ("energy", "<f4", 1, 4),
     ("level", "<u4", 1, 4), ("surface_normal", "<f4", 1, 4)], itemsize=56)

assert SYSPARAM3D_DTYPE.itemsize == 296 and SYSPARAM2D_DTYPE.itemsize == 200
assert PARTICLE3D_DTYPE.itemsize == 72 and PARTICLE2D_DTYPE.itemsize == 56
assert SYSPARAM3D_DTYPE.fields["sphere1"][1] == 204 and SYSPARAM3D_DTYPE.fields["flag0"][1] == 292
assert SYSPARAM2D_DTYPE.fields["extern_force"][1] == 168


def param_dtype(variant):
    return SYSPARAM2D_DTYPE if variant == V2D else SYSPARAM3D_DTYPE


def particle_dtype(variant):
    return PARTICLE2D_DTYPE if variant == V2D else PARTICLE3D_DTYPE


def _tuned(fn, h, what):
    f, e = _u32(), _u32()
    best = np.zeros(6, dtype=np.float32)
    _check(fn(h, ctypes.byref(f), ctypes.byref(e), best.ctypes.data), what)
    return f.value, e.value, {"force": {24: float(best[0]), 32: float(best[1]), 48: float(best[2])},
                              "energy": {16: float(best[3]), 24: float(best[4]), 32: float(best[5])}}


class SphError(RuntimeError):
    def __init__(self, code, what):
        super().__init__("%s failed: %s (%d)" % (what, STATUS_NAMES.get(code, "?"), code))
        self.code = code


_lib = None
_vp, _u32, _i32, _sz = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_size_t
_pu32 = ctypes.POINTER(ctypes.c_uint32)

_PROTOS = {
    "sph_version": ([], ctypes.c_char_p),
    "sph_param_defaults": ([_i32, _vp], _i32),
    "sph_param_derive": ([_i32, _vp], _i32),
    "sph_ctx_create": ([_i32, _vp, _u32, _u32, ctypes.POINTER(_vp)], _i32),
    "sph_ctx_destroy": ([_vp], _i32),
    "sph_set_params": ([_vp, _vp], _i32),
    "sph_get_params": ([_vp, _vp], _i32),
    "sph_upload_particles": ([_vp, _vp, _u32], _i32),
    "sph_download_particles": ([_vp, _vp, _u32], _i32),
    "sph_num_particles": ([_vp, _pu32], _i32),
    "sph_calc_hash": ([_vp], _i32),
    "sph_sort": ([_vp], _i32),
    "sph_cell_ranges": ([_vp], _i32),
    "sph_density": ([_vp], _i32),
    "sph_force": ([_vp], _i32),
    "sph_integrate": ([_vp], _i32),
    "sph_energy": ([_vp], _i32),
    "sph_split": ([_vp, _pu32, _pu32], _i32),
    "sph_merge": ([_vp, _pu32], _i32),
    "sph_compact": ([_vp, _pu32], _i32),
    "sph_density_grid": ([_vp], _i32),
    "sph_grid_normals": ([_vp], _i32),
    "sph_marching_cubes": ([_vp, _pu32], _i32),
    "sph_step": ([_vp, _u32, _pu32], _i32),
    "sph_get_u32": ([_vp, _i32, _vp, _sz], _i32),
    "sph_get_f32": ([_vp, _i32, _vp, _sz], _i32),
    "sph_count": ([_vp, _i32, ctypes.POINTER(_sz)], _i32),
    "sph_prof_enable": ([_vp, _i32], _i32),
    "sph_prof_phases": ([_vp, _u32], _i32),
    "sph_autotune": ([_vp, _u32, _u32, _u32], _i32),
    "sph_tuned_lists": ([_vp, _pu32, _pu32, _vp], _i32),
    "sph_set_lists": ([_vp, _u32, _u32], _i32),
    "sph_prof_reset": ([_vp], _i32),
    "sph_prof_read": ([_vp, _i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)], _i32),
    "sph_synchronize": ([_vp], _i32),
    "sph_copy_calibration": ([_sz, _i32], _i32),
    "sph_sort_noalloc": ([_vp, _vp, _u32, _u32, _vp], _i32),
    "sph_adj_create": ([_vp, _u32, _u32, _u32, _u32, _u32, ctypes.POINTER(_vp)], _i32),
    "sph_adj_num_particles": ([_vp, _pu32], _i32),
    "sph_adj_topology": ([_vp, _vp, _u32], _i32),
    "sph_adj_destroy": ([_vp], _i32),
    "sph_adj_upload": ([_vp, _vp, _u32], _i32),
    "sph_adj_set_target": ([_vp, _vp], _i32),
    "sph_adj_forward": ([_vp, _vp, _u32, _vp, ctypes.POINTER(ctypes.c_double)], _i32),
    "sph_adj_backward": ([_vp, _vp], _i32),
    "sph_adj_download": ([_vp, _vp, _u32], _i32),
    "sph_adj_state_grad": ([_vp, _vp, _vp, _vp], _i32),
    "sph_adj_timing": ([_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], _i32),
    "sph_batch_create": ([_i32, _u32, _vp, _vp, ctypes.POINTER(_vp)], _i32),
    "sph_batch_destroy": ([_vp], _i32),
    "sph_batch_set_params": ([_vp, _u32, _vp], _i32),
    "sph_batch_get_params": ([_vp, _u32, _vp], _i32),
    "sph_batch_upload": ([_vp, _u32, _vp, _u32], _i32),
    "sph_batch_download": ([_vp, _u32, _vp, _u32], _i32),
    "sph_batch_num_particles": ([_vp, _vp], _i32),
    "sph_batch_scene_counts": ([_vp, _vp], _i32),
    "sph_batch_step": ([_vp, _u32, _u32, _pu32], _i32),
    "sph_batch_prof_enable": ([_vp, _i32], _i32),
    "sph_batch_prof_phases": ([_vp, _u32], _i32),
    "sph_batch_autotune": ([_vp, _u32, _u32, _u32], _i32),
    "sph_batch_tuned_lists": ([_vp, _pu32, _pu32, _vp], _i32),
    "sph_batch_set_lists": ([_vp, _u32, _u32], _i32),
    "sph_batch_prof_reset": ([_vp], _i32),
    "sph_batch_prof_read": ([_vp, _i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64)], _i32),
    "sph_batch_set_graph": ([_vp, _i32], _i32),
    "sph_batch_set_launch": ([_vp, _u32], _i32),
    "sph_batch_tuned_launch": ([_vp, _pu32, _pu32, _vp], _i32),
}


def exported_names():
    return list(_PROTOS)


def load_library(path=None):
    """Load libsph_amd.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    lib = ctypes.CDLL(path or LIB_PATH)
    for name, (args, res) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if path is None:
        _lib = lib
    return lib


def _check(code, what):
    if code != OK:
        raise SphError(code, what)


def param_defaults(variant):
    """SysParam with the constants of the reference's SPHSystem constructor."""
    lib = load_library()
    p = np.zeros(1, dtype=param_dtype(variant))
    _check(lib.sph_param_defaults(variant, p.ctypes.data), "sph_param_defaults")
    return p


def param_derive(variant, p):
    """Recompute the derived fields of ``p`` (1-element SysParam array) in place."""
    lib = load_library()
    _check(lib.sph_param_derive(variant, p.ctypes.data), "sph_param_derive")
    return p


def add_extern_force(params, pos, force):
    """2D mouse force for the next step, as SPHSystem::add_extern_force
    (SPH-2D/sph_system.cpp:178-187): `pos` in [0, 1]^2 of the world, `force` per unit
    extern_coe, applied to particles within extern_len of the click; the step clears it."""
    W, D = float(params["world_width"][0]), float(params["world_depth"][0])
    params["extern_click"][0] = (np.float32(pos[0]) * np.float32(W), np.float32(pos[1]) * np.float32(D))
    params["extern_pos"][0] = (int(np.floor(np.float32(pos[0]) * np.float32(params["row_cell"][0]))),
                               int(np.floor(np.float32(pos[1]) * np.float32(params["col_cell"][0]))))
    params["extern_force"][0] = force
    params["force_waiting"] = 1


class SPHContext:
    """One scene on the GPU: owns the device buffers of libsph_amd."""

    def __init__(self, variant, params, capacity, max_triangles=None):
        self.lib = load_library()
        self.variant = variant
        p = np.ascontiguousarray(params, dtype=param_dtype(variant)).reshape(1)
        if max_triangles is None:
            max_triangles = int(p["tot_dens"][0]) if variant != V2D else 0
        self.capacity = int(capacity)
        self.max_triangles = int(max_triangles)
        h = _vp()
        _check(self.lib.sph_ctx_create(variant, p.ctypes.data, self.capacity, self.max_triangles,
                                       ctypes.byref(h)), "sph_ctx_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.sph_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ----
    def upload(self, particles):
        a = np.ascontiguousarray(particles, dtype=particle_dtype(self.variant))
        _check(self.lib.sph_upload_particles(self.h, a.ctypes.data, len(a)), "sph_upload_particles")

    def download(self):
        n = self.num_particles()
        out = np.zeros(n, dtype=particle_dtype(self.variant))
        _check(self.lib.sph_download_particles(self.h, out.ctypes.data, n), "sph_download_particles")
        return out

    def num_particles(self):
        n = _u32()
        _check(self.lib.sph_num_particles(self.h, ctypes.byref(n)), "sph_num_particles")
        return n.value

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=param_dtype(self.variant)).reshape(1)
        _check(self.lib.sph_set_params(self.h, p.ctypes.data), "sph_set_params")

    def get_params(self):
        p = np.zeros(1, dtype=param_dtype(self.variant))
        _check(self.lib.sph_get_params(self.h, p.ctypes.data), "sph_get_params")
        return p

    # ---- phases ----
    def calc_hash(self):
        _check(self.lib.sph_calc_hash(self.h), "sph_calc_hash")

    def sort(self):
        _check(self.lib.sph_sort(self.h), "sph_sort")

    def cell_ranges(self):
        _check(self.lib.sph_cell_ranges(self.h), "sph_cell_ranges")

    def density(self):
        _check(self.lib.sph_density(self.h), "sph_density")

    def force(self):
        _check(self.lib.sph_force(self.h), "sph_force")

    def integrate(self):
        _check(self.lib.sph_integrate(self.h), "sph_integrate")

    def energy(self):
        _check(self.lib.sph_energy(self.h), "sph_energy")

    def split(self):
        a, b = _u32(), _u32()
        _check(self.lib.sph_split(self.h, ctypes.byref(a), ctypes.byref(b)), "sph_split")
        return a.value, b.value

    def merge(self):
        g = _u32()
        _check(self.lib.sph_merge(self.h, ctypes.byref(g)), "sph_merge")
        return g.value

    def compact(self):
        n = _u32()
        _check(self.lib.sph_compact(self.h, ctypes.byref(n)), "sph_compact")
        return n.value

    def density_grid(self):
        _check(self.lib.sph_density_grid(self.h), "sph_density_grid")

    def grid_normals(self):
        _check(self.lib.sph_grid_normals(self.h), "sph_grid_normals")

    def marching_cubes(self):
        t = _u32()
        _check(self.lib.sph_marching_cubes(self.h, ctypes.byref(t)), "sph_marching_cubes")
        return t.value

    def step(self, flags=0):
        n = _u32()
        _check(self.lib.sph_step(self.h, flags, ctypes.byref(n)), "sph_step")
        return n.value

    def step_raw(self, flags=0):
        """sph_step returning the raw status (no exception)."""
        n = _u32()
        return self.lib.sph_step(self.h, flags, ctypes.byref(n)), n.value

    # ---- accessors ----
    def count(self, which):
        n = _sz()
        _check(self.lib.sph_count(self.h, which, ctypes.byref(n)), "sph_count")
        return n.value

    def get_u32(self, which, n=None):
        n = self.count(which) if n is None else n
        out = np.zeros(n, dtype=np.uint32)
        _check(self.lib.sph_get_u32(self.h, which, out.ctypes.data, n), "sph_get_u32")
        return out

    def get_f32(self, which, n=None):
        n = self.count(which) if n is None else n
        out = np.zeros(n, dtype=np.float32)
        _check(self.lib.sph_get_f32(self.h, which, out.ctypes.data, n), "sph_get_f32")
        return out

    def counters(self):
        c = self.get_u32(Q_COUNTERS)
        return dict(zip(COUNTER_NAMES, (int(v) for v in c)))

    # ---- timing ----
    def prof_enable(self, on=True):
        _check(self.lib.sph_prof_enable(self.h, 1 if on else 0), "sph_prof_enable")

    def prof_phases(self, names):
        """Time only the named phases (each event pair costs the stream a few us)."""
        mask = 0
        for nm in names:
            mask |= 1 << PHASES.index(nm)
        _check(self.lib.sph_prof_phases(self.h, mask), "sph_prof_phases")

    def autotune(self, rounds, after=1, every=128):
        """Restart the pair-list autotuner: `rounds` rounds of 12 tuning steps (the first
        one 6), the first after `after` steps, then one every `every` steps (rounds 0: stop
        tuning)."""
        _check(self.lib.sph_autotune(self.h, int(rounds), int(after), int(every)), "sph_autotune")

    def tuned_lists(self):
        """(force list length, energy list length, fastest ms per candidate)."""
        return _tuned(self.lib.sph_tuned_lists, self.h, "sph_tuned_lists")

    def set_lists(self, force_list, energy_list):
        """Pin the pair-list lengths (stops tuning)."""
        _check(self.lib.sph_set_lists(self.h, int(force_list), int(energy_list)), "sph_set_lists")

    def prof_reset(self):
        _check(self.lib.sph_prof_reset(self.h), "sph_prof_reset")

    def prof_read(self):
        out = {}
        for k, name in enumerate(PHASES):
            ms, cnt = ctypes.c_double(), ctypes.c_uint64()
            _check(self.lib.sph_prof_read(self.h, k, ctypes.byref(ms), ctypes.byref(cnt)), "sph_prof_read")
            out[name] = (ms.value, cnt.value)
        return out

    def synchronize(self):
        _check(self.lib.sph_synchronize(self.h), "sph_synchronize")


class SPHBatch:
    """A batch of independent scenes of one 3D variant, stepped together (C ABI
    sph_batch_*). `params` is a list of 1-element SysParam arrays, `capacities` the
    per-scene particle capacities."""

    def __init__(self, variant, params, capacities):
        self.lib = load_library()
        self.variant = variant
        self.ns = len(params)
        # filled record by record: np.concatenate may drop the records' tail padding
        recs = np.zeros(self.ns, dtype=param_dtype(variant))
        for k, p in enumerate(params):
            recs[k] = np.asarray(p, dtype=param_dtype(variant)).reshape(1)[0]
        assert recs.dtype.itemsize == param_dtype(variant).itemsize
        caps = np.asarray(capacities, dtype=np.uint32)
        assert len(caps) == self.ns
        self.h = ctypes.c_void_p()
        _check(self.lib.sph_batch_create(variant, self.ns, recs.ctypes.data, caps.ctypes.data,
                                         ctypes.byref(self.h)), "sph_batch_create")

    def close(self):
        if self.h:
            self.lib.sph_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, scene, params):
        p = np.asarray(params, dtype=param_dtype(self.variant)).reshape(1)
        _check(self.lib.sph_batch_set_params(self.h, scene, p.ctypes.data), "sph_batch_set_params")

    def get_params(self, scene):
        p = np.zeros(1, dtype=param_dtype(self.variant))
        _check(self.lib.sph_batch_get_params(self.h, scene, p.ctypes.data), "sph_batch_get_params")
        return p

    def upload(self, scene, particles):
        a = np.ascontiguousarray(particles, dtype=particle_dtype(self.variant))
        _check(self.lib.sph_batch_upload(self.h, scene, a.ctypes.data, len(a)), "sph_batch_upload")

    def num_particles(self):
        out = np.zeros(self.ns, dtype=np.uint32)
        _check(self.lib.sph_batch_num_particles(self.h, out.ctypes.data), "sph_batch_num_particles")
        return out

    def download(self, scene):
        n = int(self.num_particles()[scene])
        out = np.zeros(n, dtype=particle_dtype(self.variant))
        _check(self.lib.sph_batch_download(self.h, scene, out.ctypes.data, n), "sph_batch_download")
        return out

    def scene_counts(self):
        """Per scene, of the last step: (split requested, split done, merge groups, status)."""
        out = np.zeros((self.ns, 4), dtype=np.int32)
        _check(self.lib.sph_batch_scene_counts(self.h, out.ctypes.data), "sph_batch_scene_counts")
        return out

    def step_raw(self, flags=0, steps=1):
        n = ctypes.c_uint32()
        rc = self.lib.sph_batch_step(self.h, flags, steps, ctypes.byref(n))
        return rc, n.value

    def step(self, flags=0, steps=1):
        rc, n = self.step_raw(flags, steps)
        _check(rc, "sph_batch_step")
        return n

    def prof_enable(self, on=True):
        _check(self.lib.sph_batch_prof_enable(self.h, 1 if on else 0), "sph_batch_prof_enable")

    def prof_phases(self, names):
        """Time only the named phases (each event pair costs the stream a few us)."""
        mask = 0
        for nm in names:
            mask |= 1 << PHASES.index(nm)
        _check(self.lib.sph_batch_prof_phases(self.h, mask), "sph_batch_prof_phases")

    def autotune(self, rounds, after=1, every=128):
        _check(self.lib.sph_batch_autotune(self.h, int(rounds), int(after), int(every)),
               "sph_batch_autotune")

    def tuned_lists(self):
        return _tuned(self.lib.sph_batch_tuned_lists, self.h, "sph_batch_tuned_lists")

    def set_lists(self, force_list, energy_list):
        _check(self.lib.sph_batch_set_lists(self.h, int(force_list), int(energy_list)),
               "sph_batch_set_lists")

    def set_graph(self, on=True):
        """Graph steps (one captured hipGraph per step, counts read on the device) or stream
        steps (kernels launched one by one, sized by counts read back each step)."""
        _check(self.lib.sph_batch_set_graph(self.h, 1 if on else 0), "sph_batch_set_graph")

    def set_launch(self, groups_per_cu):
        """Pin the graph steps' workgroups per CU: 0 (launches sized by the buffers' bounds)
        or 4, 8, 16 (persistent launches)."""
        _check(self.lib.sph_batch_set_launch(self.h, int(groups_per_cu)), "sph_batch_set_launch")

    def tuned_launch(self):
        """(workgroups per CU, CUs, {candidate: mean step ms of the last round})"""
        g, cu = _u32(), _u32()
        best = np.zeros(4, dtype=np.float32)
        _check(self.lib.sph_batch_tuned_launch(self.h, ctypes.byref(g), ctypes.byref(cu),
                                               best.ctypes.data), "sph_batch_tuned_launch")
        return g.value, cu.value, {c: float(b) for c, b in zip((0, 4, 8, 16), best)}

    def prof_reset(self):
        _check(self.lib.sph_batch_prof_reset(self.h), "sph_batch_prof_reset")

    def prof_read(self):
        out = {}
        for k, name in enumerate(PHASES):
            ms, cnt = ctypes.c_double(), ctypes.c_uint64()
            _check(self.lib.sph_batch_prof_read(self.h, k, ctypes.byref(ms), ctypes.byref(cnt)),
                   "sph_batch_prof_read")
            out[name] = (ms.value, cnt.value)
        return out


def _buf_ptr(x, n=None, writable=False):
    """Address of a float32 buffer: a numpy array (host), an object with data_ptr() (a
    device tensor) or an integer address."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        assert x.dtype == np.float32 and x.flags["C_CONTIGUOUS"] and (n is None or x.size == n)
        return ctypes.c_void_p(x.ctypes.data)
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr())
    return ctypes.c_void_p(int(x))


class SPHAdjoint:
    """Differentiable rollout (C ABI sph_adj_*): the 3D-2 step, with split and merge as
    `flags` select, plus a learnable force field of `field_nodes`^3 x 3 values; loss =
    squared error of the density at a `sample_n`^3 grid against a target; the adjoint runs
    over each step's topology as the forward took it. theta, target, sample and dtheta are
    float32 numpy arrays (host) or device tensors (anything with data_ptr())."""

    def __init__(self, params, n, field_nodes, sample_n, checkpoint_every=0, flags=0):
        self.lib = load_library()
        p = np.asarray(params, dtype=SYSPARAM3D_DTYPE).reshape(1)
        self.n, self.nf, self.ns = int(n), int(field_nodes), int(sample_n)
        self.h = ctypes.c_void_p()
        _check(self.lib.sph_adj_create(p.ctypes.data, self.n, self.nf, self.ns,
                                       int(checkpoint_every), int(flags), ctypes.byref(self.h)),
               "sph_adj_create")

    def close(self):
        if self.h:
            self.lib.sph_adj_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, particles):
        a = np.ascontiguousarray(particles, dtype=PARTICLE3D_DTYPE)
        _check(self.lib.sph_adj_upload(self.h, a.ctypes.data, len(a)), "sph_adj_upload")

    def set_target(self, target):
        _check(self.lib.sph_adj_set_target(self.h, _buf_ptr(target, self.ns ** 3)), "sph_adj_set_target")

    def forward(self, theta, steps, sample=None):
        """Run `steps` steps from the uploaded state; returns the loss (float64)."""
        loss = ctypes.c_double()
        _check(self.lib.sph_adj_forward(self.h, _buf_ptr(theta, self.nf ** 3 * 3), int(steps),
                                        _buf_ptr(sample, self.ns ** 3), ctypes.byref(loss)),
               "sph_adj_forward")
        return loss.value

    def backward(self, dtheta):
        """dL/dtheta of the last forward into `dtheta`."""
        _check(self.lib.sph_adj_backward(self.h, _buf_ptr(dtheta, self.nf ** 3 * 3)), "sph_adj_backward")

    def num_particles(self):
        n = _u32()
        _check(self.lib.sph_adj_num_particles(self.h, ctypes.byref(n)), "sph_adj_num_particles")
        return n.value

    def download(self):
        n = self.num_particles()
        out = np.zeros(n, dtype=PARTICLE3D_DTYPE)
        _check(self.lib.sph_adj_download(self.h, out.ctypes.data, n), "sph_adj_download")
        return out

    def topology(self, steps):
        """Per step of the last forward: (live before the split, parents split, merge groups)."""
        out = np.zeros((steps, 3), dtype=np.uint32)
        _check(self.lib.sph_adj_topology(self.h, out.ctypes.data, int(steps)), "sph_adj_topology")
        return out

    def state_grad(self):
        g = [np.zeros((self.n, 3), dtype=np.float32) for _ in range(3)]
        _check(self.lib.sph_adj_state_grad(self.h, g[0].ctypes.data, g[1].ctypes.data,
                                           g[2].ctypes.data), "sph_adj_state_grad")
        return g

    def timing(self):
        f, b = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.sph_adj_timing(self.h, ctypes.byref(f), ctypes.byref(b)), "sph_adj_timing")
        return f.value, b.value
