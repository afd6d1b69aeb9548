"""The C ABI library: it loads without a GPU, exports every function include/sph_amd.h
declares, its record layouts match the reference's, and the host-side parameter
derivation reproduces the reference constructor's constants (no compute calls)."""
import ctypes
import math
import os
import re

import numpy as np

import sca2012_amd as sa
from sca2012_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "sph_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(sph_[a-z0-9_]+)\s*\(",
                                 txt, re.M)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the whole C ABI
    assert sorted(_abi.exported_names()) == names


def test_version_string():
    assert sa.load_library().sph_version().decode().startswith("sph_amd")


def test_record_layouts():
    assert sa.PARTICLE3D_DTYPE.itemsize == 72 and sa.PARTICLE2D_DTYPE.itemsize == 56
    assert sa.SYSPARAM3D_DTYPE.itemsize == 296 and sa.SYSPARAM2D_DTYPE.itemsize == 200
    f = sa.PARTICLE3D_DTYPE.fields
    assert [f[k][1] for k in ("pos", "vel", "ev", "acc", "dens", "press", "energy", "surface",
                              "tension", "level")] == [0, 12, 24, 36, 48, 52, 56, 60, 64, 68]
    f2 = sa.PARTICLE2D_DTYPE.fields
    assert [f2[k][1] for k in ("pos", "vel", "ev", "acc", "dens", "pres", "energy", "level",
                               "surface_normal")] == [0, 8, 16, 24, 32, 36, 40, 44, 48]


def test_param_defaults_match_reference_constructor():
    """SPH-3D-2/sph_system.cpp:6-229: world 32x16x16, cell = 1, poly6 = 315/(64 PI h^9)."""
    p = sa.param_defaults(sa.V3D2)
    assert (p["row_cell"][0], p["col_cell"][0], p["len_cell"][0]) == (32, 16, 16)
    assert p["total_cell"][0] == 32 * 16 * 16
    PI = float(np.float32(3.141592))
    for h, key in ((1.0, "large"), (float(np.float32(0.9)), "small")):
        assert p[key + "_poly6"][0] == np.float32(315.0 / (float(np.float32(64.0 * np.float32(PI))) * h ** 9))
        assert abs(p[key + "_spiky"][0] + 45.0 / (PI * h ** 6)) < 1e-5 * 45.0 / (PI * h ** 6)
        assert abs(p[key + "_lapc_poly6"][0] + 945.0 / (8 * PI * h ** 9)) < 1e-5 * 945.0 / (8 * PI * h ** 9)
    assert p["den_size"][0] == np.float32(0.25)
    assert (p["row_dens"][0], p["col_dens"][0], p["len_dens"][0]) == (129, 65, 65)
    assert p["gravity"][0][1] == np.float32(-5.0) and p["time_step"][0] == np.float32(0.1)
    assert p["split_criteria"][0] == np.float32(1.2) and p["merge_criteria"][0] == np.float32(0.4)
    assert p["sphere2"]["pos"][0][2] == np.float32(12.0)
    q = sa.param_defaults(sa.V3D4)
    assert (q["row_cell"][0], q["col_cell"][0], q["len_cell"][0]) == (40, 20, 20)
    assert q["small_mass"][0] == np.float32(0.25) and q["co_surface"][0] == np.float32(1.5)
    r = sa.param_defaults(sa.V2D)
    assert (r["row_cell"][0], r["col_cell"][0], r["total_cell"][0]) == (32, 32, 1024)
    assert r["thresh_split"][0] == np.float32(0.7) and r["extern_len"][0] == np.float32(2.0)


def test_param_derive_snaps_world_to_cells():
    p = sa.param_defaults(sa.V3D2)
    p["world_width"] = 10.5
    sa.param_derive(sa.V3D2, p)
    assert p["row_cell"][0] == 11 and p["world_width"][0] == np.float32(11.0)


def test_workload_scenes_shapes():
    p, a = scenes.adaptive_3d2(n_side=8, cells=16)
    assert len(a) == 512 and p["total_cell"][0] == 16 ** 3
    assert a["pos"].min() > 0.9 and a["pos"].max() < 16
    p2, b = scenes.dam_break_2d(n_side=8)
    assert len(b) == 64 and p2["row_cell"][0] == 50
    # the calibrated mass gives the rest density in the lattice interior (closed form)
    h = float(np.float32(0.02))
    s = scenes.lattice_density_sum(h, 0.01, 1e-12)
    assert abs(float(p2["large_mass"][0]) * float(p2["large_poly6"][0]) * s - 1000.0) < 1e-2
    # seeded PRNG is a pure function of (seed, stream, counter)
    u = scenes.uniform(7, 3, 10)
    assert np.array_equal(u, scenes.uniform(7, 3, 10)) and np.all((u >= 0) & (u < 1))
    assert np.array_equal(scenes.uniform(7, 3, 20)[:10], u)


def test_batched_workload_tanks():
    """256 tanks, N_s = round(exp(U(ln 2048, ln 65536))), each block inside its own box."""
    sizes, tanks = scenes.batched_3d2(n_scenes=8)
    assert len(tanks) == 8 and sizes.min() >= 2048 and sizes.max() <= 65536
    s2, _ = scenes.batched_3d2(n_scenes=8)
    assert np.array_equal(sizes, s2)
    for n, (p, a) in zip(sizes, tanks):
        assert len(a) == n and np.all(a["level"] == 1)
        W = np.array([p["world_width"][0], p["world_height"][0], p["world_length"][0]])
        r = float(p["large_radius"][0])
        assert np.all(a["pos"] >= r) and np.all(a["pos"] < W - r)
    big = sa.param_defaults(sa.V3D2)
    assert sa.SYSPARAM3D_DTYPE == big.dtype
