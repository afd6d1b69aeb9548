"""The C++ drop-in boundary, on the CPU (nothing here touches a GPU):

* the reference's GLUT front ends use SPHSystem through a fixed set of members; translation
  units that repeat every such use (tests/dropin_ui/, citing the reference lines) compile and
  link against libsph_amd.so through the reference's names -- SPHSystem, SysParam, Particle,
  Sphere -- for 3D-2, for 3D-4 (SPH_DROPIN_3D4) and for 2D;
* every public data member of the reference classes (SPH-3D-2/sph_system.h:26-154,
  SPH-2D/sph_system.h:24-106) is declared by the drop-in headers;
* the per-variant SysParam records keep the reference's names and byte offsets of the
  trailing bools (SPH-3D-2/sph_param.h:85-86, SPH-3D-4/sph_param.h:85-87).
"""
import os
import re
import subprocess

import pytest

import sca2012_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.dirname(sa._abi.LIB_PATH)

# public data members of the reference classes, in declaration order
REF_3D_MEMBERS = """m_host_mem m_dev_mem m_num_particle m_max_particle m_world_width m_world_height
m_world_length m_large_radius m_small_radius m_large_kernel m_small_kernel m_large_mass m_small_mass
m_cell_size m_row_cell m_col_cell m_len_cell m_total_cell m_dev_hash m_dev_index m_cell_start
m_cell_end m_dev_status m_trans_vel m_dev_split m_dev_split_index m_dev_merge m_gravity
m_wall_damping m_rest_density m_gas_constant m_time_step m_viscosity m_surface_tension m_co_surface
m_co_energy m_split_criteria m_merge_criteria m_large_poly6 m_large_spiky m_large_visco
m_large_grad_poly6 m_large_lapc_poly6 m_small_poly6 m_small_spiky m_small_visco m_small_grad_poly6
m_small_lapc_poly6 m_large_kernel_2 m_large_kernel_6 m_small_kernel_2 m_small_kernel_6
m_large_poly6_radius m_small_poly6_radius m_dens_mul m_den_size m_row_dens m_col_dens m_len_dens
m_tot_dens m_host_dens m_dev_dens m_host_dens_pos m_dev_dens_pos m_host_dens_normal m_dev_dens_normal
m_num_lines m_host_line0 m_host_line1 m_num_triangle m_host_triangle0 m_host_triangle1
m_host_triangle2 m_host_triangle_normal0 m_host_triangle_normal1 m_host_triangle_normal2 m_sphere1
m_sphere2 m_sphere3 m_sphere4 m_host_param m_sys_running m_trans_once m_use_split m_use_merge
m_disp_mode""".split()
REF_3D_METHODS = "add_box_particle add_new_particle animation add_particle marching_cube".split()
REF_2D_MEMBERS = """m_num_particle m_max_particle m_world_width m_world_depth m_large_mass
m_small_mass m_large_kernel m_small_kernel m_large_kernel2 m_small_kernel2 m_large_kernel6
m_small_kernel6 m_large_radius m_small_radius m_large_poly6 m_small_poly6 m_large_spiky m_small_spiky
m_large_visco m_small_visco m_large_grad_poly6 m_small_grad_poly6 m_large_lapc_poly6
m_small_lapc_poly6 m_gravity m_wall_damping m_rest_density m_gas_constant m_viscosity m_time_step
m_surface_normal m_sys_running m_cell_size m_row_cell m_col_cell m_total_cell m_co_surface
m_co_energy m_thresh_split m_thresh_merge m_host_mem m_dev_mem m_dev_hash m_dev_index m_cell_start
m_cell_end m_trans_vel m_dev_status m_dev_split m_dev_merge m_dev_split_index m_enable_split
m_enable_merge m_host_param""".split()
REF_2D_METHODS = "add_init_particle add_particle animate add_extern_force".split()


def _declared(path):
    txt = open(path).read()
    txt = re.sub(r"//[^\n]*", "", txt)
    return set(re.findall(r"\b(m_[a-z0-9_]+)\b", txt)), txt


def test_3d_header_declares_every_reference_member():
    names, txt = _declared(os.path.join(INC, "sph_system_amd.h"))
    missing = [m for m in REF_3D_MEMBERS if m not in names]
    assert not missing, missing
    for m in REF_3D_METHODS:
        assert re.search(r"\b%s\s*\(" % m, txt), m


def test_2d_header_declares_every_reference_member():
    names, txt = _declared(os.path.join(INC, "sph_system_2d_amd.h"))
    missing = [m for m in REF_2D_MEMBERS if m not in names]
    assert not missing, missing
    for m in REF_2D_METHODS:
        assert re.search(r"\b%s\s*\(" % m, txt), m


def _build(tmp_path, src, defs=()):
    assert os.path.exists(sa._abi.LIB_PATH), "build the library first (__graft_entry__.build)"
    out = str(tmp_path / (os.path.basename(src) + "".join(defs) + ".bin"))
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror"] + ["-D" + d for d in defs] + [
        "-o", out, os.path.join(ROOT, "tests", "dropin_ui", src), "-L" + LIBDIR, "-lsph_amd"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return out


@pytest.mark.parametrize("defs", [(), ("SPH_DROPIN_3D4",)])
def test_reference_3d_front_end_uses_compile_and_link(tmp_path, defs):
    assert os.path.exists(_build(tmp_path, "ui_3d.cpp", defs))


def test_reference_2d_front_end_uses_compile_and_link(tmp_path):
    assert os.path.exists(_build(tmp_path, "ui_2d.cpp"))


def test_param_records_keep_reference_names_and_offsets(tmp_path):
    src = tmp_path / "offsets.cpp"
    src.write_text("""
#include <cstddef>
#include "sph_amd.h"
static_assert(sizeof(SphSysParam3D2) == 296 && sizeof(SphSysParam3D4) == 296, "size");
static_assert(sizeof(SphSysParam3D) == 296 && sizeof(SphSysParam2D) == 200, "size");
static_assert(offsetof(SphSysParam3D2, force_waiting) == 292, "3D-2 force_waiting");
static_assert(offsetof(SphSysParam3D2, use_cylinder) == 293, "3D-2 use_cylinder");
static_assert(offsetof(SphSysParam3D4, force_waiting_left) == 292, "3D-4 force_waiting_left");
static_assert(offsetof(SphSysParam3D4, force_waiting_right) == 293, "3D-4 force_waiting_right");
static_assert(offsetof(SphSysParam3D4, use_cylinder) == 294, "3D-4 use_cylinder");
static_assert(offsetof(SphSysParam3D2, sphere1) == offsetof(SphSysParam3D, sphere1), "common");
static_assert(offsetof(SphSysParam3D4, merge_energy) == offsetof(SphSysParam3D, merge_energy), "common");
static_assert(offsetof(SphSysParam2D, force_waiting) == 160, "2D force_waiting");
int main() { return 0; }
""")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + INC, str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
