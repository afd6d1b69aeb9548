"""Writes tests/golden/fixtures.json from the CPU oracle.

The reference ships no tests, fixtures or golden vectors, so these fixtures pin the oracle
against regression: per-step live counts, split requests/splits and merge groups, and
sha256 digests of the full AoS state, for the reference's own start-up scenes
(SPH-3D-2 7220 particles, SPH-3D-4 41850, SPH-2D 1953; split and merge on) and for the
first steps of the 4096-particle 2D dam break. Run: python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import sca2012_amd as sa  # noqa: E402
from sca2012_amd import scenes  # noqa: E402
from oracle import orc  # noqa: E402

CASES = {
    "ref_3d2_split_merge": (sa.V3D2, lambda: scenes.reference_scene(sa.V3D2), 3, 20),
    "ref_3d4_split_merge": (sa.V3D4, lambda: scenes.reference_scene(sa.V3D4), 3, 12),
    "ref_2d_split_merge": (sa.V2D, lambda: scenes.reference_scene(sa.V2D), 3, 20),
    "dam_break_2d_4096": (sa.V2D, scenes.dam_break_2d, 0, 10),
}


def digest(a):
    """sha256 of the AoS bytes with every NaN replaced by the canonical quiet NaN (NaN
    payloads differ between x86 and the GPU; the values are otherwise compared bitwise)."""
    b = np.zeros(len(a), dtype=a.dtype)  # zeroed, so layout padding hashes as zeros
    for f in a.dtype.names:
        b[f] = a[f]
        if b[f].dtype.kind == "f":
            v = b[f]
            v[np.isnan(v)] = np.float32("nan")
    return hashlib.sha256(b.tobytes()).hexdigest()


def run_case(variant, build, flags, steps):
    p, a = build()
    o = orc.Oracle(variant, p, 200000)
    o.upload(a)
    rec = {"n0": len(a), "flags": flags, "steps": []}
    for s in range(steps):
        assert o.calc_hash() == 0
        o.sort()
        o.cell_ranges()
        o.density()
        o.force()
        o.integrate()
        o.energy()
        S = done = groups = 0
        if flags & sa.STEP_SPLIT:
            S, done = o.split()
        if flags & sa.STEP_MERGE:
            groups = o.merge()
            o.compact()
        st = o.download()
        rec["steps"].append({"n": o.num_particles(), "split_requested": S, "split_done": done,
                             "merge_groups": groups, "sha256": digest(st)})
    return rec


def main():
    out = {name: run_case(*c) for name, c in CASES.items()}
    path = os.path.join(HERE, "fixtures.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
