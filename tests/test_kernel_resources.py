"""Register budgets of the neighbour kernels, read from the gfx950 assembly (no GPU needed).

How many waves a SIMD holds is set by the kernel's registers. Measured on MI355X with
per-wave start times: a kernel reporting 80 SGPRs ran 8 waves per SIMD, one reporting 82
ran 7 (the SGPR count includes VCC and the other reserved registers the compiler adds).
VGPRs are allocated from 512 per lane in granules of 8, so 8 waves need at most 64. The
density and energy walks are written for 8 waves per SIMD and lose 10-15% of their speed
at 7; the force kernel holds four neighbour records in flight and is written for 5 waves
in a single scene (at most 96 VGPRs, where its fallback list's LDS allows 5) and 4 in a
batch (at most 128). None of them may spill to scratch: the compiler meets a wave target by
spilling when the registers do not fit, which would put scratch loads in the pair loops. A
code change that crosses one of these limits passes every parity test and only shows up as
a slower step, so it is checked here.
"""
import os
import re
import shutil
import subprocess

import pytest

from sca2012_amd import build as sbuild

HIPCC = sbuild.HIPCC
SRC = os.path.join(sbuild.CSRC, "sph_kernels.hip")

# kernel name stem -> (max SGPRs, max VGPRs): 8 waves per SIMD for the walks, 4 for force
# (whose SGPRs then only need to stay within the 102 a wave can address)
BUDGET = {
    "k_density": (80, 64),
    "k_energy": (80, 64),
    "k_force": (102, 128),
}


@pytest.fixture(scope="module")
def resources(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which(HIPCC)):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("asm") / "sph_kernels.s")
    cmd = [HIPCC] + sbuild.FLAGS + ["--cuda-device-only", "-S", SRC, "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    text = open(out).read()
    res = {}
    for m in re.finditer(r"^(_Z\w+):", text, re.M):
        name = m.group(1)
        end = text.find(".Lfunc_end", m.end())
        meta = text[end:end + 4000]
        sg = re.search(r"NumSgprs:\s*(\d+)", meta)
        vg = re.search(r"NumVgprs:\s*(\d+)", meta)
        sc = re.search(r"ScratchSize:\s*(\d+)", meta)
        if sg and vg and sc:
            res[name] = (int(sg.group(1)), int(vg.group(1)), int(sc.group(1)))
    return res


@pytest.mark.parametrize("stem", sorted(BUDGET))
def test_neighbour_kernel_registers(resources, stem):
    max_s, max_v = BUDGET[stem]
    # mangled as <length><stem>I... inside the anonymous namespace
    found = {n: r for n, r in resources.items() if re.search(r"\d%sI" % stem, n)}
    assert found, "no %s instantiation in the assembly" % stem
    over = {n: r for n, r in found.items() if r[0] > max_s or r[1] > max_v or r[2] > 0}
    assert not over, "%s over its register budget (SGPRs <= %d, VGPRs <= %d, no scratch): %s" % (
        stem, max_s, max_v, over)


def test_single_scene_force_fits_five_waves(resources):
    # k_force<V, false, L> for the fallback lengths whose LDS allows 5 blocks per CU
    found = {n: r for n, r in resources.items()
             if re.search(r"\dk_forceILi\dELb0ELi(24|32)E", n)}
    assert len(found) == 6, sorted(found)
    over = {n: r for n, r in found.items() if r[1] > 96}
    assert not over, "single-scene force over 96 VGPRs: %s" % over
