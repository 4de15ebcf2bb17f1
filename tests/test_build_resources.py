"""The register and LDS budget the stage kernels' performance rests on, checked at compile time (no GPU).

The FP64 stage kernel k_rk_stage_ring<STAGE> runs four waves per SIMD: 16 waves per CU of 160 KiB LDS
at 10 KiB of ring per wave (a block of four waves holds 40 KiB), which also needs at most 128 VGPRs per
lane. A register spill would put scratch traffic on the HBM path the kernel is bound by. The mixed
path's k_rk_stage_noalias<float> is bound by registers alone and must not spill either. The numbers
come from the compiler's resource report for gfx950 with the library's own flags.
"""
import re
import shutil
import subprocess

import pytest

from aturan_amd import build as B

pytestmark = pytest.mark.skipif(shutil.which("hipcc") is None and not B.Path("/opt/rocm/bin/hipcc").exists(),
                                reason="hipcc is needed to compile the kernels")


def resource_report(src: str) -> dict:
    """kernel name -> {"VGPRs": n, "ScratchSize": bytes, "LDS Size": bytes, "Occupancy": waves/SIMD}"""
    cmd = [B._hipcc(), *B.COMMON, "-x", "hip", *B.HIP_FLAGS, "--cuda-device-only", "-c", str(B.CSRC / src),
           "-o", "/dev/null", "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out, cur = {}, None
    for ln in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", ln)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark: +(VGPRs|ScratchSize \[bytes/lane\]|LDS Size \[bytes/block\]|Occupancy \[waves/SIMD\]): (\d+)",
                      ln)
        if m and cur is not None:
            cur[m.group(1).split(" [")[0]] = int(m.group(2))
    return out


@pytest.fixture(scope="module")
def rk():
    return resource_report("atu_rk.hip")


def test_fp64_ring_stage_fits_four_waves_per_simd(rk):
    ring = {n: v for n, v in rk.items() if "k_rk_stage_ring" in n}
    assert len(ring) == 4
    for n, v in ring.items():
        assert v["ScratchSize"] == 0, n
        assert v["VGPRs"] <= 128, (n, v["VGPRs"])
        assert v["LDS Size"] == 4 * 2 * 640 * 8, (n, v["LDS Size"])
        assert v["Occupancy"] == 4, (n, v["Occupancy"])


def test_mixed_stage_does_not_spill(rk):
    mixed = {n: v for n, v in rk.items() if "k_rk_stage_noalias" in n}
    assert len(mixed) == 4
    for n, v in mixed.items():
        assert v["ScratchSize"] == 0, n
        assert v["Occupancy"] >= 4, (n, v["Occupancy"])
