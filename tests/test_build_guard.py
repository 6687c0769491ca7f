"""The build's register guard (hermite-ugks_amd/build_lib.py): a kernel
whose compiler resource-usage remarks report scratch memory, a spilled
VGPR or all 256 VGPRs in use fails the build; SGPRs spilled into VGPR
lanes pass.  The remark lines are those hipcc -Rpass-analysis=
kernel-resource-usage prints (a block-per-face flux kernel of this
library's history: 256 VGPRs, AGPRs as spill space, then scratch)."""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "hermite-ugks_amd"))
import build_lib  # noqa: E402


def remarks(name, vgprs, agprs, sspill, vspill, scratch):
    r = "x.hip:1:1: remark: {} [-Rpass-analysis=kernel-resource-usage]\n"
    return "".join(r.format(t) for t in (
        f"Function Name: {name}", f"    TotalSGPRs: 106", f"    VGPRs: {vgprs}",
        f"    AGPRs: {agprs}", f"    ScratchSize [bytes/lane]: {scratch}",
        f"    Occupancy [waves/SIMD]: 1", f"    SGPRs Spill: {sspill}",
        f"    VGPRs Spill: {vspill}", f"    LDS Size [bytes/block]: 992"))


def test_guard_flags_register_exhaustion_and_scratch():
    text = (remarks("k_ok", 241, 0, 28, 0, 0) + remarks("k_mfma", 172, 16, 0, 0, 0) +
            remarks("k_full", 256, 216, 8, 0, 0) + remarks("k_scratch", 255, 256, 27, 8, 28))
    bad = build_lib.scratch_kernels(text)
    assert ("k_full", "256 VGPRs (the whole file)") in bad
    assert ("k_scratch", "28 bytes/lane of scratch") in bad
    assert ("k_scratch", "8 VGPRs spilled") in bad
    assert not [b for b in bad if b[0] in ("k_ok", "k_mfma")]
