"""The build's rule on register spills: a kernel that feeds LDS by direct
loads (direct-to-LDS GEMMs, the CRT and Ozaki plane products) places its
waits by hand and must not spill; other kernels may.  Checked on remark text
in the form hipcc prints with -Rpass-analysis=kernel-resource-usage."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "distributed-gemm-thesis_amd"))
import build  # noqa: E402


def remarks(name, vgpr_spill=0, sgpr_spill=0, scratch=0):
    tag = " [-Rpass-analysis=kernel-resource-usage]"
    return "\n".join([
        "x.hip:1:1: remark: Function Name: %s%s" % (name, tag),
        "x.hip:1:1: remark:     VGPRs: 200%s" % tag,
        "x.hip:1:1: remark:     ScratchSize [bytes/lane]: %d%s" % (scratch, tag),
        "x.hip:1:1: remark:     SGPRs Spill: %d%s" % (sgpr_spill, tag),
        "x.hip:1:1: remark:     VGPRs Spill: %d%s" % (vgpr_spill, tag),
    ])


DMA = "_ZN2dg12_GLOBAL__N_119sgemm_dma_xr_kernelILb0ELb1EEEvNS0_9SgemmArgsE"
CRT = "_ZN2dg12_GLOBAL__N_115crt_gemm_kernelENS0_7CrtArgsE"
PLANE = "_ZN2dg12_GLOBAL__N_115oz_plane_kernelILb1EEEvNS0_11OzPlaneArgsE"
STAGED = "_ZN2dg12_GLOBAL__N_115gemm_f64_kernelINS0_7TileCfgILi256ELi128ELi512ELi1ELi64ELi64EEELb1ELb1ELb1ELi0EEE"


@pytest.mark.parametrize("name", [DMA, CRT, PLANE])
@pytest.mark.parametrize("kind", ["vgpr", "sgpr", "scratch"])
def test_spill_in_direct_lds_kernel_fails(name, kind):
    kw = {"vgpr": dict(vgpr_spill=5), "sgpr": dict(sgpr_spill=1), "scratch": dict(scratch=24)}[kind]
    with pytest.raises(RuntimeError, match="spills"):
        build._check_spills("x.hip", remarks(name, **kw))


def test_clean_direct_lds_kernel_and_spilled_staged_kernel_pass():
    build._check_spills("x.hip", remarks(DMA) + "\n" + remarks(STAGED, vgpr_spill=2, scratch=12))


def test_spill_is_charged_to_the_right_kernel():
    # the staged kernel's spill follows a clean direct-to-LDS kernel's remarks
    build._check_spills("x.hip", remarks(CRT) + "\n" + remarks(STAGED, vgpr_spill=3))
    with pytest.raises(RuntimeError, match="crt_gemm_kernel"):
        build._check_spills("x.hip", remarks(STAGED) + "\n" + remarks(CRT, vgpr_spill=1))
