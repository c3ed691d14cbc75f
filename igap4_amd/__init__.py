"""igap4_amd -- MI355X-native residual/tangent assembly for IGAP4 patches.

The compute path is libigap_gpu.so (HIP kernels for gfx950 behind the C ABI of
include/igap_gpu.h); this package is the Python host mirror used by tests,
the benchmark and the Newton driver.
"""
from .assm import (MFMA_ROW, SVK_COEF, Assm, BC, BCClampedTip, BCDefault, MeshIGABeam, PhysGradElast,
                   PhysGradElastTime, selftest_expected, selftest_mfma,
                   selftest_unpack, smooth_state)

__all__ = ["Assm", "BC", "BCClampedTip", "BCDefault", "MeshIGABeam", "PhysGradElast",
           "PhysGradElastTime", "SVK_COEF", "MFMA_ROW", "selftest_mfma", "selftest_expected",
           "selftest_unpack", "smooth_state"]
