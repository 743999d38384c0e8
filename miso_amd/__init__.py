"""miso_amd -- MI355X (gfx950) backend for MISO's MHD time step.

The compute path is libmisohip.so (hand-written HIP kernels behind the C ABI
in include/misohip.h).  This package holds its Python host mirror
(`miso_amd.mhd`), the ctypes binding (`miso_amd.hip`) and the benchmark
initial conditions (`miso_amd.problems`).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC_DIR = os.path.join(PKG_DIR, "csrc")
LIB_DIR = os.path.join(PKG_DIR, "lib")


def build(jobs: int = 8) -> str:
    """Compile libmisohip.so for gfx950 in-tree (always recompiles)."""
    import subprocess
    subprocess.run(["make", "-C", CSRC_DIR, "clean"], check=True, capture_output=True)
    subprocess.run(["make", "-C", CSRC_DIR, f"-j{jobs}"], check=True)
    return os.path.join(LIB_DIR, "libmisohip.so")
