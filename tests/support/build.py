"""Build tests/support/libtestmult.so: the user device multipliers of
multipliers.hip, the fused multipliers of fusedmult.hip (compiled into the
library's row kernels), the C++ FFT face of fftface.cc, the inner-column and
unfused-row convolutions of innerc.cc and the reference transform
tests of residues.cc, linked against the in-tree libfftwpp_amd.so (found next to
the package at run time through $ORIGIN)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libtestmult.so")


def build(verbose=False):
    sys.path.insert(0, ROOT)
    from fftwpp_amd import build as b
    pkg = os.path.join(ROOT, "fftwpp_amd")
    cmd = [b.HIPCC] + b.HIP_FLAGS + ["-I", os.path.join(pkg, "csrc"),
                                     "-shared", os.path.join(HERE, "multipliers.hip"),
                                     os.path.join(HERE, "fusedmult.hip"),
                                     os.path.join(HERE, "fftface.cc"),
                                     os.path.join(HERE, "residues.cc"),
                                     os.path.join(HERE, "innerc.cc"),
                                     "-L", pkg, "-lfftwpp_amd",
                                     "-Wl,-rpath,$ORIGIN/../../fftwpp_amd",
                                     "-o", LIB + ".tmp"]
    b._run(cmd, verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
