"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle_bind.cpp (+ oracle.h) into oracle/_oracle*.so with g++.
The reference itself is not built: its hot path needs Eigen, spdlog and
thrust headers that the image lacks, so no oracle/_ref/ artefact exists.
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False) -> str:
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_oracle" + ext)
    cmd = [
        "g++", "-O3", "-std=c++17", "-shared", "-fPIC",
        "-ffp-contract=off",  # no FMA contraction: the reference CPU build has none
        "-I" + pybind11.get_include(),
        "-I" + sysconfig.get_paths()["include"],
        os.path.join(HERE, "oracle_bind.cpp"),
        "-o", out + ".tmp",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
