"""CPU check of the return-map vector-Jacobian products used by the adjoint
step (pyroclastmpm_amd/csrc/adjoint_math.hpp): dual-number VJPs in FP32
against central finite differences of the same map in FP64, over random
states that cover the elastic, cone and apex branches of Drucker-Prager and
linear elasticity. Compiles tests/adjoint_math_check.cpp with g++."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_return_map_vjp_matches_fp64_finite_differences():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "amc")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I",
                        os.path.join(ROOT, "pyroclastmpm_amd", "csrc"),
                        os.path.join(HERE, "adjoint_math_check.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [r.split() for r in out if r.strip()]
    branches = {r[1] for r in rows if r[0].startswith("dp")}
    assert {"elastic", "cone", "apex"} <= branches, branches
    worst = max(float(r[-1]) for r in rows)
    assert worst < 1e-5, rows
