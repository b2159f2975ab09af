"""The generated mechanism code compiled for the host (g++, no contraction)
and checked bit for bit: the slab kernel's accessor-templated functions
(right-looking factorisation with lazy assembly, Jacobian entries, solve,
rates, kinetics) against the plain generated functions and the oracle.
Source: tests/host_kernels/check_generated.cpp."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))


def test_generated_kernels_match_on_host(tmp_path):
    exe = str(tmp_path / "check_generated")
    oracle_dir = os.path.join(ROOT, "oracle")
    subprocess.run(["g++", "-O1", "-std=c++17", "-ffp-contract=off",
                    "-I", os.path.join(HERE, "host_kernels"), "-I", os.path.join(ROOT, "brams_amd", "csrc"),
                    "-o", exe, os.path.join(HERE, "host_kernels", "check_generated.cpp"),
                    "-L", oracle_dir, "-loracle", "-Wl,-rpath," + oracle_dir], check=True)
    out = subprocess.run([exe, "400", "7"], capture_output=True, text=True)
    print(out.stdout)
    assert out.returncode == 0 and out.stdout.startswith("OK"), out.stdout
