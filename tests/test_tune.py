"""The autotuner's rule on the CPU (parallel-ra_amd/csrc/pra_tune.hpp): every
variant of a knob is measured a few times on live iterations, then the
cheapest is kept for that device, knob and size band."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tuner_rule(tmp_path):
    exe = str(tmp_path / "tune_test")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-O1", "-o", exe,
                    os.path.join(ROOT, "tests", "tune_test.cpp"), "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
