"""The large-block cache's ownership rule on the CPU (fake streams, fake
allocator): a block released mid-run is reused only on its own stream, becomes
shared only after its stream is declared idle, and a flush never frees an owned
block (parallel-ra_amd/csrc/pra_blockcache.hpp)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_block_cache_ownership_rule(tmp_path):
    exe = str(tmp_path / "block_cache_test")
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-O1", "-o", exe,
                    os.path.join(ROOT, "tests", "block_cache_test.cpp"), "-lpthread"], check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
