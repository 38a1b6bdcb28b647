"""Carving of device temporaries out of slabs (maxlq_amd/csrc/mlq_slab.hpp,
host bookkeeping of the engine's allocator): tests/slab_check.cpp drives it
with 20000 random takes and returns of 1-9 grains against a model of every
grain (no overlap, blocks inside their slab, best fit, returned blocks merged
with their free neighbours, every slab released at the end)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "maxlq_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_slab_carving_against_grain_model(tmp_path):
    exe = str(tmp_path / "slab_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", CSRC,
                    os.path.join(HERE, "slab_check.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout
