"""The AHMED-style C++ surface (include/ahmed) driven the way a reference user
program drives the reference: tests/cpp/drop_in.cpp builds bbx and bemcluster
trees, assembles on the GPU (matgenGeH_gpu), checks mltaGeHVec against
direct kernel sums and solves with GMRes / CG through GeHMatrix."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_drop_in_program():
    exe = os.path.join(ROOT, "tests", "cpp", "drop_in")
    assert os.path.exists(exe), "build first: make -C ahmed_amd/csrc"
    print("[gpu] tests/cpp/drop_in: bbx N=12000, bemcluster N=9000", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    assert "drop_in OK" in p.stdout
