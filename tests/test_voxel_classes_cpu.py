"""The per-voxel classes (host/csd.h CSD, host/icls.h ICLS::Solver) on a host without a GPU: they
are built on the GPU path only, so constructing them fails with a clear error instead of falling
back to a CPU solver.  (tests/test_gpu_voxel_classes.py covers them on the GPU.)"""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "_bin", "voxel_classes")


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present: the classes construct")
def test_classes_refuse_without_gpu():
    assert os.path.exists(EXE), "voxel_classes is built by __graft_entry__.build()"
    r = subprocess.run([EXE, "nogpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 2
    assert lines[0].startswith("CSD:") and "no HIP device" in lines[0]
    assert lines[1].startswith("Solver:") and "no HIP device" in lines[1]


def test_driver_usage_errors():
    r = subprocess.run([EXE, "csd", "/nonexistent", "/tmp/x", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
