"""The C++ host interface (lsquantumed_amd/host/lsquantumed.hpp) through its
own test programs under tests/cpp: the reference's unit tests written against
the C++ mirror of its classes.  build() compiles them in-tree (tests/cpp/
_build); the CPU program is rebuilt here if a source changed (host code only),
the GPU program must already be built."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "_build")


def _run(name: str, cwd, timeout: int) -> str:
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built: run __graft_entry__.build()")
    print(f"running {name}", flush=True)
    r = subprocess.run([exe], cwd=str(cwd), capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-6000:], flush=True)
    assert r.returncode == 0, r.stdout[-6000:] + r.stderr[-3000:]
    assert "0 failed" in r.stdout
    return r.stdout


def test_cpp_host_cpu(tmp_path):
    """Environment, Basis, lattices, sublattices and RandomDisorder with the
    reference's literals (no GPU needed)."""
    subprocess.run(["make", "-s", "-C", CPP, os.path.join("_build", "test_host_cpu")], check=True)
    out = _run("test_host_cpu", tmp_path, 120)
    assert "12 cases" in out


def test_cpp_replay_schedule(tmp_path):
    """The batched solve's host bookkeeping of graph replays (which replays
    run which model; pausing a model predicted to converge): scripted and
    randomised residual sequences, no GPU."""
    subprocess.run(["make", "-s", "-C", CPP, os.path.join("_build", "test_replay_schedule")], check=True)
    out = _run("test_replay_schedule", tmp_path, 120)
    assert "3 cases" in out


@pytest.mark.gpu
def test_cpp_host_gpu(tmp_path):
    """Solver, correlations, m^2 and DSF through the C++ mirror on the GPU,
    with the reference's literal eigenvalues and correlators."""
    out = _run("test_host_gpu", tmp_path, 110)
    assert "[  FAILED  ]" not in out
    assert os.path.exists(tmp_path / "dsf1" / "dsf_d0.dat")
    assert os.path.exists(tmp_path / "dsf2" / "dsf_d0.dat")


@pytest.mark.gpu
def test_cpp_host_dumps(tmp_path):
    """The 31 build_diag / build_off_diag / build_disorder dumps of the
    reference's tests, written by the C++ mirror's MatView / VecView, equal
    the reference's .ref files character for character."""
    out = _run("test_host_dumps", tmp_path, 110)
    assert "[  FAILED  ]" not in out
