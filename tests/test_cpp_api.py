"""The C++ drop-in API (omega_playground_amd/cpp, namespace omega).

tests/cpp/reference_tests.cpp restates the reference's self-checking test
programs against the C++ classes; build() compiles it into tests/cpp/bin.
The CPU tests compile it afresh and run its host-only mesh checks; the GPU
tests run the reference's conservation and convergence checks through the
HIP kernels (double_vortex.cpp:159-167 bounds; convergence rate 2 +- 0.05,
nonlinear_swe_convergence.cpp:205, tracer_*_convergence.cpp:128-131).
"""
import os
import subprocess

import pytest

from omega_playground_amd import build as b

BIN = os.path.join(os.path.dirname(__file__), "cpp", "bin", "reference_tests")
API_BIN = os.path.join(os.path.dirname(__file__), "cpp", "bin", "api_tests")


def _run(*args, timeout=240, exe=BIN):
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: run __graft_entry__.build() first")
    print("run", exe, *args, flush=True)
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cpp_api_compiles_and_links(tmp_path, monkeypatch):
    monkeypatch.setattr(b, "CPP_TESTS", str(tmp_path))
    src = os.path.join(os.path.dirname(__file__), "cpp", "reference_tests.cpp")
    (tmp_path / "reference_tests.cpp").write_text(open(src).read())
    exe, = b.build_cpp_programs()
    assert os.path.exists(exe)


def test_cpp_additions_compile_and_link(tmp_path, monkeypatch):
    """The adjoint and ensemble classes (omega.hpp) in a driver program."""
    monkeypatch.setattr(b, "CPP_TESTS", str(tmp_path))
    src = os.path.join(os.path.dirname(__file__), "cpp", "api_tests.cpp")
    (tmp_path / "api_tests.cpp").write_text(open(src).read())
    exe, = b.build_cpp_programs()
    assert os.path.exists(exe)


def test_cpp_mesh_api_host_only():
    assert "mesh checks passed" in _run("mesh")


@pytest.mark.gpu
def test_cpp_double_vortex_conservation():
    _run("double_vortex")


@pytest.mark.gpu
def test_cpp_nonlinear_swe_convergence():
    _run("nonlinear_swe")


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["adv", "diff", "hyper"])
def test_cpp_tracer_convergence(kind):
    _run("tracer", kind)


@pytest.mark.gpu
@pytest.mark.parametrize("what", ["adjoint", "ensemble", "ensemble_adjoint", "linear", "fields"])
def test_cpp_adjoint_and_ensemble_classes(what):
    _run(what, exe=API_BIN)
