"""A C program written against the reference's interface (the uhm_*_ call
sequence of UHM/example/f90/ex1.f90:38-183) compiles against
include/uhm_amd.h and links against libuhm_amd.so unchanged."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "uhm_amd", "examples", "two_triangles.c")


def _build(lib, tmp_path):
    import uhm_amd
    libdir = os.path.dirname(uhm_amd.LIB_PATH)
    exe = str(tmp_path / "two_triangles")
    subprocess.run(["gcc", "-O1", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
                    "-o", exe, "-L", libdir, "-luhm_amd", "-Wl,-rpath," + libdir], check=True)
    return exe


def test_c_program_links_and_locks(lib, tmp_path):
    exe = _build(lib, tmp_path)
    out = subprocess.run([exe, "symbolic"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "n_dof 25" in out.stdout


@pytest.mark.gpu
def test_c_program_full_run(lib, tmp_path):
    exe = _build(lib, tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "backward_error" in out.stdout


SRC_CPP = os.path.join(ROOT, "uhm_amd", "examples", "two_triangles.cpp")


def _build_cpp(lib, tmp_path):
    """The C++ program written against include/uhm_mesh.hpp (the reference's
    uhm::Mesh_ interface, UHM/include/uhm/mesh/mesh.hxx) compiles with
    warnings as errors and links against libuhm_amd.so."""
    import uhm_amd
    libdir = os.path.dirname(uhm_amd.LIB_PATH)
    exe = str(tmp_path / "two_triangles_cpp")
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    SRC_CPP, "-o", exe, "-L", libdir, "-luhm_amd", "-Wl,-rpath," + libdir], check=True)
    return exe


def test_cpp_program_links_and_locks(lib, tmp_path):
    exe = _build_cpp(lib, tmp_path)
    out = subprocess.run([exe, "symbolic"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert "n_dof 25" in out.stdout


@pytest.mark.gpu
def test_cpp_program_full_run(lib, tmp_path):
    exe = _build_cpp(lib, tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ok" in out.stdout
