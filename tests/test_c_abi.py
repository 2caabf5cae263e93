"""The C ABI as a C program sees it: tests/c_abi/abi_check.c includes only
include/qlat_dwf.h, is compiled with warnings as errors and linked to the
built library.  On the CPU it exercises the error paths (no GPU call); on
the GPU it creates an inverter, checks Mpc^+ Mpc, reproduces the
reference's free-field propagator golden through invert_prop and destroys
the inverter."""

import os
import subprocess

import pytest

from qlattice_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_check(tmp_path):
    exe = str(tmp_path / "abi_check")
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-O1",
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c_abi", "abi_check.c"),
           "-o", exe, "-L", libdir, "-l:libqlat_dwf_hip.so", f"-Wl,-rpath,{libdir}", "-lm"]
    subprocess.run(cmd, check=True)
    return exe


def test_c_program_compiles_against_header_and_error_paths(tmp_path):
    exe = build_check(tmp_path)
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK cpu" in r.stdout


@pytest.mark.gpu
def test_c_program_on_the_gpu(tmp_path):
    print("\n[gpu] C program through the C ABI: 4^3x8 free-field propagator", flush=True)
    exe = build_check(tmp_path)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "OK gpu" in r.stdout
