"""The reference-named C API (include/hawaii_ref.h, libhawaii.so) on the CPU: every function and
variable the header declares is exported, a program written like the reference's
examples/correlation_co2_ar.cpp compiles as C99 against the header and links against the
library, and without a GPU its first engine call fails loudly (the reference's exit(1)), with
no CPU fallback."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "hawaiihybrid_amd")
SRC = os.path.join(REPO, "tests", "ref_api", "correlation_co2_ar.c")


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    lines = "\n".join(l for l in text.splitlines() if not l.startswith("typedef"))
    funcs = set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", lines, flags=re.M))
    variables = set(re.findall(r"^extern\s+\w+\s+(\w+);", text, flags=re.M))
    return funcs, variables


def exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def build_program(tmp_path):
    exe = str(tmp_path / "correlation_co2_ar")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(REPO, "include"), SRC, "-L" + PKG,
           "-lhawaii", "-lhawaii_gpu", "-Wl,-rpath," + PKG, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_library_exports_every_declared_name():
    funcs, variables = declared("hawaii_ref.h")
    assert {"init_ms", "calculate_correlation_and_save", "idct_cf_to_sf", "c_mpi_perform_integration",
            "calculate_correlation_array_and_save", "free_cfnc"} <= funcs
    assert variables == {"dipole_1", "dipole_2", "free_dipole_1", "free_dipole_2", "pes", "dpes"}
    have = exported(os.path.join(PKG, "libhawaii.so"))
    missing = sorted((funcs | variables) - have)
    assert not missing, missing


def test_examples_style_program_links(tmp_path):
    exe = build_program(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the program would run")
def test_program_fails_loudly_without_a_device(tmp_path):
    exe = build_program(tmp_path)
    r = subprocess.run([exe, "21000", "100000", "42", "64", "8", "1", "10", str(tmp_path / "cf.txt"),
                        str(tmp_path / "sf.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "ERROR" in r.stderr
    assert not os.path.exists(tmp_path / "cf.txt")
