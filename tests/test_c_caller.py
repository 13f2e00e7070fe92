"""A C program written against the reference's header names ("s2let/s2let.h",
tests/c/drop_in_caller.c) builds with gcc against include/ and links
libs2let_amd.so in place of the reference's libraries.  On the CPU it runs the
host-side API (schedule, sizes, tiling identities, random inputs, powers,
diagnostics, the SO3 parameter bridge); on the GPU it runs the reference's
round trips (test/c/test_back_and_forth.c's states) through the drop-in
host-pointer entry points."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "drop_in_caller.c")
LIBDIR = os.path.join(ROOT, "s2let_amd")


def build(tmp_path, flat=False):
    exe = str(tmp_path / "drop_in_caller")
    inc = os.path.join(ROOT, "include")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + inc, SRC, "-o", exe,
           "-L" + LIBDIR, "-ls2let_amd", "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_reference_header_names_compile_flat(tmp_path):
    """The same headers reached as "s2let.h" with -I include/s2let (the reference's
    s2let_test.c style)."""
    src = tmp_path / "flat.c"
    src.write_text('#include "s2let.h"\nint main(void) { s2let_parameters_t p = {0}; p.L = 8; p.B = 2;\n'
                   '  so3_parameters_t q; fill_so3_parameters(&q, &p); return s2let_j_max(&p) == 3 ? 0 : 1; }\n')
    exe = str(tmp_path / "flat")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include", "s2let"),
                        str(src), "-o", exe, "-L" + LIBDIR, "-ls2let_amd", "-Wl,-rpath," + LIBDIR],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert subprocess.run([exe]).returncode == 0


def test_c_caller_host_api(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


@pytest.mark.gpu
def test_c_caller_transforms(tmp_path):
    exe = build(tmp_path)
    print("C caller, transforms", flush=True)
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
