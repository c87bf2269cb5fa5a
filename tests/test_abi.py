"""The C-ABI library loads on a machine without a GPU and exports every entry point the
public headers declare; host-only argument validation works without touching a device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]
LIB = os.path.join(ROOT, "spacetime_amd", "libgpe.so")


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(gpe_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from spacetime_amd import build
        build.build()
    return ctypes.CDLL(LIB)


def test_headers_declare_the_core_api():
    names = declared_functions()
    for n in ("gpe_create", "gpe_destroy", "gpe_set_state", "gpe_get_state", "gpe_step", "gpe_itp",
              "gpe_observables", "gpe_start", "gpe_set_ramp", "gpe_set_field", "gpe_last_error"):
        assert n in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_lattice_values_host_only(lib):
    assert [lib.gpe_lattice_values(i) for i in range(5)] == [1, 12, 12, 46, 136]
    assert lib.gpe_lattice_values(7) == -22


def test_field_width_host_only(lib):
    assert [lib.gpe_field_width(k) for k in range(4)] == [1, 3, 1, 7]
    assert lib.gpe_field_width(9) == -22


def test_create_rejects_bad_config_without_device(lib):
    from spacetime_amd import gpe
    cfg = gpe.Config(lattice="bcc", dims=(0, 4, 4)).to_c()
    h = ctypes.c_void_p()
    lib.gpe_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    assert lib.gpe_create(ctypes.byref(cfg), ctypes.byref(h)) == -22
    cfg = gpe.Config(lattice="bcc", ncomp=2).to_c()
    assert lib.gpe_create(ctypes.byref(cfg), ctypes.byref(h)) == -22
    assert not h.value


def test_create_rejects_line_counts_past_32_bits_without_device(lib):
    # the kernels number lines (z ny + y) V + v in 32 bits: a box with nz ny V >= 2^31 lines must
    # be refused before any allocation (FP32 scalar BCC 2 x 13400 x 13400 blocks, ~69 GB, would
    # otherwise fit on the device and index out of bounds)
    from spacetime_amd import gpe
    lib.gpe_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h = ctypes.c_void_p()
    cfg = gpe.Config(lattice="bcc", dims=(2, 13400, 13400), ncomp=1, precision=gpe.FP32).to_c()
    assert lib.gpe_create(ctypes.byref(cfg), ctypes.byref(h)) == -22
    assert not h.value


def test_header_compiles_as_c_and_every_function_links(tmp_path, lib):
    """include/gpe.h is plain C: a C99 file that takes the address of every declared function
    compiles and links against libgpe.so (the C linkage a cgo / ctypes / JNI binding needs)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    names = declared_functions()
    src = tmp_path / "abi_check.c"
    src.write_text('#include "gpe.h"\n#include <stdio.h>\nint main(void) {\n    void *f[] = {\n' +
                   "".join("        (void *)%s,\n" % n for n in names) +
                   '    };\n    printf("%d\\n", (int)(sizeof f / sizeof f[0]));\n    return 0;\n}\n')
    out = tmp_path / "abi_check"
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-pedantic", "-Wno-pedantic", "-I" + os.path.join(ROOT, "include"),
                    str(src), "-L" + os.path.dirname(LIB), "-lgpe", "-o", str(out)], check=True)
    assert len(names) >= 58


def test_cxx_driver_compiles_and_links(tmp_path, lib):
    """The C++ host driver (tests/cxx/gpe_driver.cpp) builds against the header and the library
    alone; the GPU test runs the copy that the build step makes."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    out = tmp_path / "gpe_driver"
    subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cxx", "gpe_driver.cpp"), "-L" + os.path.dirname(LIB), "-lgpe",
                    "-o", str(out)], check=True)
    assert out.exists()
