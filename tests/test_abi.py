"""The C-ABI boundary, checked without a GPU: libsexgpu.so loads, exports every function
include/sexgpu.h declares (the reference's 13 entry points and the additions), and the
Python binding's list matches the header."""
import ctypes
import os
import re
import subprocess

import pytest

import sexgpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "sexgpu.h")
LIB = os.path.join(REPO, "gpu-sextractor_amd", "lib", "libsexgpu.so")

# the reference's entry points (src/cuda/*.h), in call order
REFERENCE_ABI = ["init_device", "back_estimate", "init_detection", "cudaFilter", "run_detection",
                 "init_deblend", "parcel_out", "clear_deblend", "analyse_gpu", "run_clean",
                 "endobject_gpu", "clear_detection", "clear_device"]


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    src = re.sub(r"^\s*#[^\n]*", "", src, flags=re.M)
    names = []
    for stmt in src.split(";"):
        stmt = stmt.strip()
        if "(" not in stmt or "typedef" in stmt or "{" in stmt.split("(")[0]:
            # a declaration may follow the closing brace of a typedef or enum
            stmt = stmt.split("}")[-1].strip()
            if "(" not in stmt or "typedef" in stmt:
                continue
        m = re.search(r"(\w+)\s*\(", stmt)
        if m and not stmt.startswith("return"):
            names.append(m.group(1))
    return names


def exported_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_reference_abi():
    names = header_functions()
    for f in REFERENCE_ABI:
        assert f in names, f
    assert len(names) == len(set(names))


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    syms = exported_symbols(LIB)
    missing = [f for f in header_functions() if f not in syms]
    assert not missing, missing


def test_library_loads_and_binding_matches_header():
    lib = ctypes.CDLL(LIB)
    declared = set(header_functions())
    for f in declared:
        assert hasattr(lib, f), f
    assert set(sexgpu.EXPORTED) == declared


def test_no_error_before_any_call():
    lib = ctypes.CDLL(LIB)
    lib.sxg_last_error.restype = ctypes.c_char_p
    assert lib.sxg_last_error() is None


def test_options_without_device():
    lib = sexgpu.gpu()
    assert lib.sxg_get_option(sexgpu.OPTIONS["deb_lds_small"]) >= 0
    assert lib.sxg_set_option(99, 1) == -1
    assert lib.sxg_get_option(99) == -1
    with pytest.raises(ValueError):
        sexgpu.set_option("deb_lds_large", 100000)
    # the level-parallel size: 0 (off) up to 2^30 pixels
    old = sexgpu.get_option("deb_glb_levels")
    try:
        assert old == 12288
        sexgpu.set_option("deb_glb_levels", 0)
        sexgpu.set_option("deb_glb_levels", 1 << 30)
        assert sexgpu.get_option("deb_glb_levels") == 1 << 30
        with pytest.raises(ValueError):
            sexgpu.set_option("deb_glb_levels", (1 << 30) + 1)
    finally:
        sexgpu.set_option("deb_glb_levels", old)


def test_outobj_layout():
    # the reference's outobjstruct with NAPER 1: 160 bytes, 16-byte aligned
    assert ctypes.sizeof(sexgpu.OutObj) == 160


def test_isorow_layout_matches_header(tmp_path):
    """The ctypes and numpy mirrors of sxg_isorow have the C layout: size and every
    field's offset, from a program compiled against include/sexgpu.h."""
    fields = [name for name, _ in sexgpu.IsoRow._fields_]
    body = "".join('printf(" %%zu", offsetof(sxg_isorow, %s));' % f for f in fields)
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "sexgpu.h"\n'
                   'int main(void) { printf("%%zu", sizeof(sxg_isorow)); %s return 0; }\n' % body)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(sexgpu.IsoRow) == sexgpu.ISOROW_DTYPE.itemsize
    assert got[1:] == [getattr(sexgpu.IsoRow, f).offset for f in fields]
    assert got[1:] == [sexgpu.ISOROW_DTYPE.fields[f][1] for f in fields]


REFSRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.isdir(REFSRC), reason="the reference's headers are not here")
def test_reference_struct_offsets(tmp_path):
    """endobject_gpu reads the reference front-end's prefstruct, obj2struct and wcsstruct
    at fixed offsets: a program compiled against the reference's own headers (src/prefs.h,
    src/types.h, NAPER 1, NAXIS 2) prints the same offsets, sizes and enum values."""
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "define.h"\n#include "globals.h"\n'
        '#include "prefs.h"\n#define O(s, f) printf(" %zu", offsetof(s, f))\n'
        'int main(void) { printf("%zu", sizeof(prefstruct)); O(prefstruct, naper);\n'
        'O(prefstruct, apert); O(prefstruct, autoparam); O(prefstruct, autoaper);\n'
        'O(prefstruct, seeing_fwhm); O(prefstruct, mag_zeropoint); O(prefstruct, mask_type);\n'
        'O(prefstruct, detect_type); printf(" %zu", sizeof(obj2struct));\n'
        'O(obj2struct, flux_aper); O(obj2struct, flux_auto); O(obj2struct, sprob);\n'
        'printf(" %d %d", (int)PHOTO, (int)MASK_CORRECT); printf(" %zu", sizeof(wcsstruct));\n'
        'O(wcsstruct, naxis); O(wcsstruct, ctype); O(wcsstruct, crval); O(wcsstruct, crpix);\n'
        'O(wcsstruct, cd); O(wcsstruct, longpole); O(wcsstruct, latpole); O(wcsstruct, lat);\n'
        'O(wcsstruct, lng); O(wcsstruct, equinox); printf("\\n"); return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-DHAVE_CONFIG_H", "-I", os.path.dirname(REFSRC), "-I", REFSRC,
                    "-I", os.path.join(REFSRC, "fits"), "-I", os.path.join(REFSRC, "wcs"), "-w",
                    "-o", str(exe), str(src)], check=True)
    want = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]
    out = (ctypes.c_long * 32)()
    n = sexgpu.gpu().sxg_reference_layout(out, 32)
    assert list(out[:n]) == want
