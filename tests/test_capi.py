"""The C ABI: the library loads and exports every entry point the public
header declares (no device needed)."""
import ctypes
import os
import re

from ion_gym_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "ion_gym_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(ig_\w+)\(",
                                 txt, re.M)))


def test_header_lists_the_abi():
    names = header_functions()
    assert "ig_fly" in names and "ig_fields_upload" in names
    assert len(names) >= 18


def test_library_exports_every_header_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_header_symbol():
    assert sorted(_lib.SIGNATURES) == header_functions()


def test_abi_version_and_error_channel():
    L = _lib.lib()
    assert L.ig_abi_version() == 2
    # a null argument is refused with a message, without touching a device
    assert L.ig_device_count(None) == -1
    assert b"null" in L.ig_last_error()


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes struct has the size and field offsets the C compiler
    gives the header's struct (a mismatch would shift later fields)."""
    import ctypes as C
    import subprocess
    structs = {"ig_fields_desc": _lib.FieldsDesc,
               "ig_flight_params": _lib.FlightParams,
               "ig_ions": _lib.Ions, "ig_results": _lib.Results,
               "ig_tune": _lib.Tune, "ig_bases_desc": _lib.BasesDesc,
               "ig_births_desc": _lib.BirthsDesc}
    lines = ["#include <stdio.h>", "#include <stddef.h>",
             '#include "ion_gym_amd.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", '
                         f'offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    want = {}
    for ln in out:
        if ln.strip():
            a, b, c = ln.split()
            want[(a, b)] = int(c)
    for cname, py in structs.items():
        assert C.sizeof(py) == want[(cname, "size")], cname
        for fname, _ in py._fields_:
            assert getattr(py, fname).offset == want[(cname, fname)], \
                (cname, fname)
