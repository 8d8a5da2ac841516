"""The ctypes mirrors of the C ABI's structs (iwpp-mp_amd/iwpp/__init__.py)
against include/iwpp.h: each struct's size and every mirrored field's offset
as the C compiler lays them out, from a probe program compiled with gcc.  A
field added to the header and not to the mirror (or the other way round, or
in another order) changes a size or an offset."""
import ctypes
import os
import shutil
import subprocess

import pytest

import iwpp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {
    "iwpp_options": iwpp.Options,
    "iwpp_stats": iwpp.Stats,
    "iwpp_image": iwpp.ImageDesc,
    "iwpp_tuning": iwpp.Tuning,
}


def _probe(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "iwpp.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "probe"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_ctypes_mirrors_match_header(tmp_path):
    c = _probe(tmp_path)
    for cname, cls in STRUCTS.items():
        for field, _ in cls._fields_:
            assert c[f"{cname}.{field}"] == getattr(cls, field).offset, (cname, field)
        assert c[cname] == ctypes.sizeof(cls), cname
