"""The ctypes mirror of the C ABI (lbmpond_amd/pond.py) against include/pond.h itself: a C
program compiled against the header prints the size and every field offset of pond_params,
pond_diag and pond_timing, which must equal the ctypes structures'.  No GPU is touched."""
import ctypes
import json
import os
import shutil
import subprocess

import pytest

from lbmpond_amd import pond as pondmod

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {"pond_params": pondmod.PondParams, "pond_diag": pondmod.PondDiag, "pond_timing": pondmod.PondTiming}


def c_layout(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "pond.h"', "int main(void) {", 'printf("{");']
    first = True
    for cname, cls in STRUCTS.items():
        for fname, _ in cls._fields_:
            sep = "" if first else ","
            first = False
            lines.append(f'printf("{sep}\\"{cname}.{fname}\\": %zu", offsetof({cname}, {fname}));')
        lines.append(f'printf(",\\"{cname}.sizeof\\": %zu", sizeof({cname}));')
    lines += ['printf("}\\n");', "return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_ctypes_structs_match_header(tmp_path):
    got = c_layout(tmp_path)
    for cname, cls in STRUCTS.items():
        for fname, _ in cls._fields_:
            assert got[f"{cname}.{fname}"] == getattr(cls, fname).offset, (cname, fname)
        assert got[f"{cname}.sizeof"] == ctypes.sizeof(cls), cname


def test_header_fields_are_all_mirrored():
    """every field the header declares in the three structs appears in the ctypes mirror"""
    import re
    txt = open(os.path.join(ROOT, "include", "pond.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    for cname, cls in STRUCTS.items():
        body = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + r"\s*;", txt, flags=re.S).group(1)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # "type a, b, c": the first name follows the type, the rest follow commas
            parts = [p.strip() for p in decl.split(",")]
            names.append(parts[0].split()[-1])
            names += parts[1:]
        assert names == [f for f, _ in cls._fields_], (cname, names)
