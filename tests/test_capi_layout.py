"""The ctypes mirrors of the C-ABI argument structs match include/vl1.h:
a small C program compiled with gcc prints sizeof and every field's
offsetof, and each must equal the ctypes structure's.  A field added to the
header but not to the mirror (or in another order) shifts every later
offset and fails here, without a GPU."""

import shutil
import subprocess
from pathlib import Path

import pytest

from volumential_amd import _capi

ROOT = Path(__file__).resolve().parents[1]

STRUCTS = {
    "vl1_payload_desc": _capi.PayloadDesc,
    "vl1_dense_payload_desc": _capi.DensePayloadDesc,
    "vl1_near_field_args": _capi.NearFieldArgs,
    "vl1_plan_desc": _capi.PlanDesc,
    "vl1_helmholtz_split_blocks": _capi.HelmholtzSplitBlocks,
    "vl1_split_opts": _capi.SplitOpts,
    "vl1_duffy_args": _capi.DuffyArgs,
    "vl1_split_args": _capi.SplitArgs,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    d = tmp_path_factory.mktemp("layout")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "vl1.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.splitlines():
        cname, fname, val = line.split()
        table[(cname, fname)] = int(val)
    return table


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_struct_matches_header(c_layout, cname):
    py = STRUCTS[cname]
    assert c_layout[(cname, "sizeof")] == _capi.C.sizeof(py), f"{cname}: size differs"
    for fname, _ in py._fields_:
        assert c_layout[(cname, fname)] == getattr(py, fname).offset, f"{cname}.{fname}"
