"""The C ABI: header, ctypes binding and shared library agree (no GPU calls)."""
import ctypes as C
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

from polarism_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pgpe.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(pgpe_\w+)\s*\(", text, re.M)))


def test_binding_declares_exactly_the_header():
    assert header_functions() == N.declared_symbols()


def test_library_exports_every_declared_symbol():
    if not N.LIB_PATH.exists():
        pytest.fail("libpgpe_hip.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(str(N.LIB_PATH))
    missing = [name for name in header_functions() if not hasattr(lib, name)]
    assert not missing


def _c_layout(struct, fields):
    src = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void){",
           f'printf("%zu\\n", sizeof({struct}));']
    src += [f'printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as tmp:
        c = Path(tmp) / "layout.c"
        exe = Path(tmp) / "layout"
        c.write_text("\n".join(src))
        subprocess.run(["gcc", str(c), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("py,cname", [(N.Member, "pgpe_member"), (N.Rk4Args, "pgpe_rk4_args"),
                                     (N.SplitArgs, "pgpe_split_args"), (N.LawsonArgs, "pgpe_lawson_args"),
                                     (N.GemmArgs, "pgpe_gemm_args"), (N.AdaptiveState, "pgpe_adaptive_state"),
                                     (N.CommitList, "pgpe_commit_list")])
def test_ctypes_struct_layout_matches_c(py, cname):
    names = [f[0] for f in py._fields_]
    got = _c_layout(cname, names)
    assert got[0] == C.sizeof(py)
    for name, off in zip(names, got[1:]):
        assert getattr(py, name).offset == off, name
