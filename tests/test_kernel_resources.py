"""Kernel resources of the built library (no GPU calls): no kernel keeps private arrays in scratch.

A private array that the compiler cannot place in registers (a struct copied through memcpy, a
run-time index, a store merged through a selected pointer) becomes per-lane scratch memory:
every access is a vector memory round trip. The gfx950 code objects of libpgpe_hip.so carry,
per kernel, the private segment size and the register spill counts in their metadata notes;
this test reads them from the offload bundles of the library's .hip_fatbin section.
"""
import os
import re
import shutil
import struct
import subprocess
import tempfile
from pathlib import Path

import pytest

from polarism_amd import _native as N

LLVM = Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "llvm" / "bin"
BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"

# kernel (mangled name) -> allowed private segment bytes per lane, with the reason
ALLOWED = {
    # 32 elements per thread on long columns: SGPR spill slots only (no array, no VGPR spill)
    "_ZN4pgpe18lawson_cols_kernelIfLi32ELi1EEEvNS_12LawsonParamsIT_EE": 32,
}


def _kernels(lib: Path):
    """(name, private bytes, VGPR spills, SGPR spills) of every gfx950 kernel in the library."""
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        fatbin = Path(tmp) / "fatbin"
        subprocess.run([str(LLVM / "llvm-objcopy"), "--dump-section", f".hip_fatbin={fatbin}", str(lib),
                        str(Path(tmp) / "stripped")], check=True, capture_output=True)
        data = fatbin.read_bytes()
        at = data.find(BUNDLE_MAGIC)
        while at >= 0:
            # bundle header: magic, entry count, then per entry offset, size, triple length, triple
            (entries,) = struct.unpack_from("<Q", data, at + len(BUNDLE_MAGIC))
            pos = at + len(BUNDLE_MAGIC) + 8
            for _ in range(entries):
                off, size, tlen = struct.unpack_from("<QQQ", data, pos)
                pos += 24
                triple = data[pos:pos + tlen].decode()
                pos += tlen
                if "gfx950" not in triple or size == 0:
                    continue
                co = Path(tmp) / "code_object"
                co.write_bytes(data[at + off:at + off + size])
                notes = subprocess.run([str(LLVM / "llvm-readelf"), "--notes", str(co)], check=True,
                                       capture_output=True, text=True).stdout
                for block in re.split(r"\n\s+- \.agpr_count:", notes)[1:]:
                    field = lambda key: re.search(r"\." + key + r":\s+(\S+)", block).group(1)
                    out.append((field("name"), int(field("private_segment_fixed_size")),
                                int(field("vgpr_spill_count")), int(field("sgpr_spill_count"))))
            at = data.find(BUNDLE_MAGIC, at + 1)
    return out


@pytest.fixture(scope="module")
def kernels():
    if not N.LIB_PATH.exists():
        pytest.fail("libpgpe_hip.so is not built (run __graft_entry__.build())")
    if not (LLVM / "llvm-readelf").exists() or not (LLVM / "llvm-objcopy").exists() or not shutil.which("c++filt"):
        pytest.skip("LLVM binary tools not found")
    return _kernels(N.LIB_PATH)


def test_library_holds_the_hot_kernels(kernels):
    names = {k[0] for k in kernels}
    assert len(names) > 500
    for stem in ("rk4_strip_pf_kernel", "rk4_step_kernel", "fft_rows_kernel", "fft_cols_half_kernel",
                 "lawson_cols_kernel", "gemm_bf16_big_kernel"):
        assert any(stem in n for n in names), stem


def test_no_kernel_keeps_private_arrays_or_spills_vector_registers(kernels):
    bad = [(name, priv, vsp) for name, priv, vsp, _ in kernels
           if vsp > 0 or priv > ALLOWED.get(name, 0)]
    if bad:
        demangled = subprocess.run(["c++filt"], input="\n".join(b[0] for b in bad), capture_output=True,
                                   text=True).stdout.splitlines()
        lines = [f"{d}: private {p} B/lane, VGPR spills {v}" for d, (_, p, v) in zip(demangled, bad)]
        pytest.fail("kernels using scratch memory:\n" + "\n".join(lines))
