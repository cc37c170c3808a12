"""Checks on the built library's gfx950 code (CPU only, no GPU calls).

The FP64 multimode units are compiled with every memory wait forced to zero
(``-amdgpu-waitcnt-forcezero``).  That option also puts an ``s_waitcnt``
between ``s_getpc_b64`` and the ``s_add_u32`` that adds a PC-relative offset
computed for the instruction right after the ``s_getpc_b64``.  Such a split
sequence addresses 4 or 8 bytes off: a call through it jumps into the middle
of an instruction (the adaptive FP64 kernels faulted with an illegal
instruction), and a constant table read through it loads shifted data.  The
kernels are written so that no such sequence is emitted (the adaptive driver
is inlined and reads its tableau through selects over immediates); this test
disassembles every code object of libfiberhip.so and fails on any split
sequence.
"""
from __future__ import annotations

import concurrent.futures
import pathlib
import re
import shutil
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB = ROOT / "fiber-raman-suppression_amd" / "libfiberhip.so"
OBJDUMP = pathlib.Path("/opt/rocm/llvm/bin/llvm-objdump")
GETPC = re.compile(r"^s_getpc_b64 s\[(\d+):\d+\]")


def _split_sequences(code_object: pathlib.Path) -> list[str]:
    """Instructions that follow an s_getpc_b64 s[n:n+1] other than s_add_u32 sn, sn, ..."""
    proc = subprocess.Popen([str(OBJDUMP), "-d", "--mcpu=gfx950", "--no-show-raw-insn", "--no-leading-addr",
                             str(code_object)], stdout=subprocess.PIPE, text=True)
    bad, want = [], None
    for line in proc.stdout:
        text = line.strip()
        if not text or text.endswith(":") or text.startswith(("Disassembly", ";")):
            continue
        if want is not None:
            if not text.startswith(want):
                bad.append(text.split("//")[0].strip())
            want = None
        m = GETPC.match(text)
        if m:
            want = f"s_add_u32 s{m.group(1)}, s{m.group(1)}, "
    proc.wait()
    assert proc.returncode == 0, f"llvm-objdump failed on {code_object.name}"
    return bad


@pytest.mark.skipif(not LIB.exists() or not OBJDUMP.exists(), reason="library or llvm-objdump not present")
def test_no_split_pc_relative_sequences(tmp_path):
    lib = tmp_path / LIB.name
    shutil.copy(LIB, lib)
    subprocess.run([str(OBJDUMP), "--offloading", str(lib)], check=True, capture_output=True, cwd=tmp_path)
    objects = sorted(tmp_path.glob(f"{LIB.name}.*gfx950"))
    assert objects, "no gfx950 code objects in the library"
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as pool:
        found = dict(zip(objects, pool.map(_split_sequences, objects)))
    bad = {p.name: v[:3] for p, v in found.items() if v}
    assert not bad, f"s_getpc_b64 not followed by its s_add_u32: {bad}"
