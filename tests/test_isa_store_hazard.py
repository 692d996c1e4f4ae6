"""ISA check of the device code (CPU only: hipcc cross-compiles gfx950 to assembly).

A VMEM store of more than 8 bytes reads its data VGPRs after it issues. For gfx950 the compiler
does not keep a VALU write of those VGPRs away from the store, and such a write one instruction
after a 16-byte store in the row kernel corrupted stored values now and then (batch members and
4096^2 steps that differed run to run). The kernels therefore store at most 8 bytes per lane per
instruction where stores run back to back; this test compiles every kernel source and fails if
a store of 12 or 16 bytes is followed within two instructions by a VALU write of its data.
"""

import importlib.util
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location(
    "s2t_build", os.path.join(ROOT, "spectral-2d-turbulence_amd", "build.py"))
build = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(build)

WIDE_STORE = re.compile(r"^(buffer|global|flat)_store_dwordx[34]\s")


def _vgprs(operand):
    m = re.match(r"v\[(\d+):(\d+)\]", operand)
    if m:
        return set(range(int(m.group(1)), int(m.group(2)) + 1))
    m = re.match(r"v(\d+)$", operand)
    return {int(m.group(1))} if m else set()


def close_writes(asm, window=2):
    """(kernel, store, writer) for VALU writes of a wide store's data within `window`
    instructions after it."""
    found = []
    funcs = set(re.findall(r"\.type\s+(_Z\S+?),\s*@function", asm))
    for name in re.findall(r"^(_Z\S+):", asm, re.M):
        if funcs and name not in funcs:
            continue  # a data object (the device arrays), not a kernel
        i = asm.index(name + ":")
        body = [l.strip() for l in asm[i:asm.index(".Lfunc_end", i)].splitlines()]
        body = [l for l in body if l and not l.startswith((";", "."))]
        for k, ins in enumerate(body):
            if not WIDE_STORE.match(ins):
                continue
            data = _vgprs(ins.split()[1].rstrip(","))
            for nxt in body[k + 1:k + 1 + window]:
                if nxt.startswith("v_") and _vgprs(nxt.split()[1].rstrip(",")) & data:
                    found.append((name, ins, nxt))
                    break
    return found


def test_detector_sees_the_pattern():
    asm = ("_Zk:\n\tbuffer_store_dwordx4 v[22:25], v28, s[20:23], s45 offen\n"
           "\tv_sub_u32_e32 v22, 0xe00, v4\n.Lfunc_end0:\n"
           "_Zj:\n\tbuffer_store_dwordx4 v[22:25], v28, s[20:23], s45 offen\n"
           "\ts_or_b64 exec, exec, s[0:1]\n\ts_nop 0\n\tv_mov_b32_e32 v23, 0\n.Lfunc_end1:\n")
    assert [f[0] for f in close_writes(asm)] == ["_Zk"]


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("src", ["step_kernels.hip", "aux_kernels.hip"])
def test_no_write_right_after_wide_store(tmp_path, src):
    out = tmp_path / (src + ".s")
    flags = [f for f in build.HIP_FLAGS if f != "-fPIC"] + build.FILE_FLAGS.get(src, [])
    proc = subprocess.run([build.HIPCC, *flags, "--cuda-device-only", "-S",
                           os.path.join(build.CSRC, src), "-o", str(out)],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert proc.returncode == 0, proc.stdout[-2000:]
    hits = close_writes(out.read_text())
    assert not hits, "\n".join("%s: %s -> %s" % h for h in hits[:10])
