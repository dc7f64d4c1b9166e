"""CPU check of the matrix-core kernels' lane exchanges in the compiled
gfx950 code: every DPP instruction of GridKernelMfma and every lane swap
(v_permlane16/32_swap) of DegridKernelMfma must run with the full EXEC
mask.  A DPP move reads its source lane even when that lane is masked off,
and then reads it as invalid (zero with bound_ctrl), so a compiler that
turned a per-lane select around an exchange into a masked branch would
give wrong sums without any error.  The nearest instruction before each
such op that writes EXEC must therefore be a restoring s_or_b64 exec (or
there must be none).  Compiles grid.hip to assembly with the library's own
flags (hipcc, no GPU needed).
"""
import os
import re
import shutil
import subprocess

import pytest

from wsclean_amd import build

HIPCC = build.HIPCC


def _exec_writer(line):
    # save-exec forms always write EXEC; other scalar ops write it when it is
    # their destination (first operand)
    return re.match(r"s_\w+_saveexec_b64 ", line) is not None or \
        re.match(r"s_\w+_b64 exec,", line) is not None


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None,
                    reason="hipcc not available")
def test_lane_exchanges_run_with_full_exec(tmp_path):
    out = tmp_path / "grid.s"
    cmd = [HIPCC] + build.COMMON + ["--cuda-device-only", "-S",
                                    os.path.join(build.CSRC, "grid.hip"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    text = out.read_text().split("\n")
    kernels = {}
    name = None
    for ln in text:
        m = re.match(r"^(_Z\w*(?:GridKernelMfma|DegridKernelMfma)\w*):", ln)
        if m:
            name = m.group(1)
            kernels[name] = []
        elif name and ln.startswith(".Lfunc_end"):
            name = None
        elif name:
            kernels[name].append(ln.strip())
    assert any("GridKernelMfma" in k for k in kernels), "no GridKernelMfma in the assembly"
    assert any("DegridKernelMfma" in k for k in kernels), "no DegridKernelMfma in the assembly"
    checked = 0
    for name, body in kernels.items():
        for i, t in enumerate(body):
            if "_dpp" not in t and "permlane" not in t:
                continue
            checked += 1
            last = next((u for u in reversed(body[:i]) if _exec_writer(u)), None)
            assert last is None or last.startswith("s_or_b64 exec, exec"), (name, t, last)
    assert checked > 0
