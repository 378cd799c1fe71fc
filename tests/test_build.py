"""Properties of the gfx950 code the library is built into (CPU only: hipcc
cross-compiles the device code here).

No kernel may call a device function.  A helper the compiler outlines gets
its LDS buffers as generic pointers and reaches them with flat loads and
stores; in one such outlined copy of the median pyramid the barrier opening
each level was emitted without a wait for the previous level's stores, and
results changed from run to run.  Every helper is __forceinline__; this test
keeps it that way."""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "heimdall_multibeam_amd", "csrc")
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _device_asm(src, out_dir):
    from heimdall_multibeam_amd.build import HIPCC, HIP_FLAGS
    out = os.path.join(out_dir, src + ".s")
    cmd = [HIPCC] + HIP_FLAGS + ["--cuda-device-only", "-S", "-x", "hip",
                                 os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    with open(out) as f:
        return f.read()


@pytest.fixture(scope="module")
def asm():
    from heimdall_multibeam_amd.build import HIPCC
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tempfile.mkdtemp(prefix="hd_isa_")
    try:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
            texts = list(ex.map(lambda s: _device_asm(s, d), SOURCES))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return dict(zip(SOURCES, texts))


def test_no_kernel_calls_a_device_function(asm):
    for src, text in asm.items():
        calls = len(re.findall(r"\bs_swappc_b64\b", text))
        assert calls == 0, "%s: %d device-function calls" % (src, calls)
        # every function in the code object is a kernel (ends in s_endpgm)
        assert "s_setpc_b64" not in text, "%s: a called (returning) function" % src


def test_every_source_targets_gfx950(asm):
    for src, text in asm.items():
        assert 'amdgcn-amd-amdhsa--gfx950' in text, src
