"""Guards on the built HIP library's own device code (built here from this
repository's sources): the packed FP32 instructions that the build switches
off with -Xclang -target-feature -Xclang -packed-fp32-ops (the flag acts on
the device pass; the host pass only warns that it does not know it) must not
appear in any gfx950 code object of salvus_amd/_lib/libsalvus_sem.so."""
import shutil
import subprocess
from pathlib import Path

import pytest

from salvus_amd import sem

OBJDUMP = Path("/opt/rocm/llvm/bin/llvm-objdump")
PACKED = ("v_pk_fma_f32", "v_pk_mul_f32", "v_pk_add_f32")


@pytest.mark.skipif(not OBJDUMP.exists() or not sem.LIB_PATH.exists(),
                    reason="llvm-objdump or the built library is missing")
def test_no_packed_fp32_in_device_code(tmp_path):
    lib = tmp_path / "lib.so"
    shutil.copy(sem.LIB_PATH, lib)
    subprocess.run([str(OBJDUMP), "--offloading", str(lib)], check=True, capture_output=True)
    objs = sorted(tmp_path.glob("lib.so.*gfx950"))
    assert objs, "no gfx950 code object in the library"
    kernels = 0
    for o in objs:
        asm = subprocess.run([str(OBJDUMP), "-d", "--mcpu=gfx950", str(o)], check=True,
                             capture_output=True, text=True).stdout
        kernels += asm.count(">:\n")
        for name in PACKED:
            assert name not in asm, "%s in %s" % (name, o.name)
    assert kernels > 10
