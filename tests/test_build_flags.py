"""The library's hipcc flags keep packed-FP32 instructions out of gfx950 code.

hipcc passes `-Xclang -target-feature -Xclang -packed-fp32-ops` to both
compilation passes; the host pass ignores it with a warning, the device
pass applies it.  A device-only build of a small float2 kernel shows the
difference (compile only, no GPU needed)."""

import os
import shutil
import subprocess

import pytest

from clifft_amd import build as B

KERNEL = r"""
#include <hip/hip_runtime.h>
__global__ void axpb(float2* a, float2 s) {
    int i = threadIdx.x;
    a[i].x = a[i].x * s.x + s.y;
    a[i].y = a[i].y * s.x + s.y;
}
"""


def device_asm(tmp_path, flags):
    src = tmp_path / "k.hip"
    src.write_text(KERNEL)
    out = tmp_path / "k.s"
    subprocess.run([B.HIPCC, *flags, "--cuda-device-only", "-S", "-o", str(out), str(src)], check=True,
                   capture_output=True)
    return out.read_text()


@pytest.mark.skipif(not os.path.exists(B.HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_packed_fp32_off_in_device_code(tmp_path):
    with_flag = device_asm(tmp_path, [f for f in B.HIP_FLAGS if f != "-fPIC"])
    assert "v_fma_f32" in with_flag and "v_pk_fma_f32" not in with_flag
    # the same kernel without the feature flag does get the packed form
    flags = [f for f in B.HIP_FLAGS if f != "-fPIC"]
    i = flags.index("-packed-fp32-ops")
    without = device_asm(tmp_path, flags[:i - 3] + flags[i + 1:])
    assert "v_pk_fma_f32" in without
