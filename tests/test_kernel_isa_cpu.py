"""The combing kernels compile to straight-line queue loops: no function calls
and no scratch.  comb_pass and the per-pair tile-height dispatch are forced
inline; a build in which the compiler outlined them (pair descriptor passed
through a generic pointer to private memory) is not one this engine runs.
Compiles comb.hip for gfx950 to assembly (hipcc, no GPU needed) and reads the
code-object metadata."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_combing_kernels_have_no_calls_and_no_scratch(tmp_path):
    out = tmp_path / "comb.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"),
                    "-o", str(out), os.path.join(ROOT, "semilocal_amd", "csrc", "comb.hip")],
                   check=True, capture_output=True)
    text = out.read_text()
    assert "s_swappc" not in text
    kernels = re.findall(r"\.name:\s+(_ZN5slhip17comb_batch_kernel\S+)\n(?:.*\n)*?\s+\.private_segment_fixed_size:\s+(\d+)",
                         text)
    # one instantiation per symbol width (1, 2 bytes) and tile set (per pair, 4, 8, 16)
    assert len(kernels) == 8, kernels
    assert all(int(size) == 0 for _, size in kernels), kernels
