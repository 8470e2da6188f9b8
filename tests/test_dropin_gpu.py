"""Runs the drop-in check program (oracle/_ref/dropin_check, built from
tests/cpp/dropin_check.cpp against the reference's headers): every reference
call site and the same call site switched to the engine must agree."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "oracle", "_ref", "dropin_check")


@pytest.mark.skipif(not os.path.exists(PROG), reason="drop-in check program not built (needs the reference tree)")
def test_reference_call_sites_switched_to_the_engine_agree(gpu):
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=600)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1].endswith(" 0 failures")
