"""CPU checks of the command-line drivers' contract (no device needed):
the permutation hash against the reference's own `hash` (test_utils.h:12-19)
compiled in place, the input-file reader, and argument / input errors that
the drivers report before touching the device."""
import os
import subprocess

import numpy as np
import pytest

from semilocal_amd import harness

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HASH_CHECK = os.path.join(ROOT, "oracle", "_ref", "hash_check")
DRIVERS = ["semi_local_hip_iterative", "semi_local_hip_hybrid_iterative", "prefix_lcs_hip",
           "prefix_lcs_hip_bit_parallel", "braid_multiplication_hip"]

needs_ref = pytest.mark.skipif(not os.access(HASH_CHECK, os.X_OK),
                               reason="oracle/_ref/hash_check needs the reference tree to build")


@needs_ref
def test_driver_hash_equals_reference_hash():
    out = subprocess.run([HASH_CHECK], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("hash ok")


@needs_ref
@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (3, 2), (64, 3), (4099, 4)])
def test_python_hash_equals_reference_hash(tmp_path, n, seed):
    p = np.random.default_rng(seed).permutation(n)
    f = tmp_path / "perm.txt"
    f.write_text(" ".join(map(str, p.tolist())))
    out = subprocess.run([HASH_CHECK, str(f)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    assert int(out.stdout) == harness.permutation_hash(p)


def test_hash_small_closed_forms():
    # identity of size 3: h = 0, then 1, then R*1 + 2 (below M)
    assert harness.permutation_hash([0]) == 0
    assert harness.permutation_hash([0, 1]) == 1
    assert harness.permutation_hash([0, 1, 2]) == 4294967279 + 2
    # size 2 swapped: columns' rows are 1, 0 -> h = 1, then R*1 + 0
    assert harness.permutation_hash([1, 0]) == 4294967279


def test_sequence_file_round_trip(tmp_path):
    f = str(tmp_path / "x.fna")
    s = np.random.default_rng(5).integers(0, 26, 777)
    harness.write_sequence(f, "/data/some name", s)
    lines = open(f).read().split("\n")
    assert lines[0] == "/data/some name" and lines[1] == "777" and lines[2].count(",") == 776
    name, got = harness.read_sequence(f)
    assert name == "/data/some name" and np.array_equal(got, s)


def _built(name):
    path = os.path.join(harness.BIN, name)
    if not os.access(path, os.X_OK):
        pytest.fail(f"{name} is not built (python -m semilocal_amd.build)")
    return path


@pytest.mark.parametrize("name", DRIVERS)
def test_drivers_report_usage(name):
    out = subprocess.run([_built(name)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "usage" in out.stderr and out.stdout == ""


def test_drivers_reject_bad_inputs(tmp_path):
    bad = tmp_path / "bad.fna"
    bad.write_text("name\n5\n1,2,3,4")
    good = tmp_path / "good.fna"
    harness.write_sequence(str(good), "g", [0, 1, 1, 0])
    out = subprocess.run([_built("semi_local_hip_iterative"), "1", str(bad), str(good)],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "disagree" in out.stderr
    out = subprocess.run([_built("prefix_lcs_hip"), "1", str(tmp_path / "missing.fna"), str(good)],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "cannot open" in out.stderr
    tri = tmp_path / "tri.fna"
    harness.write_sequence(str(tri), "t", [0, 1, 2])
    out = subprocess.run([_built("prefix_lcs_hip_bit_parallel"), "1", str(tri), str(good)],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "only the symbols 0 and 1" in out.stderr
    out = subprocess.run([_built("braid_multiplication_hip"), "1", "0", "5"], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 1 and "n must be" in out.stderr
