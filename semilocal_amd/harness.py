"""Python side of the command-line drivers' contract (semilocal_amd/bin/):
the reference's input files and stdout lines, so scripts can write inputs,
run a driver and read its result the way the reference's harness does
(utils/runners.py:16-74).

- input file (parsers.h:38-47): line 1 the name, line 2 the length, line 3
  the symbols as comma-separated integers;
- combing / LCS stdout: preprocess ms, algorithm ms, hash or score, the two
  sizes, the two names (utils/runners.py:36-74);
- braid-multiplication stdout: preprocess ms, algorithm ms, hash, n, seed
  (utils/runners.py:16-31).
"""
from __future__ import annotations

import os
import subprocess
from dataclasses import dataclass

import numpy as np

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")
_R = 4294967279
_M = 4294967291
_MASK = (1 << 64) - 1


def driver_path(name: str) -> str:
    """Path of a built driver; raises if it has not been built."""
    path = os.path.join(BIN, name)
    if not os.access(path, os.X_OK):
        raise FileNotFoundError(f"driver {name} is not built (python -m semilocal_amd.build)")
    return path


def write_sequence(path: str, name: str, symbols) -> None:
    """Writes one input file in the reference's three-line format."""
    s = np.asarray(symbols).astype(np.int64).ravel()
    with open(path, "w") as f:
        f.write(f"{name}\n{s.size}\n{','.join(map(str, s.tolist()))}")


def read_sequence(path: str) -> tuple[str, np.ndarray]:
    """Reads one input file: (name, symbols as int32).  Every listed symbol is
    kept, including the last one."""
    with open(path) as f:
        name = f.readline().rstrip("\n")
        length = int(f.readline())
        line = f.readline().strip()
    sym = np.array([int(t) for t in line.split(",") if t.strip()], np.int32) if line else np.zeros(0, np.int32)
    if sym.size != length:
        raise ValueError(f"{path}: length {length} but {sym.size} symbols")
    return name, sym


def permutation_hash(row_to_col) -> int:
    """The reference's permutation hash (test_utils.h:12-19) as the compiled
    reference computes it: h = (R h + P^-1[i]) mod M over the columns, with
    R h wrapped to 64-bit two's complement and C's truncating remainder (so
    the value can be negative)."""
    p = np.asarray(row_to_col, np.int64)
    inv = np.empty(p.size, np.int64)
    inv[p] = np.arange(p.size)
    h = 0
    for c in inv.tolist():
        t = (_R * h + c) & _MASK
        if t >= 1 << 63:
            t -= 1 << 64
        r = abs(t) % _M
        h = r if t >= 0 else -r
    return h


@dataclass
class CombingResult:
    preprocess_ms: int
    algo_ms: int
    value: int          # the permutation hash (combing drivers) or the score (LCS drivers)
    a_size: int
    b_size: int
    a_name: str
    b_name: str


@dataclass
class BraidResult:
    preprocess_ms: int
    algo_ms: int
    hash: int
    n: int
    seed: int


def _run(cmd: list[str], timeout: float) -> str:
    proc = subprocess.run(cmd, timeout=timeout, stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    if proc.returncode != 0 or proc.stderr != b"":
        raise RuntimeError(f"{os.path.basename(cmd[0])} failed ({proc.returncode}): {proc.stderr.decode()}")
    return proc.stdout.decode()


def run_combing(name: str, threads: int, a_file: str, b_file: str, timeout: float = 1200) -> CombingResult:
    """Runs a combing or LCS driver and parses its seven lines the way the
    reference's CombingRunner / LcsRunner do (utils/runners.py:36-74)."""
    pre, algo, value, a_size, b_size, a_name, b_name = _run(
        [driver_path(name), str(threads), a_file, b_file], timeout).split("\n")
    return CombingResult(int(pre), int(algo), int(value), int(a_size), int(b_size), a_name, b_name)


def run_braid(depth: int, n: int, seed: int, dump: str | None = None, timeout: float = 1200) -> BraidResult:
    """Runs braid_multiplication_hip and parses its five lines the way the
    reference's BraidMulRunner does (utils/runners.py:16-31)."""
    cmd = [driver_path("braid_multiplication_hip"), str(depth), str(n), str(seed)] + ([dump] if dump else [])
    pre, algo, h, nn, sd = _run(cmd, timeout).split("\n")
    return BraidResult(int(pre), int(algo), int(h), int(nn), int(sd))
