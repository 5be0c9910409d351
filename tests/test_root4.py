"""The fourth root of the Runge-Kutta-Fehlberg step control
(flowvc-1_amd/csrc/fvc_root4.h, the reference takes pow(x, 0.25),
ompVC/integration.c:615-618) against the C library's pow on 47M inputs: the
two agree except where the root lies within a hair of a tie between two
doubles, and at each such input the header's result is the correctly rounded
one (checked with 60-digit decimal arithmetic)."""
import os
import subprocess
from decimal import Decimal, getcontext

import cases


def test_root4_is_correctly_rounded_and_agrees_with_libm(tmp_path):
    exe = os.path.join(str(tmp_path), "root4_check")
    src = os.path.join(cases.REPO, "tests", "native", "root4_check.c")
    inc = os.path.join(cases.REPO, "flowvc-1_amd", "csrc")
    subprocess.run(["gcc", "-O2", "-std=c11", "-I" + inc, src, "-o", exe, "-lm"], check=True)
    out = subprocess.run([exe, "20000000"], capture_output=True, text=True, check=True).stdout.splitlines()
    n, bad = (int(w) for w in out[-1].split()[0::2][:2])
    assert n > 40_000_000
    assert bad < n // 1000, (bad, n)
    getcontext().prec = 60
    checked = 0
    for line in out[:-1]:
        x, mine, libm = (float.fromhex(w) for w in line.split()[1:])
        root = Decimal(x).sqrt().sqrt()
        assert abs(Decimal(mine) - root) < abs(Decimal(libm) - root), line
        checked += 1
    assert checked == min(bad, 200)
