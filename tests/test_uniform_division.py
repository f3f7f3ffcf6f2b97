"""The launch-uniform division of the fused joint kernels against IEEE FP32 division (CPU).

The fused joint iteration divides by alpha_tv + alpha_proj and lambdaf + alpha_proj, which are
the same for the whole launch, with one FMA on the binary64 reciprocal instead of the per-element
IEEE sequence (udiv in reconflow_amd/csrc/rf_device.hpp). `udiv_check.c` restates that sequence on
the host and compares it bit for bit with `n / d` in FP32 for every numerator of fourteen
binades (all subnormals included) of both signs, and the special values, for each divisor below:
the workloads' sums, divisors whose subnormal quotients include exact rounding ties (even
integers such as 98, where the unsnapped binary64 product misrounds ties), powers of two,
extremes of the FP32 range. The checker also counts the mismatches of two plausible mistakes (no
snap to the 2^-176 grid, no zero-sign fix) to show that it sees them.
"""
import os
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent

DIVISORS = [
    float(np.float32(33.3) + np.float32(60.0)),   # alpha_tv + alpha_proj of the joint workloads
    float(np.float32(0.1) + np.float32(60.0)),    # lambdaf + alpha_proj
    98.0, 6.0, 100.0,                             # even integers: subnormal quotients with exact ties
    1.0, 2.0, 0.5,                                # powers of two
    3.0, 1.5e-30, 7.3e30, 3.4028235e38, 1e-40,    # odd, tiny, huge, largest, subnormal divisors
]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp_path_factory.mktemp("udiv") / "udiv_check"
    subprocess.run([cc, "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", str(HERE / "udiv_check.c"),
                    "-o", str(exe), "-lm"], check=True)
    return exe


def run(checker, divisors):
    """One checker process per divisor, at most one per host core at a time."""
    procs, rows = [], []
    width = max(1, min(8, os.cpu_count() or 1))
    for d in divisors:
        procs.append(subprocess.Popen([str(checker), float(d).hex()], stdout=subprocess.PIPE, text=True))
        if len(procs) >= width:
            rows.append(procs.pop(0).communicate()[0])
    rows += [p.communicate()[0] for p in procs]
    res = {}
    for r in rows:
        f = r.split()
        assert len(f) == 5
        res[float.fromhex(f[0])] = tuple(int(x) for x in f[1:])
    assert len(res) == len(divisors)
    return res


def test_uniform_division_equals_ieee_for_every_numerator(checker):
    res = run(checker, DIVISORS)
    for d, (mism, checked, plain, nosign) in res.items():
        assert checked == 2 * 14 * (1 << 23) + 6
        assert mism == 0, f"d={d}: {mism} quotients differ from IEEE division"
        assert nosign >= 1, f"d={d}: the check does not see a lost zero sign"
    # the exact ties of 98's subnormal quotients need the snap to the 2^-176 grid
    assert res[98.0][2] > 1000
