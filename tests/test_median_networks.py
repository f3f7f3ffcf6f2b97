"""The fused kernels select medians with min/max networks instead of the reference's exchange
sort (medianFilter.cu:52-71, 218-239). A min/max network that returns the k-th smallest on every
input of 0s and 1s returns it on every totally ordered input (0-1 principle), so checking all
2^n binary inputs proves it. The 25-tap network is read from the header the kernels include;
NaN taps and zero results take the exchange sort on the GPU (covered by the GPU parity tests)."""
import re
from pathlib import Path

import numpy as np

CSRC = Path(__file__).resolve().parents[1] / "reconflow_amd" / "csrc"


def _binary_inputs(n):
    idx = np.arange(1 << n, dtype=np.uint32)
    return idx, [((idx >> k) & 1).astype(bool) for k in range(n)]


def test_median25_network_header_selects_the_13th_smallest():
    text = (CSRC / "rf_median25.hpp").read_text()
    ops = re.findall(r"RF_CE(_MIN|_MAX)?\((\d+), (\d+)\);", text)
    out = int(re.search(r"return m\[(\d+)\];", text).group(1))
    assert len(ops) > 50
    idx, w = _binary_inputs(25)
    for kind, a, b in ops:
        a, b = int(a), int(b)
        lo, hi = w[a] & w[b], w[a] | w[b]
        if kind in ("", "_MIN"):
            w[a] = lo
        if kind in ("", "_MAX"):
            w[b] = hi
        if kind == "_MIN":
            w[b] = None  # the max half is not computed: no later operation may read wire b
        if kind == "_MAX":
            w[a] = None
    ones = np.zeros(idx.size, np.uint8)
    for k in range(25):
        ones += ((idx >> k) & 1).astype(np.uint8)
    assert np.array_equal(w[out], ones >= 13)  # 13th smallest is 1 iff at most 12 zeros


def test_median9_column_network_selects_the_5th_smallest():
    # median9_finite (rf_device.hpp): sort the columns (m[c], m[c+3], m[c+6]) into lo / mid / hi,
    # answer med3(max of the lows, med3 of the middles, min of the highs)
    idx, w = _binary_inputs(9)

    def med3(a, b, c):
        return (a & b) | (c & (a | b))

    lo = [w[c] & w[c + 3] & w[c + 6] for c in range(3)]
    hi = [w[c] | w[c + 3] | w[c + 6] for c in range(3)]
    mi = [med3(w[c], w[c + 3], w[c + 6]) for c in range(3)]
    r = med3(lo[0] | lo[1] | lo[2], med3(*mi), hi[0] & hi[1] & hi[2])
    ones = sum(((idx >> k) & 1).astype(np.uint8) for k in range(9))
    assert np.array_equal(r, ones >= 5)


def test_generator_reproduces_the_header(tmp_path):
    import importlib.util
    import shutil

    shutil.copy(CSRC / "gen_median_network.py", tmp_path / "gen.py")
    spec = importlib.util.spec_from_file_location("gen", tmp_path / "gen.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.main()
    assert (tmp_path / "rf_median25.hpp").read_text() == (CSRC / "rf_median25.hpp").read_text()
