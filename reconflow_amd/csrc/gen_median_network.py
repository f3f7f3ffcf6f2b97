"""Writes rf_median25.hpp: a min/max network for the 13th smallest of 25 values (the 5x5 flow
median), Batcher's odd-even merge sort on 32 wires with wires 25..31 held at +inf, pruned to the
comparators (and the min / max halves of them) that reach output 12.

Run from this directory: ``python gen_median_network.py``. tests/test_median_networks.py
re-verifies the written network over all 2^25 inputs of 0s and 1s (the 0-1 principle: a
min/max network that selects correctly on every 0/1 input selects correctly on every totally
ordered input).
"""
from pathlib import Path


def odd_even_merge_sort(n):
    comps = []

    def merge(lo, hi, r):
        step = r * 2
        if step < hi - lo:
            merge(lo, hi, step)
            merge(lo + r, hi, step)
            comps.extend((i, i + r) for i in range(lo + r, hi - r, step))
        else:
            comps.append((lo, lo + r))

    def sort(lo, hi):
        if hi - lo >= 1:
            mid = lo + (hi - lo) // 2
            sort(lo, mid)
            sort(mid + 1, hi)
            merge(lo, hi, 1)

    sort(0, n - 1)
    return comps


def selection_network(n_real, n, target):
    """[(a, b, need_min, need_max)] over wires 0..n_real-1 (min to a, max to b); output wire."""
    inf = set(range(n_real, n))
    label = list(range(n))
    ops = []
    for i, j in odd_even_merge_sort(n):
        if j in inf:
            continue  # max is +inf, min is wire i: no change
        if i in inf:  # the finite value moves to i, +inf to j: relabel
            inf.discard(i)
            inf.add(j)
            label[i], label[j] = label[j], label[i]
            continue
        ops.append((label[i], label[j]))
    out = label[target]
    need, net = {out}, []
    for a, b in reversed(ops):
        na, nb = a in need, b in need
        if na or nb:
            net.append((a, b, na, nb))
            need |= {a, b}
    net.reverse()
    return net, out


def main():
    net, out = selection_network(25, 32, 12)
    lines = [
        "// Generated by gen_median_network.py: the 13th smallest of m[0..24] with a min/max network",
        "// (odd-even merge sort on 32 wires, padding at +inf, pruned to output 12).",
        f"// {len(net)} comparators, {sum(a + b for _, _, a, b in net)} min/max operations.",
        "// Included by rf_device.hpp inside namespace rf, with the RF_CE* macros defined.",
        "#pragma once",
        "",
        "__device__ __forceinline__ float median25_network(float (&m)[25]) {",
    ]
    for a, b, na, nb in net:
        if na and nb:
            lines.append(f"    RF_CE({a}, {b});")
        elif na:
            lines.append(f"    RF_CE_MIN({a}, {b});")
        else:
            lines.append(f"    RF_CE_MAX({a}, {b});")
    lines += [f"    return m[{out}];", "}", ""]
    Path(__file__).with_name("rf_median25.hpp").write_text("\n".join(lines))


if __name__ == "__main__":
    main()
