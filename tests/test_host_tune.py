"""The autotune table and the tuner's choice rule (CPU only).

The library is built with csrc/uneven_tune.h, which lineax_amd/tune_uneven.py
writes from per-bucket batch times measured on the GPU; the measured times are
kept in the table's header comment.  Here: the rule picks contiguous class
boundaries that minimise the summed time (checked against brute force over
every split of the buckets), the committed table is the rule applied to its
own recorded times, and the library reports the table's values."""

import itertools
import pathlib
import re

from lineax_amd import _kernels as K
from lineax_amd import tune_uneven as T

TABLE = pathlib.Path(T.__file__).resolve().parent / "csrc" / "uneven_tune.h"


def _recorded():
    text = TABLE.read_text()
    times = {}
    for line in text.splitlines():
        m = re.match(r"//\s+n =\s*(\d+): (.*)", line)
        if m:
            times[int(m.group(1))] = {k: float(v) * 1e-3 for k, v in re.findall(r"(\d+)x\d+ ([0-9.]+) ms", m.group(2))}
    max4 = int(re.search(r"kTunedMax4 = (\d+)", text).group(1))
    max8 = int(re.search(r"kTunedMax8 = (\d+)", text).group(1))
    return times, max4, max8


def _brute(times):
    ns = sorted(times)
    best = None
    for i, j in itertools.combinations_with_replacement(range(len(ns) + 1), 2):
        # buckets [0, i) on 4x4, [i, j) on 8x8, the rest on 16x16
        if i and ns[i - 1] > 64 or j and ns[j - 1] > 128:
            continue
        total = sum(times[n]["4" if k < i else ("8" if k < j else "16")] for k, n in enumerate(ns))
        if best is None or total < best[0] - 1e-15:
            best = (total, i, j)
    return best[0]


def _total(times, m4, m8):
    return sum(t["4" if n <= m4 else ("8" if n <= m8 else "16")] for n, t in times.items())


def test_committed_table_is_the_rule_on_its_recorded_times():
    times, max4, max8 = _recorded()
    assert len(times) == 16 and 0 <= max4 <= max8 <= 128 and max4 <= 64
    assert T.choose(times) == (max4, max8)
    assert abs(_total(times, max4, max8) - _brute(times)) < 1e-12


def test_rule_matches_brute_force_on_synthetic_times():
    ns = list(range(8, 129, 8))
    cases = [
        {n: {"16": 1.0, "8": 2.0, **({"4": 3.0} if n <= 64 else {})} for n in ns},   # 16x16 best everywhere
        {n: {"16": 3.0, "8": 1.0, **({"4": 2.0} if n <= 64 else {})} for n in ns},   # 8x8 best up to 128
        {n: {"16": 3.0, "8": 2.0 + (n == 40), **({"4": 1.0 + 2.5 * (n == 40)} if n <= 64 else {})} for n in ns},
    ]
    for times in cases:
        m4, m8 = T.choose(times)
        assert abs(_total(times, m4, m8) - _brute(times)) < 1e-12
    assert T.choose(cases[0]) == (0, 0)
    assert T.choose(cases[1]) == (0, 128)


def test_library_reports_the_table():
    _, max4, max8 = _recorded()
    assert K.uneven_tuned_classes() == (max4, max8)
