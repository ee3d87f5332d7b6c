"""Shared test setup: import paths for the package and the oracle, the
`gpu` marker (tests that need an MI355X; everything else runs on the CPU),
and oracle values of the larger test matrices read from a golden fixture."""
import hashlib
import json
import os
import sys
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "matrix-permanent-calculator_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X); run with -m gpu")


_CACHE = None


def matrix_key(mat) -> str:
    """Key of a matrix in tests/golden/oracle_cache.json: dtype, shape, bytes."""
    a = np.ascontiguousarray(mat)
    h = hashlib.sha256()
    h.update(str(a.dtype).encode())
    h.update(str(a.shape).encode())
    h.update(a.tobytes())
    return h.hexdigest()[:32]


def oracle_value(mat) -> float:
    """The permanent of `mat` from the oracle: read from the golden cache
    (written by tests/golden/make_golden.py --cache: exact for n <= 26 and
    for integer matrices, double-double Ryser otherwise) when the matrix is
    listed there, else computed with the exact oracle."""
    global _CACHE
    if _CACHE is None:
        path = os.path.join(GOLDEN, "oracle_cache.json")
        _CACHE = json.load(open(path)) if os.path.exists(path) else {}
    rec = _CACHE.get(matrix_key(mat))
    if rec is not None:
        return float(Fraction(rec["exact"])) if "exact" in rec else rec["dd_hi"] + rec["dd_lo"]
    import oracle as O

    return float(O.exact_permanent(mat))


def assert_target(r, ex, tol: float = 1e-10, ctx=None):
    """The accuracy contract of a result of the automatic mode: a permanent
    that is exactly 0 comes back as exactly 0; any other within `tol`
    relative of the oracle value `ex` and within its own stated error
    (err_bound for integer matrices, else the larger of err_bound and
    err_est), and marked certified."""
    ex = float(ex)
    if ex == 0.0:
        assert r.value == 0.0, (ctx, r, ex)
        return
    err = abs(r.value - ex)
    assert err <= tol * abs(ex), (ctx, r.value, ex, err / abs(ex))
    assert err <= r.stated or (r.stated == 0.0 and r.value == ex), (ctx, r, ex)
    assert r.certified, (ctx, r)


def cli_result(stdout: str):
    """(name, value) of a `perman` run: the name from the reference-format
    `Result: <name> <perm> in <s>` line, the value at full precision from the
    `Value:` line that follows it."""
    import re

    m = re.search(r"^Result: (\S+) \S+ in \S+$", stdout, re.M)
    v = re.search(r"^Value: (\S+)$", stdout, re.M)
    assert m and v, stdout
    return m.group(1), float(v.group(1))
