"""GPU regression pin: the plain and energy-traced steps reach the same bits
as when tests/golden/state_sha.json was written (tests/golden/make_state_sha.py,
run on an MI355X with the kernels of that time).

Every column-pass length class of the FP32 plain step (two-field 128, 256,
512, 1024; general plan 384, 768), 2-D and 3-D grids (both packed slots of
the column pass), batches of one and two instances, and the FP64
field-split and four-field passes: 20 plain steps, then 5 energy-traced
steps; the digest of each instance's (phi, con) after each phase and the
energy trace as exact doubles must match.  A change to the step kernels
that is meant to keep their arithmetic (register use, load order, where the
per-mode coefficients come from) passes only if every bit is kept.
"""
import importlib.util
import json
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "state_sha.json")


def _gen():
    spec = importlib.util.spec_from_file_location("make_state_sha", os.path.join(HERE, "golden", "make_state_sha.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["make_state_sha"] = mod
    spec.loader.exec_module(mod)
    return mod


GEN = _gen()


@pytest.mark.parametrize("case", GEN.CASES, ids=[GEN.case_key(c) for c in GEN.CASES])
def test_state_bits_pinned(hip, case):
    gold = json.load(open(GOLD))
    assert gold["plain_steps"] == GEN.PLAIN and gold["traced_steps"] == GEN.TRACED
    want = gold["cases"][GEN.case_key(case)]
    print("case", GEN.case_key(case), flush=True)
    got = GEN.run_case(hip, case)
    assert got["plain"] == want["plain"], "state after the plain steps differs"
    assert got["traced"] == want["traced"], "state after the energy-traced steps differs"
    assert got["energy_trace"] == want["energy_trace"], "energy trace differs"
