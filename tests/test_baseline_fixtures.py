"""The benchmark frames against the committed reference fixture
(tests/golden/baseline_ref_hashes.json, written by make_baseline_hashes.py
from the compiled reference): the frozen generator reproduces every input
plane bit for bit (it uses the host's libm, so this pins the inputs on each
machine), and the oracle's packet and decode hashes equal the reference's at
full size for the product flag sets (qm 4 + deblock at 1080p, qm 0 + deblock
+ refined qp table at 4K, 4444 XQ 12-bit + 16-bit alpha, RAW 12-bit) and the
plain C defaults."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import baseline_cases as bc  # noqa: E402
from helpers import orc_decode, orc_encode  # noqa: E402

CASES = bc.cases()
# frames the oracle encodes and decodes here (the rest are covered on the GPU)
ORACLE_IDS = {bc.case_id(c) for c in CASES
              if c[0] == "cpu_1080p" or (c[0] == "hq_4k" and c[4] < 2) or c[4] in (100, 104, 108, 111)}


@pytest.fixture(scope="module")
def fixture():
    return bc.load_fixture()


@pytest.mark.parametrize("case", CASES, ids=[bc.case_id(c) for c in CASES])
def test_frame_matches_reference_fixture(pkg, abi, syn, oracle, fixture, case):
    e = fixture[bc.case_id(case)]
    planes = bc.make_frame(pkg, syn, case)
    assert [bc.sha(p) for p in planes] == e["input_sha256"]
    if bc.case_id(case) not in ORACLE_IDS:
        return
    cfg = bc.make_config(abi, case)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    assert len(pkt) == e["packet_bytes"] and bc.sha(pkt) == e["packet_sha256"]
    rc, dec, _, _ = orc_decode(oracle, pkt, [(p.shape[1], p.shape[0]) for p in planes])
    assert rc == 0 and [bc.sha(p) for p in dec] == e["decode_sha256"]
