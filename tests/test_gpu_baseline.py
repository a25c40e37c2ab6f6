"""The benchmark workloads on the GPU at full size against the compiled
reference's committed hashes (tests/golden/baseline_ref_hashes.json): every
frame of the 1080p CPU case, the 4K 422 HQ batch (seeds 0-7), the 4096x2160
4444 XQ + alpha frames and the uneven batch (720p Proxy, 1080p LT, 4K HQ, 8K
HQ, RAW 4096x3072) is generated here (input hashes checked first: the
generator uses this host's libm), encoded in one device-resident batch per
group, and its packets and decoded planes hashed."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import baseline_cases as bc  # noqa: E402

GROUPS = ["cpu_1080p", "hq_4k", "xq_4096", "uneven"]


@pytest.mark.gpu
@pytest.mark.parametrize("group", GROUPS)
def test_group_matches_reference_hashes(pkg, abi, syn, codec, group):
    fx = bc.load_fixture()
    cases = [c for c in bc.cases() if c[0] == group]
    frames, ins, outs, cfgs = [], [], [], []
    for c in cases:
        planes = bc.make_frame(pkg, syn, c)
        assert [bc.sha(p) for p in planes] == fx[bc.case_id(c)]["input_sha256"], "generator differs on this host"
        bufs = []
        for p in planes:
            b = codec.alloc(p.nbytes)
            b.upload(p)
            bufs.append(b)
        frames.append(planes)
        ins.append(bufs)
        outs.append(codec.alloc(sum(p.nbytes for p in planes) * 2 + (1 << 20)))
        cfgs.append(bc.make_config(abi, c))
    st, sizes, rcs, _ = codec.encode_batch_dev(cfgs, [[b.addr for b in bb] for bb in ins], outs)
    assert st == 0, codec.last_error()
    pkts = [outs[i].download(sizes[i]).tobytes() for i in range(len(cases))]
    for c, pkt in zip(cases, pkts):
        e = fx[bc.case_id(c)]
        assert (len(pkt), bc.sha(pkt)) == (e["packet_bytes"], e["packet_sha256"]), bc.case_id(c)
    dec = [[codec.alloc(p.nbytes) for p in planes] for planes in frames]
    st, rcs, infos = codec.decode_batch_dev(pkts, outs, [[b.addr for b in d] for d in dec])
    assert st == 0, codec.last_error()
    for c, planes, d in zip(cases, frames, dec):
        got = [b.download(p.nbytes, np.uint16).reshape(p.shape) for b, p in zip(d, planes)]
        assert [bc.sha(g) for g in got] == fx[bc.case_id(c)]["decode_sha256"], bc.case_id(c)
    for bb in ins + dec:
        for b in bb:
            b.close()
    for b in outs:
        b.close()
