"""Target-size encode on the GPU against the oracle's restatement of the
reference search: the C ABI (tc_frame_encode_sized) in every search mode over
targets inside, below and above the reachable range, the process-wide seed
over a call sequence, the device-resident batch (one seed state per request),
a 1080p frame, the minimality property at 4K (the chosen qp fits, the one
below does not) and the BUFFER_TOO_SMALL rules."""
import ctypes

import numpy as np
import pytest

from helpers import frame_for, orc_encode, orc_sized
from test_sized_oracle import FRAMES, _frame

pytestmark = pytest.mark.gpu


def _sizes(pkg, abi, syn, oracle, spec):
    out = {}
    for q in range(10, 96, 6):
        cfg, planes = _frame(pkg, abi, syn, spec, q)
        out[q] = len(orc_encode(oracle, cfg, planes)[1])
    return out


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_sized_matches_oracle(pkg, abi, syn, oracle, codec, mode):
    codec.sized_set_mode(mode)
    oracle.orc_sized_set_mode(mode)
    try:
        for spec in FRAMES:
            sizes = _sizes(pkg, abi, syn, oracle, spec)
            lo, hi = min(sizes.values()), max(sizes.values())
            cfg, planes = _frame(pkg, abi, syn, spec, 40)
            for target in (0, lo - 1, lo, (lo + hi) // 2, sizes[34], sizes[70] + 1, hi, 2 * hi):
                for (qmin, qmax, qbase) in ((0, 63, 40), (20, 95, 90), (30, 34, 10), (5, 80, 5)):
                    cfg.qp_base = qbase
                    codec.sized_reset()
                    oracle.orc_sized_reset()
                    st, pkt, stats, q = codec.encode_sized_status(cfg, planes, target, qmin, qmax)
                    o = orc_sized(oracle, cfg, planes, target, qmin, qmax)
                    assert (st, q) == (o[0], o[1]), (spec[:3], target, qmin, qmax, st, q, o[:2], codec.last_error())
                    assert pkt == o[2] and stats.packet_size == o[3], (spec[:3], target, qmin, qmax)
    finally:
        codec.sized_set_mode(0)
        oracle.orc_sized_set_mode(0)


def test_sized_seed_sequence(pkg, abi, syn, oracle, codec):
    codec.sized_reset()
    oracle.orc_sized_reset()
    for spec_i, targets in ((0, (9000, 9000, 6000, 6000, 12000, 3000, 9000)), (2, (4000, 4000, 2000, 7000))):
        cfg, planes = _frame(pkg, abi, syn, FRAMES[spec_i], 40)
        for t in targets:
            st, pkt, stats, q = codec.encode_sized_status(cfg, planes, t, 0, 95)
            o = orc_sized(oracle, cfg, planes, t, 0, 95)
            assert (st, q, pkt) == (o[0], o[1], o[2]), (spec_i, t, st, q, o[:2])


def _batch_req(abi, codec, cfg, planes, target, qmin, qmax, stream, keep):
    r = abi.TcxSizedReq()
    r.cfg = cfg
    for i, p in enumerate(planes):
        b = codec.alloc(p.nbytes)
        b.upload(np.ascontiguousarray(p))
        keep.append(b)
        r.dev_planes[i] = b.addr
    cap = codec.packet_bound(cfg)
    o = codec.alloc(cap)
    keep.append(o)
    r.dev_out, r.dev_out_cap = o.addr, cap
    r.target_bytes, r.qp_min, r.qp_max = target, qmin, qmax
    r.stream = ctypes.pointer(stream) if stream is not None else None
    return r, o


def test_sized_batch_streams(pkg, abi, syn, oracle, codec):
    """Requests of different frames, targets and qp ranges in one call, each with
    its own seed state; two calls in a row so the seeds are used."""
    oracle.orc_sized_set_mode(0)
    specs = [(FRAMES[0], 9000, 0, 95), (FRAMES[1], 3000, 10, 90), (FRAMES[2], 4000, 0, 63),
             (FRAMES[3], 9000, 0, 95), (FRAMES[0], 5000, 20, 60), (FRAMES[2], 1, 0, 95)]
    streams = [abi.TcxSizedStream() for _ in specs]
    oracle_state = [(ctypes.c_uint32 * 5)() for _ in specs]
    keep = []
    frames = [_frame(pkg, abi, syn, s[0], 40) for s in specs]
    for call in range(2):
        reqs, outs = [], []
        for k, ((spec, t, qmin, qmax), (cfg, planes)) in enumerate(zip(specs, frames)):
            r, o = _batch_req(abi, codec, cfg, planes, t + call * 500, qmin, qmax,
                              streams[k] if k != 5 else None, keep)
            reqs.append(r)
            outs.append(o)
        st = codec.encode_sized_batch_dev(reqs)
        assert st == 0, codec.last_error()
        for k, ((spec, t, qmin, qmax), (cfg, planes)) in enumerate(zip(specs, frames)):
            if k != 5:
                oracle.orc_sized_put_state(oracle_state[k])
            else:
                oracle.orc_sized_reset()
            o = orc_sized(oracle, cfg, planes, t + call * 500, qmin, qmax)
            if k != 5:
                oracle.orc_sized_get_state(oracle_state[k])
                s = streams[k]
                assert [s.qp, s.valid, s.bytes, s.prev, s.prev_bytes] == list(oracle_state[k])
            r = reqs[k]
            assert (r.rc, r.qp_used, r.packet_size, r.encodes) == (o[0], o[1], o[3], o[4]), (call, k)
            assert outs[k].download(r.packet_size).tobytes() == o[2], (call, k)


def test_sized_1080p_matches_oracle(pkg, abi, syn, oracle, codec):
    cfg = abi.make_config(1920, 1080, 62, profile=3, qmatrix_id=4, features=1, slice_rows=8, entropy_mode=8)
    planes = pkg.synth_frame422(syn, 1920, 1080, 21)
    codec.sized_reset()
    oracle.orc_sized_reset()
    for target in (700_000, 700_000, 400_000):
        st, pkt, stats, q = codec.encode_sized_status(cfg, planes, target, 0, 95)
        o = orc_sized(oracle, cfg, planes, target, 0, 95)
        assert (st, q, stats.packet_size) == (o[0], o[1], o[3]) and pkt == o[2], (target, st, q, o[:2])


def test_sized_4k_minimal_qp(pkg, abi, syn, codec):
    cfg = abi.make_config(3840, 2160, 62, profile=3, features=3, slice_rows=8, entropy_mode=8)
    planes = pkg.synth_frame422(syn, 3840, 2160, 22)
    codec.sized_reset()
    target = 6_000_000
    st, pkt, stats, q = codec.encode_sized_status(cfg, planes, target, 0, 95)
    assert st == 0 and len(pkt) <= target and stats.packet_size == len(pkt), (st, q, codec.last_error())
    at = abi.make_config(3840, 2160, q, profile=3, features=3, slice_rows=8, entropy_mode=8)
    assert codec.encode(at, planes) == pkt
    if q > 0:
        below = abi.make_config(3840, 2160, q - 1, profile=3, features=3, slice_rows=8, entropy_mode=8)
        assert len(codec.encode(below, planes)) > target


def test_sized_buffer_too_small(pkg, abi, syn, oracle, codec):
    cfg, planes = _frame(pkg, abi, syn, FRAMES[0], 40)
    try:
        for mode in (0, 1):
            codec.sized_set_mode(mode)
            codec.sized_reset()
            ok, pkt, stats, q = codec.encode_sized_status(cfg, planes, 9000, 0, 95)
            assert ok == 0 and stats.packet_size == len(pkt)
            codec.sized_reset()
            st, _, st2, q2 = codec.encode_sized_status(cfg, planes, 9000, 0, 95, out_cap=len(pkt) - 1)
            # the same search; its final size reported; strict reports qp 0, legacy its qp
            assert st == abi.TC_ERR_BUFFER_TOO_SMALL and st2.packet_size == len(pkt)
            assert q2 == (0 if mode == 0 else q)
            codec.sized_reset()
            st, _, st3, _ = codec.encode_sized_status(cfg, planes, 9000, 0, 95, out_cap=0)
            assert st == abi.TC_ERR_BUFFER_TOO_SMALL and st3.packet_size == len(pkt)
    finally:
        codec.sized_set_mode(0)
