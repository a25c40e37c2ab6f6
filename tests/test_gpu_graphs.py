"""Launch sequences replayed from captured HIP graphs: every replay of a
captured encode / decode batch must give the packets and samples of the eager
launch sequence and of the oracle, also when eager work and other graph
shapes run between the replays."""
import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

MIX = [
    ("422", 640, 360, 31, dict(qp=75, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("4444", 256, 144, 32, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                                alpha_bit_depth=16)),
    ("raw", 320, 240, 33, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 352, 288, 34, dict(qp=40, profile=3, slice_rows=4, entropy_mode=1)),
]


def _cfg(abi, w, h, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    em = kw.pop("entropy_mode", 8)
    return abi.make_config(w, h, qp, entropy_mode=em, **kw)


def _dims(abi, cfg):
    return abi.plane_dims(cfg.visible_width, cfg.visible_height, cfg.pixel_format, cfg.alpha_mode)


@pytest.fixture(scope="module")
def mix(pkg, abi, syn, oracle, codec):
    out = []
    for kind, w, h, seed, kw in MIX:
        planes = frame_for(pkg, syn, kind, w, h, seed)
        cfg = _cfg(abi, w, h, kw)
        rc, pkt = orc_encode(oracle, cfg, planes)
        assert rc == 0
        rc, dec, _, _ = orc_decode(oracle, pkt, _dims(abi, cfg))
        assert rc == 0
        bufs = []
        for p in planes:
            b = codec.alloc(p.nbytes)
            b.upload(np.ascontiguousarray(p))
            bufs.append(b)
        pk = codec.alloc(len(pkt))
        pk.upload(np.frombuffer(pkt, np.uint8))
        out.append(dict(cfg=cfg, pkt=pkt, dec=dec, dev_in=bufs, dev_pkt=pk))
    return out


def _encode(abi, codec, items, outs):
    st, sizes, rcs, _ = codec.encode_batch_dev([m["cfg"] for m in items], [[b.addr for b in m["dev_in"]] for m in items],
                                               outs)
    assert st == 0 and rcs == [0] * len(items), codec.last_error()
    return [o.download(n).tobytes() for o, n in zip(outs, sizes)]


def _decode(abi, codec, items, planes):
    st, rcs, _ = codec.decode_batch_dev([m["pkt"] for m in items], [m["dev_pkt"] for m in items],
                                        [[b.addr for b in p] for p in planes])
    assert st == 0 and rcs == [0] * len(items), codec.last_error()
    res = []
    for m, p in zip(items, planes):
        res.append([b.download(w * h * 2, np.uint16).reshape(h, w) for b, (w, h) in zip(p, _dims(abi, m["cfg"]))])
    return res


@pytest.mark.gpu
def test_graph_replays_match_eager_and_oracle(abi, codec, mix):
    outs = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in mix]
    planes = [[codec.alloc(w * h * 2) for (w, h) in _dims(abi, m["cfg"])] for m in mix]
    codec.set_graphs(False)
    eager_pk = _encode(abi, codec, mix, outs)
    eager_dec = _decode(abi, codec, mix, planes)
    assert eager_pk == [m["pkt"] for m in mix]
    codec.set_graphs(True)
    try:
        h0, b0 = codec.graph_stats()
        for rep in range(4):
            # clear the outputs between replays so a replay that wrote nothing fails
            for o in outs:
                o.upload(np.zeros(o.nbytes, np.uint8))
            for p in planes:
                for b in p:
                    b.upload(np.zeros(b.nbytes, np.uint8))
            assert _encode(abi, codec, mix, outs) == eager_pk, f"encode replay {rep}"
            got = _decode(abi, codec, mix, planes)
            for m, g, e in zip(mix, got, eager_dec):
                assert all((a == b).all() for a, b in zip(g, e)), f"decode replay {rep}"
                assert all((a == b).all() for a, b in zip(g, m["dec"])), f"decode replay {rep} vs oracle"
            # other shapes between the replays: a sub-batch through a graph, one frame eagerly
            assert _encode(abi, codec, mix[1:3], outs[1:3]) == eager_pk[1:3]
            codec.set_graphs(False)
            assert _encode(abi, codec, mix[:1], outs[:1]) == eager_pk[:1]
            codec.set_graphs(True)
        h1, b1 = codec.graph_stats()
        assert b1 > b0, "no launch sequence was captured"
        assert h1 - h0 >= 6, "captured sequences were not replayed"
    finally:
        codec.set_graphs(True)


@pytest.mark.gpu
def test_graph_per_frame_api_replays(abi, pkg, mix):
    # the drop-in calls reuse one captured sequence per frame shape
    c = pkg.DeviceCodec()
    m = mix[0]
    h0, _ = c.graph_stats()
    for _ in range(3):
        rc, planes, _ = c.decode(m["pkt"], strict=False)
        assert rc == 0
        assert all((a == b).all() for a, b in zip(planes, m["dec"]))
    h1, _ = c.graph_stats()
    assert h1 - h0 >= 2
