"""Persistent rANS2 entropy launches and the autotuner: every launch shape
(waves per workgroup, grid size down to a single workgroup that works through
the whole list) gives the oracle's packets and samples, and the autotuner's
run over all candidates leaves the output unchanged."""
import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

MIX = [
    ("422", 1280, 720, 41, dict(qp=68, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("4444", 256, 144, 42, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                                alpha_bit_depth=16)),
    ("raw", 320, 240, 43, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 77, 45, 44, dict(qp=30, profile=3, slice_rows=2)),
    ("422", 352, 288, 45, dict(qp=40, profile=3, slice_rows=4, entropy_mode=0)),
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


def _roundtrip(abi, codec, mix):
    outs = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in mix]
    st, sizes, rcs, _ = codec.encode_batch_dev([m["cfg"] for m in mix], [[b.addr for b in m["dev_in"]] for m in mix],
                                               outs)
    assert st == 0 and rcs == [0] * len(mix), codec.last_error()
    pkts = [o.download(n).tobytes() for o, n in zip(outs, sizes)]
    planes = [[codec.alloc(w * h * 2) for (w, h) in _dims(abi, m["cfg"])] for m in mix]
    st, rcs, _ = codec.decode_batch_dev([m["pkt"] for m in mix], [m["dev_pkt"] for m in mix],
                                        [[b.addr for b in p] for p in planes])
    assert st == 0 and rcs == [0] * len(mix), codec.last_error()
    decs = [[b.download(w * h * 2, np.uint16).reshape(h, w) for b, (w, h) in zip(p, _dims(abi, m["cfg"]))]
            for m, p in zip(mix, planes)]
    return pkts, decs


@pytest.mark.gpu
@pytest.mark.parametrize("wpg,groups", [(1, 0), (2, 0), (4, 0), (1, 1), (2, 1), (4, 3), (1, 7)])
def test_launch_shapes_match_oracle(abi, codec, mix, wpg, groups):
    codec.tune_force(wpg, groups)
    try:
        pkts, decs = _roundtrip(abi, codec, mix)
    finally:
        codec.tune_force(0)
    for i, m in enumerate(mix):
        assert pkts[i] == m["pkt"], f"frame {i} packet, shape {wpg}x{groups}"
        assert all((a == b).all() for a, b in zip(decs[i], m["dec"])), f"frame {i} decode, shape {wpg}x{groups}"


@pytest.mark.gpu
def test_autotune_keeps_output(abi, codec, mix):
    outs = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in mix]
    st, rep = codec.autotune_batch([m["cfg"] for m in mix], [[b.addr for b in m["dev_in"]] for m in mix], outs)
    assert st == 0, codec.last_error()
    assert rep.candidates == 6
    assert rep.enc_wpg in (1, 2, 4) and rep.dec_wpg in (1, 2, 4)
    assert 0 < rep.enc_ms_best <= rep.enc_ms_default and 0 < rep.dec_ms_best <= rep.dec_ms_default
    assert rep.fdct_tpw in (1, 2, 4) and rep.idct_tpw in (1, 2, 4)
    assert 0 < rep.fdct_ms_best <= rep.fdct_ms_default and 0 < rep.idct_ms_best <= rep.idct_ms_default
    for i, m in enumerate(mix):
        assert outs[i].download(len(m["pkt"])).tobytes() == m["pkt"]
    # the tuned shape is used from now on
    pkts, decs = _roundtrip(abi, codec, mix)
    assert pkts == [m["pkt"] for m in mix]
    codec.tune_reset()


@pytest.mark.gpu
@pytest.mark.parametrize("tpw", [1, 2, 4])
def test_transform_tilings_match_oracle(abi, codec, oracle, mix, tpw):
    codec.tune_force_tiles(tpw)
    try:
        pkts, decs = _roundtrip(abi, codec, mix)
        # a damaged slice: concealed tiles inside multi-tile workgroups
        m = mix[0]
        pkt = bytearray(m["pkt"])
        count = int.from_bytes(pkt[30:32], "big")
        off = 53
        for k in range(count):
            size = int.from_bytes(pkt[off:off + 4], "big")
            if k == 3:
                pkt[off + 17 + size // 2] ^= 0x3C
            off += 17 + size
        pkt = bytes(pkt)
        rc, planes, info = codec.decode(pkt, strict=False)
    finally:
        codec.tune_force_tiles(0)
    for i, mm in enumerate(mix):
        assert pkts[i] == mm["pkt"], f"frame {i} packet, {tpw} tiles per workgroup"
        assert all((a == b).all() for a, b in zip(decs[i], mm["dec"])), f"frame {i} decode, {tpw} tiles"
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, pkt, _dims(abi, m["cfg"]))
    assert rc == o_rc == abi.TC_WARN_CONCEALED
    assert list(info.slice_status[:count]) == o_status[:count]
    assert all((a == b).all() for a, b in zip(planes, o_planes))
