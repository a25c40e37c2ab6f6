"""The benchmark's batch configurations at their real sizes, against the oracle:
an uneven batch of 720p Proxy, 1080p LT, 4K and 8K 422 HQ frames and 12-bit
RAW phase planes in one launch sequence, and 4096x2160 4444 XQ 12-bit frames
with a 16-bit alpha plane. Packets byte for byte, decodes sample for sample."""
import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

# (kind, w, h, seed, cfg kwargs): product settings per tier (fixed qp)
UNEVEN = [
    ("422", 1280, 720, 51, dict(qp=75, profile=3, qmatrix_id=4, features=1, slice_rows=8)),   # Proxy
    ("422", 1920, 1080, 52, dict(qp=68, profile=3, qmatrix_id=4, features=1, slice_rows=8)),  # LT
    ("422", 3840, 2160, 53, dict(qp=62, profile=3, qmatrix_id=0, features=3, slice_rows=8)),  # HQ
    ("422", 8192, 4320, 54, dict(qp=62, profile=3, qmatrix_id=0, features=3, slice_rows=8)),  # HQ 8K
    ("raw", 4096, 3072, 55, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
]
XQ = [("4444", 4096, 2160, 56 + i, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4,
                                         alpha_mode=1, alpha_bit_depth=16)) for i in range(2)]


def _cfg(abi, w, h, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    return abi.make_config(w, h, qp, entropy_mode=8, **kw)


def _dims(abi, cfg):
    return abi.plane_dims(cfg.visible_width, cfg.visible_height, cfg.pixel_format, cfg.alpha_mode)


def _prepare(pkg, abi, syn, oracle, codec, spec):
    out = []
    for kind, w, h, seed, kw in spec:
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
        out.append(dict(cfg=cfg, pkt=pkt, dec=dec, dev_in=bufs, dev_pkt=pk, name=f"{kind} {w}x{h}"))
    return out


def _check_batch(abi, codec, items):
    outs = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in items]
    st, sizes, rcs, _ = codec.encode_batch_dev([m["cfg"] for m in items],
                                               [[b.addr for b in m["dev_in"]] for m in items], outs)
    assert st == 0 and rcs == [0] * len(items), codec.last_error()
    for m, o, n in zip(items, outs, sizes):
        assert n == len(m["pkt"]) and o.download(n).tobytes() == m["pkt"], m["name"]
    planes = [[codec.alloc(w * h * 2) for (w, h) in _dims(abi, m["cfg"])] for m in items]
    st, rcs, _ = codec.decode_batch_dev([m["pkt"] for m in items], [m["dev_pkt"] for m in items],
                                        [[b.addr for b in p] for p in planes])
    assert st == 0 and rcs == [0] * len(items), codec.last_error()
    for m, p in zip(items, planes):
        for k, (b, (w, h)) in enumerate(zip(p, _dims(abi, m["cfg"]))):
            got = b.download(w * h * 2, np.uint16).reshape(h, w)
            assert (got == m["dec"][k]).all(), f"{m['name']} plane {k}"


@pytest.mark.gpu
def test_uneven_batch_full_size(pkg, abi, syn, oracle, codec):
    items = _prepare(pkg, abi, syn, oracle, codec, UNEVEN)
    _check_batch(abi, codec, items)
    # the same frames in the reverse order: slice order in the work list is by length
    _check_batch(abi, codec, list(reversed(items)))


@pytest.mark.gpu
def test_4444xq_alpha_full_size(pkg, abi, syn, oracle, codec):
    items = _prepare(pkg, abi, syn, oracle, codec, XQ)
    _check_batch(abi, codec, items)
