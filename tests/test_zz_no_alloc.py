"""The no-allocation decode entry point (tcx_decode_batch_dev_noalloc)
against tcx_decode_batch_dev: two calls in a row, the second with a smaller
batch, each compared sample for sample and slice status for slice status
(a damaged slice included), plus the limit and residency rejections."""
import numpy as np
import pytest

from helpers import frame_for, orc_encode


def _streams(pkg, abi, syn, oracle, specs):
    out = []
    for kind, w, h, seed, kw in specs:
        kw = dict(kw)
        qp = kw.pop("qp")
        cfg = abi.make_config(w, h, qp, entropy_mode=kw.pop("em", 8), **kw)
        rc, pkt = orc_encode(oracle, cfg, frame_for(pkg, syn, kind, w, h, seed))
        assert rc == 0
        out.append((pkt, abi.plane_dims(w, h, kw.get("pixel_format", 0), kw.get("alpha_mode", 0))))
    return out


def _run(codec, fn, streams):
    dev_pk, planes = [], []
    for pkt, dims in streams:
        b = codec.alloc(len(pkt))
        b.upload(np.frombuffer(pkt, np.uint8))
        dev_pk.append(b)
        planes.append([codec.alloc(w * h * 2) for (w, h) in dims])
    st, rcs, infos = fn([p for p, _ in streams], dev_pk, [[b.addr for b in pl] for pl in planes])
    got = [[b.download(w * h * 2, np.uint16).reshape(h, w) for b, (w, h) in zip(pl, dims)]
           for pl, (_, dims) in zip(planes, streams)]
    status = [list(infos[i].slice_status)[:infos[i].slice_count] for i in range(len(streams))]
    for b in dev_pk + [b for pl in planes for b in pl]:
        b.close()
    return st, rcs, got, status


@pytest.mark.gpu
def test_noalloc_decode_equals_batch_decode(pkg, abi, syn, oracle, codec):
    first = _streams(pkg, abi, syn, oracle, [
        ("422", 3840, 2160, 1, dict(qp=62, profile=3, qmatrix_id=0, features=3, slice_rows=8)),
        ("4444", 128, 64, 2, dict(qp=30, profile=6, pixel_format=2, bit_depth=12, slice_rows=2, alpha_mode=1,
                                  alpha_bit_depth=16)),
        ("raw", 320, 256, 3, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
        ("422", 64, 48, 4, dict(qp=20, profile=3, slice_rows=2, em=0)),
    ])   # 259,200 + 512 + 1,280 + 96 = 261,088 blocks, just under the 262,144 limit
    # a damaged colour slice in the second call: concealed in both paths
    pkt, dims = first[3]
    b = bytearray(pkt)
    b[53 + 17 + 10] ^= 0x21
    second = [first[2], (bytes(b), dims)]
    for batch in (first, second):            # two calls in a row, the second smaller
        ref = _run(codec, codec.decode_batch_dev, batch)
        got = _run(codec, codec.decode_batch_dev_noalloc, batch)
        assert got[0] == ref[0] and got[1] == ref[1]
        assert got[3] == ref[3]
        for a, r in zip(got[2], ref[2]):
            assert all((x == y).all() for x, y in zip(a, r))
    assert ref[0] == abi.TC_WARN_CONCEALED


@pytest.mark.gpu
def test_noalloc_decode_rejects_beyond_limits(pkg, abi, syn, oracle, codec):
    big = _streams(pkg, abi, syn, oracle, [("422", 4096, 2160, 5, dict(qp=62, profile=3, slice_rows=8))])
    st, rcs, _, _ = _run(codec, codec.decode_batch_dev_noalloc, big)
    assert st == abi.TC_ERR_LIMIT_EXCEEDED        # 138,240 + 2 x 69,120 = 276,480 blocks > 262,144
    small = _streams(pkg, abi, syn, oracle, [("422", 64, 32, 6, dict(qp=30, profile=3))])
    st, rcs, _, _ = _run(codec, codec.decode_batch_dev_noalloc, small * 17)
    assert st == abi.TC_ERR_LIMIT_EXCEEDED        # 17 frames > 16
