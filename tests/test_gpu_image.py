"""GPU decode of .toos files through tc_image_decode: the reference's golden
image files and envelopes around oracle-encoded packets of every entropy
generation and plane layout; planes and slice status equal the oracle's
decode of the PIXL packet."""
import glob
import os
import zlib

import pytest

from helpers import frame_for, orc_decode, orc_encode, toos_layout, toos_repair, toos_wrap

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
IMAGES = sorted(glob.glob(os.path.join(GOLDEN, "golden_image_v1*.bin")))


def _pixl(data):
    _, _, ents = toos_layout(data)
    (off, size), = [(e[2], e[3]) for e in ents if e[0] == 0x5049584C]
    return data[off:off + size]


def _dims(info_planes):
    return [(a.shape[1], a.shape[0]) for a in info_planes]


@pytest.mark.gpu
@pytest.mark.parametrize("path", IMAGES, ids=[os.path.basename(p)[:-4] for p in IMAGES])
def test_golden_image_decode(pkg, oracle, path):
    ic = pkg.ToposImageCodec()
    data = open(path, "rb").read()
    rc, planes, out = ic.decode_status(data)
    assert rc == 0, ic._detail(rc)
    assert out.concealed_slices == 0
    o_rc, o_planes, _, _ = orc_decode(oracle, _pixl(data), _dims(planes))
    assert o_rc == 0
    for a, b in zip(planes, o_planes):
        assert (a == b).all()


CASES = [
    # (kind, w, h, image profile, cfg kwargs)
    ("422", 1920, 1080, 0, dict(qp=24, profile=3, qmatrix_id=4, slice_rows=8)),
    ("4444", 512, 272, 2, dict(qp=30, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                              alpha_bit_depth=16, color_matrix=0)),
    ("444", 256, 128, 1, dict(qp=20, profile=5, pixel_format=2, bit_depth=10, slice_rows=8, color_matrix=0)),
    ("raw", 512, 384, 3, dict(qp=20, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("em", [0, 1, 8], ids=["v1-rice", "v2-vlc", "rans2"])
@pytest.mark.parametrize("kind,w,h,iprof,kw", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}" for c in CASES])
def test_image_decode_matches_oracle(pkg, abi, syn, oracle, em, kind, w, h, iprof, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    src = kind
    planes = frame_for(pkg, syn, src, w, h, 70 + em)
    if kw.get("bit_depth", 10) == 10 and kind == "444":
        planes = [(p >> 2).astype(p.dtype) for p in planes]  # the generator's 12-bit range to 10 bits
    cfg = abi.make_config(w, h, qp, entropy_mode=em, **kw)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    data = toos_wrap(pkt, iprof)
    ic = pkg.ToposImageCodec()
    d_rc, d_planes, out = ic.decode_status(data)
    assert d_rc == 0, ic._detail(d_rc)
    o_rc, o_planes, _, _ = orc_decode(oracle, pkt, _dims(d_planes))
    assert o_rc == 0
    assert len(d_planes) == len(planes)
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()


@pytest.mark.gpu
@pytest.mark.parametrize("em", [0, 8], ids=["v1-rice", "rans2"])
def test_image_decode_conceals_like_oracle(pkg, abi, syn, oracle, em):
    w, h = 320, 176
    planes = frame_for(pkg, syn, "422", w, h, 81)
    cfg = abi.make_config(w, h, 30, profile=3, entropy_mode=em, slice_rows=4)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    pkt = bytearray(pkt)
    count = int.from_bytes(pkt[30:32], "big")
    off = 53
    for k in range(count):
        size = int.from_bytes(pkt[off:off + 4], "big")
        if k in (2, 7):
            pay = off + 17
            pkt[pay + size - 1] ^= 0xFF if em == 0 else 0x00
            pkt[pay + size // 2] ^= 0x11
            pkt[off + 13:off + 17] = zlib.crc32(bytes(pkt[pay:pay + size])).to_bytes(4, "big")
        off += 17 + size
    pkt = bytes(pkt)
    data = toos_repair(toos_wrap(pkt, 0))
    ic = pkg.ToposImageCodec()
    d_rc, d_planes, out = ic.decode_status(data)
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, pkt, _dims(d_planes))
    assert d_rc == o_rc
    assert out.concealed_slices == o_nconc
    assert list(out.slice_status[:count]) == o_status[:count]
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()
