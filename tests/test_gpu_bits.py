"""GPU parity of the bit-packed colour entropy modes, V1 bounded Rice and V2
canonical VLC: the reference's golden packets (encode byte-identical, decode
fingerprint equal), larger frames against the oracle in every plane layout,
a mixed batch of all three generations, and concealment of damaged slices."""
import struct
import zlib

import pytest

from helpers import frame_for, orc_decode, orc_encode
from test_oracle_golden import GOLDEN_SETS, golden_cases, golden_config, golden_dims, golden_inputs, plane_fold

BITS_SETS = [n for n, cases in GOLDEN_SETS.items() if any(c[13] in (0, 1) for c in cases)]


@pytest.mark.gpu
@pytest.mark.parametrize("case,packet,fold", golden_cases(BITS_SETS))
def test_golden_packets_on_gpu(abi, oracle, codec, case, packet, fold):
    cfg = golden_config(abi, case)
    st, pkt, _ = codec.encode_status(cfg, golden_inputs(oracle, case))
    assert st == 0, codec.last_error()
    assert pkt == packet
    rc, planes, info = codec.decode(packet, strict=False)
    assert rc == 0 and info.concealed_slices == 0
    assert plane_fold(planes[:len(golden_dims(abi, case))]) == struct.unpack(">Q", fold)[0]


CASES = [
    # (kind, w, h, seed, cfg kwargs)
    ("422", 1920, 1080, 41, dict(qp=24, profile=3, qmatrix_id=4, slice_rows=8)),
    ("422", 333, 187, 42, dict(qp=2, profile=3, slice_rows=3, qp_delta_chroma=4)),
    ("4444", 512, 272, 43, dict(qp=30, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                               alpha_bit_depth=16)),
    ("444", 200, 120, 44, dict(qp=12, profile=5, pixel_format=1, bit_depth=12, slice_rows=2, qmatrix_id=1)),
    ("raw", 512, 384, 45, dict(qp=20, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 64, 64, 46, dict(qp=0, profile=3, slice_rows=1)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("em", [0, 1], ids=["v1-rice", "v2-vlc"])
@pytest.mark.parametrize("kind,w,h,seed,kw", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}" for c in CASES])
def test_bits_modes_match_oracle(pkg, abi, syn, oracle, codec, em, kind, w, h, seed, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    planes = frame_for(pkg, syn, kind, w, h, seed)
    cfg = abi.make_config(w, h, qp, entropy_mode=em, **kw)
    o_rc, o_pkt = orc_encode(oracle, cfg, planes)
    assert o_rc == 0
    st, pkt, _ = codec.encode_status(cfg, planes)
    assert st == 0, codec.last_error()
    assert len(pkt) == len(o_pkt)
    assert pkt == o_pkt
    dims = abi.plane_dims(w, h, kw.get("pixel_format", 0), kw.get("alpha_mode", 0))
    r_rc, r_planes, _, _ = orc_decode(oracle, o_pkt, dims)
    d_rc, d_planes, info = codec.decode(o_pkt, strict=False)
    assert d_rc == r_rc == 0
    for a, b in zip(d_planes, r_planes):
        assert (a == b).all()


@pytest.mark.gpu
def test_mixed_generation_batch(pkg, abi, syn, oracle, codec):
    # one batch holding V1, V2 and rANS2 frames of different sizes
    frames = []
    for i, (em, w, h) in enumerate([(0, 640, 360), (8, 320, 180), (1, 480, 272), (0, 64, 48), (1, 1280, 720)]):
        planes = frame_for(pkg, syn, "422", w, h, 60 + i)
        cfg = abi.make_config(w, h, 28, profile=3, entropy_mode=em, slice_rows=8)
        rc, pkt = orc_encode(oracle, cfg, planes)
        assert rc == 0
        frames.append((cfg, planes, pkt, abi.plane_dims(w, h, 0, 0)))
    in_bufs = []
    for _, planes, _, _ in frames:
        bs = []
        for p in planes:
            b = codec.alloc(p.nbytes)
            b.upload(p)
            bs.append(b)
        in_bufs.append(bs)
    outs = [codec.alloc(len(f[2]) + (1 << 20)) for f in frames]
    st, sizes, rcs, _ = codec.encode_batch_dev([f[0] for f in frames], [[b.addr for b in bs] for bs in in_bufs],
                                               outs)
    assert st == 0 and rcs == [0] * len(frames), codec.last_error()
    for i, f in enumerate(frames):
        assert outs[i].download(sizes[i]).tobytes() == f[2], f"frame {i}"
    for cfg, planes, pkt, dims in frames:
        o_rc, o_planes, _, _ = orc_decode(oracle, pkt, dims)
        d_rc, d_planes, _ = codec.decode(pkt, strict=False)
        assert d_rc == o_rc == 0
        for a, b in zip(d_planes, o_planes):
            assert (a == b).all()


@pytest.mark.gpu
@pytest.mark.parametrize("em", [0, 1], ids=["v1-rice", "v2-vlc"])
@pytest.mark.parametrize("damage", ["flip", "truncate", "trailing", "crc"])
def test_bits_damaged_slices_conceal_like_oracle(pkg, abi, syn, oracle, codec, em, damage):
    w, h = 320, 176
    planes = frame_for(pkg, syn, "422", w, h, 31)
    cfg = abi.make_config(w, h, 30, profile=3, entropy_mode=em, slice_rows=4)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    count = int.from_bytes(pkt[30:32], "big")
    off, table = 53, []
    for _ in range(count):
        size = int.from_bytes(pkt[off:off + 4], "big")
        table.append((off, size))
        off += 17 + size
    out, prev = bytearray(pkt[:53]), 53
    for k, (o, n) in enumerate(table):
        hdr = bytearray(pkt[o:o + 17])
        pay = bytearray(pkt[o + 17:o + 17 + n])
        if k in (1, 4, 9):
            if damage == "flip":
                pay[n // 3] ^= 0x24
            elif damage == "truncate":
                pay[-3:] = b"\xff\xff\xff"
            elif damage == "trailing":
                pay += b"\x00\x00"              # whole bytes past the last symbol
            else:
                pay[n // 2] ^= 0x01             # CRC left stale
        if damage != "crc":
            hdr[13:17] = zlib.crc32(bytes(pay)).to_bytes(4, "big")
        hdr[0:4] = len(pay).to_bytes(4, "big")
        out += hdr + pay
    # frame size field and header CRC
    out[41:45] = len(out).to_bytes(4, "big")
    out[49:53] = zlib.crc32(bytes(out[:49])).to_bytes(4, "big")
    pkt = bytes(out)
    dims = abi.plane_dims(w, h, 0, 0)
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, pkt, dims)
    if damage != "flip":
        assert o_nconc == 3
    d_rc, d_planes, info = codec.decode(pkt, strict=False)
    assert d_rc == o_rc
    assert list(info.slice_status[:count]) == o_status[:count]
    assert info.concealed_slices == o_nconc
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()
