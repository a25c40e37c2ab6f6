"""The oracle against the compiled reference (oracle/_ref, built from the
reference sources when they are present) on frames larger than the golden
vectors: packets byte-identical and decodes bit-exact for every entropy
generation the path codes (V1 Rice, V2 VLC, rANS2) and every plane layout
(4:2:2, 4:4:4 / GBR with alpha, CFA), including a damaged packet's
concealment."""
import pytest

from helpers import frame_for, orc_decode, orc_encode, ref_decode, ref_encode

CASES = [
    # (kind, w, h, seed, cfg kwargs)
    ("422", 640, 360, 21, dict(qp=24, profile=3, qmatrix_id=4, slice_rows=8)),
    ("422", 333, 187, 22, dict(qp=5, profile=3, slice_rows=3, qp_delta_chroma=4)),
    ("4444", 256, 136, 23, dict(qp=30, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                               alpha_bit_depth=16)),
    ("444", 200, 120, 24, dict(qp=12, profile=5, pixel_format=1, bit_depth=12, slice_rows=2, qmatrix_id=1)),
    ("raw", 256, 192, 25, dict(qp=20, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
]
MODES = [0, 1, 8]


@pytest.mark.parametrize("em", MODES, ids=["v1-rice", "v2-vlc", "rans2"])
@pytest.mark.parametrize("kind,w,h,seed,kw", CASES, ids=[f"{c[0]}-{c[1]}x{c[2]}" for c in CASES])
def test_oracle_equals_reference(pkg, abi, syn, oracle, ref, em, kind, w, h, seed, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    planes = frame_for(pkg, syn, kind, w, h, seed)
    cfg = abi.make_config(w, h, qp, entropy_mode=em, **kw)
    r_rc, r_pkt, _ = ref_encode(ref, abi, cfg, planes)
    o_rc, o_pkt = orc_encode(oracle, cfg, planes)
    assert r_rc == o_rc == 0
    assert o_pkt == r_pkt
    dims = abi.plane_dims(w, h, kw.get("pixel_format", 0), kw.get("alpha_mode", 0))
    r_rc, r_planes, _ = ref_decode(ref, abi, r_pkt, dims)
    o_rc, o_planes, _, _ = orc_decode(oracle, o_pkt, dims)
    assert r_rc == o_rc == 0
    for a, b in zip(r_planes, o_planes):
        assert (a == b).all()


@pytest.mark.parametrize("em", MODES, ids=["v1-rice", "v2-vlc", "rans2"])
@pytest.mark.parametrize("damage", ["flip", "truncate"])
def test_oracle_conceals_like_reference(pkg, abi, syn, oracle, ref, em, damage):
    import zlib
    w, h = 320, 176
    planes = frame_for(pkg, syn, "422", w, h, 31)
    cfg = abi.make_config(w, h, 30, profile=3, entropy_mode=em, slice_rows=4)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    pkt = bytearray(pkt)
    count = int.from_bytes(pkt[30:32], "big")
    off, table = 53, []
    for _ in range(count):
        size = int.from_bytes(pkt[off:off + 4], "big")
        table.append((off + 17, size))
        off += 17 + size
    for k in (1, 4, 9):
        o, n = table[k]
        if damage == "flip":
            pkt[o + n // 3] ^= 0x24
        else:
            for j in range(o + n - 3, o + n):
                pkt[j] = 0xFF if em != 8 else 0
        pkt[o - 4:o] = zlib.crc32(bytes(pkt[o:o + n])).to_bytes(4, "big")
    pkt = bytes(pkt)
    dims = abi.plane_dims(w, h, 0, 0)
    r_rc, r_planes, info = ref_decode(ref, abi, pkt, dims)
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, pkt, dims)
    assert r_rc == o_rc
    if damage == "truncate":
        assert o_nconc == 3
    assert info.concealed_slices == o_nconc
    assert list(info.slice_status[:count]) == o_status[:count]
    for a, b in zip(r_planes, o_planes):
        assert (a == b).all()
