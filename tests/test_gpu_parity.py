"""GPU parity: packets byte-identical to the oracle encoder, decodes
bit-exact to the oracle decoder (which is itself pinned to the reference)."""
import pytest

from helpers import frame_for, orc_decode, orc_encode

CASES_422 = [
    # (w, h, qp, qmatrix, features, slice_rows)
    (64, 48, 62, 0, 0, 0),
    (73, 39, 30, 1, 0, 2),
    (16, 16, 0, 0, 0, 1),
    (1, 1, 40, 0, 0, 0),
    (200, 120, 62, 4, 1, 8),
    (256, 144, 24, 0, 3, 8),
    (1920, 1080, 62, 4, 1, 8),
]


@pytest.mark.gpu
@pytest.mark.parametrize("w,h,qp,qm,feat,sr", CASES_422)
def test_encode_decode_422(pkg, abi, syn, oracle, codec, w, h, qp, qm, feat, sr):
    planes = frame_for(pkg, syn, "422", w, h, seed=w * 7 + h)
    cfg = abi.make_config(w, h, qp, profile=3, qmatrix_id=qm, slice_rows=sr, entropy_mode=8, features=feat)
    orc_rc, orc_pkt = orc_encode(oracle, cfg, planes)
    assert orc_rc == 0
    st, pkt, _ = codec.encode_status(cfg, planes)
    assert st == 0, codec.last_error()
    assert len(pkt) == len(orc_pkt)
    assert pkt == orc_pkt
    dims = abi.plane_dims(w, h, 0, 0)
    o_rc, o_planes, o_status, _ = orc_decode(oracle, orc_pkt, dims)
    d_rc, d_planes, info = codec.decode(orc_pkt, strict=False)
    assert d_rc == o_rc == 0
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()


CASES_OTHER = [
    # (kind, w, h, cfg kwargs)
    ("4444", 73, 39, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1, alpha_bit_depth=16)),
    ("4444", 256, 128, dict(qp=40, profile=6, pixel_format=2, bit_depth=12, slice_rows=2, alpha_mode=2, alpha_bit_depth=8)),
    ("444", 64, 40, dict(qp=30, profile=3, pixel_format=1, bit_depth=12, slice_rows=1, qmatrix_id=1)),
    ("raw", 128, 96, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 96, 64, dict(qp=8, profile=3, bit_depth=16, slice_rows=3)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,kw", CASES_OTHER, ids=[f"{c[0]}-{c[1]}x{c[2]}-{i}" for i, c in enumerate(CASES_OTHER)])
def test_encode_decode_other_formats(pkg, abi, syn, oracle, codec, kind, w, h, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    planes = frame_for(pkg, syn, kind, w, h, seed=w + 3 * h)
    if kw.get("bit_depth") == 16:
        planes = [p.astype("uint16") * 64 for p in planes]
    cfg = abi.make_config(w, h, qp, entropy_mode=8, **kw)
    orc_rc, orc_pkt = orc_encode(oracle, cfg, planes)
    st, pkt, stats = codec.encode_status(cfg, planes)
    assert st == orc_rc, codec.last_error()
    if st != 0:
        return
    assert pkt == orc_pkt
    dims = abi.plane_dims(w, h, kw.get("pixel_format", 0), kw.get("alpha_mode", 0))
    o_rc, o_planes, _, _ = orc_decode(oracle, orc_pkt, dims)
    d_rc, d_planes, info = codec.decode(orc_pkt, strict=False)
    assert d_rc == o_rc
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()


def _slice_flags(pkt):
    """(plane, first payload byte) of every non-empty slice; for rANS2 colour
    slices that byte holds the level model (bits 0-1) and the DC model (bit 2)."""
    count = int.from_bytes(pkt[30:32], "big")
    out, off = set(), 53
    for _ in range(count):
        size = int.from_bytes(pkt[off:off + 4], "big")
        if size:
            out.add((pkt[off + 4], pkt[off + 17]))
        off += 17 + size
    return out


# content and sizes for which the encoder picks each conditional DC model
# variant: flags 4 (levels unconditioned), 5 (position), 6 (previous level)
DC_MODEL_CASES = [
    ("444", 1024, 512, 3, {4, 5}, dict(qp=95, profile=3, pixel_format=1, bit_depth=12, slice_rows=32)),
    ("422", 1024, 512, 3, {6}, dict(qp=80, profile=3, slice_rows=64)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("kind,w,h,seed,want,kw", DC_MODEL_CASES, ids=["dc-model-4-5", "dc-model-6"])
def test_dc_context_model_variants(pkg, abi, syn, oracle, codec, kind, w, h, seed, want, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    planes = frame_for(pkg, syn, kind, w, h, seed=seed)
    cfg = abi.make_config(w, h, qp, entropy_mode=8, **kw)
    orc_rc, orc_pkt = orc_encode(oracle, cfg, planes)
    assert orc_rc == 0
    assert want <= {f for _, f in _slice_flags(orc_pkt)}, "content no longer selects the DC model variants"
    st, pkt, _ = codec.encode_status(cfg, planes)
    assert st == 0, codec.last_error()
    assert pkt == orc_pkt
    dims = abi.plane_dims(w, h, cfg.pixel_format, 0)
    o_rc, o_planes, _, _ = orc_decode(oracle, orc_pkt, dims)
    d_rc, d_planes, _ = codec.decode(orc_pkt, strict=False)
    assert d_rc == o_rc == 0
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()


def _extreme_plane(M, w, h, maxv, invert):
    """Blocks whose samples are 0 or maxv in the sign pattern of one 2-D basis
    function (u, v) of the transform matrix: the forward transform's sums and
    products reach their worst case for that output frequency."""
    import numpy as np
    p = np.zeros((h, w), np.uint16)
    k = 0
    for by in range(h // 8):
        for bx in range(w // 8):
            u, v = k % 8, (k // 8) % 8
            k += 1
            pos = (M[u][:, None] * M[v][None, :]) >= 0
            if invert ^ (k % 2 == 0):
                pos = ~pos
            p[by * 8:(by + 1) * 8, bx * 8:(bx + 1) * 8] = np.where(pos, maxv, 0)
    return p


@pytest.mark.gpu
@pytest.mark.parametrize("fmt,bd,qp", [("422", 10, 0), ("422", 10, 62), ("444", 12, 0), ("444", 12, 30)])
def test_extreme_contrast_transforms(pkg, abi, oracle, codec, fmt, bd, qp):
    """Worst-case transform ranges (the 24-bit forward-transform multiplies,
    the exact f64 inverse with clamped coefficients): samples at 0 and the
    largest value in the sign patterns of every basis function. GPU packet
    equals the oracle's; the decode of it equals the oracle's decode."""
    import ctypes
    import numpy as np
    M = np.ctypeslib.as_array((ctypes.c_int16 * 64).in_dll(oracle, "orc_M")).astype(np.int64).reshape(8, 8)
    w, h = 256, 64
    maxv = (1 << bd) - 1
    pf = 0 if fmt == "422" else 1
    dims = abi.plane_dims(w, h, pf, 0)
    planes = [_extreme_plane(M, pw, ph, maxv, i == 1) for i, (pw, ph) in enumerate(dims)]
    cfg = abi.make_config(w, h, qp, profile=3, pixel_format=pf, bit_depth=bd, slice_rows=2, entropy_mode=8)
    orc_rc, orc_pkt = orc_encode(oracle, cfg, planes)
    assert orc_rc == 0
    st, pkt, _ = codec.encode_status(cfg, planes)
    assert st == 0, codec.last_error()
    assert pkt == orc_pkt
    o_rc, o_planes, _, _ = orc_decode(oracle, orc_pkt, dims)
    d_rc, d_planes, _ = codec.decode(orc_pkt, strict=False)
    assert d_rc == o_rc == 0
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()
