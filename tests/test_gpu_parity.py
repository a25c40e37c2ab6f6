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
