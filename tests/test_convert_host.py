"""Packed RGB(A) -> planar conversion and planar 4:4:4 -> 4:2:2, CPU side:
the oracle's restatement pinned to the reference unit test's known values
(native/tests/unit/test_color_convert.c), to invariants of the arithmetic, and
to the compiled reference (scalar and AVX2 rows, oracle/_ref) over every
format, output mode, depth, range and alpha setting; the product's argument
checks (they run before any GPU work) against the reference's."""
import ctypes

import numpy as np
import pytest

from helpers import cvt_call, cvt_source, p422_call

FORMATS = range(8)
MODES = (0, 1, 2)


def orc_cvt(oracle):
    return oracle.orc_convert_packed_rgb


def test_known_values(oracle, abi):
    """test_color_convert.c:46-86: white, black, pure red (BGRA32, BT.709 limited 10-bit)."""
    w, h = 34, 9
    src = np.full((h, w * 4), 255, np.uint8)
    rc, (y, u, v, a), _ = cvt_call(orc_cvt(oracle), abi, src, w, h, abi.CVT_BGRA32, 0, with_alpha=True)
    assert rc == 0 and (y == 940).all() and (u == 512).all() and (v == 512).all() and (a == 65535).all()
    src[:] = 0
    rc, (y, u, v, a), _ = cvt_call(orc_cvt(oracle), abi, src, w, h, abi.CVT_BGRA32, 0, with_alpha=True)
    assert rc == 0 and (y == 64).all() and (u == 512).all() and (v == 512).all() and (a == 0).all()
    px = src.reshape(h, w, 4)
    px[..., 2] = 255
    px[..., 3] = 255
    rc, (y, u, v, a), _ = cvt_call(orc_cvt(oracle), abi, src, w, h, abi.CVT_BGRA32, 0, with_alpha=True)
    assert rc == 0 and abs(int(y[0, 0]) - 250) <= 1 and abs(int(u[0, 0]) - 409) <= 1
    assert (v == 960).all() and (a == 65535).all()


def test_gbr_passthrough_formula(oracle, abi):
    """test_color_convert.c:88-119: GBR planes = clip(rint((x / 255) * 1023))."""
    w, h = 34, 9
    i = np.arange(w * h, dtype=np.uint32)
    src = np.stack([(i * 7 + 1) & 255, (i * 13 + 2) & 255, (i * 29 + 3) & 255], 1).astype(np.uint8)
    src = src.reshape(h, w * 3)
    rc, (g, b, r), _ = cvt_call(orc_cvt(oracle), abi, src, w, h, abi.CVT_BGR24, 2)
    assert rc == 0

    def expect(x):
        return np.clip(np.rint((x.astype(np.float32) / np.float32(255)) * np.float32(1023)), 0, 1023)

    px = src.reshape(h, w, 3)
    assert (g == expect(px[..., 1])).all() and (b == expect(px[..., 0])).all() and (r == expect(px[..., 2])).all()


def test_422_is_box_of_444(oracle, abi):
    """4:2:2 chroma = rint((a + b) / 2) of the quantised 4:4:4 chroma pair; an odd
    width's last column equals the last pixel's 4:4:4 value
    (test_color_convert.c:145-211); Y is the same in both modes."""
    rng = np.random.default_rng(3)
    for fmt in (abi.CVT_BGRA32, abi.CVT_RGB48):
        for w in (33, 34):
            src, _ = cvt_source(rng, fmt, abi, w, 7)
            rc, (y2, u2, v2), _ = cvt_call(orc_cvt(oracle), abi, src, w, 7, fmt, 0)
            rc4, (y4, u4, v4), _ = cvt_call(orc_cvt(oracle), abi, src, w, 7, fmt, 1)
            assert rc == 0 and rc4 == 0 and (y2 == y4).all()
            for c2, c4 in ((u2, u4), (v2, v4)):
                c4p = np.concatenate([c4, c4[:, -1:]], 1) if w % 2 else c4
                box = np.rint((c4p[:, 0::2].astype(np.float64) + c4p[:, 1::2]) * 0.5)
                assert (c2 == box).all()


def test_alpha_rules(oracle, abi):
    """alpha = rint((a / norm) * (0xFFFF >> shift)) << shift; opaque or 3-channel
    input writes 0xFFFF << shift (color_convert.c:83-94)."""
    rng = np.random.default_rng(4)
    w, h = 20, 3
    for fmt in (abi.CVT_RGBA32, abi.CVT_BGRA64):
        src, _ = cvt_source(rng, fmt, abi, w, h)
        for shift in (0, 4, 6, 15, 20):
            s = min(shift, 15)
            rc, planes, _ = cvt_call(orc_cvt(oracle), abi, src, w, h, fmt, 1, alpha_shift=shift, with_alpha=True)
            assert rc == 0
            if abi.cvt_sample_bytes(fmt) == 1:
                av = src.reshape(h, w, 4)[..., 3].astype(np.float32) / np.float32(255)
            else:
                av = src.view("<u2").reshape(h, w, 4)[..., 3].astype(np.float32) / np.float32(65535)
            nmax = np.float32(0xFFFF >> s)
            exp = (np.clip(np.rint(av * nmax), 0, nmax).astype(np.uint32) << s).astype(np.uint16)
            assert (planes[3] == exp).all(), (fmt, shift)
            rc, planes, _ = cvt_call(orc_cvt(oracle), abi, src, w, h, fmt, 1, alpha_shift=shift, alpha_opaque=1,
                                     with_alpha=True)
            assert rc == 0 and (planes[3] == ((0xFFFF << s) & 0xFFFF)).all()
    src, _ = cvt_source(rng, abi.CVT_BGR24, abi, w, h)
    rc, planes, _ = cvt_call(orc_cvt(oracle), abi, src, w, h, abi.CVT_BGR24, 0, alpha_shift=6, with_alpha=True)
    assert rc == 0 and (planes[3] == 0xFFC0).all()


def _ref_modes(ref):
    ref.tc_dev_cvt_simd_set.argtypes = [ctypes.c_int32]
    ref.tc_dev_cvt_simd_set.restype = None
    ref.tc_cvt_have_avx2.restype = ctypes.c_int
    return [0, 1] if ref.tc_cvt_have_avx2() else [0]


def _grid(abi):
    for fmt in FORMATS:
        for mode in MODES:
            for w in (1, 2, 33, 34, 67):
                yield fmt, mode, w


def test_oracle_matches_reference_all_formats(oracle, ref, abi):
    rng = np.random.default_rng(5)
    simd = _ref_modes(ref)
    try:
        for fmt, mode, w in _grid(abi):
            h = 5
            for kind in ("random", "max", "zero"):
                src, stride = cvt_source(rng, fmt, abi, w, h, row_pad=3 if w == 33 else 0, kind=kind)
                for bd, fr, shift, opq, with_a, kr, kb in (
                        (10, 0, 0, 0, True, 0.2126, 0.0722), (12, 1, 4, 0, True, 0.299, 0.114),
                        (10, 1, 6, 1, True, 0.2627, 0.0593), (12, 0, 0, 0, False, 0.2126, 0.0722)):
                    kw = dict(bit_depth=bd, kr=kr, kb=kb, full_range=fr, alpha_shift=shift, alpha_opaque=opq,
                              with_alpha=with_a, src_stride=stride, out_pad=2 if w == 67 else 0)
                    rc_o, po, fo = cvt_call(orc_cvt(oracle), abi, src, w, h, fmt, mode, **kw)
                    for m in simd:
                        ref.tc_dev_cvt_simd_set(m)
                        rc_r, pr, fr_ = cvt_call(ref.tc_convert_packed_rgb, abi, src, w, h, fmt, mode, **kw)
                        assert rc_o == 0 and rc_r == 0
                        for k, (a, b) in enumerate(zip(fo, fr_)):
                            assert (a == b).all(), (fmt, mode, w, kind, bd, fr, shift, opq, m, k)
    finally:
        ref.tc_dev_cvt_simd_set(-1)


def test_oracle_matches_reference_every_8bit_colour(oracle, ref, abi):
    """All 2^24 8-bit RGB triples through 4:4:4 and GBR at 10 and 12 bits, limited
    and full range: the oracle equals the compiled reference's scalar rows."""
    _ref_modes(ref)
    v = np.arange(1 << 24, dtype=np.uint32)
    src = np.stack([v & 255, (v >> 8) & 255, v >> 16], 1).astype(np.uint8).reshape(4096, 4096 * 3)
    ref.tc_dev_cvt_simd_set(0)
    try:
        for mode, bd, fr in ((1, 10, 0), (1, 12, 1), (2, 12, 0)):
            rc_o, _, fo = cvt_call(orc_cvt(oracle), abi, src, 4096, 4096, abi.CVT_RGB24, mode, bit_depth=bd,
                                   full_range=fr)
            rc_r, _, fr_ = cvt_call(ref.tc_convert_packed_rgb, abi, src, 4096, 4096, abi.CVT_RGB24, mode,
                                    bit_depth=bd, full_range=fr)
            assert rc_o == 0 and rc_r == 0
            for a, b in zip(fo, fr_):
                assert np.array_equal(a, b), (mode, bd, fr)
    finally:
        ref.tc_dev_cvt_simd_set(-1)


def test_planar_422_oracle(oracle, ref, abi):
    rng = np.random.default_rng(6)
    for w in (1, 2, 3, 64, 257):
        h = 4
        u = rng.integers(0, 65536, size=(h, w + 5), dtype=np.uint16)
        v = rng.integers(0, 65536, size=(h, w + 5), dtype=np.uint16)
        rc, uo, vo = p422_call(oracle.orc_convert_444_422, abi, u, v, w, h, bit_depth=16, in_stride=w + 5,
                               out_pad=1)
        assert rc == 0
        up = np.concatenate([u[:, :w], u[:, w - 1:w]], 1) if w % 2 else u[:, :w]
        assert (uo == ((up[:, 0::2].astype(np.uint32) + up[:, 1::2] + 1) >> 1)).all()
        rc_r, uo_r, vo_r = p422_call(ref.tc_convert_yuv444_to_422, abi, u, v, w, h, bit_depth=16, in_stride=w + 5,
                                     out_pad=1)
        assert rc_r == 0 and (uo == uo_r).all() and (vo == vo_r).all()


def test_product_argument_checks_match_reference(pkg, ref, abi):
    """test_color_convert.c:214-243 and the planar rules: the product rejects the
    same arguments with the same status, before any device work."""
    lib = pkg.load_library()
    w = 17
    src = np.zeros((1, 34 * 4), np.uint8)
    bad = abi.TC_ERR_INVALID_ARGUMENT
    cases = [dict(format=99), dict(bit_depth=8), dict(bit_depth=16), dict(out_mode=7), dict(src_stride=1),
             dict(width=0), dict(height=0), dict(width=40000), dict(full_range=2), dict(alpha_opaque=2),
             dict(struct_size=0)]
    for case in cases:
        for fn in (lib.tc_convert_packed_rgb, ref.tc_convert_packed_rgb, lib.tcx_convert_packed_rgb_dev):
            p = abi.CvtParams()
            p.struct_size = ctypes.sizeof(p)
            p.format, p.width, p.height, p.out_mode, p.bit_depth = abi.CVT_BGRA32, w, 1, 0, 10
            p.src = src.ctypes.data
            y = np.zeros(64, np.uint16)
            p.y_out = p.u_out = p.v_out = y.ctypes.data
            for k, val in case.items():
                setattr(p, k, val)
            assert fn(ctypes.byref(p)) == bad, (case, fn)
    for fn in (lib.tc_convert_packed_rgb, ref.tc_convert_packed_rgb):
        assert fn(None) == bad
    u = np.zeros((2, 8), np.uint16)
    for case in (dict(bit_depth=8), dict(width=0), dict(u_in_stride=3), dict(v_out_stride=1), dict(struct_size=4)):
        for fn in (lib.tc_convert_yuv444_to_422, ref.tc_convert_yuv444_to_422, lib.tcx_convert_yuv444_to_422_dev):
            p = abi.Planar422Params()
            p.struct_size = ctypes.sizeof(p)
            p.width, p.height, p.bit_depth = 8, 2, 10
            p.u_in = p.v_in = p.u_out = p.v_out = u.ctypes.data
            for k, val in case.items():
                setattr(p, k, val)
            assert fn(ctypes.byref(p)) == bad, (case, fn)


def test_binding_checks(pkg, abi):
    c = pkg.ToposCodec()
    with pytest.raises(pkg.ToposCodecError):
        c.convert_packed_rgb(np.zeros((2, 12), np.uint8), 4, 2, fmt=abi.CVT_BGR48)
    with pytest.raises(pkg.ToposCodecError):
        c.convert_packed_rgb(np.zeros((2, 12), np.uint8), 5, 2, fmt=abi.CVT_BGR24)
    with pytest.raises(pkg.ToposCodecError):
        c.convert_yuv444_to_422(np.zeros((2, 4), np.uint16), np.zeros((2, 5), np.uint16))
