"""Packed RGB(A) -> planar conversion and planar 4:4:4 -> 4:2:2 on the GPU
against the oracle: every format, output mode, depth, range and alpha setting
at small sizes, odd widths and padded strides, 4K frames of 8-bit and 16-bit
input through both the host-buffer and the device-resident calls, and a
converted frame fed to the batched encoder."""
import ctypes

import numpy as np
import pytest

from helpers import cvt_call, cvt_source, orc_encode, p422_call

pytestmark = pytest.mark.gpu


def test_convert_grid_matches_oracle(oracle, abi, codec):
    rng = np.random.default_rng(11)
    lib = codec._lib
    n = 0
    for fmt in range(8):
        for mode in (0, 1, 2):
            for w in (1, 2, 33, 34, 67, 300):
                h = 5
                src, stride = cvt_source(rng, fmt, abi, w, h, row_pad=3 if w in (33, 300) else 0)
                for bd, fr, shift, opq, with_a, kr, kb in (
                        (10, 0, 0, 0, True, 0.2126, 0.0722), (12, 1, 4, 0, True, 0.299, 0.114),
                        (10, 1, 6, 1, True, 0.2627, 0.0593), (12, 0, 0, 0, False, 0.2126, 0.0722)):
                    kw = dict(bit_depth=bd, kr=kr, kb=kb, full_range=fr, alpha_shift=shift, alpha_opaque=opq,
                              with_alpha=with_a, src_stride=stride, out_pad=2 if w == 67 else 0)
                    rc_o, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, w, h, fmt, mode, **kw)
                    rc_g, _, fg = cvt_call(lib.tc_convert_packed_rgb, abi, src, w, h, fmt, mode, **kw)
                    assert rc_o == 0 and rc_g == 0, codec.last_error()
                    for k, (a, b) in enumerate(zip(fo, fg)):
                        assert np.array_equal(a, b), (fmt, mode, w, bd, fr, shift, opq, k)
                    n += 1
    assert n == 8 * 3 * 6 * 4


def test_convert_every_8bit_colour(oracle, abi, codec):
    """All 2^24 8-bit RGB triples, 4:4:4 BT.709 limited 10-bit and GBR 12-bit."""
    v = np.arange(1 << 24, dtype=np.uint32)
    src = np.stack([v & 255, (v >> 8) & 255, v >> 16], 1).astype(np.uint8).reshape(4096, 4096 * 3)
    for mode, bd, fr in ((1, 10, 0), (1, 12, 1), (2, 12, 0)):
        rc_o, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, 4096, 4096, abi.CVT_RGB24, mode,
                               bit_depth=bd, full_range=fr)
        rc_g, _, fg = cvt_call(codec._lib.tc_convert_packed_rgb, abi, src, 4096, 4096, abi.CVT_RGB24, mode,
                               bit_depth=bd, full_range=fr)
        assert rc_o == 0 and rc_g == 0
        for a, b in zip(fo, fg):
            assert np.array_equal(a, b), (mode, bd, fr)


def test_convert_every_16bit_sample(oracle, abi, codec):
    """Every 16-bit value in every channel (R, G, B, A permuted differently) of
    an RGBA64 frame, all three output modes with alpha, against the oracle."""
    v = np.arange(1 << 16, dtype=np.uint32)
    px = np.stack([v, (v * 40503) & 0xFFFF, 0xFFFF - v, (v * 7919) & 0xFFFF], 1).astype("<u2")
    src = px.view(np.uint8).reshape(256, 256 * 8)
    for mode, bd, fr in ((0, 10, 0), (1, 12, 1), (2, 12, 0)):
        kw = dict(bit_depth=bd, full_range=fr, with_alpha=True)
        rc_o, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, 256, 256, abi.CVT_RGBA64, mode, **kw)
        rc_g, _, fg = cvt_call(codec._lib.tc_convert_packed_rgb, abi, src, 256, 256, abi.CVT_RGBA64, mode, **kw)
        assert rc_o == 0 and rc_g == 0
        for k, (a, b) in enumerate(zip(fo, fg)):
            assert np.array_equal(a, b), (mode, bd, fr, k)


@pytest.mark.parametrize("fmt", [2, 7])
def test_convert_4k_binding_and_device(oracle, abi, codec, fmt):
    rng = np.random.default_rng(12 + fmt)
    w, h = 3840, 2160
    src, _ = cvt_source(rng, fmt, abi, w, h)
    rc, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, w, h, fmt, 0, with_alpha=True)
    assert rc == 0
    arr = src if abi.cvt_sample_bytes(fmt) == 1 else src.view("<u2")
    got = codec.convert_packed_rgb(arr, w, h, fmt=fmt, out_mode=0)
    for a, b in zip(fo, got):
        assert np.array_equal(a, b)
    # device-resident: padded source pitch and padded output rows
    nch, smp = abi.cvt_channels(fmt), abi.cvt_sample_bytes(fmt)
    pitch = w * nch * smp + 512
    dsrc = codec.alloc(pitch * h)
    padded = np.zeros((h, pitch), np.uint8)
    padded[:, : w * nch * smp] = src
    dsrc.upload(padded)
    shapes = abi.cvt_plane_shapes(w, h, 0, True)
    outs = [codec.alloc((s[1] + 64) * s[0] * 2) for s in shapes]
    p = abi.CvtParams()
    p.format, p.width, p.height, p.out_mode, p.bit_depth = fmt, w, h, 0, 10
    p.kr, p.kb = 0.2126, 0.0722
    p.src, p.src_stride = dsrc.addr, pitch
    p.y_out, p.y_stride = outs[0].addr, w + 64
    p.u_out, p.u_stride = outs[1].addr, shapes[1][1] + 64
    p.v_out, p.v_stride = outs[2].addr, shapes[2][1] + 64
    p.a_out, p.a_stride = outs[3].addr, w + 64
    assert codec.convert_packed_rgb_dev(p) == 0, codec.last_error()
    for b, s, a in zip(outs, shapes, fo):
        got = b.download((s[1] + 64) * s[0] * 2, np.uint16).reshape(s[0], s[1] + 64)[:, : s[1]]
        assert np.array_equal(got, a)


# (format, width, row padding): rows of 270 / 8012 / 534 bytes are read sample
# by sample; 544 / 4016 / 4016 bytes take the vector loads (odd and even widths)
@pytest.mark.parametrize("fmt,w,pad", [(2, 67, 2), (6, 1001, 4), (7, 66, 6), (7, 67, 8), (3, 1001, 12),
                                       (2, 1000, 16)])
def test_convert_device_strides(oracle, abi, codec, fmt, w, pad):
    """Device-resident 4-channel sources whose row stride is not a multiple of
    16 bytes (read sample by sample) and one that is (vector loads), odd and
    even widths, every output mode, against the oracle."""
    rng = np.random.default_rng(40 + fmt + w)
    h = 6
    src, _ = cvt_source(rng, fmt, abi, w, h)
    nch, smp = abi.cvt_channels(fmt), abi.cvt_sample_bytes(fmt)
    pitch = w * nch * smp + pad
    assert (pitch % 16 == 0) == (pad in (8, 12, 16))
    padded = np.zeros((h, pitch), np.uint8)
    padded[:, : w * nch * smp] = src
    dsrc = codec.alloc(pitch * h)
    dsrc.upload(padded)
    for mode in (0, 1, 2):
        rc, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, w, h, fmt, mode, with_alpha=True)
        assert rc == 0
        shapes = abi.cvt_plane_shapes(w, h, mode, True)
        outs = [codec.alloc(s[1] * s[0] * 2) for s in shapes]
        p = abi.CvtParams()
        p.format, p.width, p.height, p.out_mode, p.bit_depth = fmt, w, h, mode, 10
        p.kr, p.kb = 0.2126, 0.0722
        p.src, p.src_stride = dsrc.addr, pitch
        p.y_out, p.u_out, p.v_out, p.a_out = (b.addr for b in outs)
        assert codec.convert_packed_rgb_dev(p) == 0, codec.last_error()
        for b, sh, want in zip(outs, shapes, fo):
            got = b.download(sh[1] * sh[0] * 2, np.uint16).reshape(sh[0], sh[1])
            assert np.array_equal(got, want), (fmt, w, pad, mode)


def test_planar_444_to_422(oracle, abi, codec):
    rng = np.random.default_rng(13)
    for w, h in ((1, 3), (3, 2), (7, 2), (8, 2), (9, 2), (15, 2), (16, 2), (17, 2), (257, 5), (3840, 2160)):
        u = rng.integers(0, 65536, size=(h, w + 7), dtype=np.uint16)
        v = rng.integers(0, 65536, size=(h, w + 7), dtype=np.uint16)
        rc_o, uo, vo = p422_call(oracle.orc_convert_444_422, abi, u, v, w, h, bit_depth=16, in_stride=w + 7,
                                 out_pad=3)
        rc_g, ug, vg = p422_call(codec._lib.tc_convert_yuv444_to_422, abi, u, v, w, h, bit_depth=16,
                                 in_stride=w + 7, out_pad=3)
        assert rc_o == 0 and rc_g == 0
        assert np.array_equal(uo, ug) and np.array_equal(vo, vg), (w, h)
    u = rng.integers(0, 1024, size=(64, 130), dtype=np.uint16)
    v = rng.integers(0, 1024, size=(64, 130), dtype=np.uint16)
    got = codec.convert_yuv444_to_422(u, v, bit_depth=10)
    _, uo, vo = p422_call(oracle.orc_convert_444_422, abi, u, v, 130, 64, bit_depth=10)
    assert np.array_equal(got[0], uo) and np.array_equal(got[1], vo)


def test_converted_frame_encodes_like_oracle(oracle, abi, codec):
    """BGRA 1080p -> 4:2:2 10-bit planes on the device -> batched encode: the packet
    equals the oracle's encode of the oracle's conversion."""
    rng = np.random.default_rng(14)
    w, h = 1920, 1080
    src, _ = cvt_source(rng, abi.CVT_BGRA32, abi, w, h)
    src = np.ascontiguousarray(np.repeat(src.reshape(h // 8, 8, -1)[:, :1], 8, 1).reshape(h, -1))  # smoother rows
    rc, (y, u, v), _ = cvt_call(oracle.orc_convert_packed_rgb, abi, src, w, h, abi.CVT_BGRA32, 0)
    assert rc == 0
    cfg = abi.make_config(w, h, 62, profile=3, entropy_mode=8)
    rc, pkt = orc_encode(oracle, cfg, [np.ascontiguousarray(y), np.ascontiguousarray(u), np.ascontiguousarray(v)])
    assert rc == 0
    dsrc = codec.alloc(src.nbytes)
    dsrc.upload(src)
    planes = [codec.alloc(w * h * 2), codec.alloc((w // 2) * h * 2), codec.alloc((w // 2) * h * 2)]
    p = abi.CvtParams()
    p.format, p.width, p.height, p.out_mode, p.bit_depth = abi.CVT_BGRA32, w, h, 0, 10
    p.kr, p.kb = 0.2126, 0.0722
    p.src = dsrc.addr
    p.y_out, p.u_out, p.v_out = planes[0].addr, planes[1].addr, planes[2].addr
    assert codec.convert_packed_rgb_dev(p) == 0, codec.last_error()
    out = codec.alloc(len(pkt) + (1 << 20))
    st, sizes, rcs, _ = codec.encode_batch_dev([cfg], [[b.addr for b in planes]], [out])
    assert st == 0 and rcs == [0], codec.last_error()
    assert sizes[0] == len(pkt) and out.download(sizes[0]).tobytes() == pkt


@pytest.mark.parametrize("fmt", [0, 2, 5, 7])
@pytest.mark.parametrize("w", [1001, 1002, 1003, 1004])
def test_convert_device_quads_and_row_tails(oracle, abi, codec, fmt, w):
    """Device-resident buffers that take four pixels per thread (16-byte
    aligned source rows, output rows padded to 8 elements): 3- and 4-channel,
    8- and 16-bit sources, widths ending in a full quad and in a 1-, 2- or
    3-pixel tail, every output mode, against the oracle."""
    rng = np.random.default_rng(70 + fmt + w)
    h = 5
    src, _ = cvt_source(rng, fmt, abi, w, h)
    nch, smp = abi.cvt_channels(fmt), abi.cvt_sample_bytes(fmt)
    pitch = (w * nch * smp + 15) // 16 * 16 + 16
    padded = np.zeros((h, pitch), np.uint8)
    padded[:, : w * nch * smp] = src
    dsrc = codec.alloc(pitch * h)
    dsrc.upload(padded)
    for mode in (0, 1, 2):
        rc, _, fo = cvt_call(oracle.orc_convert_packed_rgb, abi, src, w, h, fmt, mode, with_alpha=True)
        assert rc == 0
        shapes = abi.cvt_plane_shapes(w, h, mode, True)
        strides = [(s[1] + 7) // 8 * 8 + 8 for s in shapes]
        outs = [codec.alloc(st * s[0] * 2) for st, s in zip(strides, shapes)]
        for b, st, s in zip(outs, strides, shapes):
            b.upload(np.full(st * s[0], 0xBEEF, np.uint16))
        p = abi.CvtParams()
        p.format, p.width, p.height, p.out_mode, p.bit_depth = fmt, w, h, mode, 10
        p.kr, p.kb = 0.2126, 0.0722
        p.src, p.src_stride = dsrc.addr, pitch
        p.y_out, p.u_out, p.v_out, p.a_out = (b.addr for b in outs)
        p.y_stride, p.u_stride, p.v_stride, p.a_stride = strides
        assert codec.convert_packed_rgb_dev(p) == 0, codec.last_error()
        for b, st, sh, want in zip(outs, strides, shapes, fo):
            got = b.download(st * sh[0] * 2, np.uint16).reshape(sh[0], st)
            assert np.array_equal(got[:, : sh[1]], want), (fmt, w, mode)
            assert (got[:, sh[1]:] == 0xBEEF).all(), (fmt, w, mode)


@pytest.mark.parametrize("w", [1, 9, 16, 17, 1000, 1001, 1003])
@pytest.mark.parametrize("aligned", [True, False])
def test_planar_444_to_422_device(oracle, abi, codec, w, aligned):
    """Device-resident planes: rows 16-byte aligned in and 8-byte aligned out
    (four output samples per thread, row tails sample by sample) and odd
    strides (one thread per sample), against the oracle; padding untouched."""
    rng = np.random.default_rng(90 + w)
    h = 4
    cw = (w + 1) // 2
    ist = (w + 7) // 8 * 8 + 8 if aligned else w + 3
    ost = (cw + 3) // 4 * 4 + 4 if aligned else cw + 1
    u = rng.integers(0, 65536, size=(h, ist), dtype=np.uint16)
    v = rng.integers(0, 65536, size=(h, ist), dtype=np.uint16)
    rc, uo, vo = p422_call(oracle.orc_convert_444_422, abi, u, v, w, h, bit_depth=16, in_stride=ist)
    assert rc == 0
    bufs = [codec.alloc(a.nbytes) for a in (u, v)]
    for b, a in zip(bufs, (u, v)):
        b.upload(a)
    outs = [codec.alloc(ost * h * 2) for _ in range(2)]
    for b in outs:
        b.upload(np.full(ost * h, 0xBEEF, np.uint16))
    p = abi.Planar422Params()
    p.width, p.height, p.bit_depth = w, h, 16
    p.u_in, p.u_in_stride, p.v_in, p.v_in_stride = bufs[0].addr, ist, bufs[1].addr, ist
    p.u_out, p.u_out_stride, p.v_out, p.v_out_stride = outs[0].addr, ost, outs[1].addr, ost
    assert codec.convert_yuv444_to_422_dev(p) == 0, codec.last_error()
    for b, want in zip(outs, (uo, vo)):
        got = b.download(ost * h * 2, np.uint16).reshape(h, ost)
        assert np.array_equal(got[:, :cw], want), (w, aligned)
        assert (got[:, cw:] == 0xBEEF).all(), (w, aligned)
