"""Reduced / scaled / request decode on the GPU against the oracle: every
request mode, scale and target over rANS2, V1 Rice, V2 VLC, 4:4:4 + alpha and
RAW streams, deblocked exact 2:1 targets, a concealed slice, 4K AUTO_2K, the
batch and device-resident forms with padded strides, the .toos preview, and
a full-size decode into padded device strides of a deblocking stream."""
import ctypes

import numpy as np
import pytest

from helpers import lib_decode_request, orc_decode, orc_scaled, toos_wrap
from test_preview_oracle import PACKETS, _packet, _requests, oracle_request

pytestmark = pytest.mark.gpu


def _check(codec, oracle, abi, pkt, cfg, req, what):
    rc, got, info = lib_decode_request(codec._lib, abi, pkt, req)
    orc_rc, want, nc = oracle_request(oracle, abi, pkt, cfg, req)
    assert rc == orc_rc, (what, rc, orc_rc, codec.last_error())
    assert info.concealed_slices == nc, what
    assert len(got) == len(want), what
    for k, (a, b) in enumerate(zip(got, want)):
        assert a.shape == b.shape and np.array_equal(a, b), (what, k)


@pytest.mark.parametrize("i", range(len(PACKETS)))
def test_request_modes_match_oracle(pkg, abi, syn, oracle, codec, i):
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[i])
    for req in _requests(abi, cfg):
        _check(codec, oracle, abi, pkt, cfg, req,
               (PACKETS[i][:3], req.mode, req.scale, req.target_width, req.target_height, req.flags))


def test_reduced_and_scaled_entries(pkg, abi, syn, oracle, codec):
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[0])
    lib = codec._lib
    W, H = cfg.visible_width, cfg.visible_height
    for s in range(1, 5):
        d = abi.SCALE_DIVISOR[s]
        tw, th = -(-W // d), -(-H // d)
        dims = abi.target_plane_dims(tw, th, 0, 3)
        outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
        ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None]))
        info = abi.FrameOutput()
        assert lib.tc_frame_decode_reduced(pkt, len(pkt), s, ptrs, None, ctypes.byref(info)) == 0
        assert (info.visible_width, info.visible_height) == (tw, th)
        rc, want, _, _ = orc_scaled(oracle, pkt, tw, th, abi.SCALE_LIMIT[s], dims)
        assert rc == 0 and all(np.array_equal(a, b) for a, b in zip(outs, want)), s
    for (tw, th) in ((100, 50), (W // 2, H // 2), (W, H)):
        dims = abi.target_plane_dims(tw, th, 0, 3)
        outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
        ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None]))
        info = abi.FrameOutput()
        assert lib.tc_frame_decode_scaled(pkt, len(pkt), tw, th, ptrs, None, ctypes.byref(info)) == 0
        rc, want, _, _ = orc_scaled(oracle, pkt, tw, th, 63, dims)
        assert rc == 0 and all(np.array_equal(a, b) for a, b in zip(outs, want)), (tw, th)
    info = abi.FrameOutput()
    assert lib.tc_frame_decode_scaled(pkt, len(pkt), W + 1, H, None, None, ctypes.byref(info)) == \
        abi.TC_ERR_INVALID_ARGUMENT


def test_concealed_slice(pkg, abi, syn, oracle, codec):
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[0])
    bad = bytearray(pkt)
    bad[53 + 17 + 5] ^= 0x5A
    bad = bytes(bad)
    for req in (abi.make_request(abi.DECODE_MODE_REDUCED, 1), abi.make_request(abi.DECODE_MODE_SCALED, target=(50, 30)),
                abi.make_request(abi.DECODE_MODE_REDUCED, 2)):
        _check(codec, oracle, abi, bad, cfg, req, ("conceal", req.mode, req.scale))


def test_auto2k_4k(pkg, abi, syn, oracle, codec):
    from test_preview_oracle import _packet as mk
    spec = ("422", 3840, 2160, 37, dict(qp=62, profile=3, features=3, slice_rows=8))
    pkt, cfg = mk(pkg, abi, syn, oracle, spec)
    for req in (abi.make_request(abi.DECODE_MODE_AUTO_2K), abi.make_request(abi.DECODE_MODE_REDUCED, 4),
                abi.make_request(abi.DECODE_MODE_SCALED, target=(1280, 720))):
        _check(codec, oracle, abi, pkt, cfg, req, ("4k", req.mode, req.scale))


def test_batch_reduced(pkg, abi, syn, oracle, codec):
    items = [_packet(pkg, abi, syn, oracle, PACKETS[i]) for i in (0, 2, 3, 4)]
    n = len(items)
    for s in (0, 1, 3):
        pk = (abi.BatchPacket * n)()
        bufs = [np.frombuffer(p, np.uint8) for p, _ in items]
        for i, b in enumerate(bufs):
            pk[i].data, pk[i].size = b.ctypes.data, b.nbytes
        infos = (abi.FrameOutput * n)()
        assert codec._lib.tc_frame_decode_batch_reduced(pk, n, s, None, None, infos) == 0
        outs, ptrs = [], (ctypes.c_void_p * (4 * n))()
        for i in range(n):
            dims = abi.target_plane_dims(infos[i].visible_width, infos[i].visible_height, infos[i].pixel_format,
                                         infos[i].plane_count)
            o = [np.zeros((h, w), np.uint16) for (w, h) in dims]
            outs.append(o)
            for p, a in enumerate(o):
                ptrs[4 * i + p] = a.ctypes.data
        assert codec._lib.tc_frame_decode_batch_reduced(pk, n, s, ptrs, None, infos) == 0, codec.last_error()
        for i, (pkt, cfg) in enumerate(items):
            req = abi.make_request(abi.DECODE_MODE_REDUCED, s) if s else abi.make_request(abi.DECODE_MODE_FULL)
            _, want, _ = oracle_request(oracle, abi, pkt, cfg, req)
            assert all(np.array_equal(a, b) for a, b in zip(outs[i], want)), (s, i)


def test_device_request_padded_strides(pkg, abi, syn, oracle, codec):
    """tcx_decode_batch_dev_request with output strides 40 samples wider than the
    target planes, and a full-size decode of a deblocking stream into padded
    device strides."""
    items = [_packet(pkg, abi, syn, oracle, PACKETS[i]) for i in (0, 1, 4)]
    for req in (abi.make_request(abi.DECODE_MODE_REDUCED, 1), abi.make_request(abi.DECODE_MODE_FULL),
                abi.make_request(abi.DECODE_MODE_SCALED, target=(60, 30)),
                abi.make_request(abi.DECODE_MODE_REDUCED, 2, flags=abi.DECODE_FLAG_DROP_ALPHA)):
        n = len(items)
        host = (ctypes.c_void_p * n)()
        sizes = (ctypes.c_size_t * n)()
        dpk = (ctypes.c_uint64 * n)()
        planes = (ctypes.c_uint64 * (4 * n))()
        strides = (ctypes.c_size_t * (4 * n))()
        infos = (abi.FrameOutput * n)()
        rcs = (ctypes.c_int32 * n)()
        keep, bufs, wants = [], [], []
        for i, (pkt, cfg) in enumerate(items):
            a = np.frombuffer(pkt, np.uint8)
            keep.append(a)
            host[i], sizes[i] = a.ctypes.data, len(pkt)
            b = codec.alloc(len(pkt))
            b.upload(a)
            keep.append(b)
            dpk[i] = b.addr
            _, want, _ = oracle_request(oracle, abi, pkt, cfg, req)
            wants.append(want)
            row = []
            for p, w in enumerate(want):
                st = w.shape[1] + 40
                d = codec.alloc(st * w.shape[0] * 2)
                keep.append(d)
                planes[4 * i + p] = d.addr
                strides[4 * i + p] = st
                row.append((d, st))
            bufs.append(row)
        st = codec._lib.tcx_decode_batch_dev_request(n, host, sizes, dpk, planes, strides, infos, rcs,
                                                     ctypes.byref(req))
        assert st == 0 and list(rcs) == [0] * n, (req.mode, codec.last_error())
        for i in range(n):
            assert infos[i].plane_count == len(wants[i])
            for (d, stp), w in zip(bufs[i], wants[i]):
                got = d.download(stp * w.shape[0] * 2, np.uint16).reshape(w.shape[0], stp)[:, : w.shape[1]]
                assert np.array_equal(got, w), (req.mode, req.scale, i)


@pytest.mark.parametrize("pad,shift", [(8, 0), (4, 8), (1, 2)])
def test_deblock_plane_alignment(pkg, abi, syn, oracle, codec, pad, shift):
    """Full-size decode of a deblocking stream into device planes of each
    alignment class of the in-place filter: base and row pitch 16-byte aligned
    (aligned row chunks), 8-byte aligned (half-row vectors) and 2-byte aligned
    (sample by sample)."""
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[0])
    req = abi.make_request(abi.DECODE_MODE_FULL)
    _, want, _ = oracle_request(oracle, abi, pkt, cfg, req)
    a = np.frombuffer(pkt, np.uint8)
    b = codec.alloc(len(pkt))
    b.upload(a)
    host = (ctypes.c_void_p * 1)(a.ctypes.data)
    sizes = (ctypes.c_size_t * 1)(len(pkt))
    dpk = (ctypes.c_uint64 * 1)(b.addr)
    planes = (ctypes.c_uint64 * 4)()
    strides = (ctypes.c_size_t * 4)()
    infos = (abi.FrameOutput * 1)()
    rcs = (ctypes.c_int32 * 1)()
    bufs = []
    for p, w in enumerate(want):
        st = w.shape[1] + pad
        d = codec.alloc(st * w.shape[0] * 2 + 16)
        d.upload(np.full(st * w.shape[0] + 8, 0x5A5A, np.uint16))
        bufs.append((d, st))
        planes[p] = d.addr + shift
        strides[p] = st
    st = codec._lib.tcx_decode_batch_dev_request(1, host, sizes, dpk, planes, strides, infos, rcs,
                                                 ctypes.byref(req))
    assert st == 0 and rcs[0] == 0, codec.last_error()
    for (d, stp), w in zip(bufs, want):
        raw = d.download(stp * w.shape[0] * 2 + 16)
        got = raw[shift: shift + stp * w.shape[0] * 2].view(np.uint16).reshape(w.shape[0], stp)
        assert np.array_equal(got[:, : w.shape[1]], w), (pad, shift)
        # the row padding and the bytes before the plane keep their contents
        assert (got[:, w.shape[1]:] == 0x5A5A).all() and (raw[:shift] == 0x5A).all(), (pad, shift)


def test_image_preview(pkg, abi, syn, oracle, codec):
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[0])
    data = toos_wrap(pkt, 0)
    img = pkg.ToposImageCodec()
    W, H = cfg.visible_width, cfg.visible_height
    for (tw, th) in ((W // 2, H // 2), (100, 40), (W // 4, H // 4)):
        dims = abi.target_plane_dims(tw, th, 0, 3)
        outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
        st, info = img.decode_preview_status(data, tw, th, outs)
        assert st == 0, (tw, th, st)
        s = next((s for s in range(1, 5) if (-(-W // abi.SCALE_DIVISOR[s]), -(-H // abi.SCALE_DIVISOR[s])) == (tw, th)),
                 None)
        limit = abi.SCALE_LIMIT[s] if s else 63
        rc, want, _, _ = orc_scaled(oracle, pkt, tw, th, limit, dims)
        assert rc == 0 and all(np.array_equal(a, b) for a, b in zip(outs, want)), (tw, th)
        assert (info.visible_width, info.visible_height) == (tw, th)
