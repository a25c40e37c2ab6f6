"""Reduced / scaled / request decode (tc_frame_decode_reduced, _scaled,
_request), CPU side: the oracle's restatement against the compiled reference
(native/src/codec/codec.c:9100-9640) over every scale, arbitrary targets,
AUTO_2K, DROP_ALPHA, rANS2 / V1 Rice / V2 VLC / RAW / 4:4:4 + alpha streams,
deblocked streams on exact 2:1 targets and a concealed slice."""
import ctypes

import numpy as np
import pytest

from helpers import frame_for, lib_decode_request, orc_encode, orc_scaled

PACKETS = [
    ("422", 256, 128, 31, dict(qp=62, profile=3, features=3, slice_rows=2)),
    ("422", 250, 130, 32, dict(qp=70, profile=3, features=1, slice_rows=4)),
    ("422", 96, 40, 33, dict(qp=50, profile=3, qmatrix_id=1, slice_rows=1, entropy_mode=0)),
    ("444", 64, 48, 34, dict(qp=58, profile=5, pixel_format=1, slice_rows=2, entropy_mode=1)),
    ("4444", 64, 32, 35, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=2, alpha_mode=1,
                              alpha_bit_depth=16)),
    ("raw", 128, 64, 36, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=2)),
]


def _packet(pkg, abi, syn, oracle, spec):
    kind, w, h, seed, kw = spec
    kw = dict(kw)
    qp, em = kw.pop("qp"), kw.pop("entropy_mode", 8)
    cfg = abi.make_config(w, h, qp, entropy_mode=em, **kw)
    rc, pkt = orc_encode(oracle, cfg, frame_for(pkg, syn, kind, w, h, seed))
    assert rc == 0
    return pkt, cfg


def _ref(ref):
    ref.tc_frame_decode_request.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p]
    return ref


def oracle_request(oracle, abi, pkt, cfg, req):
    """The request plan (codec.c:9277-9357) in Python, then the oracle decode."""
    W, H = cfg.visible_width, cfg.visible_height
    pf, am = cfg.pixel_format, cfg.alpha_mode
    drop = 1 if (req.flags & abi.DECODE_FLAG_DROP_ALPHA) and am else 0
    nplanes = (4 if pf == 3 else 3 + (1 if am else 0)) - drop
    mode, tw, th, limit, scaled = req.mode, W, H, 63, False
    if mode == abi.DECODE_MODE_SCALED:
        tw, th, scaled = req.target_width, req.target_height, True
    elif mode == abi.DECODE_MODE_REDUCED and req.scale:
        d = abi.SCALE_DIVISOR[req.scale]
        tw, th, limit, scaled = -(-W // d), -(-H // d), abi.SCALE_LIMIT[req.scale], True
    elif mode == abi.DECODE_MODE_AUTO_2K:
        if req.target_width:
            tw, th = req.target_width, req.target_height
            scaled = (tw, th) != (W, H)
        elif W > 2048 or H > 2048:
            for s in (1, 2, 3, 4):
                d = abi.SCALE_DIVISOR[s]
                if -(-W // d) <= 2048 and -(-H // d) <= 2048:
                    tw, th, limit, scaled = -(-W // d), -(-H // d), abi.SCALE_LIMIT[s], True
                    break
    dims = abi.target_plane_dims(tw, th, pf, nplanes)
    if not scaled:
        outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
        ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
        st = (ctypes.c_uint8 * 512)()
        nc = ctypes.c_uint32(0)
        rc = oracle.orc_frame_decode_ex(pkt, ctypes.c_size_t(len(pkt)), ptrs, None, st, ctypes.byref(nc), drop)
        return rc, outs, nc.value
    rc, outs, _, nc = orc_scaled(oracle, pkt, tw, th, limit, dims, drop)
    return rc, outs, nc


def _requests(abi, cfg):
    W, H = cfg.visible_width, cfg.visible_height
    reqs = [abi.make_request(abi.DECODE_MODE_FULL), abi.make_request(abi.DECODE_MODE_AUTO_2K)]
    reqs += [abi.make_request(abi.DECODE_MODE_REDUCED, s) for s in range(5)]
    for t in ((W // 2, H // 2), (W // 3 + 1, H // 3), (37, 21), (W, H), (1, 1), (W - 1, H)):
        reqs.append(abi.make_request(abi.DECODE_MODE_SCALED, target=t))
    reqs.append(abi.make_request(abi.DECODE_MODE_AUTO_2K, target=(W // 2, H // 2)))
    if cfg.alpha_mode:
        reqs += [abi.make_request(m, s, flags=abi.DECODE_FLAG_DROP_ALPHA)
                 for (m, s) in ((abi.DECODE_MODE_FULL, 0), (abi.DECODE_MODE_REDUCED, 1), (abi.DECODE_MODE_REDUCED, 3))]
    return reqs


def undefined_samples(abi, cfg, plane, dw, dh):
    """Target samples the reference leaves undefined: on an exact 2:1 colour
    plane whose width is not a multiple of 8, the right-edge blocks are
    clipped to nx < 4 target columns, but the half-sample kernels
    (native/src/simd/transform_avx2.c:527-568, dispatch.c:279-307) fill only
    the first nx * ny entries of the block's row-major 4x4 value array and
    native/src/codec/color_store.h:315-333 reads entry iy * 4 + ix, so the
    entries past nx * ny come from uninitialised stack memory. Here those
    samples are the nearest source sample, like every other target sample."""
    mask = np.zeros((dh, dw), bool)
    W, H = cfg.visible_width, cfg.visible_height
    dims = abi.plane_dims(W, H, cfg.pixel_format, cfg.alpha_mode)
    if plane >= len(dims) or (cfg.alpha_mode and plane == 3):
        return mask
    vw, vh = dims[plane]
    if dw * 2 != vw or dh * 2 != vh or vw % 8 == 0:
        return mask
    px0 = 8 * ((vw - 1) // 8)
    nx = (vw - px0) // 2
    for py0 in range(0, vh, 8):
        ny = min(8, vh - py0) // 2
        for iy in range(ny):
            for ix in range(nx):
                if iy * 4 + ix >= nx * ny:
                    mask[py0 // 2 + iy, px0 // 2 + ix] = True
    return mask


def _compare(oracle, ref, abi, pkt, cfg, req, what):
    rr, rp, rinfo = lib_decode_request(ref, abi, pkt, req)
    orc_rc, op, onc = oracle_request(oracle, abi, pkt, cfg, req)
    assert rr == orc_rc, (what, rr, orc_rc)
    assert rinfo.concealed_slices == onc, what
    assert len(rp) == len(op), what
    for k, (a, b) in enumerate(zip(rp, op)):
        assert a.shape == b.shape, (what, k)
        m = undefined_samples(abi, cfg, k, a.shape[1], a.shape[0])
        assert np.array_equal(a[~m], b[~m]), (what, k)


@pytest.mark.parametrize("i", range(len(PACKETS)))
def test_preview_oracle_matches_reference(pkg, abi, syn, oracle, ref, i):
    _ref(ref)
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[i])
    for req in _requests(abi, cfg):
        _compare(oracle, ref, abi, pkt, cfg, req,
                 (PACKETS[i][:3], req.mode, req.scale, req.target_width, req.target_height, req.flags))


def test_preview_concealed_slice(pkg, abi, syn, oracle, ref):
    _ref(ref)
    pkt, cfg = _packet(pkg, abi, syn, oracle, PACKETS[0])
    bad = bytearray(pkt)
    bad[53 + 17 + 5] ^= 0x5A          # first slice payload: its CRC fails
    bad = bytes(bad)
    for req in (abi.make_request(abi.DECODE_MODE_REDUCED, 1), abi.make_request(abi.DECODE_MODE_SCALED, target=(50, 30)),
                abi.make_request(abi.DECODE_MODE_REDUCED, 2)):
        _compare(oracle, ref, abi, bad, cfg, req, ("conceal", req.mode, req.scale))


def test_preview_auto2k_4k(pkg, abi, syn, oracle, ref):
    _ref(ref)
    spec = ("422", 3840, 2160, 37, dict(qp=62, profile=3, features=3, slice_rows=8))
    pkt, cfg = _packet(pkg, abi, syn, oracle, spec)
    for req in (abi.make_request(abi.DECODE_MODE_AUTO_2K), abi.make_request(abi.DECODE_MODE_REDUCED, 4)):
        _compare(oracle, ref, abi, pkt, cfg, req, ("4k", req.mode, req.scale))


def test_scale_dimensions_and_limits(ref, abi):
    """tc_decode_scale_dimensions = ceil(source / divisor) (codec.c:9174-9194)."""
    ref.tc_decode_scale_dimensions.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                               ctypes.c_void_p, ctypes.c_void_p]
    for (w, h) in ((3840, 2160), (1, 1), (1921, 1081), (8191, 4319)):
        for s in range(5):
            tw, th = ctypes.c_uint32(0), ctypes.c_uint32(0)
            assert ref.tc_decode_scale_dimensions(w, h, s, ctypes.byref(tw), ctypes.byref(th)) == 0
            d = abi.SCALE_DIVISOR[s]
            assert (tw.value, th.value) == (-(-w // d), -(-h // d))
