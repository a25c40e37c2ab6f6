"""GPU parity of the public batch and decoder-context entry points a drop-in
caller uses (reference topos_codec.h:374-411, 502-571): tc_frame_decode_batch
with valid, concealed, rejected and structurally broken packets, and the
tc_decoder_* contexts (prepare, decode, batch, request and surface forms,
one context reused across frames of different sizes). Every decoded plane
and slice status is compared with the oracle's decode of the same packet."""
import ctypes

import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

vp = ctypes.c_void_p
SENTINEL = 0xBEEF

FRAMES = [
    # (kind, w, h, seed, cfg kwargs)
    ("422", 640, 360, 11, dict(qp=40, profile=3, qmatrix_id=4, slice_rows=4, features=1)),
    ("4444", 256, 144, 12, dict(qp=30, profile=6, pixel_format=2, bit_depth=12, slice_rows=2, alpha_mode=1,
                               alpha_bit_depth=16)),
    ("raw", 320, 256, 13, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 77, 45, 14, dict(qp=10, profile=3, slice_rows=1)),
]


def slice_payloads(pkt):
    """(payload offset, payload size) of every slice of a packet."""
    off, out = 53, []
    while off < len(pkt):
        size = int.from_bytes(pkt[off:off + 4], "big")
        out.append((off + 17, size))
        off += 17 + size
    return out


def damage_slice(pkt, k):
    b = bytearray(pkt)
    off, size = slice_payloads(pkt)[k]
    b[off + size // 2] ^= 0x5A
    return bytes(b)


@pytest.fixture(scope="module")
def streams(pkg, abi, syn, oracle):
    out = []
    for kind, w, h, seed, kw in FRAMES:
        kw = dict(kw)
        qp = kw.pop("qp")
        cfg = abi.make_config(w, h, qp, entropy_mode=8, **kw)
        rc, pkt = orc_encode(oracle, cfg, frame_for(pkg, syn, kind, w, h, seed))
        assert rc == 0
        dims = abi.plane_dims(w, h, kw.get("pixel_format", 0), kw.get("alpha_mode", 0))
        out.append((pkt, dims))
    return out


def expect(oracle, pkt, dims):
    rc, planes, status, nconc = orc_decode(oracle, pkt, dims)
    return rc, planes, status[:len(slice_payloads(pkt))], nconc


def lib_of(pkg):
    lib = pkg.load_library()
    lib.tc_frame_decode_batch.argtypes = [vp, ctypes.c_uint32, vp, vp, vp]
    lib.tc_decoder_create.argtypes = [vp, vp]
    lib.tc_decoder_destroy.argtypes = [vp]
    lib.tc_decoder_destroy.restype = None
    lib.tc_decoder_set_max_slice_workers.argtypes = [vp, ctypes.c_uint32]
    lib.tc_decoder_prepare.argtypes = [vp, vp, ctypes.c_size_t, vp]
    lib.tc_decoder_decode.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
    lib.tc_decoder_decode_request.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp]
    lib.tc_decoder_decode_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp]
    lib.tc_decoder_decode_batch_request.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, vp]
    lib.tc_decoder_decode_surface.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
    lib.tc_decoder_decode_surface_request.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp]
    return lib


def batch_args(abi, pkts, dims_list, null_plane=None):
    n = len(pkts)
    bp = (abi.BatchPacket * n)()
    keep = [np.frombuffer(p, np.uint8) for p in pkts]
    for i, b in enumerate(keep):
        bp[i].data, bp[i].size = b.ctypes.data, len(pkts[i])
    outs = [[np.full((h, w), SENTINEL, np.uint16) for (w, h) in dims] for dims in dims_list]
    ptrs = (ctypes.c_void_p * (4 * n))()
    for i, o in enumerate(outs):
        for p, a in enumerate(o):
            ptrs[4 * i + p] = a.ctypes.data
    if null_plane is not None:
        ptrs[4 * null_plane[0] + null_plane[1]] = None
    return bp, keep, outs, ptrs, (abi.FrameOutput * n)()


def check_frame(oracle, pkt, dims, planes, info):
    rc, ref, status, nconc = expect(oracle, pkt, dims)
    for a, b in zip(planes, ref):
        assert (a == b).all()
    assert list(info.slice_status)[:len(status)] == status
    assert info.concealed_slices == nconc
    return rc


@pytest.mark.gpu
def test_decode_batch_valid_frames(pkg, abi, oracle, streams):
    lib = lib_of(pkg)
    pkts = [s[0] for s in streams]
    bp, keep, outs, ptrs, infos = batch_args(abi, pkts, [s[1] for s in streams])
    assert lib.tc_frame_decode_batch(bp, len(pkts), None, None, infos) == 0     # query mode
    assert [infos[i].visible_width for i in range(len(pkts))] == [f[1] for f in FRAMES]
    assert lib.tc_frame_decode_batch(bp, len(pkts), ptrs, None, infos) == 0
    for i, (pkt, dims) in enumerate(streams):
        assert check_frame(oracle, pkt, dims, outs[i], infos[i]) == 0


@pytest.mark.gpu
def test_decode_batch_concealed_frame_delivered(pkg, abi, oracle, streams):
    lib = lib_of(pkg)
    pkts = [s[0] for s in streams]
    pkts[1] = damage_slice(pkts[1], 2)          # one colour slice fails its CRC
    pkts[3] = damage_slice(pkts[3], 0)
    bp, keep, outs, ptrs, infos = batch_args(abi, pkts, [s[1] for s in streams])
    rc = lib.tc_frame_decode_batch(bp, len(pkts), ptrs, None, infos)
    assert rc == abi.TC_WARN_CONCEALED           # first non-OK status in frame order
    rcs = [check_frame(oracle, pkts[i], streams[i][1], outs[i], infos[i]) for i in range(len(pkts))]
    assert rcs == [0, abi.TC_WARN_CONCEALED, 0, abi.TC_WARN_CONCEALED]
    assert infos[1].concealed_slices == 1 and infos[1].slice_status[2] == 1


@pytest.mark.gpu
def test_decode_batch_structure_failure_rejects_whole_batch(pkg, abi, streams):
    lib = lib_of(pkg)
    pkts = [s[0] for s in streams]
    broken = bytearray(pkts[2])
    broken[30] ^= 0x01                           # frame header CRC mismatch
    pkts[2] = bytes(broken)
    bp, keep, outs, ptrs, infos = batch_args(abi, pkts, [s[1] for s in streams])
    assert lib.tc_frame_decode_batch(bp, len(pkts), ptrs, None, infos) == abi.TC_ERR_CHECKSUM_MISMATCH
    assert all((p == SENTINEL).all() for o in outs for p in o)     # nothing produced
    pkts[2] = streams[2][0][:len(streams[2][0]) - 5]                # truncated tail
    bp, keep, outs, ptrs, infos = batch_args(abi, pkts, [s[1] for s in streams])
    assert lib.tc_frame_decode_batch(bp, len(pkts), ptrs, None, infos) < 0
    assert all((p == SENTINEL).all() for o in outs for p in o)


@pytest.mark.gpu
def test_decode_batch_bad_frame_rejected_alone(pkg, abi, oracle, streams):
    lib = lib_of(pkg)
    pkts = [s[0] for s in streams]
    bp, keep, outs, ptrs, infos = batch_args(abi, pkts, [s[1] for s in streams], null_plane=(1, 2))
    assert lib.tc_frame_decode_batch(bp, len(pkts), ptrs, None, infos) == abi.TC_ERR_INVALID_ARGUMENT
    for i in (0, 2, 3):
        assert check_frame(oracle, pkts[i], streams[i][1], outs[i], infos[i]) == 0
    assert all((p == SENTINEL).all() for p in outs[1][:2])


@pytest.mark.gpu
def test_decoder_context_forms_match_oracle(pkg, abi, oracle, streams):
    lib = lib_of(pkg)
    d = ctypes.c_void_p()
    cfg = abi.DecoderConfig()
    cfg.struct_size, cfg.abi_version, cfg.max_slice_workers = ctypes.sizeof(cfg), 2, 4
    assert lib.tc_decoder_create(ctypes.byref(cfg), ctypes.byref(d)) == 0
    try:
        full = abi.make_request(abi.DECODE_MODE_FULL)
        for rep in range(2):                     # the same context across frames of different sizes, twice
            for k, (pkt, dims) in enumerate(streams):
                info = abi.FrameOutput()
                assert lib.tc_decoder_prepare(d, pkt, len(pkt), ctypes.byref(info)) == 0
                assert info.visible_width == FRAMES[k][1]
                outs = [np.full((h, w), SENTINEL, np.uint16) for (w, h) in dims]
                views = abi.plane_views(outs)
                assert lib.tc_decoder_decode(d, pkt, len(pkt), views, ctypes.byref(info)) == 0
                check_frame(oracle, pkt, dims, outs, info)
                outs2 = [np.full((h, w), SENTINEL, np.uint16) for (w, h) in dims]
                assert lib.tc_decoder_decode_request(d, pkt, len(pkt), ctypes.byref(full), abi.plane_views(outs2),
                                                     ctypes.byref(info)) == 0
                assert all((a == b).all() for a, b in zip(outs, outs2))
                for mem in (abi.TC_DECODE_MEMORY_CPU, abi.TC_DECODE_MEMORY_HOST_VISIBLE_GPU):
                    outs3 = [np.full((h, w), SENTINEL, np.uint16) for (w, h) in dims]
                    s = abi.DecodeSurface()
                    s.struct_size, s.abi_version, s.memory_type = ctypes.sizeof(s), 2, mem
                    v3 = abi.plane_views(outs3)
                    for i in range(4):
                        s.planes[i] = v3[i]
                    assert lib.tc_decoder_decode_surface(d, pkt, len(pkt), ctypes.byref(s), ctypes.byref(info)) == 0
                    assert all((a == b).all() for a, b in zip(outs, outs3))
                    outs3[0][:] = SENTINEL
                    full.memory_type = mem
                    assert lib.tc_decoder_decode_surface_request(d, pkt, len(pkt), ctypes.byref(full),
                                                                 ctypes.byref(s), ctypes.byref(info)) == 0
                    full.memory_type = 0
                    assert (outs3[0] == outs[0]).all()
            assert lib.tc_decoder_set_max_slice_workers(d, 1 + rep) == 0
        # batch forms: damaged frame delivered concealed, a bad view rejects its frame alone
        pkts = [s[0] for s in streams]
        pkts[0] = damage_slice(pkts[0], 1)
        n = len(pkts)
        bp = (abi.BatchPacket * n)()
        keep = [np.frombuffer(p, np.uint8) for p in pkts]
        for i, b in enumerate(keep):
            bp[i].data, bp[i].size = b.ctypes.data, len(pkts[i])
        infos = (abi.FrameOutput * n)()
        assert lib.tc_decoder_decode_batch(d, bp, n, None, infos) == 0      # query
        outs = [[np.full((h, w), SENTINEL, np.uint16) for (w, h) in s[1]] for s in streams]
        views = (abi.PlaneView * (4 * n))()
        for i, o in enumerate(outs):
            v = abi.plane_views(o)
            for p in range(4):
                views[4 * i + p] = v[p]
        views[4 * 2 + 1].abi_version = 1
        rc = lib.tc_decoder_decode_batch(d, bp, n, views, infos)
        assert rc == abi.TC_WARN_CONCEALED
        for i in (0, 1, 3):
            check_frame(oracle, pkts[i], streams[i][1], outs[i], infos[i])
        assert (outs[2][0] == SENTINEL).all()
        for o in outs:
            for p in o:
                p[:] = SENTINEL
        views[4 * 2 + 1].abi_version = 2
        rc = lib.tc_decoder_decode_batch_request(d, bp, n, ctypes.byref(full), views, infos)
        assert rc == abi.TC_WARN_CONCEALED
        for i in range(n):
            check_frame(oracle, pkts[i], streams[i][1], outs[i], infos[i])
    finally:
        lib.tc_decoder_destroy(d)
