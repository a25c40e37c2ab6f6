"""Helpers shared by the tests: oracle / reference calls through ctypes."""
import ctypes

import numpy as np


def orc_encode(oracle, cfg, planes, cap=None):
    ptrs = (ctypes.c_void_p * 4)(*([p.ctypes.data for p in planes] + [None] * (4 - len(planes))))
    n = ctypes.c_size_t(0)
    if cap is None:
        cap = sum(p.nbytes for p in planes) * 4 + 65536
    out = np.empty(cap, np.uint8)
    rc = oracle.orc_frame_encode(ctypes.byref(cfg), ptrs, None, ctypes.c_void_p(out.ctypes.data),
                                 ctypes.c_size_t(cap), ctypes.byref(n))
    return rc, out[: n.value].tobytes() if rc == 0 else b""


def orc_decode(oracle, packet, dims):
    outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    st = (ctypes.c_uint8 * 512)()
    nc = ctypes.c_uint32(0)
    rc = oracle.orc_frame_decode(packet, ctypes.c_size_t(len(packet)), ptrs, None, st, ctypes.byref(nc))
    return rc, outs, list(st), nc.value


def ref_encode(ref, abi, cfg, planes):
    inp = abi.FrameInput()
    inp.struct_size = ctypes.sizeof(inp)
    inp.abi_version = 2
    for i, p in enumerate(planes):
        inp.planes[i] = p.ctypes.data
    bound = ref.tc_frame_packet_bound(ctypes.byref(cfg))
    out = ctypes.create_string_buffer(max(bound, 1))
    st = abi.FrameStats()
    rc = ref.tc_frame_encode(ctypes.byref(cfg), ctypes.byref(inp), out, ctypes.c_size_t(bound), ctypes.byref(st))
    return rc, out.raw[: st.packet_size] if rc == 0 else b"", st


def ref_decode(ref, abi, packet, dims):
    outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
    info = abi.FrameOutput()
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    rc = ref.tc_frame_decode(packet, ctypes.c_size_t(len(packet)), ptrs, None, ctypes.byref(info))
    return rc, outs, info


def frame_for(pkg, syn, kind, w, h, seed):
    if kind == "422":
        return pkg.synth_frame422(syn, w, h, seed)
    if kind == "4444":
        return pkg.synth_frame4444(syn, w, h, seed, alpha=True)
    if kind == "444":
        return pkg.synth_frame4444(syn, w, h, seed, alpha=False)
    if kind == "raw":
        return pkg.synth_raw12(syn, w, h, seed)
    raise ValueError(kind)


# ---------------------------------------------------------------- .toos files
# A minimal TPIM envelope writer for tests: preamble, two-entry directory,
# IDSC and PIXL, following docs/image/topos_image_file_spec_v0.md sections
# 3, 4 and 6 (the reference writer is native/src/image/image_writer.c).
import struct as _struct
import zlib as _zlib


def idsc_for_packet(packet, image_profile):
    """IDSC bytes derived from a packet header as the reference derives them
    (native/src/image/image_reader.c:736-815)."""
    h = packet
    pf, bd, profile = h[11], h[12], h[10]
    am, abd, flags = h[13], h[14], int.from_bytes(h[8:10], "big")
    vw, vh = int.from_bytes(h[23:25], "big"), int.from_bytes(h[25:27], "big")
    sar_num, sar_den = int.from_bytes(h[37:39], "big"), int.from_bytes(h[39:41], "big")
    sar_num, sar_den = sar_num or 1, sar_den or 1
    model = 0 if pf == 2 else (3 if pf == 3 else 1)
    container = bd if bd >= 12 else 10
    presence = 0 if am == 0 else (2 if flags & 1 else 1)
    sub = 1 if pf == 0 else (2 if pf == 3 else 0)
    b = bytearray(128)
    b[0:4] = b"IDSC"
    _struct.pack_into(">HHII", b, 4, 1, 0, 128, 1)  # version 1.0, size, orientation normalised
    _struct.pack_into(">iiiiiiii", b, 16, 0, 0, vw - 1, vh - 1, 0, 0, vw - 1, vh - 1)
    _struct.pack_into(">III", b, 48, 1, sar_num, sar_den)
    fields = [model, h[27], 0, bd, container, 1, sub, h[36], presence, am, abd, 0, 0,
              h[6], h[7], image_profile, profile, pf, h[33], h[34], h[35], h[32], 0, 0]
    b[60:84] = bytes(fields)
    b[84:88] = h[49:53]
    return bytes(b)


def toos_wrap(packet, image_profile, idsc=None):
    """.toos bytes: preamble (64) | directory (2 x 32) | IDSC (128) | PIXL."""
    idsc = idsc if idsc is not None else idsc_for_packet(packet, image_profile)
    dir_off = 64
    idsc_off = dir_off + 2 * 32
    pixl_off = idsc_off + len(idsc)
    total = pixl_off + len(packet)
    d = bytearray()
    for typ, off, data in ((b"IDSC", idsc_off, idsc), (b"PIXL", pixl_off, packet)):
        d += typ + _struct.pack(">IQQII", 1, off, len(data), _zlib.crc32(data), 0)
    p = bytearray(64)
    p[0:4] = b"TPIM"
    _struct.pack_into(">HHIIQQIIIIII", p, 4, 1, 0, 64, 0, total, dir_off, 32, 2, 0, 0, 0, _zlib.crc32(bytes(d)))
    _struct.pack_into(">I", p, 48, _zlib.crc32(bytes(p)))
    return bytes(p) + bytes(d) + idsc + bytes(packet)


def toos_layout(data):
    """(directory offset, count, [(type, flags, offset, size)])"""
    dir_off, = _struct.unpack(">Q", data[24:32])
    count, = _struct.unpack(">I", data[36:40])
    if count > 64 or dir_off + 32 * count > len(data):
        return dir_off, 0, []
    ents = []
    for i in range(count):
        e = data[dir_off + 32 * i: dir_off + 32 * (i + 1)]
        typ, fl, off, size, _crc, _r = _struct.unpack(">IIQQII", e)
        ents.append((typ, fl, off, size))
    return dir_off, count, ents


def toos_repair(data, fix_packet=True):
    """Recompute, in dependency order, the packet header CRC and the IDSC
    binding digest, every chunk CRC, the directory CRC and the preamble CRC."""
    b = bytearray(data)
    dir_off, count, ents = toos_layout(bytes(b))
    pix = [e for e in ents if e[0] == 0x5049584C]
    ids = [e for e in ents if e[0] == 0x49445343]
    if fix_packet and pix and ids and pix[0][3] >= 53 and ids[0][3] == 128:
        po, io_ = pix[0][2], ids[0][2]
        if po + 53 <= len(b) and io_ + 88 <= len(b):
            b[po + 49:po + 53] = _zlib.crc32(bytes(b[po:po + 49])).to_bytes(4, "big")
            b[io_ + 84:io_ + 88] = b[po + 49:po + 53]
    for i, (typ, fl, off, size) in enumerate(ents):
        if off + size <= len(b):
            e = dir_off + 32 * i
            b[e + 24:e + 28] = _zlib.crc32(bytes(b[off:off + size])).to_bytes(4, "big")
    if dir_off + 32 * count <= len(b):
        b[52:56] = _zlib.crc32(bytes(b[dir_off:dir_off + 32 * count])).to_bytes(4, "big")
    b[48:52] = b"\0\0\0\0"
    b[48:52] = _zlib.crc32(bytes(b[:64])).to_bytes(4, "big")
    return bytes(b)


def cvt_call(fn, abi, src, width, height, fmt, out_mode, *, bit_depth=10, kr=0.2126, kb=0.0722, full_range=0,
             alpha_shift=0, alpha_opaque=0, with_alpha=False, src_stride=0, out_pad=0):
    """Calls a tc_convert_packed_rgb-shaped function (oracle, compiled reference
    or the product) with host buffers; out_pad extra elements per output row
    exercise the output strides. Returns (rc, planes cropped to their widths)."""
    shapes = abi.cvt_plane_shapes(width, height, out_mode, with_alpha)
    full = [np.full((h, w + out_pad), 0xBEEF, np.uint16) for (h, w) in shapes]
    p = abi.CvtParams()
    p.struct_size = ctypes.sizeof(p)
    p.format, p.width, p.height, p.out_mode, p.bit_depth = fmt, width, height, out_mode, bit_depth
    p.kr, p.kb, p.full_range, p.alpha_shift, p.alpha_opaque = kr, kb, full_range, alpha_shift, alpha_opaque
    p.src, p.src_stride = src.ctypes.data, src_stride
    p.y_out, p.y_stride = full[0].ctypes.data, (width + out_pad) if out_pad else 0
    p.u_out, p.u_stride = full[1].ctypes.data, (shapes[1][1] + out_pad) if out_pad else 0
    p.v_out, p.v_stride = full[2].ctypes.data, (shapes[2][1] + out_pad) if out_pad else 0
    if with_alpha:
        p.a_out, p.a_stride = full[3].ctypes.data, (width + out_pad) if out_pad else 0
    rc = fn(ctypes.byref(p))
    return rc, [f[:, : s[1]] for f, s in zip(full, shapes)], full


def p422_call(fn, abi, u, v, width, height, *, bit_depth=10, in_stride=0, out_pad=0):
    cw = (width + 1) // 2
    uo = np.full((height, cw + out_pad), 0xBEEF, np.uint16)
    vo = np.full((height, cw + out_pad), 0xBEEF, np.uint16)
    p = abi.Planar422Params()
    p.struct_size = ctypes.sizeof(p)
    p.width, p.height, p.bit_depth = width, height, bit_depth
    p.u_in, p.u_in_stride, p.v_in, p.v_in_stride = u.ctypes.data, in_stride, v.ctypes.data, in_stride
    p.u_out, p.v_out = uo.ctypes.data, vo.ctypes.data
    p.u_out_stride = p.v_out_stride = (cw + out_pad) if out_pad else 0
    rc = fn(ctypes.byref(p))
    return rc, uo[:, :cw], vo[:, :cw]


def cvt_source(rng, fmt, abi, width, height, row_pad=0, kind="random"):
    """Packed source rows (uint8 view of the bytes) with optional row padding."""
    nch, smp = abi.cvt_channels(fmt), abi.cvt_sample_bytes(fmt)
    row = width * nch * smp
    buf = np.zeros((height, row + row_pad), np.uint8)
    if kind == "random":
        buf[:, :row] = rng.integers(0, 256, size=(height, row), dtype=np.uint8)
    elif kind == "max":
        buf[:, :row] = 255
    return buf, row + row_pad if row_pad else 0


# ---------------------------------------------------------------- sized encode
def orc_sized(oracle, cfg, planes, target, qp_min, qp_max, cap=None):
    """Oracle target-size encode: (rc, qp_used, packet, size, encodes counted)."""
    ptrs = (ctypes.c_void_p * 4)(*([p.ctypes.data for p in planes] + [None] * (4 - len(planes))))
    if cap is None:
        cap = sum(p.nbytes for p in planes) * 4 + 65536
    out = np.empty(max(cap, 1), np.uint8)
    q, n, cnt = ctypes.c_uint8(0), ctypes.c_size_t(0), ctypes.c_uint32(0)
    rc = oracle.orc_frame_encode_sized(ctypes.byref(cfg), ptrs, None, ctypes.c_uint32(target), ctypes.c_uint8(qp_min),
                                       ctypes.c_uint8(qp_max), ctypes.byref(q), ctypes.c_void_p(out.ctypes.data),
                                       ctypes.c_size_t(cap), ctypes.byref(n), ctypes.byref(cnt))
    return rc, q.value, out[: n.value].tobytes() if rc == 0 else b"", n.value, cnt.value


def lib_sized(lib, abi, cfg, planes, target, qp_min, qp_max, cap=None):
    """tc_frame_encode_sized of a library (the compiled reference or the
    product): (rc, qp_used, packet, stats.packet_size)."""
    inp = abi.FrameInput()
    inp.struct_size = ctypes.sizeof(inp)
    inp.abi_version = 2
    for i, p in enumerate(planes):
        inp.planes[i] = p.ctypes.data
    if cap is None:
        cap = lib.tc_frame_packet_bound(ctypes.byref(cfg))
    out = ctypes.create_string_buffer(max(cap, 1))
    st = abi.FrameStats()
    q = ctypes.c_uint8(0)
    rc = lib.tc_frame_encode_sized(ctypes.byref(cfg), ctypes.byref(inp), ctypes.c_uint32(target),
                                   ctypes.c_uint8(qp_min), ctypes.c_uint8(qp_max), ctypes.byref(q),
                                   out if cap else None, ctypes.c_size_t(cap), ctypes.byref(st))
    return rc, q.value, out.raw[: st.packet_size] if rc == 0 else b"", st.packet_size


# ---------------------------------------------------------------- preview decode
def orc_scaled(oracle, packet, tw, th, limit, dims, skip_alpha=0):
    """Oracle target-size decode into planes of dims [(w, h)]: (rc, planes, status, concealed)."""
    outs = [np.zeros((h, w), np.uint16) for (w, h) in dims]
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    st = (ctypes.c_uint8 * 512)()
    nc = ctypes.c_uint32(0)
    rc = oracle.orc_frame_decode_scaled(packet, ctypes.c_size_t(len(packet)), ctypes.c_uint32(tw),
                                        ctypes.c_uint32(th), ctypes.c_uint32(limit), ctypes.c_int32(skip_alpha),
                                        ptrs, None, st, ctypes.byref(nc))
    return rc, outs, list(st), nc.value


def lib_decode_request(lib, abi, packet, request):
    """tc_frame_decode_request of a library: query, then decode into tight
    planes of the reported geometry. (rc, planes, info)."""
    info = abi.FrameOutput()
    rc = lib.tc_frame_decode_request(packet, ctypes.c_size_t(len(packet)), ctypes.byref(request), None, None,
                                     ctypes.byref(info))
    if rc < 0:
        return rc, [], info
    dims = abi.target_plane_dims(info.visible_width, info.visible_height, info.pixel_format, info.plane_count)
    outs = [np.full((h, w), 0xABCD, np.uint16) for (w, h) in dims]
    ptrs = (ctypes.c_void_p * 4)(*([o.ctypes.data for o in outs] + [None] * (4 - len(outs))))
    rc = lib.tc_frame_decode_request(packet, ctypes.c_size_t(len(packet)), ctypes.byref(request), ptrs, None,
                                     ctypes.byref(info))
    return rc, outs, info


# ---------------------------------------------------------------- robustness corpus
import os as _os

_GOLDEN = _os.path.join(_os.path.dirname(_os.path.abspath(__file__)), "golden")


def golden_bitstream_records():
    """The 56 records of golden_bitstream_v1.bin (reference
    native/tests/conformance/golden_bitstream.c:209-230: 8 frame headers, 30
    Rice streams, 12 block payloads, 6 complete packets) and the stored fold."""
    data = open(_os.path.join(_GOLDEN, "golden_bitstream_v1.bin"), "rb").read()
    assert data[:4] == b"TPB1" and int.from_bytes(data[4:8], "big") == 1
    count = int.from_bytes(data[8:12], "big")
    fold = int.from_bytes(data[16:24], "big")
    off, recs = 24, []
    for _ in range(count):
        n = int.from_bytes(data[off:off + 4], "big")
        recs.append(data[off + 4:off + 4 + n])
        off += 4 + n
    assert off == len(data)
    return recs, fold


def splitmix_fold(records):
    """The record fold of golden_bitstream.c:31-49 (splitmix finaliser per byte)."""
    M = (1 << 64) - 1

    def mix(h, v):
        h = (h + 0x9E3779B97F4A7C15) & M
        z = v
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return h ^ (z ^ (z >> 31))

    g = 0x5A5A5A5A5A5A5A5A
    for r in records:
        g = mix(g, len(r))
        for b in r:
            g = mix(g, b)
    return g


def corpus_packets():
    """(name, bytes) of the reference's bitstream robustness corpus
    (native/tests/fuzz/seed_corpus_bitstream: valid seeds, bit flips,
    truncations, forged and over-long fields) and the six complete packets of
    golden_bitstream_v1.bin."""
    d = _os.path.join(_GOLDEN, "seed_corpus_bitstream")
    out = [(f, open(_os.path.join(d, f), "rb").read()) for f in sorted(_os.listdir(d))]
    recs, _ = golden_bitstream_records()
    out += [(f"golden_bitstream_packet_{i}", r) for i, r in enumerate(recs[50:])]
    return out
