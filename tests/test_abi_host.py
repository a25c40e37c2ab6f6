"""CPU-only checks of the C-ABI library: it loads without a GPU, exports every
function the public headers declare, and its host-side rules (config
validation, packet bound, structure parse / query decode, error codes) agree
with the compiled reference where that is built."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import orc_encode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("topos_codec.h", "topos_gpu.h", "topos_image.h"):
        path = os.path.join(ROOT, "include", h)
        if not os.path.exists(path):
            continue
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?!typedef\b)(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([a-z_][a-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_all_declared_symbols(pkg):
    lib = pkg.load_library()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_no_gpu_fails_loudly(pkg, abi, syn):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    c = pkg.ToposCodec()
    cfg = abi.make_config(64, 48, 62, profile=3)
    st, pkt, _ = c.encode_status(cfg, pkg.synth_frame422(syn, 64, 48, 1))
    assert st == abi.TC_ERR_NOT_IMPLEMENTED and pkt == b""
    assert "GPU" in c.last_error()


def _cfgs(abi):
    base = dict(width=1920, height=1080, qp=62, profile=3, entropy_mode=8)
    variants = []
    def mk(**kw):
        d = dict(base); d.update(kw)
        w, h, qp = d.pop("width"), d.pop("height"), d.pop("qp")
        return abi.make_config(w, h, qp, **d)
    variants.append(mk())
    variants.append(mk(features=3))
    variants.append(mk(features=3, bit_depth=0))          # features need explicit 10-bit
    variants.append(mk(features=4))
    variants.append(mk(profile=0, features=1))
    variants.append(mk(qmatrix_id=4))
    variants.append(mk(qmatrix_id=4, bit_depth=12))
    variants.append(mk(qmatrix_id=5))
    variants.append(mk(qmatrix_id=2, pixel_format=2, bit_depth=12, profile=6))
    variants.append(mk(qmatrix_id=2))
    variants.append(mk(qp=96))
    variants.append(mk(qp=70))
    variants.append(mk(width=0))
    variants.append(mk(width=16385))
    variants.append(mk(slice_rows=1, height=8000))
    variants.append(mk(entropy_mode=3))
    variants.append(mk(entropy_mode=12))
    variants.append(mk(entropy_mode=0))
    variants.append(mk(entropy_mode=1, bit_depth=16))
    variants.append(mk(alpha_mode=1, alpha_bit_depth=16))
    variants.append(mk(alpha_mode=1, alpha_bit_depth=8))
    variants.append(mk(alpha_mode=2, alpha_bit_depth=9))
    variants.append(mk(alpha_mode=3))
    variants.append(mk(pixel_format=3, profile=7, bit_depth=12, qmatrix_id=0))
    variants.append(mk(pixel_format=3, profile=7, bit_depth=12, alpha_mode=1, alpha_bit_depth=16))
    variants.append(mk(pixel_format=3, profile=0, bit_depth=12, qp_delta_chroma=2))
    variants.append(mk(pixel_format=3, profile=3, bit_depth=12))
    variants.append(mk(pixel_format=4))
    variants.append(mk(profile=6, pixel_format=2, bit_depth=10))
    variants.append(mk(profile=5, pixel_format=0))
    variants.append(mk(profile=2))
    variants.append(mk(bit_depth=14))
    variants.append(mk(color_matrix=7))
    variants.append(mk(pixel_format=2, color_matrix=1))
    variants.append(mk(color_primaries=2))
    variants.append(mk(color_transfer=20))
    variants.append(mk(chroma_siting=3))
    variants.append(mk(color_range=2))
    variants.append(mk(sar_num=1))
    variants.append(mk(struct_size=12))
    variants.append(mk(abi_version=7))
    c = mk(); c.reserved[1] = 2; variants.append(c)
    c = mk(); c.reserved[6] = 1; variants.append(c)
    c = mk(entropy_mode=10); c.reserved[3] = 2; variants.append(c)
    return variants


def test_config_validate_matches_reference(pkg, abi, ref):
    lib = pkg.load_library()
    for i, cfg in enumerate(_cfgs(abi)):
        a = lib.tc_frame_config_validate(ctypes.byref(cfg))
        b = ref.tc_frame_config_validate(ctypes.byref(cfg))
        assert a == b, (i, abi.config_fields(cfg), a, b)
        assert lib.tc_frame_packet_bound(ctypes.byref(cfg)) == ref.tc_frame_packet_bound(ctypes.byref(cfg))


def _corruptions(pkt):
    out = [("ok", pkt)]
    b = bytearray(pkt); b[0] ^= 1; out.append(("magic", bytes(b)))
    b = bytearray(pkt); b[20] ^= 1; out.append(("hdr-crc", bytes(b)))
    out.append(("short", pkt[:40]))
    out.append(("trunc", pkt[:-5]))
    out.append(("extra", pkt + b"\0"))
    b = bytearray(pkt); b[53 + 4] = 7; out.append(("slice-plane", bytes(b)))
    b = bytearray(pkt); b[53 + 10] = 3; out.append(("slice-k", bytes(b)))
    b = bytearray(pkt); b[53 + 10] = 15; out.append(("slice-k15", bytes(b)))
    b = bytearray(pkt); b[53 + 6] ^= 1; out.append(("slice-y0", bytes(b)))
    b = bytearray(pkt); b[53 + 9] = 0; out.append(("slice-qp", bytes(b)))
    b = bytearray(pkt); b[60] ^= 0x40; out.append(("payload", bytes(b)))
    return out


def _refix_header_crc(b):
    import zlib
    b = bytearray(b)
    c = zlib.crc32(bytes(b[:49])) & 0xFFFFFFFF
    b[49:53] = c.to_bytes(4, "big")
    return bytes(b)


def test_query_decode_matches_reference(pkg, abi, syn, oracle, ref):
    lib = pkg.load_library()
    planes = pkg.synth_frame422(syn, 96, 64, 5)
    cfg = abi.make_config(96, 64, 40, profile=3, slice_rows=2, entropy_mode=8, features=1)
    rc, pkt = orc_encode(oracle, cfg, planes)
    assert rc == 0
    cases = _corruptions(pkt)
    # header field edits with a recomputed header CRC
    for off, val in [(6, 3), (6, 10), (7, 1), (10, 4), (11, 5), (12, 11), (13, 3), (27, 5), (28, 9),
                     (29, 99), (30, 0), (32, 2), (33, 4), (34, 3), (35, 2), (36, 3), (45, 6), (46, 1), (48, 1),
                     (8, 0x10), (15, 1)]:
        b = bytearray(pkt); b[off] = val
        cases.append((f"field{off}={val}", _refix_header_crc(b)))
    for name, p in cases:
        ia, ib = abi.FrameOutput(), abi.FrameOutput()
        a = lib.tc_frame_decode(p, ctypes.c_size_t(len(p)), None, None, ctypes.byref(ia))
        b = ref.tc_frame_decode(p, ctypes.c_size_t(len(p)), None, None, ctypes.byref(ib))
        assert a == b, (name, a, b, lib.tc_last_error())
        assert bytes(ia) == bytes(ib), name


def test_plane_geometry_matches_reference(pkg, abi, ref):
    lib = pkg.load_library()
    for pf, am, w, h in [(0, 0, 73, 39), (1, 0, 35, 21), (2, 1, 64, 48), (3, 0, 101, 77)]:
        info = abi.FrameOutput()
        info.visible_width, info.visible_height, info.pixel_format = w, h, pf
        info.plane_count = 4 if (pf == 3 or am) else 3
        for p in range(5):
            wa, ha, wb, hb = (ctypes.c_uint32() for _ in range(4))
            ra = lib.tc_frame_plane_geometry(ctypes.byref(info), p, ctypes.byref(wa), ctypes.byref(ha))
            rb = ref.tc_frame_plane_geometry(ctypes.byref(info), p, ctypes.byref(wb), ctypes.byref(hb))
            assert (ra, wa.value, ha.value) == (rb, wb.value, hb.value)
