"""CPU checks of the drop-in boundary: the exported symbol table, symbol
binding when another Topos library is loaded globally, version / capability
negotiation, and the argument rules of the batch and decoder-context calls
(every case here is decided before any GPU work, so the return codes compare
directly with the compiled reference's)."""
import ctypes
import os
import re
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from helpers import orc_encode

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TREE = "/root/reference/native"
LIB = os.path.join(ROOT, "topos-codec_amd", "libtopos_codec_amd.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libtopos_ref.so")

# pinned reference exports (native/tests/abi/public_symbols_v1.txt) this
# library does not provide: scalable (V7-B) encode and base-layer reads, the
# .toos writer / metadata / HALF helpers, and MOV multi-track audio; none is
# on the intra frame-coding path
NOT_PROVIDED = {
    "tc_frame_encode_scalable", "tc_movie_packet_base", "tc_movie_packet_base_batch",
    "tc_image_write", "tc_image_read_metadata", "tc_image_query_capabilities", "tc_image_derive_idsc",
    "tc_image_half_to_codes", "tc_image_codes_to_half",
    "tc_mux_add_audio_track", "tc_mux_set_audio_track_asc", "tc_mux_set_audio_track_priming", "tc_mux_add_audio_to",
}


def dynamic_exports(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return sorted(line.split()[-1] for line in out.splitlines() if line.strip())


def header_functions(paths):
    names = set()
    for path in paths:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        names.update(re.findall(r"(?<!\(\*)\b(tcx?_[A-Za-z0-9_]+)\s*\(", text))
    return names


def test_exports_are_exactly_the_declared_functions(pkg):
    pkg.load_library()
    inc = os.path.join(ROOT, "include")
    declared = header_functions([os.path.join(inc, h) for h in os.listdir(inc) if h.endswith(".h")])
    exported = dynamic_exports(LIB)
    assert sorted(declared) == exported
    # no internal helper leaks (engine, arena, format rules, kernels' host side)
    assert not [n for n in exported if n.startswith(("tcg_", "tc_engine", "tc_arena", "_Z"))]


@pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference tree not present")
def test_exports_follow_the_reference_interface(pkg):
    exported = set(dynamic_exports(LIB))
    ref_decl = header_functions([os.path.join(REF_TREE, "include", h)
                                 for h in ("topos_codec.h", "topos_image.h")])
    public = {n for n in exported if not n.startswith("tcx_")}
    assert public - ref_decl == set(), "exports without a reference declaration"
    pinned = set(open(os.path.join(REF_TREE, "tests", "abi", "public_symbols_v1.txt")).read().split())
    assert pinned - exported == NOT_PROVIDED


_BIND_SCRIPT = textwrap.dedent("""
    import ctypes, sys
    ref = ctypes.CDLL(sys.argv[1], mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(sys.argv[2], mode=ctypes.RTLD_LOCAL)
    pkt = open(sys.argv[3], "rb").read()
    class PV(ctypes.Structure):
        _fields_ = [("s", ctypes.c_uint32), ("a", ctypes.c_uint32), ("p", ctypes.c_void_p), ("st", ctypes.c_size_t)]
    w, h = int(sys.argv[4]), int(sys.argv[5])
    bufs = [ctypes.create_string_buffer(2 * w * h) for _ in range(3)]
    views = (PV * 4)()
    for i in range(3):
        views[i].s, views[i].a, views[i].p, views[i].st = ctypes.sizeof(PV), 2, ctypes.addressof(bufs[i]), w
    info = ctypes.create_string_buffer(4096)
    dec = ctypes.c_void_p()
    assert lib.tc_decoder_create(None, ctypes.byref(dec)) == 0
    rc = lib.tc_decoder_decode(dec, pkt, ctypes.c_size_t(len(pkt)), views, info)
    lib.tc_decoder_destroy(dec)
    print("RC", rc)
""")


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="compiled reference (oracle/_ref) not built")
def test_internal_calls_never_bind_to_another_topos_library(tmp_path, pkg, abi, syn, oracle):
    """A host that links the reference globally and opens this library
    RTLD_LOCAL: this library's calls between its own exports (the decoder
    context -> tc_frame_decode_request -> engine) must stay inside it. Without
    a GPU this library's decode fails with TC_ERR_NOT_IMPLEMENTED, while a
    call diverted into the reference would decode on the CPU and return OK."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present: the no-GPU status is the probe")
    w, h = 64, 32
    cfg = abi.make_config(w, h, 40, profile=3, entropy_mode=8)
    rc, pkt = orc_encode(oracle, cfg, pkg.synth_frame422(syn, w, h, 3))
    assert rc == 0
    pp = tmp_path / "p.bin"
    pp.write_bytes(pkt)
    sp = tmp_path / "bind.py"
    sp.write_text(_BIND_SCRIPT)
    env = dict(os.environ, LD_DEBUG="bindings", LD_DEBUG_OUTPUT=str(tmp_path / "ld"))
    r = subprocess.run([sys.executable, str(sp), REF_LIB, LIB, str(pp), str(w), str(h)], env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"RC {abi.TC_ERR_NOT_IMPLEMENTED}" in r.stdout
    logs = "".join(open(tmp_path / f).read() for f in os.listdir(tmp_path) if f.startswith("ld"))
    crossed = [l for l in logs.splitlines()
               if "libtopos_codec_amd.so" in l.split(" to ")[0] and "libtopos_ref.so" in l.split(" to ")[-1]]
    assert crossed == []


# ---------------------------------------------------------------- negotiation

def _setup(lib):
    lib.tc_scalable_status_name.restype = ctypes.c_char_p
    lib.tc_scalable_status_name.argtypes = [ctypes.c_int32]
    lib.tc_query_support.argtypes = [ctypes.c_uint32] * 4
    return lib


def test_version_and_capabilities(pkg, abi):
    lib = _setup(pkg.load_library())
    v = abi.VersionInfo()
    assert lib.tc_version(ctypes.byref(v)) == 0
    assert (v.struct_size, v.abi_version) == (ctypes.sizeof(v), 2)
    assert (v.version_major, v.version_minor, v.version_patch) == (0, 1, 0)
    assert v.git_commit and v.build_target
    assert lib.tc_version(None) == abi.TC_ERR_INVALID_ARGUMENT
    c = abi.CodecCapabilities()
    assert lib.tc_query_capabilities(ctypes.byref(c)) == 0
    assert c.struct_size == ctypes.sizeof(c) and c.max_width == 16384 and c.auto2k_max_dim == 2048
    assert c.flags & 0x1F == 0x1F and c.flags & (1 << 7) and not c.flags & ((1 << 5) | (1 << 6))
    f = abi.CpuFeatures()
    assert lib.tc_query_cpu_features(ctypes.byref(f)) == 0 and f.struct_size == ctypes.sizeof(f)
    assert lib.tc_query_capabilities(None) == lib.tc_query_cpu_features(None) == abi.TC_ERR_INVALID_ARGUMENT
    s = ctypes.c_int32(-1)
    assert lib.tc_query_scalable_status(ctypes.byref(s)) == 0 and s.value == 0
    assert [lib.tc_scalable_status_name(i) for i in (0, 1, 2, 3, 4, -1)] == [
        b"experimental", b"reference", b"eligible", b"default", b"unknown", b"unknown"]


def test_negotiation_matches_reference(pkg, abi, ref):
    lib = _setup(pkg.load_library())
    _setup(ref)
    for fn in ("tc_version", "tc_query_capabilities", "tc_query_cpu_features"):
        cls = {"tc_version": abi.VersionInfo, "tc_query_capabilities": abi.CodecCapabilities,
               "tc_query_cpu_features": abi.CpuFeatures}[fn]
        a, b = cls(), cls()
        assert getattr(lib, fn)(ctypes.byref(a)) == getattr(ref, fn)(ctypes.byref(b)) == 0
        assert (a.struct_size, a.abi_version) == (b.struct_size, b.abi_version)
        if fn == "tc_version":
            assert (a.version_major, a.version_minor, a.version_patch) == (b.version_major, b.version_minor,
                                                                            b.version_patch)
        if fn == "tc_query_cpu_features":
            assert a.flags == b.flags
        if fn == "tc_query_capabilities":
            assert (a.flags, a.max_width, a.max_height, a.auto2k_max_dim) == (
                b.flags, b.max_width, b.max_height, b.auto2k_max_dim)
    for profile in (0, 1, 2, 3, 4, 5, 6, 7, 8, 255):
        for pf in (0, 1, 2, 3, 4):
            for bd in (8, 10, 12, 14, 16):
                for am in (0, 1, 2, 3):
                    assert lib.tc_query_support(profile, pf, bd, am) == ref.tc_query_support(profile, pf, bd, am), \
                        (profile, pf, bd, am)


# ---------------------------------------------------------------- argument rules

@pytest.fixture(scope="module")
def packets(pkg, abi, syn, oracle):
    w, h = 64, 32
    cfg = abi.make_config(w, h, 40, profile=3, entropy_mode=8, slice_rows=2)
    rc, good = orc_encode(oracle, cfg, pkg.synth_frame422(syn, w, h, 5))
    assert rc == 0
    bad_hdr = bytearray(good)
    bad_hdr[20] ^= 0x40               # frame header CRC no longer matches
    short = good[:40]                 # shorter than a frame header
    return w, h, good, bytes(bad_hdr), short


class _Pkt(ctypes.Structure):
    _fields_ = [("data", ctypes.c_char_p), ("size", ctypes.c_size_t)]


def _batch(pkts):
    bp = (_Pkt * max(1, len(pkts)))()
    for i, p in enumerate(pkts):
        bp[i].data, bp[i].size = p, len(p)
    return bp


def test_frame_decode_batch_rules_match_reference(pkg, abi, ref, packets):
    w, h, good, bad_hdr, short = packets
    ours = pkg.load_library()
    for lib in (ours, ref):
        lib.tc_frame_decode_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]
    def run(lib, pkts, count=None, infos=True, planes=False, packets_null=False):
        bp = _batch(pkts)
        inf = (abi.FrameOutput * max(1, len(pkts)))()
        outs = [np.zeros((h, w), np.uint16) for _ in range(3 * len(pkts))]
        ptrs = None
        if planes:
            ptrs = (ctypes.c_void_p * (4 * len(pkts)))()
            for i in range(len(pkts)):
                for p in range(3):
                    ptrs[4 * i + p] = outs[3 * i + p].ctypes.data
        rc = lib.tc_frame_decode_batch(None if packets_null else bp, len(pkts) if count is None else count,
                                       ptrs, None, inf if infos else None)
        return rc, inf
    cases = [
        dict(pkts=[good], count=0),
        dict(pkts=[good], count=70000, packets_null=True),
        dict(pkts=[good], infos=False),
        dict(pkts=[good], packets_null=True),
        dict(pkts=[good, bad_hdr]),                  # query mode: structure failure rejects the batch
        dict(pkts=[good, short], planes=True),       # decode mode: same, before any frame
        dict(pkts=[bad_hdr, good], planes=True),
    ]
    for c in cases:
        ra, _ = run(ours, **c)
        rb, _ = run(ref, **c)
        assert ra == rb, c
        assert ra != 0 or c.get("count") == 0
    # query mode: every frame's info equals the reference's
    ra, ia = run(ours, [good, good])
    rb, ib = run(ref, [good, good])
    assert ra == rb == 0
    assert bytes(ia) == bytes(ib)


def test_decoder_context_rules_match_reference(pkg, abi, ref, packets):
    w, h, good, bad_hdr, short = packets
    vp = ctypes.c_void_p
    ours = pkg.load_library()
    for lib in (ours, ref):
        lib.tc_decoder_create.argtypes = [vp, vp]
        lib.tc_decoder_destroy.argtypes = [vp]
        lib.tc_decoder_destroy.restype = None
        lib.tc_decoder_set_max_slice_workers.argtypes = [vp, ctypes.c_uint32]
        lib.tc_decoder_prepare.argtypes = [vp, vp, ctypes.c_size_t, vp]
        lib.tc_decoder_decode.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
        lib.tc_decoder_decode_request.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp]
        lib.tc_decoder_decode_batch.argtypes = [vp, vp, ctypes.c_uint32, vp, vp]
        lib.tc_decoder_decode_surface.argtypes = [vp, vp, ctypes.c_size_t, vp, vp]
        lib.tc_decoder_decode_surface_request.argtypes = [vp, vp, ctypes.c_size_t, vp, vp, vp]
    outs = [np.zeros((h, w), np.uint16) for _ in range(3)]
    good_views = abi.plane_views(outs)
    bad_size = abi.plane_views(outs)
    bad_size[1].struct_size = 0
    no_pixels = abi.plane_views(outs)
    no_pixels[2].pixels = None

    def surface(views, mem=0, size=None):
        s = abi.DecodeSurface()
        s.struct_size = ctypes.sizeof(s) if size is None else size
        s.abi_version = 2
        s.memory_type = mem
        for i in range(4):
            s.planes[i] = views[i]
        return s

    def results(lib):
        out = []
        d = ctypes.c_void_p()
        cfg = abi.DecoderConfig()
        out.append(lib.tc_decoder_create(None, None))
        cfg.struct_size, cfg.abi_version = 0, 2
        out.append(lib.tc_decoder_create(ctypes.byref(cfg), ctypes.byref(d)))
        out.append(d.value is None)
        cfg.struct_size, cfg.abi_version = ctypes.sizeof(cfg), 1
        out.append(lib.tc_decoder_create(ctypes.byref(cfg), ctypes.byref(d)))
        cfg.abi_version = 2
        cfg.max_slice_workers = 3
        out.append(lib.tc_decoder_create(ctypes.byref(cfg), ctypes.byref(d)))
        out.append(lib.tc_decoder_set_max_slice_workers(None, 2))
        out.append(lib.tc_decoder_set_max_slice_workers(d, 0))
        info = abi.FrameOutput()
        pk, n = good, len(good)
        out.append(lib.tc_decoder_prepare(None, pk, n, ctypes.byref(info)))
        out.append(lib.tc_decoder_prepare(d, pk, n, None))
        out.append(lib.tc_decoder_prepare(d, None, n, ctypes.byref(info)))
        out.append(lib.tc_decoder_prepare(d, bad_hdr, n, ctypes.byref(info)))
        out.append(lib.tc_decoder_prepare(d, pk, n, ctypes.byref(info)))
        out.append(bytes(info))
        out.append(lib.tc_decoder_decode(None, pk, n, good_views, ctypes.byref(info)))
        out.append(lib.tc_decoder_decode(d, pk, n, good_views, None))
        out.append(lib.tc_decoder_decode(d, None, n, good_views, ctypes.byref(info)))
        out.append(lib.tc_decoder_decode(d, pk, n, None, ctypes.byref(info)))
        out.append(lib.tc_decoder_decode(d, short, len(short), good_views, ctypes.byref(info)))
        out.append(lib.tc_decoder_decode(d, pk, n, bad_size, ctypes.byref(info)))
        out.append(lib.tc_decoder_decode(d, pk, n, no_pixels, ctypes.byref(info)))
        req = abi.make_request(0)
        out.append(lib.tc_decoder_decode_request(d, pk, n, None, good_views, ctypes.byref(info)))
        req.reserved[0] = 1
        out.append(lib.tc_decoder_decode_request(d, pk, n, ctypes.byref(req), good_views, ctypes.byref(info)))
        req.reserved[0] = 0
        out.append(lib.tc_decoder_decode_request(d, pk, n, ctypes.byref(req), bad_size, ctypes.byref(info)))
        s_ok = surface(good_views)
        out.append(lib.tc_decoder_decode_surface(d, pk, n, None, ctypes.byref(info)))
        s_mem = surface(good_views, mem=2)
        out.append(lib.tc_decoder_decode_surface(d, pk, n, ctypes.byref(s_mem), ctypes.byref(info)))
        s_sz = surface(good_views, size=8)
        out.append(lib.tc_decoder_decode_surface(d, pk, n, ctypes.byref(s_sz), ctypes.byref(info)))
        s_bad = surface(bad_size)
        out.append(lib.tc_decoder_decode_surface(d, pk, n, ctypes.byref(s_bad), ctypes.byref(info)))
        out.append(lib.tc_decoder_decode_surface_request(d, pk, n, None, ctypes.byref(s_ok), ctypes.byref(info)))
        req.memory_type = 1
        out.append(lib.tc_decoder_decode_surface_request(d, pk, n, ctypes.byref(req), ctypes.byref(s_ok),
                                                         ctypes.byref(info)))
        req.memory_type = 0
        out.append(lib.tc_decoder_decode_surface_request(d, pk, n, ctypes.byref(req), ctypes.byref(s_bad),
                                                         ctypes.byref(info)))
        bp = _batch([good, bad_hdr])
        infos = (abi.FrameOutput * 2)()
        out.append(lib.tc_decoder_decode_batch(None, bp, 2, None, infos))
        out.append(lib.tc_decoder_decode_batch(d, bp, 0, None, infos))
        out.append(lib.tc_decoder_decode_batch(d, bp, 2, None, infos))
        bp2 = _batch([good, good])
        out.append(lib.tc_decoder_decode_batch(d, bp2, 2, None, infos))
        out.append(bytes(infos))
        lib.tc_decoder_destroy(d)
        lib.tc_decoder_destroy(None)
        return out

    a, b = results(ours), results(ref)
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x == y, (i, x, y)
