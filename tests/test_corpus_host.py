"""CPU checks over the reference's robustness corpus (256 files of
native/tests/fuzz/seed_corpus_bitstream and the records of
golden_bitstream_v1.bin): the fixture files are intact (stored fold), this
library's structure parse gives the reference's status for every file, and
the oracle's full decode (status, slice status, planes) equals the compiled
reference's for every file that parses."""
import ctypes

import pytest

from helpers import corpus_packets, golden_bitstream_records, orc_decode, ref_decode, splitmix_fold


def test_golden_bitstream_fixture_intact():
    recs, fold = golden_bitstream_records()
    assert len(recs) == 56
    assert splitmix_fold(recs) == fold
    for h in recs[:8]:                    # frame headers: magic, CRC over bytes 0..48
        import zlib
        assert len(h) == 53 and zlib.crc32(h[:49]) == int.from_bytes(h[49:53], "big")


def test_corpus_query_status_matches_reference(pkg, abi, ref):
    c = pkg.ToposCodec()
    kinds = {}
    for name, pkt in corpus_packets():
        st, info = c.query(pkt)
        r_info = abi.FrameOutput()
        r_st = ref.tc_frame_decode(pkt, ctypes.c_size_t(len(pkt)), None, None, ctypes.byref(r_info))
        assert st == r_st, name
        if st == 0:
            for f in ("visible_width", "visible_height", "plane_count", "bit_depth", "profile", "pixel_format",
                      "alpha_mode", "slice_count"):
                assert getattr(info, f) == getattr(r_info, f), (name, f)
        kinds[st] = kinds.get(st, 0) + 1
    assert kinds.get(0, 0) >= 80 and len(kinds) >= 3   # decodable, and several rejection codes


def test_corpus_oracle_decode_matches_reference(pkg, abi, oracle, ref):
    c = pkg.ToposCodec()
    decoded = concealed = 0
    for name, pkt in corpus_packets():
        st, info = c.query(pkt)
        if st != 0:
            continue
        dims = abi.plane_dims(info.visible_width, info.visible_height, info.pixel_format, info.alpha_mode)
        r_rc, r_planes, r_info = ref_decode(ref, abi, pkt, dims)
        o_rc, o_planes, o_status, o_nc = orc_decode(oracle, pkt, dims)
        assert o_rc == r_rc, name
        assert o_status[:info.slice_count] == list(r_info.slice_status)[:info.slice_count], name
        assert o_nc == r_info.concealed_slices, name
        assert all((a == b).all() for a, b in zip(o_planes, r_planes)), name
        decoded += 1
        concealed += o_rc == abi.TC_WARN_CONCEALED
    assert decoded >= 80 and concealed >= 60
