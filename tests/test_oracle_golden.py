"""Pin the oracle against the reference's own conformance vectors
(native/tests/conformance/golden_transform_v1.bin and the golden_codec_*.bin
packet sets, copied verbatim into tests/golden/ as data):

 - transform/quant/dequant/reconstruction: 1380 records {depth, qp, px, F, q,
   Fp, rec} (layout native/tests/conformance/golden_util.h:1-38);
 - intra packets and decode fingerprints of every entropy generation the
   path codes: V1 Rice (golden_codec_v1*.bin, 39 cases incl. 12-bit, 4:4:4,
   GBR, profiles 5/6, CFA), V2 canonical VLC (golden_codec_v2_vlc.bin, 8) and
   rANS2 (golden_codec_v7r2_intra.bin, 8, and the two em=8 CFA cases);
   inputs regenerated with the oracle's restatement of the reference test
   image synthesiser (native/tests/conformance/golden_codec.c:163-238).
"""
import ctypes
import os
import struct

import numpy as np
import pytest

from helpers import orc_decode, orc_encode

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
M64 = (1 << 64) - 1


def mix64(h, v):
    h = (h + 0x9E3779B97F4A7C15) & M64
    z = v & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return h ^ (z ^ (z >> 31))


def plane_fold(planes):
    fold = 0xDED0C0DE00C0FFEE
    for p, a in enumerate(planes):
        fold = mix64(fold, p)
        for v in a.reshape(-1).tolist():
            fold = mix64(fold, int(v))
    return fold


def read_tpc1(path):
    d = open(path, "rb").read()
    assert d[:4] == b"TPC1"
    _ver, count = struct.unpack(">II", d[4:12])
    off = 24  # magic, version, count, pad, fold
    recs = []
    for _ in range(count):
        (n,) = struct.unpack(">I", d[off:off + 4])
        recs.append(d[off + 4: off + 4 + n])
        off += 4 + n
    return recs


def test_golden_transform_records(oracle):
    d = open(os.path.join(GOLDEN, "golden_transform_v1.bin"), "rb").read()
    assert d[:4] == b"TPG1"
    rs = 2 + 128 + 256 * 3 + 128
    n = (len(d) - 24) // rs
    assert n == 1380
    flat = (ctypes.c_uint16 * 64)(*([16] * 64))
    oracle.orc_qstep.restype = ctypes.c_uint32
    F = (ctypes.c_int32 * 64)()
    q = (ctypes.c_int32 * 64)()
    xh = (ctypes.c_int32 * 64)()
    x = (ctypes.c_int32 * 64)()
    for i in range(n):
        r = d[24 + i * rs: 24 + (i + 1) * rs]
        depth, qp = r[0], r[1]
        px = np.frombuffer(r[2:130], "<u2")
        gF = np.frombuffer(r[130:386], "<i4")
        gq = np.frombuffer(r[386:642], "<i4")
        gFp = np.frombuffer(r[642:898], "<i4")
        grec = np.frombuffer(r[898:1026], "<u2")
        mid = 1 << (depth - 1)
        for k in range(64):
            x[k] = int(px[k]) - mid
        oracle.orc_fdct(x, depth, F)
        assert list(F) == gF.tolist(), f"record {i} F"
        oracle.orc_quant(F, flat, qp, 0, depth, q)
        assert list(q) == gq.tolist(), f"record {i} q"
        Q = oracle.orc_qstep(0, 16, qp)
        fp = [max(-(1 << 25), min(1 << 25, v * Q)) for v in gq.tolist()]
        assert fp == gFp.tolist(), f"record {i} Fp"
        oracle.orc_dequant_idct(q, flat, qp, 0, depth, xh)
        rec = [max(0, min((1 << depth) - 1, v + mid)) for v in xh]
        assert rec == grec.tolist(), f"record {i} rec"


# Cases in the field order of native/tests/conformance/golden_codec.c:75-90
# (seed, w, h, kind, qmatrix, qp, qp_delta_chroma, slice_rows (0 = 32),
#  alpha_mode, alpha_bit_depth, bit_depth (0 = 10), pixel_format, profile,
#  entropy mode). Synthesiser kinds: 0 flat, 1 gradient, 2 grain, 3 detail,
# 4 mixed.
FLAT, GRADIENT, GRAIN, DETAIL, MIXED = range(5)
GOLDEN_SETS = {
    # golden_codec.c:94-103
    "golden_codec_v1.bin": [
        (0x0102030405060708, 8, 8, FLAT, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0),
        (0x1112131415161718, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 0, 0, 0, 0),
        (0x2122232425262728, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 0, 0, 0, 0),
        (0x3132333435363738, 48, 48, DETAIL, 1, 12, 0, 2, 2, 12, 0, 0, 0, 0),
        (0x4142434445464748, 96, 64, GRAIN, 1, 24, 6, 0, 0, 0, 0, 0, 0, 0),
        (0x5152535455565758, 128, 72, GRADIENT, 0, 8, 0, 8, 0, 0, 0, 0, 0, 0),
        (0x6162636465666768, 16, 16, FLAT, 1, 63, 0, 0, 0, 0, 0, 0, 0, 0),
        (0x7172737475767778, 64, 48, DETAIL, 1, 42, 0, 1, 2, 8, 0, 0, 0, 0),
    ],
    # :107-113
    "golden_codec_v1_bd12.bin": [
        (0x8182838485868788, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 12, 0, 0, 0),
        (0x9192939495969798, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 12, 0, 0, 0),
        (0xA1A2A3A4A5A6A7A8, 48, 48, DETAIL, 1, 12, 0, 2, 2, 12, 12, 0, 0, 0),
        (0xB1B2B3B4B5B6B7B8, 96, 64, GRADIENT, 1, 24, 6, 0, 0, 0, 12, 0, 0, 0),
        (0xC1C2C3C4C5C6C7C8, 16, 16, FLAT, 1, 63, 0, 0, 2, 8, 12, 0, 0, 0),
    ],
    # :117-124
    "golden_codec_v1_pf444.bin": [
        (0xD1D2D3D4D5D6D7D8, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 0, 1, 0, 0),
        (0xE1E2E3E4E5E6E7E8, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 0, 1, 0, 0),
        (0xF1F2F3F4F5F6F7F8, 48, 48, DETAIL, 1, 12, 0, 2, 2, 12, 0, 1, 0, 0),
        (0x0F1E2D3C4B5A6978, 96, 64, GRADIENT, 1, 24, 6, 0, 0, 0, 12, 1, 0, 0),
        (0x1F2E3D4C5B6A7988, 16, 16, FLAT, 1, 63, 0, 0, 2, 8, 0, 1, 0, 0),
        (0x2F3E4D5C6B7A8999, 64, 48, GRAIN, 0, 28, 3, 0, 0, 0, 12, 1, 0, 0),
    ],
    # :128-135
    "golden_codec_v1_gbr.bin": [
        (0x3F4E5D6C7B8A9901, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 0, 2, 0, 0),
        (0x4F5E6D7C8B9A0112, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 0, 2, 0, 0),
        (0x5F6E7D8C9BA02234, 48, 48, DETAIL, 1, 12, 0, 2, 2, 12, 0, 2, 0, 0),
        (0x6F7E8D9CAB103456, 96, 64, GRADIENT, 1, 24, 6, 0, 0, 0, 12, 2, 0, 0),
        (0x7F8E9DACB0215678, 16, 16, FLAT, 1, 63, 0, 0, 2, 8, 0, 2, 0, 0),
        (0x8F9EADBCD0327899, 64, 48, GRAIN, 0, 28, 3, 0, 0, 0, 12, 2, 0, 0),
    ],
    # :139-146
    "golden_codec_v1_profiles.bin": [
        (0x9A8B7C6D5E4F3021, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 0, 1, 5, 0),
        (0xA9B8C7D6E5F41332, 48, 48, DETAIL, 1, 12, 0, 2, 2, 12, 12, 1, 5, 0),
        (0xB8C7D6E5F4A42443, 96, 64, GRADIENT, 1, 24, 6, 0, 0, 0, 12, 2, 5, 0),
        (0xC7D6E5F4A3B53554, 16, 16, FLAT, 1, 63, 0, 0, 0, 0, 12, 1, 6, 0),
        (0xD6E5F4A3B2C64665, 64, 48, GRAIN, 0, 28, 3, 0, 0, 0, 12, 2, 6, 0),
        (0xE5F4A3B2C1D75776, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 12, 2, 6, 0),
    ],
    # :151-160
    "golden_codec_v2_vlc.bin": [
        (0x0A1B2C3D4E5F6071, 64, 48, GRAIN, 1, 20, 0, 0, 0, 0, 0, 0, 0, 1),
        (0x1A2B3C4D5E6F7082, 48, 48, DETAIL, 1, 12, 0, 2, 0, 0, 0, 0, 0, 1),
        (0x2A3B4C5D6E7F8093, 36, 20, MIXED, 0, 33, 0, 3, 0, 0, 12, 0, 0, 1),
        (0x3A4B5C6D7E8F90A4, 96, 64, GRAIN, 1, 24, 6, 0, 0, 0, 12, 1, 0, 1),
        (0x4A5B6C7D8E9FA0B5, 16, 16, FLAT, 1, 63, 0, 0, 0, 0, 0, 1, 0, 1),
        (0x5A6B7C8D9EAFB0C6, 64, 48, GRADIENT, 0, 28, 3, 0, 0, 0, 0, 2, 0, 1),
        (0x6A7B8C9DAFB0C1D7, 36, 20, MIXED, 0, 33, 0, 3, 1, 16, 0, 0, 0, 1),
        (0x7A8B9DAFB0C1D2E8, 64, 48, DETAIL, 1, 12, 0, 2, 2, 12, 12, 1, 0, 1),
    ],
    # :172-181 (em 8 grid of the retired generation 3 cases)
    "golden_codec_v7r2_intra.bin": [
        (0x301, 73, 39, GRADIENT, 1, 24, 0, 2, 0, 0, 10, 0, 0, 8),
        (0x302, 35, 21, DETAIL, 1, 18, 0, 1, 1, 16, 12, 0, 0, 8),
        (0x303, 73, 39, MIXED, 0, 31, 0, 2, 2, 12, 10, 1, 0, 8),
        (0x304, 35, 21, GRADIENT, 1, 24, 0, 1, 0, 0, 12, 1, 0, 8),
        (0x305, 73, 39, DETAIL, 1, 18, 0, 2, 1, 16, 10, 2, 0, 8),
        (0x306, 35, 21, MIXED, 0, 31, 0, 1, 2, 8, 12, 2, 0, 8),
        (0x307, 1, 1, FLAT, 0, 0, 0, 1, 0, 0, 10, 0, 0, 8),
        (0x308, 16, 16, GRAIN, 1, 63, 0, 1, 0, 0, 12, 2, 0, 8),
    ],
    # :187-194 (CFA phase planes, V1 Rice and em 8)
    "golden_codec_v1_traw.bin": [
        (0x7219000000000001, 64, 48, GRAIN, 0, 20, 0, 0, 0, 0, 12, 3, 7, 0),
        (0x7219000000000002, 36, 20, GRADIENT, 0, 33, 0, 3, 0, 0, 12, 3, 7, 0),
        (0x7219000000000003, 48, 48, DETAIL, 0, 12, 0, 0, 0, 0, 12, 3, 7, 0),
        (0x7219000000000004, 64, 48, MIXED, 0, 24, 0, 2, 0, 0, 12, 3, 7, 0),
        (0x7219000000000005, 64, 48, GRAIN, 0, 20, 0, 0, 0, 0, 12, 3, 7, 8),
        (0x7219000000000006, 48, 48, DETAIL, 0, 63, 0, 0, 0, 0, 12, 3, 7, 8),
    ],
}


def golden_inputs(oracle, case):
    seed, w, h, kind, qm, qp, qpdc, sr, am, abd, bd, pf, prof, em = case
    bd = bd or 10
    u16p = ctypes.c_void_p
    if pf == 3:
        pw, ph = w // 2, h // 2
        planes = []
        for i in range(4):
            a = np.zeros((ph, pw), np.uint16)
            oracle.orc_image_synth(ctypes.c_uint64(seed + i * (1 << 32)), pw, ph, kind, bd, 0,
                                   u16p(a.ctypes.data), None, None, None)
            planes.append(a)
        return planes
    cw = w if pf != 0 else (w + 1) // 2
    y = np.zeros((h, w), np.uint16)
    u = np.zeros((h, cw), np.uint16)
    v = np.zeros((h, cw), np.uint16)
    a = np.zeros((h, w), np.uint16) if am else None
    oracle.orc_image_synth(ctypes.c_uint64(seed), w, h, kind, bd, 1 if pf != 0 else 0,
                           u16p(y.ctypes.data), u16p(u.ctypes.data), u16p(v.ctypes.data),
                           u16p(a.ctypes.data) if a is not None else None)
    return [y, u, v] + ([a] if a is not None else [])


def golden_config(abi, case):
    """golden_codec.c:212-238: slice rows 0 -> 32, matrix identity for GBR /
    CFA, primaries 1, transfer 1 (CFA: 20 at 12 bits), range 0."""
    seed, w, h, kind, qm, qp, qpdc, sr, am, abd, bd, pf, prof, em = case
    bd = bd or 10
    return abi.make_config(w, h, qp, profile=prof if prof else (7 if pf == 3 else 3), pixel_format=pf,
                           bit_depth=bd, qmatrix_id=qm, slice_rows=sr if sr else 32, entropy_mode=em,
                           alpha_mode=am, alpha_bit_depth=abd, qp_delta_chroma=qpdc,
                           color_matrix=0 if pf in (2, 3) else 1, color_primaries=1,
                           color_transfer=(20 if pf == 3 else 1), color_range=0)


def golden_dims(abi, case):
    seed, w, h, kind, qm, qp, qpdc, sr, am, abd, bd, pf, prof, em = case
    if pf == 3:
        return [((w + 1) // 2, (h + 1) // 2)] * 4
    return abi.plane_dims(w, h, pf, am)


def golden_cases(names=None):
    out = []
    for name, cases in GOLDEN_SETS.items():
        if names is not None and name not in names:
            continue
        recs = read_tpc1(os.path.join(GOLDEN, name))
        assert len(recs) == 2 * len(cases)
        tag = name[len("golden_codec_"):-len(".bin")]
        for i, c in enumerate(cases):
            out.append(pytest.param(c, recs[2 * i], recs[2 * i + 1], id=f"{tag}-{i}"))
    return out


@pytest.mark.parametrize("case,packet,fold", golden_cases())
def test_golden_codec_oracle(abi, oracle, case, packet, fold):
    dims = golden_dims(abi, case)
    rc, planes, status, conc = orc_decode(oracle, packet, dims)
    assert rc == 0 and conc == 0
    assert plane_fold(planes) == struct.unpack(">Q", fold)[0]
    cfg = golden_config(abi, case)
    rc, pkt = orc_encode(oracle, cfg, golden_inputs(oracle, case))
    assert rc == 0
    assert pkt == packet
