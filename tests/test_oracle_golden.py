"""Pin the oracle against the reference's own conformance vectors
(native/tests/conformance/golden_transform_v1.bin, golden_codec_v7r2_intra.bin,
golden_codec_v1_traw.bin, copied verbatim into tests/golden/ as data):

 - transform/quant/dequant/reconstruction: 1380 records {depth, qp, px, F, q,
   Fp, rec} (layout native/tests/conformance/golden_util.h:1-38);
 - rANS2 intra packets and decode fingerprints: the 8 em=8 cases of
   native/tests/conformance/golden_codec.c:135-144 and the two em=8 CFA cases
   :149-156; inputs regenerated with the oracle's restatement of the
   reference test image synthesiser.
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


# (seed, w, h, kind, qmatrix, qp, slice_rows, alpha_mode, alpha_bd, bit_depth, pixel_format, profile)
V7R2_CASES = [
    (0x301, 73, 39, 1, 1, 24, 2, 0, 0, 10, 0, 0),
    (0x302, 35, 21, 3, 1, 18, 1, 1, 16, 12, 0, 0),
    (0x303, 73, 39, 4, 0, 31, 2, 2, 12, 10, 1, 0),
    (0x304, 35, 21, 1, 1, 24, 1, 0, 0, 12, 1, 0),
    (0x305, 73, 39, 3, 1, 18, 2, 1, 16, 10, 2, 0),
    (0x306, 35, 21, 4, 0, 31, 1, 2, 8, 12, 2, 0),
    (0x307, 1, 1, 0, 0, 0, 1, 0, 0, 10, 0, 0),
    (0x308, 16, 16, 2, 1, 63, 1, 0, 0, 12, 2, 0),
]
TRAW_CASES = [  # record pairs 4 and 5 of golden_codec_v1_traw.bin
    (0x7219000000000005, 64, 48, 2, 0, 20, 0, 0, 0, 12, 3, 7),
    (0x7219000000000006, 48, 48, 3, 0, 63, 0, 0, 0, 12, 3, 7),
]


def golden_inputs(oracle, case):
    seed, w, h, kind, qm, qp, sr, am, abd, bd, pf, prof = case
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
    seed, w, h, kind, qm, qp, sr, am, abd, bd, pf, prof = case
    return abi.make_config(w, h, qp, profile=prof if prof else (7 if pf == 3 else 3), pixel_format=pf,
                           bit_depth=bd, qmatrix_id=qm, slice_rows=sr if sr else 32, entropy_mode=8,
                           alpha_mode=am, alpha_bit_depth=abd,
                           color_matrix=0 if pf in (2, 3) else 1, color_primaries=1,
                           color_transfer=(20 if pf == 3 else 1), color_range=0)


def golden_cases():
    out = []
    recs = read_tpc1(os.path.join(GOLDEN, "golden_codec_v7r2_intra.bin"))
    for i, c in enumerate(V7R2_CASES):
        out.append(pytest.param(c, recs[2 * i], recs[2 * i + 1], id=f"v7r2-{c[0]:#x}"))
    recs = read_tpc1(os.path.join(GOLDEN, "golden_codec_v1_traw.bin"))
    for j, c in enumerate(TRAW_CASES):
        i = 4 + j
        out.append(pytest.param(c, recs[2 * i], recs[2 * i + 1], id=f"traw-{c[0] & 0xFF}"))
    return out


@pytest.mark.parametrize("case,packet,fold", golden_cases())
def test_golden_codec_oracle(abi, oracle, case, packet, fold):
    seed, w, h, kind, qm, qp, sr, am, abd, bd, pf, prof = case
    dims = abi.plane_dims(w, h, pf, am)
    if pf == 3:
        dims = [((w + 1) // 2, (h + 1) // 2)] * 4
    rc, planes, status, conc = orc_decode(oracle, packet, dims)
    assert rc == 0 and conc == 0
    assert plane_fold(planes) == struct.unpack(">Q", fold)[0]
    cfg = golden_config(abi, case)
    rc, pkt = orc_encode(oracle, cfg, golden_inputs(oracle, case))
    assert rc == 0
    assert pkt == packet
