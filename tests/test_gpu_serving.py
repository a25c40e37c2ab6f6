"""GPU serving paths: mixed-format batches, the concurrent encode + decode
serving step, concurrent callers of the drop-in API on several host threads,
a frame coded alone vs inside a mixed batch, and slice concealment of damaged
packets. Every packet is compared byte for byte with the oracle encoder and
every decode with the oracle decoder (slice status included)."""
import threading

import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

# (kind, w, h, seed, cfg kwargs): an uneven mix of sizes and formats
MIX = [
    ("422", 640, 360, 11, dict(qp=75, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("4444", 192, 108, 12, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                                alpha_bit_depth=16)),
    ("raw", 320, 240, 13, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 1920, 1080, 14, dict(qp=68, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("422", 77, 45, 15, dict(qp=30, profile=3, slice_rows=2)),
]


def _cfg(abi, w, h, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    return abi.make_config(w, h, qp, entropy_mode=8, **kw)


def _dims(abi, cfg):
    return abi.plane_dims(cfg.visible_width, cfg.visible_height, cfg.pixel_format, cfg.alpha_mode)


def _upload_frame(codec, planes):
    bufs = []
    for p in planes:
        b = codec.alloc(p.nbytes)
        b.upload(np.ascontiguousarray(p))
        bufs.append(b)
    return bufs


@pytest.fixture(scope="module")
def mix(pkg, abi, syn, oracle):
    out = []
    for kind, w, h, seed, kw in MIX:
        planes = frame_for(pkg, syn, kind, w, h, seed)
        cfg = _cfg(abi, w, h, kw)
        rc, pkt = orc_encode(oracle, cfg, planes)
        assert rc == 0
        rc, dec, status, nconc = orc_decode(oracle, pkt, _dims(abi, cfg))
        assert rc == 0
        out.append(dict(planes=planes, cfg=cfg, pkt=pkt, dec=dec))
    return out


@pytest.mark.gpu
def test_serve_step_mixed_batches(abi, codec, mix):
    n = len(mix)
    in_bufs = [_upload_frame(codec, m["planes"]) for m in mix]
    enc_out = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in mix]
    dec_pk = []
    for m in mix:
        b = codec.alloc(len(m["pkt"]))
        b.upload(np.frombuffer(m["pkt"], np.uint8))
        dec_pk.append(b)
    # decode the batch in reverse order so the two directions see different frames
    order = list(reversed(range(n)))
    dec_planes = [[codec.alloc(w * h * 2) for (w, h) in _dims(abi, mix[i]["cfg"])] for i in order]
    st, er, dr = codec.serve_step(
        enc=([m["cfg"] for m in mix], [[b.addr for b in bs] for bs in in_bufs], enc_out),
        dec=([mix[i]["pkt"] for i in order], [dec_pk[i] for i in order],
             [[b.addr for b in o] for o in dec_planes]))
    assert st == 0, codec.last_error()
    sizes, rcs, _ = er
    assert rcs == [0] * n
    for i, m in enumerate(mix):
        assert sizes[i] == len(m["pkt"])
        assert enc_out[i].download(sizes[i]).tobytes() == m["pkt"], f"frame {i}"
    drcs, infos = dr
    assert drcs == [0] * n
    for j, i in enumerate(order):
        for p, (w, h) in enumerate(_dims(abi, mix[i]["cfg"])):
            got = dec_planes[j][p].download(w * h * 2, np.uint16).reshape(h, w)
            assert (got == mix[i]["dec"][p]).all(), f"frame {i} plane {p}"


@pytest.mark.gpu
def test_frame_alone_equals_frame_in_mix(abi, codec, mix):
    # the 1080p frame coded alone and as one member of the mixed batch
    k = 3
    alone_in = _upload_frame(codec, mix[k]["planes"])
    alone_out = codec.alloc(len(mix[k]["pkt"]) + (1 << 20))
    st, sizes, rcs, _ = codec.encode_batch_dev([mix[k]["cfg"]], [[b.addr for b in alone_in]], [alone_out])
    assert st == 0 and rcs == [0]
    alone = alone_out.download(sizes[0]).tobytes()
    in_bufs = [_upload_frame(codec, m["planes"]) for m in mix]
    outs = [codec.alloc(len(m["pkt"]) + (1 << 20)) for m in mix]
    st, sizes, rcs, _ = codec.encode_batch_dev([m["cfg"] for m in mix], [[b.addr for b in bs] for bs in in_bufs],
                                               outs)
    assert st == 0
    assert outs[k].download(sizes[k]).tobytes() == alone == mix[k]["pkt"]


@pytest.mark.gpu
def test_concurrent_host_threads(pkg, abi, mix):
    # the drop-in calls from several threads at once (distinct device contexts)
    errors = []

    def worker(t):
        try:
            c = pkg.ToposCodec()
            for rep in range(2):
                for i in range(t % 2, len(mix), 2):
                    m = mix[i]
                    st, pkt, _ = c.encode_status(m["cfg"], m["planes"])
                    if st != 0 or pkt != m["pkt"]:
                        errors.append((t, rep, i, "encode", st))
                    rc, planes, _ = c.decode(m["pkt"], strict=False)
                    if rc != 0 or not all((a == b).all() for a, b in zip(planes, m["dec"])):
                        errors.append((t, rep, i, "decode", rc))
        except Exception as e:  # noqa: BLE001  (reported below)
            errors.append((t, repr(e)))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert errors == []


def _slice_table(pkt):
    """(payload offset, payload size) of every slice: 53-byte frame header,
    then 17-byte slice headers each followed by its payload."""
    count = int.from_bytes(pkt[30:32], "big")
    out, off = [], 53
    for _ in range(count):
        size = int.from_bytes(pkt[off:off + 4], "big")
        out.append((off + 17, size))
        off += 17 + size
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["crc", "payload_crc_fixed", "truncated_stream"])
def test_damaged_slices_conceal_like_oracle(abi, codec, oracle, mix, case):
    import zlib
    m = mix[0]
    pkt = bytearray(m["pkt"])
    table = _slice_table(bytes(pkt))
    assert len(table) > 3
    off, size = table[2]
    if case == "crc":
        pkt[off + size // 2] ^= 0x5A                    # payload damaged, slice CRC now wrong
    elif case == "payload_crc_fixed":
        pkt[off + size // 2] ^= 0x5A                    # damaged payload with a matching CRC:
        crc = zlib.crc32(bytes(pkt[off:off + size]))    # the entropy decoder itself must reject
        pkt[off - 4:off] = crc.to_bytes(4, "big")       # or decode it exactly as the oracle does
    else:
        for j in range(off + size - 8, off + size):     # tail of the rANS stream zeroed
            pkt[j] = 0
        crc = zlib.crc32(bytes(pkt[off:off + size]))
        pkt[off - 4:off] = crc.to_bytes(4, "big")
    pkt = bytes(pkt)
    dims = _dims(abi, m["cfg"])
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, pkt, dims)
    d_rc, d_planes, info = codec.decode(pkt, strict=False)
    assert d_rc == o_rc
    assert list(info.slice_status[:len(table)]) == o_status[:len(table)]
    assert info.concealed_slices == o_nconc
    for a, b in zip(d_planes, o_planes):
        assert (a == b).all()
