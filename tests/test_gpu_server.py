"""GPU serving loop (tcx_server_*): encode and decode requests of mixed sizes
and formats submitted as they arrive, continuous batching over several lanes,
and admission against an HBM budget. Every result is compared with the same
request run alone (one-frame batch call) and with the oracle: packets byte for
byte, decodes sample for sample, slice status of damaged packets included."""
import threading
import time
import zlib

import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode

# (kind, w, h, seed, cfg kwargs): uneven sizes, formats and entropy modes
REQS = [
    ("422", 640, 360, 21, dict(qp=75, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("4444", 256, 144, 22, dict(qp=69, profile=6, pixel_format=2, bit_depth=12, slice_rows=4, alpha_mode=1,
                                alpha_bit_depth=16)),
    ("raw", 320, 240, 23, dict(qp=64, profile=7, pixel_format=3, bit_depth=12, slice_rows=8)),
    ("422", 1280, 720, 24, dict(qp=68, profile=3, qmatrix_id=4, features=1, slice_rows=8)),
    ("422", 77, 45, 25, dict(qp=30, profile=3, slice_rows=2)),
    ("422", 352, 288, 26, dict(qp=40, profile=3, slice_rows=4, entropy_mode=0)),
]


def _cfg(abi, w, h, kw):
    kw = dict(kw)
    qp = kw.pop("qp")
    em = kw.pop("entropy_mode", 8)
    return abi.make_config(w, h, qp, entropy_mode=em, **kw)


def _dims(abi, cfg):
    return abi.plane_dims(cfg.visible_width, cfg.visible_height, cfg.pixel_format, cfg.alpha_mode)


def _upload(codec, arr):
    b = codec.alloc(arr.nbytes)
    b.upload(np.ascontiguousarray(arr))
    return b


@pytest.fixture(scope="module")
def cases(pkg, abi, syn, oracle, codec):
    out = []
    for kind, w, h, seed, kw in REQS:
        planes = frame_for(pkg, syn, kind, w, h, seed)
        cfg = _cfg(abi, w, h, kw)
        rc, pkt = orc_encode(oracle, cfg, planes)
        assert rc == 0
        rc, dec, _, _ = orc_decode(oracle, pkt, _dims(abi, cfg))
        assert rc == 0
        dev_in = [_upload(codec, p) for p in planes]
        dev_pkt = _upload(codec, np.frombuffer(pkt, np.uint8))
        out.append(dict(planes=planes, cfg=cfg, pkt=pkt, dec=dec, dev_in=dev_in, dev_pkt=dev_pkt))
    return out


def _alone(abi, codec, c):
    """The request alone: a one-frame batch call each way."""
    out = codec.alloc(len(c["pkt"]) + (1 << 20))
    st, sizes, rcs, _ = codec.encode_batch_dev([c["cfg"]], [[b.addr for b in c["dev_in"]]], [out])
    assert st == 0 and rcs == [0], codec.last_error()
    pkt = out.download(sizes[0]).tobytes()
    planes = [codec.alloc(w * h * 2) for (w, h) in _dims(abi, c["cfg"])]
    st, rcs, _ = codec.decode_batch_dev([c["pkt"]], [c["dev_pkt"]], [[b.addr for b in planes]])
    assert st == 0 and rcs == [0], codec.last_error()
    dec = [b.download(w * h * 2, np.uint16).reshape(h, w) for b, (w, h) in zip(planes, _dims(abi, c["cfg"]))]
    return pkt, dec


def _download_result(codec, pkg, abi, comp, cfg):
    if comp.kind == abi.TCX_REQ_ENCODE:
        buf = np.empty(comp.packet_size, np.uint8)
        assert codec._lib.tcx_d2h(pkg._ptr(buf), comp.dev_packet, comp.packet_size) == 0
        return buf.tobytes()
    dec = []
    for p, (w, h) in enumerate(_dims(abi, cfg)):
        assert comp.strides[p] == w
        a = np.empty((h, w), np.uint16)
        assert codec._lib.tcx_d2h(pkg._ptr(a), comp.dev_planes[p], w * h * 2) == 0
        dec.append(a)
    return dec


def _drain(srv, n_expected, timeout_s=60.0, on_done=None):
    got = []
    t_end = time.time() + timeout_s
    while len(got) < n_expected:
        assert time.time() < t_end, f"{len(got)} of {n_expected} completions"
        for c in srv.poll(64, timeout_s=1.0):
            got.append(c)
            if on_done is not None:
                on_done(c)
    return got


@pytest.mark.gpu
def test_server_results_equal_alone_and_oracle(pkg, abi, codec, cases):
    alone = [_alone(abi, codec, c) for c in cases]
    for c, (pkt, dec) in zip(cases, alone):
        assert pkt == c["pkt"]
        assert all((a == b).all() for a, b in zip(dec, c["dec"]))
    srv = pkg.Server(codec, 2 << 30, max_batch=4, max_inflight=3)
    try:
        submitted = {}
        k = 0
        for rep in range(3):
            for i, c in enumerate(cases):
                # alternate resident and host-staged inputs on both directions
                if (i + rep) % 2 == 0:
                    st, rid = srv.submit_encode(c["cfg"], dev_planes=[b.addr for b in c["dev_in"]], user=k)
                else:
                    st, rid = srv.submit_encode(c["cfg"], host_planes=c["planes"], user=k)
                assert st == 0, codec.last_error()
                submitted[rid] = ("enc", i)
                k += 1
                dev_pkt = c["dev_pkt"].addr if (i + rep) % 2 == 1 else 0
                st, rid = srv.submit_decode(c["pkt"], dev_packet=dev_pkt, user=k)
                assert st == 0, codec.last_error()
                submitted[rid] = ("dec", i)
                k += 1
        comps = _drain(srv, len(submitted))
        batched = 0
        for comp in comps:
            kind, i = submitted[comp.id]
            assert comp.status == 0, (kind, i, codec.last_error())
            assert comp.t_submit_ns <= comp.t_admit_ns <= comp.t_launch_ns <= comp.t_done_ns
            res = _download_result(codec, pkg, abi, comp, cases[i]["cfg"])
            if kind == "enc":
                assert res == alone[i][0], f"encode {i}"
            else:
                assert all((a == b).all() for a, b in zip(res, alone[i][1])), f"decode {i}"
            batched += comp.batch_frames > 1
            assert srv.release(comp.id) == 0
        assert batched > 0
        st = srv.stats()
        assert st.submitted == st.completed == len(submitted)
        assert st.held_bytes == 0 and st.queued == 0 and st.inflight == 0
        assert st.batch_frames == len(submitted)
    finally:
        srv.close()


@pytest.mark.gpu
def test_server_gathers_bursts_into_equal_batches(pkg, abi, codec, cases):
    """Arrival coalescing and equal batches: a burst submitted faster than
    the coalescing gap is not started as a one-request batch; once arrivals
    stop, nine waiting requests with max_batch 4 run as 3 + 3 + 3 (not
    4 + 4 + 1). Every result equals the request alone."""
    c = cases[4]
    srv = pkg.Server(codec, 1 << 30, max_batch=4, max_inflight=1, gather_us=20000)
    try:
        for k in range(9):
            st, _ = srv.submit_encode(c["cfg"], dev_planes=[b.addr for b in c["dev_in"]], user=k)
            assert st == 0, codec.last_error()
        comps = _drain(srv, 9)
        assert sorted(comp.batch_frames for comp in comps) == [3] * 9
        for comp in comps:
            assert comp.status == 0
            assert _download_result(codec, pkg, abi, comp, c["cfg"]) == c["pkt"]
            assert srv.release(comp.id) == 0
        assert srv.stats().batches == 3
    finally:
        srv.close()


@pytest.mark.gpu
def test_server_admission_holds_budget(pkg, abi, codec, cases):
    # room for the output planes and staged packet of two 1280x720 decodes only
    c = cases[3]
    one = 0
    for (pw, ph) in _dims(abi, c["cfg"]):
        one += (pw * ph * 2 + 255) // 256 * 256
    one += (len(c["pkt"]) + 1024 + 255) // 256 * 256
    budget = 2 * one + 4096
    srv = pkg.Server(codec, budget, max_batch=8, max_inflight=4)
    try:
        ids = []
        for k in range(8):
            st, rid = srv.submit_decode(c["pkt"], user=k)
            assert st == 0
            ids.append(rid)
        waits = []

        def done(comp):
            assert comp.status == 0
            res = _download_result(codec, pkg, abi, comp, c["cfg"])
            assert all((a == b).all() for a, b in zip(res, c["dec"]))
            waits.append(comp.t_admit_ns - comp.t_submit_ns)
            assert srv.release(comp.id) == 0

        _drain(srv, 8, on_done=done)
        st = srv.stats()
        assert st.peak_bytes <= budget
        assert st.held_bytes == 0
        # six of the eight had to wait for earlier results to be released
        assert sorted(waits)[2] > 1_000_000
    finally:
        srv.close()


@pytest.mark.gpu
def test_server_rejects_what_can_never_fit(pkg, abi, codec, cases):
    srv = pkg.Server(codec, 1 << 20, max_batch=2, max_inflight=1)
    try:
        st, rid = srv.submit_decode(cases[3]["pkt"])
        assert st == abi.TC_ERR_OUT_OF_MEMORY and rid is None
        assert "budget" in codec.last_error()
        # bad requests fail at submission; the server keeps serving
        st, _ = srv.submit_decode(b"\x00" * 100)
        assert st < 0
        bad = _cfg(abi, 64, 48, dict(qp=62, profile=3))
        bad.bit_depth = 9
        st, _ = srv.submit_encode(bad, host_planes=cases[4]["planes"])
        assert st < 0
        st, rid = srv.submit_encode(cases[4]["cfg"], host_planes=cases[4]["planes"])
        assert st == 0
        (comp,) = _drain(srv, 1)
        assert comp.status == 0
        assert _download_result(codec, pkg, abi, comp, cases[4]["cfg"]) == cases[4]["pkt"]
        assert srv.release(comp.id) == 0
        assert srv.release(comp.id) < 0
    finally:
        srv.close()


@pytest.mark.gpu
def test_server_poisson_arrivals_and_concealment(pkg, abi, codec, cases, oracle):
    # a damaged packet among the arrivals: the slice CRC of slice 2 no longer matches
    c0 = cases[0]
    dmg = bytearray(c0["pkt"])
    count = int.from_bytes(dmg[30:32], "big")
    off = 53
    table = []
    for _ in range(count):
        size = int.from_bytes(dmg[off:off + 4], "big")
        table.append((off + 17, size))
        off += 17 + size
    po, ps = table[2]
    dmg[po + ps // 2] ^= 0x5A
    dmg = bytes(dmg)
    o_rc, o_planes, o_status, o_nconc = orc_decode(oracle, dmg, _dims(abi, c0["cfg"]))
    assert o_rc == abi.TC_WARN_CONCEALED and zlib.crc32(dmg) != zlib.crc32(c0["pkt"])
    rng = np.random.default_rng(7)
    plan = [(int(rng.integers(0, len(cases))), bool(rng.random() < 0.6)) for _ in range(40)]
    plan[17] = (-1, True)
    gaps = rng.exponential(0.004, size=len(plan))
    srv = pkg.Server(codec, 4 << 30, max_batch=6, max_inflight=4)
    infos = {}
    submitted = {}
    errors = []

    def arrivals():
        try:
            for (i, dec), g in zip(plan, gaps):
                time.sleep(float(g))
                if i < 0:
                    info = abi.FrameOutput()
                    st, rid = srv.submit_decode(dmg, info=info)
                    infos[rid] = info
                elif dec:
                    st, rid = srv.submit_decode(cases[i]["pkt"], dev_packet=cases[i]["dev_pkt"].addr)
                else:
                    st, rid = srv.submit_encode(cases[i]["cfg"], dev_planes=[b.addr for b in cases[i]["dev_in"]])
                if st != 0:
                    errors.append(st)
                    return
                submitted[rid] = (i, dec)
        except Exception as e:  # noqa: BLE001  (reported below)
            errors.append(repr(e))

    try:
        th = threading.Thread(target=arrivals)
        th.start()
        comps = _drain(srv, len(plan))
        th.join()
        assert errors == []
        for comp in comps:
            i, dec = submitted[comp.id]
            if i < 0:
                info = infos[comp.id]
                assert comp.status == abi.TC_WARN_CONCEALED
                assert list(info.slice_status[:count]) == o_status[:count]
                assert info.concealed_slices == o_nconc
                res = _download_result(codec, pkg, abi, comp, c0["cfg"])
                assert all((a == b).all() for a, b in zip(res, o_planes))
            else:
                assert comp.status == 0
                res = _download_result(codec, pkg, abi, comp, cases[i]["cfg"])
                if dec:
                    assert all((a == b).all() for a, b in zip(res, cases[i]["dec"]))
                else:
                    assert res == cases[i]["pkt"]
            assert srv.release(comp.id) == 0
        assert srv.stats().held_bytes == 0
    finally:
        srv.close()
