"""MOV container, video track (host C, no GPU): the demuxer's index and the
muxer's bytes pinned to the reference's own conformance fixture
(native/tests/conformance/golden_mov_v1.bin, generated by
native/tests/conformance/golden_mov.c: five movies with their index
summaries), and to the compiled reference (oracle/_ref) for the metadata atoms,
the faststart rewrite, error codes of damaged files and the mux's rejections."""
import ctypes
import os
import struct

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "golden_mov_v1.bin")


def golden_records():
    """golden_mov.c:291-330: 'TPM1' | version | count | fold | {u32be len | bytes}."""
    data = open(GOLDEN, "rb").read()
    assert data[:4] == b"TPM1"
    version, count = struct.unpack(">II", data[4:12])
    assert version == 1 and count == 10
    recs, off = [], 24
    for _ in range(count):
        (n,) = struct.unpack(">I", data[off:off + 4])
        recs.append(data[off + 4:off + 4 + n])
        off += 4 + n
    assert off == len(data)
    return recs


def golden_fold(recs):
    """golden_mov.c:30-38: the splitmix fold over lengths and bytes."""
    m = (1 << 64) - 1

    def mix(h, v):
        h = (h + 0x9E3779B97F4A7C15) & m
        z = v
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return h ^ (z ^ (z >> 31))

    h = 0x5A5A5A5A5A5A5A5A
    for r in recs:
        h = mix(h, len(r))
        for b in r:
            h = mix(h, b)
    return h


@pytest.fixture(scope="module")
def lib(pkg):
    return pkg.load_library()


@pytest.fixture(scope="module")
def reflib(ref, pkg):
    pkg.movie_signatures(ref)
    return ref


@pytest.fixture(scope="module")
def recs():
    return golden_records()


def summary(pkg, lib, data):
    """golden_mov.c:239-289: n, index_bytes, faststart, timescale, then per
    sample pts u64, dur u32, size u32, sync u8 (all big-endian)."""
    with pkg.ToposMovieFile(lib, data) as mv:
        info = mv.info()
        out = struct.pack(">IIII", info.sample_count, info.index_bytes, info.faststart, info.timescale)
        for i in range(info.sample_count):
            need = ctypes.c_size_t(0)
            assert lib.tc_movie_packet(mv.handle, i, None, 0, ctypes.byref(need)) == -15
            pts, dur = mv.pts(i)
            out += struct.pack(">QII", pts, dur, need.value) + bytes([mv.is_sync(i)])
        return out


def config_from(abi, info):
    """The movie config the file was written with, from its tpcC and sample entry."""
    names = [f for f, _ in abi._MOVIE_FIELDS if f not in ("struct_size", "abi_version")]
    return abi.sized(abi.MovieConfig, **{k: getattr(info, k) for k in names})


def remux(pkg, lib, abi, data, faststart=None):
    """Demux a movie and mux its samples again (same config, pts, durations)."""
    with pkg.ToposMovieFile(lib, data) as mv:
        info = mv.info()
        cfg = config_from(abi, info)
        samples = [(mv.packet(i), *mv.pts(i)) for i in range(info.sample_count)]
        fs = info.faststart if faststart is None else faststart
    with pkg.ToposMuxFile(lib, None, cfg) as mx:
        for pkt, pts, dur in samples:
            mx.add_packet(pkt, pts, dur)
        mx.finish()
        out = mx.data
    return pkg.movie_faststart(lib, out) if fs else out


def test_golden_file_fold(recs):
    data = open(GOLDEN, "rb").read()
    assert golden_fold(recs) == struct.unpack(">Q", data[16:24])[0]


@pytest.mark.parametrize("case", range(5))
def test_golden_demux_summary(pkg, lib, recs, case):
    """The index of every golden movie equals the reference's stored summary."""
    assert summary(pkg, lib, recs[2 * case]) == recs[2 * case + 1]


@pytest.mark.parametrize("case", range(5))
def test_golden_remux_byte_exact(pkg, lib, abi, recs, case):
    """Muxing the golden movie's samples again gives the stored file byte for byte
    (standard layout; case 1 through the faststart rewrite)."""
    assert remux(pkg, lib, abi, recs[2 * case]) == recs[2 * case]


def test_golden_faststart_of_standard(pkg, lib, recs):
    """Cases 0 and 1 hold the same samples: faststart(case 0) is case 1."""
    fs = pkg.movie_faststart(lib, recs[0])
    assert fs == recs[2]
    assert pkg.movie_faststart(lib, fs) == b""  # already moov-first: nothing written


def test_golden_fixture_cases(pkg, lib, recs):
    """golden_mov.c:100-106: sizes, frame counts, alpha, sar and variable durations."""
    want = [(64, 40, 3, 0, (1, 1)), (64, 40, 3, 0, (1, 1)), (36, 20, 5, 1, (1, 1)), (64, 40, 4, 2, (5, 4)),
            (64, 40, 6, 0, (1, 1))]
    for case, (w, h, n, am, sar) in enumerate(want):
        with pkg.ToposMovieFile(lib, recs[2 * case]) as mv:
            info = mv.info()
            assert (info.visible_width, info.visible_height, info.sample_count, info.alpha_mode) == (w, h, n, am)
            assert (info.sar_num, info.sar_den) == sar and info.timescale == 24 and info.reserved[0] == 1
            assert info.faststart == (case == 1)
            durs = [mv.pts(i)[1] for i in range(n)]
            assert durs == ([1 + i % 3 for i in range(n)] if case == 4 else [1] * n)
            assert all(mv.is_sync(i) and mv.prev_sync(i) == i for i in range(n))


@pytest.mark.parametrize("case", [0, 4])
def test_packet_batch_matches_single(pkg, lib, recs, case):
    with pkg.ToposMovieFile(lib, recs[2 * case]) as mv:
        n = mv.info().sample_count
        singles = [mv.packet(i) for i in range(n)]
        arena, offs, sizes = mv.packets_into_batch(0, n)
        assert arena == b"".join(singles) and sizes == [len(p) for p in singles]
        assert offs == [sum(sizes[:i]) for i in range(n)]
        arena, offs, sizes = mv.packets_into_batch(1, n - 1)
        assert arena == b"".join(singles[1:])
        need = ctypes.c_size_t(7)
        assert lib.tc_movie_packet_batch(mv.handle, 0, 0, None, 0, None, None, ctypes.byref(need)) == 0
        assert need.value == 7  # count 0: nothing touched
        assert lib.tc_movie_packet_batch(mv.handle, n - 1, 2, None, 0, None, None, None) == -1
        buf = ctypes.create_string_buffer(len(arena))
        assert lib.tc_movie_packet_batch(mv.handle, 1, n - 1, buf, len(arena) - 1, None, None, None) == -15


def test_open_fd_path(pkg, lib, recs, tmp_path):
    p = tmp_path / "a.mov"
    p.write_bytes(recs[6])
    with pkg.ToposMovieFile(lib, str(p)) as mv:
        assert mv.info().sample_count == 4
        arena, _, _ = mv.packets_into_batch(0, 4)
    assert summary(pkg, lib, recs[6]) == recs[7]
    with pkg.ToposMovieFile(lib, recs[6]) as mv:
        assert mv.packets_into_batch(0, 4)[0] == arena


def _mux_both(pkg, libs, abi, data, meta=None, budget=None, raw=None, edit=None):
    outs = []
    for L in libs:
        with pkg.ToposMovieFile(L, data) as mv:
            info = mv.info()
            cfg = config_from(abi, info)
            if edit:
                edit(cfg)
            samples = [(mv.packet(i), *mv.pts(i)) for i in range(info.sample_count)]
        with pkg.ToposMuxFile(L, None, cfg) as mx:
            if meta:
                mx.set_movie_meta(**meta)
            if budget:
                mx.set_alpha_budget(**budget)
            if raw:
                mx.set_raw_meta(raw)
            for pkt, pts, dur in samples:
                mx.add_packet(pkt, pts, dur)
            mx.finish()
            outs.append(mx.data)
    return outs


@pytest.mark.parametrize("case", range(5))
def test_metadata_atoms_match_reference(pkg, lib, reflib, abi, recs, case):
    """tpcD (tier, vendor, label), tpcB (alpha movies) and trwm: the same bytes
    as the reference writer, read back identically by both readers, and kept
    through the faststart rewrite."""
    budget = dict(target_ratio_bp=2500, actual_ratio_bp=2471, max_abs_error=3, flags=4,
                  frame_count=[3, 3, 5, 4, 6][case]) if case in (2, 3) else None
    meta = dict(tier_id=4, vendor="TOPOS", label="Topos 422 HQ")
    raw = bytes(range(40))
    ours, theirs = _mux_both(pkg, [lib, reflib], abi, recs[2 * case], meta=meta, budget=budget, raw=raw)
    assert ours == theirs
    fs_ours, fs_theirs = pkg.movie_faststart(lib, ours), pkg.movie_faststart(reflib, ours)
    assert fs_ours == fs_theirs
    for data in (ours, fs_ours):
        for L in (lib, reflib):
            with pkg.ToposMovieFile(L, data) as mv:
                info = mv.info()
                assert info.reserved[1] == 4 and mv.read_raw_meta() == raw
                bi = mv.alpha_budget()
                if budget is None:
                    assert bi is None
                else:
                    assert {k: getattr(bi, k) for k in budget} == budget
        assert summary(pkg, lib, data) == summary(pkg, reflib, data)


def test_sar_zero_and_timescale_default(pkg, lib, reflib, abi, recs):
    """timescale 0 -> 24000 at create; sar 0/0 stays in tpcC, reads back as 1/1."""
    def edit(cfg):
        cfg.timescale = 0
    ours, theirs = _mux_both(pkg, [lib, reflib], abi, recs[0], edit=edit)
    assert ours == theirs
    with pkg.ToposMovieFile(lib, ours) as mv:
        assert mv.info().timescale == 24000


def _open_rc(pkg, L, data):
    """Status of open, then of every packet read when it opens; the summary."""
    src = pkg.MemoryIo(data)
    h = ctypes.c_void_p()
    rc = int(L.tc_movie_open(src.ref(), ctypes.byref(h)))
    if rc != 0:
        return (rc,)
    info = pkg.abi.sized(pkg.abi.MovieInfo)
    L.tc_movie_info(h, ctypes.byref(info))
    out = [rc, bytes(info)]
    for i in range(info.sample_count):
        need = ctypes.c_size_t(0)
        L.tc_movie_packet(h, i, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(max(need.value, 1))
        out.append((int(L.tc_movie_packet(h, i, buf, need.value, None)), need.value))
    L.tc_movie_close(h)
    return tuple(out)


def _mutations(data, seed, count):
    """Deterministic damage: byte flips weighted to the moov, truncations."""
    import random
    rnd = random.Random(seed)
    moov = data.find(b"moov") - 4
    out = []
    for _ in range(count):
        d = bytearray(data)
        if rnd.random() < 0.8 and moov > 0:
            pos = rnd.randrange(moov, len(d)) if moov > 40 else rnd.randrange(moov, moov + 900)
            pos = min(pos, len(d) - 1)
        else:
            pos = rnd.randrange(len(d))
        d[pos] ^= 1 << rnd.randrange(8) if rnd.random() < 0.7 else rnd.randrange(1, 256)
        out.append(bytes(d))
    for cut in sorted(rnd.sample(range(1, len(data)), min(40, len(data) - 1))):
        out.append(data[:cut])
    return out


@pytest.mark.parametrize("case", range(5))
def test_damaged_files_match_reference(pkg, lib, reflib, recs, case):
    """Open / info / packet statuses of damaged and cut files agree with the
    reference reader (flips in the moov's tables, atom sizes and tpcC, cut
    downloads, damaged packet headers)."""
    data = recs[2 * case]
    for i, d in enumerate(_mutations(data, 1000 + case, 250)):
        a, b = _open_rc(pkg, lib, d), _open_rc(pkg, reflib, d)
        assert a == b, f"mutation {i}: ours {a[:1]} reference {b[:1]}"


def test_mux_rejections_match_reference(pkg, lib, reflib, abi, recs):
    """pts gaps, zero durations, packets whose header differs from the movie
    config, calls after finish: the same statuses as the reference writer."""
    with pkg.ToposMovieFile(lib, recs[0]) as mv:
        cfg = config_from(abi, mv.info())
        pkts = [mv.packet(i) for i in range(3)]
    with pkg.ToposMovieFile(lib, recs[4]) as mv:
        alpha_pkt = mv.packet(0)

    def run(L):
        sink = pkg.MemorySink()
        h = ctypes.c_void_p()
        codes = [int(L.tc_mux_create(ctypes.byref(cfg), sink.ref(), ctypes.byref(h)))]
        codes.append(int(L.tc_mux_add_packet(h, pkts[0], len(pkts[0]), 1, 1)))      # pts must start at 0
        codes.append(int(L.tc_mux_add_packet(h, pkts[0], len(pkts[0]), 0, 0)))      # dur 0
        codes.append(int(L.tc_mux_add_packet(h, alpha_pkt, len(alpha_pkt), 0, 1)))  # other config
        codes.append(int(L.tc_mux_add_packet(h, pkts[0][:-1], len(pkts[0]) - 1, 0, 1)))  # cut packet
        codes.append(int(L.tc_mux_add_packet(h, pkts[0], len(pkts[0]), 0, 2)))
        codes.append(int(L.tc_mux_add_packet(h, pkts[1], len(pkts[1]), 1, 1)))      # gap
        codes.append(int(L.tc_mux_add_packet(h, pkts[1], len(pkts[1]), 2, 1)))
        codes.append(int(L.tc_mux_finish(h)))
        codes.append(int(L.tc_mux_finish(h)))
        codes.append(int(L.tc_mux_add_packet(h, pkts[2], len(pkts[2]), 3, 1)))
        L.tc_mux_free(h)
        return codes, bytes(sink.data)

    a, b = run(lib), run(reflib)
    assert a == b
    assert a[0][:1] == [0] and a[0][-3:] == [0, -12, -12]


def test_mux_config_rules_match_reference(pkg, lib, reflib, abi):
    """tc_mux_create's validation of the movie config against the reference's."""
    cases = []
    for kw in [dict(), dict(timescale=1000001), dict(profile=4), dict(bit_depth=9), dict(width=0),
               dict(qmatrix=7), dict(alpha_mode=1, alpha_bit_depth=16), dict(alpha_mode=3),
               dict(color_matrix=5), dict(pixel_format=1, profile=5), dict(chroma_siting=9),
               dict(sar_num=0, sar_den=0), dict(sar_num=3, sar_den=0)]:
        w = kw.pop("width", 64)
        cases.append(abi.movie_config(w, 40, **kw))
    half = abi.movie_config(64, 40)
    half.reserved[2] = 2
    cases.append(half)
    for cfg in cases:
        rcs = []
        for L in (lib, reflib):
            sink = pkg.MemorySink()
            h = ctypes.c_void_p()
            rcs.append(int(L.tc_mux_create(ctypes.byref(cfg), sink.ref(), ctypes.byref(h))))
            if rcs[-1] == 0:
                L.tc_mux_free(h)
        assert rcs[0] == rcs[1], (rcs, cfg.profile, cfg.bit_depth)


def test_audio_and_timecode_not_implemented(pkg, lib, abi):
    cfg = abi.movie_config(64, 40)
    cfg.reserved[0] = 1
    cfg.reserved[1] = 48000
    cfg.reserved[2] = 2
    cfg.reserved[4] = 16
    sink = pkg.MemorySink()
    h = ctypes.c_void_p()
    assert lib.tc_mux_create(ctypes.byref(cfg), sink.ref(), ctypes.byref(h)) == abi.TC_ERR_NOT_IMPLEMENTED
    assert lib.tc_mux_set_timecode(None, 1, 0, 0, 0, 24, 0) == abi.TC_ERR_NOT_IMPLEMENTED


def test_metadata_setter_rules(pkg, lib, reflib, abi, recs):
    """tc_mux_set_movie_meta / set_alpha_budget / set_raw_meta argument rules."""
    with pkg.ToposMovieFile(lib, recs[4]) as mv:
        alpha_cfg = config_from(abi, mv.info())
    plain_cfg = abi.movie_config(64, 40)

    def run(L):
        codes = []
        for cfg in (plain_cfg, alpha_cfg):
            sink = pkg.MemorySink()
            h = ctypes.c_void_p()
            assert L.tc_mux_create(ctypes.byref(cfg), sink.ref(), ctypes.byref(h)) == 0
            for kw in [dict(tier_id=9), dict(tier_id=8), dict(reserved=(ctypes.c_uint8 * 3)(0, 1, 0))]:
                codes.append(int(L.tc_mux_set_movie_meta(h, ctypes.byref(abi.sized(abi.MovieMeta, **kw)))))
            for kw in [dict(target_ratio_bp=10001), dict(target_ratio_bp=0xFFFF), dict(target_ratio_bp=10000)]:
                codes.append(int(L.tc_mux_set_alpha_budget(h, ctypes.byref(abi.sized(abi.AlphaBudgetInfo, **kw)))))
            for n in (0, 1, 256, 257):
                rm = abi.sized(abi.RawMeta, payload_size=n)
                codes.append(int(L.tc_mux_set_raw_meta(h, ctypes.byref(rm))))
            L.tc_mux_free(h)
        return codes

    assert run(lib) == run(reflib)


def test_large_movie_parallel_reads(pkg, lib, reflib, abi, syn, oracle, tmp_path):
    """A 1080p movie over 8 MiB opened by descriptor: the batch read splits into
    parallel preads; bytes and statuses equal the callback source and the
    reference reader."""
    from helpers import frame_for, orc_encode
    cfg = abi.make_config(1920, 1080, 4, profile=3, entropy_mode=8, slice_rows=8)
    rc, pkt = orc_encode(oracle, cfg, frame_for(pkg, syn, "422", 1920, 1080, 7))
    assert rc == 0
    st, header = pkg.ToposCodec.query(pkg.ToposCodec(), pkt)
    assert st == 0
    n = max(3, (9 << 20) // len(pkt) + 1)
    with pkg.ToposMuxFile(lib, str(tmp_path / "big.mov"), abi.movie_config_for(cfg, header)) as mx:
        for i in range(n):
            mx.add_packet(pkt, i, 1)
        mx.finish()
        data = mx.data
    assert len(data) > (8 << 20)
    with pkg.ToposMovieFile(lib, str(tmp_path / "big.mov")) as mv:
        arena, offs, sizes = mv.packets_into_batch(0, n)
        tail = mv.packets_into_batch(1, n - 1)[0]
    assert arena == pkt * n and tail == pkt * (n - 1) and sizes == [len(pkt)] * n
    with pkg.ToposMovieFile(reflib, str(tmp_path / "big.mov")) as mv:
        assert mv.packets_into_batch(0, n)[0] == arena
