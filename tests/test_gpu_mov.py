"""MOV samples decoded on the GPU as one batch (tcx_movie_decode_batch_dev):
the five movies of the reference's golden_mov_v1 fixture (odd sizes, alpha
modes 1 and 2, sar, variable durations) and a 1080p 422 HQ movie, each sample
against the oracle's decode of the same packet, sample for sample."""
import ctypes

import numpy as np
import pytest

from helpers import frame_for, orc_decode, orc_encode
from test_mov_host import golden_records

pytestmark = pytest.mark.gpu


def _decode_movie(pkg, abi, codec, mv, start, count):
    info = mv.info()
    dims = abi.plane_dims(info.visible_width, info.visible_height, info.pixel_format, info.alpha_mode)
    arena_bytes = len(mv.packets_into_batch(start, count)[0])
    arena = codec.alloc(arena_bytes)
    planes = [[codec.alloc(w * h * 2) for (w, h) in dims] for _ in range(count)]
    print(f"movie {info.visible_width}x{info.visible_height} samples {start}+{count}", flush=True)
    st, rcs, infos = codec.decode_movie_batch_dev(mv, start, count, arena, [[b.addr for b in p] for p in planes])
    assert st == 0 and rcs == [0] * count, codec.last_error()
    out = [[b.download(w * h * 2).view(np.uint16).reshape(h, w) for b, (w, h) in zip(p, dims)] for p in planes]
    return dims, out, infos


@pytest.mark.parametrize("case", range(5))
def test_golden_movie_batch(pkg, abi, oracle, codec, case):
    data = golden_records()[2 * case]
    lib = codec._lib
    with pkg.ToposMovieFile(lib, data) as mv:
        n = mv.info().sample_count
        dims, out, infos = _decode_movie(pkg, abi, codec, mv, 0, n)
        for i in range(n):
            rc, want, _, _ = orc_decode(oracle, mv.packet(i), dims)
            assert rc == 0
            for p in range(len(dims)):
                assert np.array_equal(out[i][p], want[p]), (case, i, p)
            assert infos[i].visible_width == dims[0][0]


def test_hq_1080p_movie_range(pkg, abi, syn, oracle, codec):
    """Eight 1080p 422 HQ samples (four distinct frames, qp varying by frame)
    muxed in memory; samples 3..7 decoded as one batch."""
    lib = codec._lib
    pkts, cfg0 = [], None
    for i in range(4):
        cfg = abi.make_config(1920, 1080, 60 + i, profile=3, qmatrix_id=0, features=3, slice_rows=8,
                              entropy_mode=8)
        cfg0 = cfg0 or cfg
        rc, pkt = orc_encode(oracle, cfg, frame_for(pkg, syn, "422", 1920, 1080, 300 + i))
        assert rc == 0
        pkts.append(pkt)
    st, header = codec.query(pkts[0])
    assert st == 0
    mcfg = abi.movie_config_for(cfg0, header, timescale=24000)
    with pkg.ToposMuxFile(lib, None, mcfg) as mx:
        for k in range(8):
            mx.add_packet(pkts[k % 4], 1001 * k, 1001)
        mx.finish()
        data = mx.data
    with pkg.ToposMovieFile(lib, data) as mv:
        assert mv.info().sample_count == 8
        dims, out, _ = _decode_movie(pkg, abi, codec, mv, 3, 5)
    for j in range(5):
        rc, want, _, _ = orc_decode(oracle, pkts[(3 + j) % 4], dims)
        assert rc == 0
        for p in range(3):
            assert np.array_equal(out[j][p], want[p]), (j, p)


def test_arena_too_small(pkg, abi, codec):
    data = golden_records()[0]
    lib = codec._lib
    with pkg.ToposMovieFile(lib, data) as mv:
        total = len(mv.packets_into_batch(0, 3)[0])
        arena = codec.alloc(total)
        arena.nbytes = total - 1
        planes = [[codec.alloc(64 * 40 * 2) for _ in range(3)] for _ in range(3)]
        st, _, _ = codec.decode_movie_batch_dev(mv, 0, 3, arena, [[b.addr for b in p] for p in planes])
        arena.nbytes = total
        assert st == abi.TC_ERR_BUFFER_TOO_SMALL
        st, _, _ = codec.decode_movie_batch_dev(mv, 2, 2, arena, [[b.addr for b in p] for p in planes[:2]])
        assert st == abi.TC_ERR_INVALID_ARGUMENT
