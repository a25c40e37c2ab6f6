"""Host-side checks of the serving loop: the HBM range allocator against a
byte-ownership model, the Python mirrors of the server structs against the C
header, and the loud failure without a GPU."""
import ctypes
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "topos-codec_amd", "csrc", "host")


def _build(tmp, src, name, extra=()):
    out = str(tmp / name)
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-Wall", "-Wextra", "-I", HOST, "-I",
                    os.path.join(ROOT, "include"), *extra, src, "-o", out], check=True)
    return out


@pytest.fixture(scope="module")
def arena(tmp_path_factory):
    t = tmp_path_factory.mktemp("arena")
    lib = ctypes.CDLL(_build(t, os.path.join(ROOT, "tests", "csrc", "arenacheck.c"), "libarena.so", ["-shared"]))
    lib.ac_run.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
    return lib


@pytest.mark.parametrize("seed,units", [(1, 64), (2, 257), (3, 1000), (4, 7), (5, 4096)])
def test_arena_matches_model(arena, seed, units):
    assert arena.ac_run(seed, 4000, units) == 0


def test_server_structs_match_header(abi, tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "topos_gpu.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(tcx_request), sizeof(tcx_completion),
           sizeof(tcx_server_stats), sizeof(tcx_server_config), offsetof(tcx_request, packet),
           offsetof(tcx_request, stats), offsetof(tcx_completion, t_done_ns), offsetof(tcx_server_stats, inflight));
    return 0;
}
""")
    exe = _build(tmp_path, str(src), "sz")
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(abi.TcxRequest), ctypes.sizeof(abi.TcxCompletion), ctypes.sizeof(abi.TcxServerStats),
            ctypes.sizeof(abi.TcxServerConfig), abi.TcxRequest.packet.offset, abi.TcxRequest.stats.offset,
            abi.TcxCompletion.t_done_ns.offset, abi.TcxServerStats.inflight.offset]
    assert got == want


def test_sized_struct_layouts(abi, tmp_path):
    src = tmp_path / "sized_sz.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "topos_gpu.h"
int main(void) {
    printf("%zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(tcx_sized_req), sizeof(tcx_sized_stream),
           offsetof(tcx_sized_req, target_bytes), offsetof(tcx_sized_req, stream), offsetof(tcx_sized_req, rc),
           offsetof(tcx_sized_req, packet_size), offsetof(tcx_sized_req, encodes), offsetof(tcx_sized_req, stats));
    return 0;
}
""")
    exe = _build(tmp_path, str(src), "sized_sz")
    got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    R = abi.TcxSizedReq
    want = [ctypes.sizeof(R), ctypes.sizeof(abi.TcxSizedStream), R.target_bytes.offset, R.stream.offset,
            R.rc.offset, R.packet_size.offset, R.encodes.offset, R.stats.offset]
    assert got == want


def test_server_without_gpu_fails_loudly(pkg):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    codec = pkg.DeviceCodec()
    with pytest.raises(pkg.ToposCodecError) as e:
        pkg.Server(codec, 1 << 20)
    assert e.value.status == pkg.abi.TC_ERR_NOT_IMPLEMENTED
