"""CPU checks of the .toos envelope read path (include/topos_image.h): the
reference's golden image files probe and validate, and the host logic agrees
with the compiled reference (oracle/_ref) on every golden file and on
systematic byte mutations of the preamble, directory, IDSC and packet header,
with and without repaired checksums (return codes and probe results)."""
import ctypes
import glob
import os

import pytest

from helpers import idsc_for_packet, orc_encode, toos_layout, toos_repair, toos_wrap, frame_for

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
IMAGES = sorted(glob.glob(os.path.join(GOLDEN, "golden_image_v1*.bin")))


def _ref_image(ref):
    vp = ctypes.c_void_p
    ref.tc_image_probe.argtypes = [vp, vp]
    ref.tc_image_validate.argtypes = [vp, ctypes.c_uint32, vp]
    ref.tc_image_derive_idsc.argtypes = [vp, ctypes.c_size_t, ctypes.c_uint8, vp]
    return ref


def _probe(lib, pkg, data):
    io = pkg.MemoryIo(data)
    info = pkg.abi.ImageInfo()
    rc = lib.tc_image_probe(io.ref(), ctypes.byref(info))
    return rc, bytes(info)


def _validate(lib, pkg, data, flags):
    io = pkg.MemoryIo(data)
    info = pkg.abi.ImageInfo()
    return lib.tc_image_validate(io.ref(), flags, ctypes.byref(info))


@pytest.mark.parametrize("path", IMAGES, ids=[os.path.basename(p)[:-4] for p in IMAGES])
def test_golden_images_probe_and_validate(pkg, path):
    ic = pkg.ToposImageCodec()
    info = ic.validate(path, deep=True)
    assert info.pixl_size > 53 and info.idsc.payload_major == 1
    assert info.preamble.file_size == os.path.getsize(path)


@pytest.mark.parametrize("path", IMAGES, ids=[os.path.basename(p)[:-4] for p in IMAGES])
def test_golden_images_match_reference(pkg, ref, path):
    ref = _ref_image(ref)
    lib = pkg.load_library()
    data = open(path, "rb").read()
    assert _probe(lib, pkg, data) == _probe(ref, pkg, data)
    for flags in (0, 1):
        assert _validate(lib, pkg, data, flags) == _validate(ref, pkg, data, flags) == 0


def _mutation_offsets(data):
    dir_off, count, ents = toos_layout(data)
    offs = list(range(64)) + list(range(dir_off, dir_off + 32 * count))
    for typ, _fl, off, size in ents:
        if typ == 0x49445343:
            offs += list(range(off, off + min(size, 128)))
        if typ == 0x5049584C:
            offs += list(range(off, off + min(size, 53)))
    return sorted(set(o for o in offs if o < len(data)))


@pytest.mark.parametrize("name", ["golden_image_v1.bin", "golden_image_v1_alpha.bin", "golden_image_v1_meta.bin",
                                  "golden_image_v1_traw.bin"])
def test_envelope_mutations_match_reference(pkg, ref, name):
    ref = _ref_image(ref)
    lib = pkg.load_library()
    base = open(os.path.join(GOLDEN, name), "rb").read()
    seen = set()
    n = 0
    for off in _mutation_offsets(base):
        for bit in (0x01, 0x80):
            m = bytearray(base)
            m[off] ^= bit
            for variant in (bytes(m), toos_repair(bytes(m))):
                r_mine, i_mine = _probe(lib, pkg, variant)
                r_ref, i_ref = _probe(ref, pkg, variant)
                assert r_mine == r_ref, (off, bit)
                if r_ref == 0:
                    assert i_mine == i_ref, (off, bit)
                    assert _validate(lib, pkg, variant, 1) == _validate(ref, pkg, variant, 1), (off, bit)
                seen.add(r_ref)
                n += 1
    assert n > 500
    # the sweep reaches the checks of every layer
    assert {-100, -101, -102, -104, 0} <= seen


def test_truncated_and_resized_files_match_reference(pkg, ref):
    ref = _ref_image(ref)
    lib = pkg.load_library()
    base = open(os.path.join(GOLDEN, "golden_image_v1_gbr12.bin"), "rb").read()
    for cut in (0, 10, 63, 64, 100, 200, len(base) - 1):
        d = base[:cut]
        assert _probe(lib, pkg, d)[0] == _probe(ref, pkg, d)[0], cut
    d = base + b"\0" * 7
    assert _probe(lib, pkg, d)[0] == _probe(ref, pkg, d)[0]


@pytest.mark.parametrize("em", [0, 1, 8])
def test_envelope_writer_matches_reference_idsc(pkg, abi, syn, oracle, ref, em):
    # the tests' envelope writer derives the IDSC exactly as the reference does
    ref = _ref_image(ref)
    for kind, w, h, prof, kw in (("422", 160, 90, 0, dict(profile=3)),
                                 ("4444", 96, 64, 2, dict(profile=6, pixel_format=2, bit_depth=12, alpha_mode=1,
                                                          alpha_bit_depth=16, color_matrix=0)),
                                 ("raw", 128, 96, 3, dict(profile=7, pixel_format=3, bit_depth=12))):
        planes = frame_for(pkg, syn, kind, w, h, 5)
        cfg = abi.make_config(w, h, 30, entropy_mode=em, slice_rows=4, **kw)
        rc, pkt = orc_encode(oracle, cfg, planes)
        assert rc == 0
        out = abi.ImageIdsc()
        assert ref.tc_image_derive_idsc(pkt, len(pkt), prof, ctypes.byref(out)) == 0
        mine = idsc_for_packet(pkt, prof)
        io = pkg.MemoryIo(toos_wrap(pkt, prof))
        info = abi.ImageInfo()
        lib = pkg.load_library()
        assert lib.tc_image_probe(io.ref(), ctypes.byref(info)) == 0
        assert bytes(info.idsc) == bytes(out)
        assert len(mine) == 128
