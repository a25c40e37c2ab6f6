"""The integer helpers the kernels share with the host (tc_math.h), checked on
the CPU: exact rANS state division, exact quantiser division, GF(2) CRC
combination against zlib, and the rANS2 row tables against the compiled
reference's own row codec (native/src/entropy/rans.c:283-362)."""
import ctypes
import os
import subprocess
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "csrc", "mathcheck.c")


@pytest.fixture(scope="module")
def mc(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("mc") / "libmathcheck.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "topos-codec_amd", "csrc", "kernels"), SRC, "-o", out, "-lm"], check=True)
    lib = ctypes.CDLL(out)
    lib.mc_rans_div.restype = ctypes.c_uint32
    lib.mc_qdiv_magic.restype = ctypes.c_uint64
    lib.mc_qdiv_apply.restype = ctypes.c_uint32
    lib.mc_qdiv_apply.argtypes = [ctypes.c_uint32, ctypes.c_uint64]
    lib.mc_crc_chunked.restype = ctypes.c_uint32
    lib.mc_gf2_mulmod.restype = ctypes.c_uint32
    lib.mc_enc_update_check.restype = ctypes.c_uint32
    lib.mc_zigzag_pos.restype = ctypes.c_uint32
    lib.mc_qdiv32_check.restype = ctypes.c_uint32
    lib.mc_qdiv32_all.restype = ctypes.c_uint32
    lib.mc_qdiv32_all.argtypes = [ctypes.c_uint32]
    lib.mc_ctx_tables_check.restype = ctypes.c_uint32
    lib.mc_dct8_check.restype = ctypes.c_uint32
    lib.mc_dct8_check.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64]
    return lib


def test_rans_division_exact(mc):
    rng = np.random.default_rng(1)
    for f in range(1, 4097):
        xmax = (1 << 19) * f  # the state is below 2^19 * f after renormalisation
        xs = [0, 1, f - 1, f, f + 1, xmax - 1, xmax - f, xmax - f - 1, (1 << 31) - 1 if xmax > (1 << 31) - 1 else xmax - 2]
        xs += rng.integers(0, min(xmax, 1 << 31), 24).tolist()
        for x in xs:
            x = int(min(x, xmax - 1, (1 << 31) - 1))
            assert mc.mc_rans_div(x, f) == x // f, (x, f)


def test_rans_encoder_update_exact(mc):
    # x' = (x / f) * 4096 + x % f + cum through the reciprocal constants the
    # encode kernel uses, for every frequency and states up to the bound
    assert mc.mc_enc_update_check(2000) == 0


def test_butterfly_transforms_exact(mc, oracle):
    # the kernels' even/odd 8-point transforms equal the direct matrix products
    # (the oracle's matrix), forward in int32 for inputs up to 2^22 and inverse
    # in int64 and f64 for inputs up to 2^45 (all sums stay below 2^53)
    m = np.ctypeslib.as_array((ctypes.c_int16 * 64).in_dll(oracle, "orc_M")).astype(np.int32)
    arr = (ctypes.c_int32 * 64)(*m.tolist())
    assert mc.mc_dct8_check(arr, 200000, (1 << 22), ctypes.c_int64(1 << 45)) == 0


def test_context_lookup_tables(mc):
    # packed-nibble context buckets of the decode kernel vs the bucket rules
    # of the V7-R2 conditional models (position, previous category, DC)
    assert mc.mc_ctx_tables_check() == 0


def test_zigzag_closed_form(mc):
    # the scan position the decode kernel computes per lane, against the
    # zigzag walk built step by step (right / down-left / down / up-right)
    order, r, c, up = [], 0, 0, True
    while len(order) < 64:
        order.append(8 * r + c)
        if up:
            if c == 7: r, up = r + 1, False
            elif r == 0: c, up = c + 1, False
            else: r, c = r - 1, c + 1
        else:
            if r == 7: c, up = c + 1, True
            elif c == 0: r, up = r + 1, True
            else: r, c = r + 1, c - 1
    inv = [0] * 64
    for i, n in enumerate(order):
        inv[n] = i
    assert order[:6] == [0, 1, 8, 16, 9, 2]
    assert [mc.mc_zigzag_pos(n) for n in range(64)] == inv


def test_quant_division_mulhi_exact(mc, oracle):
    # every step Q = max(1, (qm * qp_scale + 128) >> 8) the quantiser tables can
    # produce (5 matrices x 2 classes x 64 positions x 2 qp tables x 96 qp)
    qm = np.ctypeslib.as_array((ctypes.c_uint16 * (5 * 2 * 64)).in_dll(oracle, "orc_qm")).astype(np.int64)
    sc = np.ctypeslib.as_array((ctypes.c_uint32 * (2 * 96)).in_dll(oracle, "orc_qp_scale")).astype(np.int64)
    qs = np.unique(np.maximum(1, (qm[:, None] * sc[None, :] + 128) >> 8)).astype(np.uint32)
    extra = np.array([1, 2, 3, 7, (1 << 21) - 1, 1 << 21, (1 << 24) + 1, (1 << 25) - 3, (1 << 31) - 1, 1 << 31],
                     np.uint32)
    qs = np.unique(np.concatenate([qs, extra]))
    assert qs.max() > (1 << 24)
    arr = (ctypes.c_uint32 * len(qs))(*qs.tolist())
    assert mc.mc_qdiv32_check(arr, len(qs), 64) == 0
    # the transform kernel adds the bias Q/2 - dz (dz = Q/4 on AC) without
    # saturating: it is never negative for any step the tables produce
    q64 = qs.astype(np.int64)
    assert ((q64 >> 1) - q64 // 4 >= 0).all()


@pytest.mark.parametrize("q", [3, 1448, 27555])
def test_quant_division_mulhi_exhaustive(mc, q):
    # all 2^31 numerators for a few divisors (a few seconds each in C)
    assert mc.mc_qdiv32_all(q) == 0


def test_quant_division_exact(mc):
    rng = np.random.default_rng(2)
    divisors = [1, 2, 3, 5, 7, 16, 1448, 4096, 27555, 18045205] + rng.integers(1, 1 << 25, 200).tolist()
    for d in divisors:
        m = mc.mc_qdiv_magic(int(d))
        ns = [0, 1, d - 1, d, d + 1, (1 << 26) - 1, (1 << 25) + (d >> 1)] + rng.integers(0, 1 << 26, 50).tolist()
        for n in ns:
            n = int(n)
            if n * d >= (1 << 51):
                continue
            assert mc.mc_qdiv_apply(n, m) == n // d, (n, d)


def crc_tables():
    tab = []
    for i in range(256):
        c = i
        for _ in range(8):
            c = (c >> 1) ^ 0xEDB88320 if c & 1 else c >> 1
        tab.append(c)
    return tab


def test_crc_combine_matches_zlib(mc, pkg):
    lib = pkg.load_library()
    t = (ctypes.c_uint32 * 256)()
    x2n = (ctypes.c_uint32 * 32)()
    lib.tcx_crc_tables(t, x2n)
    assert list(t) == crc_tables()
    rng = np.random.default_rng(3)
    for n in [0, 1, 2, 63, 64, 65, 255, 1000, 4097, 80000]:
        data = rng.integers(0, 256, n, dtype=np.uint8)
        for parts in (1, 64, 256):
            got = mc.mc_crc_chunked(data.ctypes.data_as(ctypes.c_void_p), n, t, x2n, parts)
            assert got == zlib.crc32(data.tobytes()) & 0xFFFFFFFF, (n, parts)


class RansModel(ctypes.Structure):
    _fields_ = [("syms", ctypes.c_uint32), ("freq", ctypes.c_uint16 * 64), ("cum", ctypes.c_uint16 * 65)]


def test_row_tables_match_reference(mc, ref):
    rng = np.random.default_rng(4)
    for trial in range(400):
        nsym = [28, 29, 64][trial % 3]
        kind = trial % 5
        if kind == 0:
            cnt = np.zeros(nsym, np.uint32)
        elif kind == 1:
            cnt = rng.integers(0, 3, nsym).astype(np.uint32)
        elif kind == 2:
            cnt = (rng.pareto(1.0, nsym) * 50).astype(np.uint32)
        elif kind == 3:
            cnt = np.zeros(nsym, np.uint32); cnt[rng.integers(0, nsym)] = rng.integers(1, 10**6)
        else:
            cnt = rng.integers(0, 10**6, nsym).astype(np.uint32)
            cnt[rng.random(nsym) < 0.5] = 0
        c = (ctypes.c_uint32 * nsym)(*cnt.tolist())
        a, b = (ctypes.c_uint8 * 64)(), (ctypes.c_uint8 * 64)()
        mc.mc_row_bytes(a, nsym, c)
        ref.tc_rans2_row_encode(b, nsym, c)
        assert bytes(a)[:nsym] == bytes(b)[:nsym]
        f, cu = (ctypes.c_uint16 * 64)(), (ctypes.c_uint16 * 65)()
        rc_a = mc.mc_model(a, nsym, f, cu)
        m = RansModel()
        rc_b = ref.tc_rans2_row_decode(b, nsym, ctypes.byref(m))
        assert (rc_a == 0) == (rc_b == 0)
        if rc_a == 0:
            assert list(f)[:nsym] == list(m.freq)[:nsym]
            assert list(cu)[:nsym + 1] == list(m.cum)[:nsym + 1]


def test_ctx_log2_matches_reference(pkg, ref):
    lib = pkg.load_library()
    lib.tcx_ctx_log2.restype = ctypes.c_double
    lib.tcx_ctx_log2.argtypes = [ctypes.c_double]
    ref.tc_ctx_log2.restype = ctypes.c_double
    ref.tc_ctx_log2.argtypes = [ctypes.c_double]
    for f in range(1, 4097):
        a, b = lib.tcx_ctx_log2(float(f)), ref.tc_ctx_log2(float(f))
        assert a.hex() == b.hex(), f
    # closed form sanity: exact at powers of two, close to log2 elsewhere
    import math
    for k in range(13):
        assert lib.tcx_ctx_log2(float(1 << k)) == float(k)
    for f in range(1, 4097, 37):
        assert abs(lib.tcx_ctx_log2(float(f)) - math.log2(f)) < 1e-7  # series to z^13


def _round_f32(x):
    """Nearest float32 (ties to even) of an exact rational; normal range only."""
    from fractions import Fraction
    if x == 0:
        return np.float32(0.0)
    sign = -1 if x < 0 else 1
    x = abs(x)
    e = x.numerator.bit_length() - x.denominator.bit_length()
    while Fraction(2) ** e > x:
        e -= 1
    while Fraction(2) ** (e + 1) <= x:
        e += 1
    assert -126 <= e <= 127
    m = x / Fraction(2) ** (e - 23)
    n = m.numerator // m.denominator
    rem = m - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(sign * n * 2.0 ** (e - 23))


@pytest.mark.parametrize("d", [255, 65535])
def test_conversion_normalisation_exact(d):
    """The conversion kernel's c / d for integer samples (tc_convert.hip cvt_norm):
    q0 = c * r, res = fma(-q0, d, c), q = fma(res, r, q0) with r = 1 / d
    correctly rounded, each fma rounded once, equals the IEEE quotient for every
    c in 0..d."""
    from fractions import Fraction
    df = np.float32(d)
    r = np.float32(1.0) / df
    c = np.arange(d + 1, dtype=np.float32)
    want = c / df
    q0 = c * r
    fr, fd = Fraction(float(r)), Fraction(d)
    for i in range(d + 1):
        ci, q = Fraction(int(c[i])), Fraction(float(q0[i]))
        res = _round_f32(ci - q * fd)
        got = _round_f32(Fraction(float(res)) * fr + q)
        assert got == want[i], (d, i)


@pytest.mark.parametrize("kr,kb", [(0.2126, 0.0722), (0.299, 0.114), (0.2627, 0.0593)])
def test_conversion_chroma_division_exact(kr, kb):
    """The conversion kernel's chroma division (tc_convert.hip cvt_div): the
    binary64 product a * RN64(1 / d), rounded once to binary32, equals the IEEE
    binary32 quotient a / d for d = 2 (1 - kb) and 2 (1 - kr) as the host
    rounds them. Numerators: random binary32 values over every binade from
    2^-40 to 1 (both signs), the (b - y) and (r - y) of random 8- and 16-bit
    colours in the kernel's operation order, and the binary32 neighbours of
    m d for quotients m just below, at and above binary32 midpoints (the
    hardest cases). The binary32 reciprocal product fails on this set."""
    rng = np.random.default_rng(5)
    n = 1 << 20
    nums = [np.ldexp(rng.uniform(1.0, 2.0, n), rng.integers(-40, 0, n)) * rng.choice([-1.0, 1.0], n)]
    for norm in (255.0, 65535.0):
        c = (rng.integers(0, int(norm) + 1, (3, n)).astype(np.float32) / np.float32(norm)).astype(np.float32)
        rf, gf, bf = c
        kf = np.float32(1.0 - kr - kb)
        y = (np.float32(kr) * rf + kf * gf).astype(np.float32)
        y = (y + np.float32(kb) * bf).astype(np.float32)
        nums += [bf - y, rf - y]
    for dd in (2.0 * (1.0 - kb), 2.0 * (1.0 - kr)):
        d = np.float32(dd)
        # quotients near midpoints: m = (2j + 1) 2^(e - 24) in [2^e, 2^(e+1))
        e = rng.integers(-20, 0, n)
        j = rng.integers(1 << 22, 1 << 23, n)
        m = np.ldexp((2 * j + 1).astype(np.float64), e - 24)
        a = (m * np.float64(d)).astype(np.float32)
        nums += [a, np.nextafter(a, np.float32(-2)), np.nextafter(a, np.float32(2))]
    a_all = np.concatenate([np.asarray(v, np.float32) for v in nums])
    for dd in (2.0 * (1.0 - kb), 2.0 * (1.0 - kr)):
        d = np.float32(dd)
        want = a_all / d
        rd = 1.0 / np.float64(d)
        got = (a_all.astype(np.float64) * rd).astype(np.float32)
        assert np.array_equal(got, want), int(np.count_nonzero(got != want))
        assert np.count_nonzero(a_all * (np.float32(1.0) / d) != want) > 1000


def _idct8_f64(v):
    """The kernels' f64 inverse butterfly (tc_math.h TC_IDCT8_BODY), in the same
    operation order, on Python floats (IEEE double)."""
    p, q = 13.0 * v[0] + 9.0 * v[4], 13.0 * v[0] - 9.0 * v[4]
    r, t = 12.0 * v[2] + 5.0 * v[6], 5.0 * v[2] - 12.0 * v[6]
    e0, e3, e1, e2 = p + r, p - r, q + t, q - t
    o0 = 13.0 * v[1] + 11.0 * v[3] + 8.0 * v[5] + 1.0 * v[7]
    o1 = 11.0 * v[1] - 4.0 * v[3] - 13.0 * v[5] - 7.0 * v[7]
    o2 = 7.0 * v[1] - 13.0 * v[3] + 4.0 * v[5] + 11.0 * v[7]
    o3 = 1.0 * v[1] - 8.0 * v[3] + 11.0 * v[5] - 13.0 * v[7]
    return [e0 + o0, e1 + o1, e2 + o2, e3 + o3, e3 - o3, e2 - o2, e1 - o1, e0 - o0]


def _ref_round(v):
    # reference sign-split rounding of an exact integer (transform.c:69-117)
    return (v + (1 << 31)) >> 32 if v >= 0 else -((-v + (1 << 31)) >> 32)


def test_f64_inverse_transform_worst_case(oracle):
    """Dequantised coefficients at the clamp (|G| = f_clamp W, f_clamp = 2^25
    for bit depths <= 12) with the signs that maximise each output sample:
    every f64 step of the kernel's dequant / IDCT / rounding (columns first,
    then rows) stays an exact integer below 2^53, so it equals the exact
    integer transform and the reference rounding (the kernel's one-FMA
    rounding is checked too)."""
    import math
    M = np.ctypeslib.as_array((ctypes.c_int16 * 64).in_dll(oracle, "orc_M")).astype(np.int64).reshape(8, 8)
    W = np.ctypeslib.as_array((ctypes.c_uint32 * 64).in_dll(oracle, "orc_W")).astype(np.int64).reshape(8, 8)
    fc = 1 << 25
    worst = 0.0
    for y in range(8):
        for x in range(8):
            su = np.sign(M[:, y]).astype(np.int64)
            sv = np.sign(M[:, x]).astype(np.int64)
            G = (su[:, None] * sv[None, :]) * fc * W          # exact integers (< 2^39)
            # the kernel's order: pass 1 (columns v): T[y'][v] = sum_u M[u][y'] G[u][v],
            # pass 2 (rows y'): X[y'][x'] = sum_v T[y'][v] M[v][x']
            Tc = [_idct8_f64([float(G[u][v]) for u in range(8)]) for v in range(8)]   # Tc[v][y']
            B = [[Tc[v][yy] for v in range(8)] for yy in range(8)]                    # B[y'][v]
            rows = [_idct8_f64(B[yy]) for yy in range(8)]                             # rows[y'][x']
            exact = M.T @ G @ M                                   # X[y][x] = sum_uv M[u][y] G[u][v] M[v][x]
            for yy in range(8):
                for xx in range(8):
                    v = rows[yy][xx]
                    worst = max(worst, abs(v))
                    assert v == float(exact[yy, xx]) and abs(v) < 2.0 ** 53
                    r = int(math.floor(abs(v) * 2.0 ** -32 + 0.5))  # the kernel's FMA form (exact, unfused too)
                    assert (r if v >= 0 else -r) == _ref_round(int(exact[yy, xx]))
            assert all(abs(b) < 2.0 ** 53 for row in B for b in row)
    assert worst > 2.0 ** 45    # the test reaches the range the argument is about


def _fdct8_multiplicands(v):
    """The butterfly form of the forward 8-point transform (k_fdct_quant's
    fdct8_i24, tc_fdct8_i32) over symbolic inputs: each value is its integer
    coefficient vector over the 8 inputs. Returns (multiplicands, outputs)."""
    s = [v[k] + v[7 - k] for k in range(4)]
    d = [v[k] - v[7 - k] for k in range(4)]
    e0, e1, f0, f1 = s[0] + s[3], s[1] + s[2], s[0] - s[3], s[1] - s[2]
    mult = [e0 + e1, e0 - e1, f0, f1, d[0], d[1], d[2], d[3]]
    out = [13 * (e0 + e1), 13 * d[0] + 11 * d[1] + 7 * d[2] + d[3], 12 * f0 + 5 * f1,
           11 * d[0] - 4 * d[1] - 13 * d[2] - 8 * d[3], 9 * (e0 - e1), 8 * d[0] - 13 * d[1] + 4 * d[2] + 11 * d[3],
           5 * f0 - 12 * f1, d[0] - 7 * d[1] + 11 * d[2] - 13 * d[3]]
    return mult, out


def test_fdct_24bit_multiplies_exact(oracle):
    """k_fdct_quant multiplies with 24-bit operands (v_mul_i32_i24) for bit
    depths <= 12: every multiplicand must stay below 2^23 in magnitude and
    every product and output within 32 bits, for the worst-case inputs. Level-
    shifted 12-bit samples lie in [-2048, 2047]; a linear form's largest
    magnitude over that box is sum |coef| * 2048. The column pass's inputs are
    the row pass's outputs of one frequency v, each bounded alike."""
    eye = [np.eye(8, dtype=np.int64)[k] for k in range(8)]
    mult, out = _fdct8_multiplicands(eye)
    M = np.ctypeslib.as_array((ctypes.c_int16 * 64).in_dll(oracle, "orc_M")).astype(np.int64).reshape(8, 8)
    assert all((np.array(out) == M).ravel())          # the butterfly is the codec's matrix M
    xmax = 2048
    row_mult = max(int(np.abs(m).sum()) * xmax for m in mult)
    row_out = [int(np.abs(o).sum()) * xmax for o in out]          # bound of the row output of frequency v
    assert row_mult < 2 ** 23 and max(row_out) < 2 ** 23
    col_mult = max(int(np.abs(m).sum()) * b for m in mult for b in row_out)
    col_prod = max(13 * int(np.abs(m).sum()) * b for m in mult for b in row_out)
    col_out = max(int(np.abs(o).sum()) * b for o in out for b in row_out)
    assert col_mult < 2 ** 23, col_mult
    assert col_prod < 2 ** 31 and col_out < 2 ** 31
    # 16-bit samples (the int32 path) would not fit: the bound is tight enough to matter
    assert max(int(np.abs(m).sum()) * b * 16 for m in mult for b in row_out) >= 2 ** 23
