"""The oracle's SEG-Y sample decode against what the SEG-Y standard fixes:
IBM System/360 floats with known values (100, -118.625, 1, 0.5, zero, the
largest and the smallest normalized word), every IBM word of a random sample
against exact rational arithmetic (fractions) rounded once, and the integer
and IEEE formats against Python's own big-endian unpacking.  parseSEG, which
decodes on the reference's side, is not vendored: the decode is pinned by
the standard's definitions, not by parseSEG's outputs."""
import math
import struct
from fractions import Fraction

import numpy as np

import oracle as O


def ibm_exact(w: int) -> Fraction:
    s, e, f = w >> 31, (w >> 24) & 0x7F, w & 0xFFFFFF
    v = Fraction(f, 1 << 24) * Fraction(16) ** (e - 64)
    return -v if s else v


def words(ws, fmt="u4"):
    return np.array(ws, dtype=">" + fmt).view(np.uint8)


def test_ibm_known_values():
    ws = [0x42640000, 0xC276A000, 0x41100000, 0x40800000, 0x00000000, 0x80000000,
          0x7FFFFFFF, 0x00100000]
    y64 = O.decode_samples(words(ws), 1, len(ws), 1, offset=0, dtype=np.float64)[0]
    assert list(y64[:4]) == [100.0, -118.625, 1.0, 0.5]
    assert y64[4] == 0 and y64[5] == 0 and math.copysign(1, y64[5]) == -1
    assert y64[6] == float(ibm_exact(0x7FFFFFFF)) and 7.2e75 < y64[6] < 7.3e75
    assert y64[7] == 16.0 ** -65
    y32 = O.decode_samples(words(ws), 1, len(ws), 1, offset=0)[0]
    assert list(y32[:4]) == [100.0, -118.625, 1.0, 0.5]
    assert y32[6] == np.inf and y32[7] == 0, "beyond float32: Inf, and 0 below its subnormals"


def test_ibm_random_words_exact():
    w = np.random.default_rng(5).integers(0, 1 << 32, 4000, dtype=np.uint64).astype(np.uint32)
    raw = w.astype(">u4").view(np.uint8)
    y64 = O.decode_samples(raw, 1, w.size, 1, offset=0, dtype=np.float64)[0]
    y32 = O.decode_samples(raw, 1, w.size, 1, offset=0, dtype=np.float32)[0]
    for i, wi in enumerate(w.tolist()):
        ex = float(ibm_exact(wi))          # exact: the value fits a float64
        assert y64[i] == ex
        with np.errstate(over="ignore"):
            assert y32[i] == np.float32(ex) or (np.isinf(y32[i]) and abs(ex) > 3.4e38)


def test_integer_and_ieee_formats():
    rng = np.random.default_rng(9)
    i32 = rng.integers(-2 ** 31, 2 ** 31, 500, dtype=np.int64)
    i16 = rng.integers(-2 ** 15, 2 ** 15, 500, dtype=np.int64)
    i8 = rng.integers(-128, 128, 500, dtype=np.int64)
    f32 = rng.standard_normal(500).astype(np.float32) * 1e3
    cases = [(2, struct.pack(">500i", *i32.tolist()), i32),
             (3, struct.pack(">500h", *i16.tolist()), i16),
             (8, struct.pack(">500b", *i8.tolist()), i8),
             (5, struct.pack(">500f", *f32.tolist()), f32.astype(np.float64))]
    for fmt, b, want in cases:
        raw = np.frombuffer(b, np.uint8)
        got = O.decode_samples(raw, fmt, 500, 1, offset=0, dtype=np.float64)[0]
        assert np.array_equal(got, want.astype(np.float64)), fmt
        got32 = O.decode_samples(raw, fmt, 500, 1, offset=0)[0]
        assert np.array_equal(got32, want.astype(np.float32)), fmt


def test_trace_layout():
    """Traces of a 240-byte header then the samples: the headers are skipped."""
    rows, ns = 3, 5
    vals = np.arange(rows * ns, dtype=np.int64).reshape(rows, ns) - 7
    blob = b"".join(bytes(range(240)) + struct.pack(f">{ns}h", *r.tolist()) for r in vals)
    got = O.decode_samples(np.frombuffer(blob, np.uint8), 3, ns, rows)
    assert np.array_equal(got, vals.astype(np.float32))


def test_segd_formats():
    """SEG-D demultiplexed formats against Python's big-endian unpacking:
    8036 24-bit two's complement (the rails +8388607 and -8388608, -1,
    random words), 8038 int32, 8058 IEEE float32, 8080 IEEE float64 (NaN
    payloads and infinities kept in float64, rounded once to float32)."""
    rng = np.random.default_rng(11)
    i24 = np.concatenate([[8388607, -8388608, -1, 0, 1],
                          rng.integers(-2 ** 23, 2 ** 23, 495, dtype=np.int64)])
    b24 = b"".join(int(v).to_bytes(3, "big", signed=True) for v in i24)
    got = O.decode_samples(np.frombuffer(b24, np.uint8), 8036, 500, 1, offset=0,
                           dtype=np.float64)[0]
    assert np.array_equal(got, i24.astype(np.float64))
    got32 = O.decode_samples(np.frombuffer(b24, np.uint8), 8036, 500, 1, offset=0)[0]
    assert np.array_equal(got32.astype(np.float64), i24.astype(np.float64))  # exact in float32
    i32 = rng.integers(-2 ** 31, 2 ** 31, 500, dtype=np.int64)
    got = O.decode_samples(np.frombuffer(struct.pack(">500i", *i32.tolist()), np.uint8), 8038,
                           500, 1, offset=0, dtype=np.float64)[0]
    assert np.array_equal(got, i32.astype(np.float64))
    f32 = rng.standard_normal(500).astype(np.float32)
    got = O.decode_samples(np.frombuffer(struct.pack(">500f", *f32.tolist()), np.uint8), 8058,
                           500, 1, offset=0)[0]
    assert np.array_equal(got, f32)
    f64 = np.concatenate([[math.inf, -math.inf, 1e300, -1e-300, 0.0],
                          rng.standard_normal(495) * 1e3])
    raw64 = np.frombuffer(struct.pack(">500d", *f64.tolist()), np.uint8)
    got = O.decode_samples(raw64, 8080, 500, 1, offset=0, dtype=np.float64)[0]
    assert np.array_equal(got, f64)
    with np.errstate(over="ignore"):
        assert np.array_equal(O.decode_samples(raw64, 8080, 500, 1, offset=0)[0],
                              f64.astype(np.float32))
