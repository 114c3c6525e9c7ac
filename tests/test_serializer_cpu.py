"""Frame I/O in the reference's serializer format, on the CPU (host code of
the library; no device needed).

The text is checked byte for byte against the layout of SaveSphParticleSet /
PushParticleSetToFile (src/third/serializer.cpp:822-911), numbers printed
with C's %g (Python's '%g' is the same conversion).  The number parser is
pinned on values whose result the reference's digit-accumulating ParseFloat
(src/third/obj_loader.cpp:55-170) fixes: exact dyadic values, its documented
failures, and "0.3", which that parser reads as 3 * 0.1 =
0.30000000000000004 where a correctly rounded parser gives 0.3.
"""
import math

import numpy as np
import pytest

from bubbles_amd import abi


def expected_block(flags, spacing, pos=None, vel=None, dens=None, mass=0.0, boundary=None):
    letters = "".join(c for bit, c in [(abi.SER_POSITION, "p"), (abi.SER_VELOCITY, "v"),
                                        (abi.SER_DENSITY, "d"), (abi.SER_MASS, "m"),
                                        (abi.SER_BOUNDARY, "b"), (abi.SER_BOUNDARY_EXCLUSIVE, "o")]
                      if flags & bit)
    n = len(next(a for a in (pos, vel, dens, boundary) if a is not None))
    keep = [i for i in range(n)
            if not ((flags & abi.SER_BOUNDARY_EXCLUSIVE) and boundary is not None and not boundary[i])]
    lines = ["FluidBegin", '\t"Type" particles', f'\t"Count" {len(keep)}', f'\t"Format" {letters}',
             "\t\"Spacing\" %g" % spacing, "\tDataBegin"]
    for i in keep:
        f = []
        if flags & abi.SER_POSITION:
            f += ["%g" % v for v in pos[i]]
        if flags & abi.SER_VELOCITY:
            f += ["%g" % v for v in vel[i]]
        if flags & abi.SER_DENSITY:
            f.append("%g" % dens[i])
        if flags & abi.SER_MASS:
            f.append("%g" % mass)
        if flags & abi.SER_BOUNDARY:
            f.append("%d" % boundary[i])
        lines.append("\t\t" + " ".join(f))
    lines += ["\tDataEnd", "FluidEnd"]
    return "\n".join(lines) + "\n"


def test_writer_matches_reference_layout(tmp_path):
    rng = np.random.default_rng(0)
    n = 50
    pos = rng.normal(size=(n, 3)) * [1e-3, 1.0, 1e4]
    vel = rng.normal(size=(n, 3))
    dens = 1000.0 + rng.normal(size=n)
    pos[0] = [0.0, -0.0, 1e-5]
    path = tmp_path / "frame.txt"
    flags = abi.SER_POSITION | abi.SER_VELOCITY | abi.SER_DENSITY | abi.SER_MASS
    abi.serializer_write(path, flags, 0.02, pos, vel, dens, mass=1.234e-3)
    assert path.read_text() == expected_block(flags, 0.02, pos, vel, dens, 1.234e-3)


def test_writer_appends_and_reader_takes_first_block(tmp_path):
    path = tmp_path / "two.txt"
    a = np.array([[0.5, 0.25, 0.125], [1.0, 2.0, 3.0]])
    b = np.array([[9.0, 9.0, 9.0]])
    abi.serializer_write(path, abi.SER_POSITION, 0.01, a)
    abi.serializer_write(path, abi.SER_POSITION, 0.01, b)
    text = path.read_text()
    assert text == expected_block(abi.SER_POSITION, 0.01, a) + expected_block(abi.SER_POSITION, 0.01, b)
    r = abi.serializer_read(path)
    assert r["flags"] == abi.SER_POSITION
    np.testing.assert_array_equal(r["position"], a)  # dyadic values round-trip exactly


def test_boundary_exclusive_rule(tmp_path):
    pos = np.arange(12, dtype=np.float64).reshape(4, 3)
    bnd = np.array([1, 0, 2, 0], dtype=np.int32)
    flags = abi.SER_POSITION | abi.SER_BOUNDARY | abi.SER_BOUNDARY_EXCLUSIVE
    path = tmp_path / "b.txt"
    abi.serializer_write(path, flags, 0.02, pos, boundary=bnd)
    assert path.read_text() == expected_block(flags, 0.02, pos, boundary=bnd)
    r = abi.serializer_read(path)
    np.testing.assert_array_equal(r["position"], pos[[0, 2]])
    np.testing.assert_array_equal(r["boundary"], [1, 2])


def test_large_frame_is_identical_to_serial_layout(tmp_path):
    # above the size where formatting is split over host threads
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, 2, size=(30011, 3))
    path = tmp_path / "big.txt"
    abi.serializer_write(path, abi.SER_POSITION, 0.01, pos)
    assert path.read_text() == expected_block(abi.SER_POSITION, 0.01, pos)


@pytest.mark.parametrize("text,value", [
    ("0.5", 0.5), ("-2", -2.0), ("+3.25", 3.25), ("1e3", 1000.0), ("2.5E+1", 25.0),
    ("12abc", 12.0), ("7.", 7.0), ("0", 0.0),
    # failures of the reference parser read as 0
    (".5", 0.0), ("1e", 0.0), ("1e+", 0.0), ("abc", 0.0), ("-", 0.0), ("nan", 0.0),
])
def test_parse_number_known_answers(text, value):
    got = abi.serializer_parse_number(text)
    assert got == value and math.copysign(1.0, got) == math.copysign(1.0, value)


def test_parse_number_accumulates_digits_like_the_reference():
    # fraction digits are added one by one, d * 10^-k: "0.3" -> 3 * 0.1
    assert abi.serializer_parse_number("0.3") == 3 * 0.1 != 0.3
    assert abi.serializer_parse_number("0.125") == ((0.1 + 2 * 0.01) + 5 * 0.001)
    # a negative exponent goes through m * 5^e * 2^e
    assert abi.serializer_parse_number("1e-3") == math.ldexp(5.0 ** -3, -3)


def test_parse_number_is_within_an_ulp_of_correct_rounding():
    rng = np.random.default_rng(2)
    for v in rng.normal(size=2000) * 10.0 ** rng.integers(-6, 6, size=2000):
        t = "%g" % v
        got = abi.serializer_parse_number(t)
        ref = float(t)
        assert abs(got - ref) <= 2 * math.ulp(ref), (t, got, ref)


def test_reader_skips_comments_blank_lines_and_crlf(tmp_path):
    path = tmp_path / "c.txt"
    path.write_bytes(b'FluidBegin\r\n\t"Type" particles\r\n\t"Count" 2\r\n\t"Format" pv\r\n'
                     b'\tDataBegin\r\n# comment\r\n\r\n\t\t1 2 3 4 5 6\r\n\t\t0.5 0.25 0 -1 -2 -3\r\n'
                     b'\tDataEnd\r\nFluidEnd\r\n')
    r = abi.serializer_read(path)
    assert r["flags"] == abi.SER_POSITION | abi.SER_VELOCITY
    np.testing.assert_array_equal(r["position"], [[1, 2, 3], [0.5, 0.25, 0]])
    np.testing.assert_array_equal(r["velocity"], [[4, 5, 6], [-1, -2, -3]])


def test_reader_errors(tmp_path):
    p1 = tmp_path / "noend.txt"
    p1.write_text('FluidBegin\n\t"Count" 1\n\t"Format" p\n\tDataBegin\n\t\t1 2 3\n')
    with pytest.raises(abi.BbxError):
        abi.serializer_read(p1)
    p2 = tmp_path / "badfmt.txt"
    p2.write_text('FluidBegin\n\t"Format" pq\n\tDataBegin\n\tDataEnd\nFluidEnd\n')
    with pytest.raises(abi.BbxError):
        abi.serializer_read(p2)
    with pytest.raises(abi.BbxError):
        abi.serializer_write(tmp_path / "n.txt", abi.SER_NORMAL, 0.02, np.zeros((1, 3)))
