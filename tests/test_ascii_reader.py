"""capow_read_ascii (the host's parallel catalogue reader) against a
restatement of the reference's serial loop (src/read_data.c:32-88):
fgets with a 2000-character buffer (so a physical line is read in pieces of
at most 1999 characters, the last one keeping the newline), pieces starting
with '#' skipped, sscanf("%lf" x 6) per piece (the longest numeric prefix of
each field, extra columns ignored), v_los = column 4/5/6 for LOS 1/2/else,
and for LOS = 0 the first particle outside [MIN, MAX] is fatal.  Rows must be
bitwise equal for every thread count, and errors must name the first failing
piece in file order.  No GPU needed."""
import re

import numpy as np
import pytest

from capow_amd import api

_NUM = re.compile(rb"[ \t\n\v\f\r]*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)")


def reference_read(path, los=0, lo=(0.0, 0.0, 0.0), hi=(1000.0, 1000.0, 1000.0)):
    """(rows, error): the reference's loop on decimal data."""
    data = open(path, "rb").read()
    rows = []
    pos = 0
    while pos < len(data):
        end = data.find(b"\n", pos)
        stop = len(data) if end < 0 else end + 1
        piece = data[pos:min(stop, pos + 1999)]
        pos += len(piece)
        if piece.startswith(b"#"):
            continue
        vals = []
        s = 0
        for _ in range(6):
            m = _NUM.match(piece, s)
            if not m:
                break
            vals.append(float(m.group(1)))
            s = m.end()
        if len(vals) < 6:
            return rows, ("parse", piece)
        if los == 0 and any(vals[a] < lo[a] or vals[a] > hi[a] for a in range(3)):
            return rows, ("box", tuple(vals[:3]))
        tvr = vals[3] if los == 1 else (vals[4] if los == 2 else vals[5])
        rows.append((vals[0], vals[1], vals[2], tvr))
    return rows, None


def _catalogue(path, n, seed, comments=True):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, 1000, size=(n, 3))
    vel = rng.normal(0, 300, size=(n, 3))
    lines = []
    for i in range(n):
        if comments and i % 97 == 0:
            lines.append("# comment %d 1 2 3" % i)
        x, y, z = pos[i]
        vx, vy, vz = vel[i]
        form = i % 5
        if form == 0:
            lines.append("%.17g %.17g %.17g %.17g %.17g %.17g" % (x, y, z, vx, vy, vz))
        elif form == 1:
            lines.append("  %.6e\t%.6e %.6e %.3f %.3f %.3f  extra 7 8" % (x, y, z, vx, vy, vz))
        elif form == 2:
            lines.append("%.9f %.9f %.9f %+.5E %+.5E %+.5E" % (x, y, z, vx, vy, vz))
        elif form == 3:
            lines.append("%.3f %.3f %.3f %.1f %.1f %.1f 9abc" % (x, y, z, vx, vy, vz))
        else:
            lines.append("%d %d %d %d %d %d" % (x, y, z, vx, vy, vz))
    open(path, "w").write("\n".join(lines) + "\n")


@pytest.mark.parametrize("los", [0, 1, 2, 3])
def test_rows_match_reference_loop_for_every_thread_count(tmp_path, los):
    path = str(tmp_path / "cat.txt")
    _catalogue(path, 120000, seed=los + 1)   # ~6 MB: several 1 MB thread ranges
    ref, err = reference_read(path, los)
    assert err is None
    ref = np.array(ref)
    for nt in (1, 2, 3, 7, 16, 0):
        got = api.read_ascii(path, los=los, nthreads=nt)
        assert got.shape == ref.shape
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), nt
    got32 = api.read_ascii(path, los=los, precision=api.FP32, nthreads=5)
    assert np.array_equal(got32.view(np.uint32), ref.astype(np.float32).view(np.uint32))


def test_first_malformed_line_in_file_order(tmp_path):
    path = str(tmp_path / "bad.txt")
    _catalogue(path, 150000, seed=9, comments=False)
    lines = open(path).read().split("\n")
    lines[100000] = "1 2 3 4 5"            # five numbers
    lines[140000] = "junk"
    lines[30] = "# a comment is fine"
    open(path, "w").write("\n".join(lines))
    _, err = reference_read(path)
    assert err == ("parse", b"1 2 3 4 5\n")
    for nt in (1, 4, 16):
        with pytest.raises(api.CapowError) as e:
            api.read_ascii(path, nthreads=nt)
        assert e.value.code == api.CAPOW_ERR_PARSE
        assert str(e.value).endswith("error reading file: 1 2 3 4 5\n")


def test_blank_line_is_an_error(tmp_path):
    path = str(tmp_path / "blank.txt")
    open(path, "w").write("1 2 3 4 5 6\n\n7 8 9 1 2 3\n")
    assert reference_read(path)[1] == ("parse", b"\n")
    with pytest.raises(api.CapowError) as e:
        api.read_ascii(path, nthreads=2)
    assert e.value.code == api.CAPOW_ERR_PARSE


def test_first_particle_out_of_box_for_los0_only(tmp_path):
    path = str(tmp_path / "oob.txt")
    _catalogue(path, 90000, seed=4, comments=False)
    lines = open(path).read().split("\n")
    lines[70000] = "1000.5 1 1 0 0 0"
    lines[20000] = "5 -0.25 5 0 0 0"
    open(path, "w").write("\n".join(lines))
    _, err = reference_read(path, 0)
    assert err == ("box", (5.0, -0.25, 5.0))
    for nt in (1, 8):
        with pytest.raises(api.CapowError) as e:
            api.read_ascii(path, los=0, nthreads=nt)
        assert e.value.code == api.CAPOW_ERR_OUT_OF_BOX
        assert "x=5.000000, y=-0.250000, z=5.000000" in str(e.value)
    # other lines of sight keep such particles (the deposit drops them)
    ref, err = reference_read(path, 3)
    assert err is None
    got = api.read_ascii(path, los=3, nthreads=8)
    assert np.array_equal(got, np.array(ref))


def test_long_lines_are_read_in_fgets_pieces(tmp_path):
    path = str(tmp_path / "long.txt")
    # 6 numbers, then spaces past the 1999-character piece: the second piece
    # holds no number -> error, as fgets + sscanf in the reference
    open(path, "w").write("1 2 3 4 5 6" + " " * 2500 + "\n")
    assert reference_read(path)[1][0] == "parse"
    with pytest.raises(api.CapowError) as e:
        api.read_ascii(path)
    assert e.value.code == api.CAPOW_ERR_PARSE
    # a line of exactly 1999 characters leaves its newline as a piece of its own
    line = "1 2 3 4 5 6"
    line = line + " " * (1999 - len(line))
    open(path, "w").write(line + "\n")
    assert reference_read(path)[1] == ("parse", b"\n")
    with pytest.raises(api.CapowError):
        api.read_ascii(path)
    # one character shorter fits in one piece with its newline
    open(path, "w").write(line[:-1] + "\n" + "7 8 9 1 2 3")
    ref, err = reference_read(path)
    assert err is None and len(ref) == 2
    got = api.read_ascii(path, nthreads=3)
    assert np.array_equal(got, np.array(ref))
    # a comment longer than a piece: its second piece is parsed as data
    open(path, "w").write("#" + "1 2 3 4 5 6 " * 200 + "\n")
    ref, err = reference_read(path)
    got = api.read_ascii(path)
    assert err is None and np.array_equal(got, np.array(ref))


def test_empty_missing_and_unterminated(tmp_path):
    path = str(tmp_path / "empty.txt")
    open(path, "w").write("")
    assert api.read_ascii(path).shape == (0, 4)
    with pytest.raises(api.CapowError) as e:
        api.read_ascii(str(tmp_path / "missing.txt"))
    assert e.value.code == api.CAPOW_ERR_IO
    open(path, "w").write("# only comments\n# more\n")
    assert api.read_ascii(path, nthreads=4).shape == (0, 4)
    open(path, "w").write("1 2 3 4 5 6")      # no final newline
    assert np.array_equal(api.read_ascii(path), np.array([[1.0, 2.0, 3.0, 6.0]]))
