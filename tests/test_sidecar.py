"""The binary sidecar of the periodic catalogue (capow_write_sidecar /
capow_read_sidecar): reading it must give what capow_read_ascii gives on the
text -- rows bitwise equal for every LOS and both precisions, the same first
particle outside the box for LOS = 0 -- checked against the restatement of
the reference's serial loop (src/read_data.c:32-88) in test_ascii_reader;
a sidecar whose text changed is refused as stale, a damaged one as not a
sidecar.  No counterpart in the reference (it parses the text every run).
No GPU needed."""
import os
import struct

import numpy as np
import pytest

from capow_amd import api
from tests.test_ascii_reader import _catalogue, reference_read


@pytest.mark.parametrize("los", [0, 1, 2, 3])
def test_sidecar_rows_equal_text_rows(tmp_path, los):
    text = str(tmp_path / "cat.txt")
    side = str(tmp_path / "cat.capowbin")
    _catalogue(text, 60000, seed=40 + los)
    ref, err = reference_read(text, los)
    assert err is None
    ref = np.array(ref)
    api.write_sidecar(text, side, nthreads=3)
    for nt in (1, 4, 0):
        got = api.read_sidecar(side, text, los=los, nthreads=nt)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), nt
    got32 = api.read_sidecar(side, text, los=los, precision=api.FP32)
    txt32 = api.read_ascii(text, los=los, precision=api.FP32)
    assert np.array_equal(got32.view(np.uint32), txt32.view(np.uint32))
    # header: magic, version, columns, rows; then 6 doubles per line
    raw = open(side, "rb").read()
    assert raw[:8] == b"CAPOWCAT"
    version, ncols, nrows = struct.unpack_from("<IIQ", raw, 8)
    assert (version, ncols, nrows) == (1, 6, len(ref))
    assert len(raw) == 64 + 48 * nrows


def test_sidecar_first_particle_out_of_box(tmp_path):
    text = str(tmp_path / "cat.txt")
    side = str(tmp_path / "cat.capowbin")
    lines = ["%d %d %d 1 2 3" % (10 * i, 20, 30) for i in range(50)]
    lines[17] = "500 2000.5 30 1 2 3"     # y above MAX
    lines[31] = "-1 20 30 1 2 3"          # x below MIN
    open(text, "w").write("\n".join(lines) + "\n")
    ref, err = reference_read(text, 0)
    assert err == ("box", (500.0, 2000.5, 30.0))
    api.write_sidecar(text, side)
    with pytest.raises(api.CapowError) as e:
        api.read_sidecar(side, text, los=0)
    assert e.value.code == api.CAPOW_ERR_OUT_OF_BOX
    assert "x=500.000000, y=2000.500000, z=30.000000" in str(e.value)
    # a line of sight other than 0 keeps every particle
    assert api.read_sidecar(side, text, los=3).shape == (50, 4)


def test_sidecar_stale_damaged_and_empty(tmp_path):
    text = str(tmp_path / "cat.txt")
    side = str(tmp_path / "cat.capowbin")
    _catalogue(text, 2000, seed=9)
    api.write_sidecar(text, side)
    assert api.read_sidecar(side, text).shape == (2000, 4)
    # the text rewritten (size and time change): refused; without the text
    # path the sidecar is read as it is
    with open(text, "a") as f:
        f.write("1 2 3 4 5 6\n")
    with pytest.raises(api.CapowError) as e:
        api.read_sidecar(side, text)
    assert e.value.code == api.CAPOW_ERR_STALE
    assert api.read_sidecar(side).shape[0] == len(reference_read(text)[0]) - 1
    # same size, other modification time: refused
    api.write_sidecar(text, side)
    st = os.stat(text)
    os.utime(text, ns=(st.st_atime_ns, st.st_mtime_ns + 1000))
    with pytest.raises(api.CapowError) as e:
        api.read_sidecar(side, text)
    assert e.value.code == api.CAPOW_ERR_STALE
    # damaged: truncated rows, wrong magic, too short for a header
    api.write_sidecar(text, side)
    raw = open(side, "rb").read()
    for bad in (raw[:-8], b"XAPOWCAT" + raw[8:], raw[:40]):
        open(side, "wb").write(bad)
        with pytest.raises(api.CapowError) as e:
            api.read_sidecar(side)
        assert e.value.code == api.CAPOW_ERR_PARSE
    # a malformed text line fails the write, naming the line
    open(text, "a").write("1 2 oops\n")
    with pytest.raises(api.CapowError) as e:
        api.write_sidecar(text, side)
    assert e.value.code == api.CAPOW_ERR_PARSE and "1 2 oops" in str(e.value)
    # empty text: an empty sidecar
    open(text, "w").close()
    api.write_sidecar(text, side)
    assert api.read_sidecar(side, text).shape == (0, 4)
    with pytest.raises(api.CapowError) as e:
        api.read_sidecar(str(tmp_path / "missing.capowbin"))
    assert e.value.code == api.CAPOW_ERR_IO
