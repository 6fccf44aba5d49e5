"""The capow host program's input handling and failure paths, which exit
before any GPU work: argv (reference src/main.c:39-47), the parameter parser
(src/read_params.c:30-290), check_inputs (src/read_params.c:294-428) and
the catalogue reader (src/read_data.c:51-84).  Every failure prints the
reference's message and "Fatal Error at line L in routine R" (src/main.c:1298)
and exits with status 1."""
import os
import subprocess

import numpy as np
import pytest

from tests.hostfiles import write_catalogue, write_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "capow_amd", "bin", "capow")


def run(*args, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, env=e, timeout=60)
    return r.returncode, r.stdout


def test_usage_when_arguments_missing():
    rc, out = run("a", "b")
    assert rc == 1
    assert "Input parameters not found" in out
    assert "Call with <ParameterFile, DataFile, RandomFile, OutputFile>" in out


def test_missing_parameter_file(tmp_path):
    rc, out = run(str(tmp_path / "nope.dat"), "d", "r", "o")
    assert rc == 1
    assert "ERROR: Parameter file '%s' not found." % (tmp_path / "nope.dat") in out
    assert "Fatal Error at line 330 in routine read_param.c" in out


def test_missing_tag(tmp_path):
    pf = write_params(str(tmp_path / "p.dat"), drop=("NK",))
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1
    assert "ERROR: I miss a value for tag 'NK' in parameter file" in out


def test_duplicate_and_unknown_tags(tmp_path):
    pf = write_params(str(tmp_path / "p.dat"), extra=(("NX", 64), ("Bogus", 3)))
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1
    assert "Tag 'NX' not allowed or multiple defined." in out
    assert "Tag 'Bogus' not allowed or multiple defined." in out


def test_legacy_example_tags_tolerated(tmp_path):
    """The reference's own example file carries five unregistered tags
    (files/run_parameters.dat:7-12); they are accepted with a warning."""
    pf = write_params(str(tmp_path / "p.dat"),
                      extra=(("InputDir", "/x/"), ("FileBase", "f.dat"), ("InputFormat", 1)))
    rc, out = run(pf, str(tmp_path / "missing.txt"), "r", "o")
    assert "WARNING: tag 'InputDir' is not used" in out
    # parsing succeeded: the run proceeds to the catalogue, which is missing
    assert "Task 0 cannot open input file" in out
    assert "Fatal Error at line 96 in routine read_data" in out


@pytest.mark.parametrize("over,line,text", [
    (dict(NumFilesInParallel=0), 297, "`NumFilesInParallel' is less than 1"),
    (dict(LOS=4), 313, "`LOS' is neither 0 (real-space), 1, 2 nor 3."),
    (dict(InterpOrder=0), 329, "`InterpOrder' is neither 1, 2 nor 3."),
    (dict(DoInterlacing=2), 337, "`DoInterlacing' is neither 0 nor 1."),
    (dict(Output2D=3), 345, "`Output2D' is neither 0 nor 1."),
    (dict(Output2D=1, NMU=0), 354, "`NMU' is less than 1."),
    (dict(OutputLog=-1), 363, "`OutputLog' is neither 0 nor 1."),
    (dict(NX=0), 371, "`NX' is less than 1."),
    (dict(NY=0), 379, "`NY' is less than 1."),
    (dict(NZ=0), 387, "`NZ' is less than 1."),
    (dict(NK=0), 395, "`NK' is less than 1."),
    (dict(XMAX=-1.0), 403, "`XMAX' is less than `XMIN'."),
    (dict(YMAX=-1.0), 411, "`XMAX' is less than `XMIN'."),
    (dict(ZMAX=-1.0), 419, "`XMAX' is less than `XMIN'."),
    (dict(Maxk=-0.5), 427, "`Maxk' is less than `Mink'."),
])
def test_check_inputs(over, line, text, tmp_path):
    pf = write_params(str(tmp_path / "p.dat"), **over)
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1
    assert text in out
    assert "Fatal Error at line %d in routine read_param.c" % line in out


def test_catalogue_errors(tmp_path):
    pf = write_params(str(tmp_path / "p.dat"))
    bad = tmp_path / "bad.txt"
    bad.write_text("# header\n1 2 3 4 5 6\n1 2 3\n")
    rc, out = run(pf, str(bad), "r", "o")
    assert rc == 1
    assert "Task 0 has error reading file: 1 2 3" in out
    assert "Fatal Error at line 102 in routine read_data" in out
    oob = write_catalogue(str(tmp_path / "oob.txt"), np.array([[1.0, 2.0, 1000.5]]),
                          np.zeros((1, 3)))
    rc, out = run(pf, oob, "r", "o")
    assert rc == 1
    assert "has particle out of grid bounds" in out
    assert "Fatal Error at line 110 in routine read_data" in out


def test_catalogue_errors_through_the_sidecar(tmp_path):
    """CAPOW_SIDECAR=1: the catalogue errors are the text reader's (the
    first malformed line, the first particle out of the box), and a
    catalogue whose parse succeeds leaves a current sidecar beside it (its
    box error then comes from the sidecar on the next run)."""
    from capow_amd import api
    pf = write_params(str(tmp_path / "p.dat"))
    env = {"CAPOW_SIDECAR": "1"}
    bad = tmp_path / "bad.txt"
    bad.write_text("# header\n1 2 3 4 5 6\n1 2 3\n")
    rc, out = run(pf, str(bad), "r", "o", env=env)
    assert rc == 1
    assert "Task 0 has error reading file: 1 2 3" in out
    assert "Fatal Error at line 102 in routine read_data" in out
    assert not os.path.exists(str(bad) + ".capowbin")
    oob = write_catalogue(str(tmp_path / "oob.txt"), np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 1000.5]]),
                          np.zeros((2, 3)))
    for attempt in range(2):   # the first run writes the sidecar, the second reads it
        rc, out = run(pf, oob, "r", "o", env=env)
        assert rc == 1
        assert "has particle out of grid bounds: x=1.000000, y=2.000000, z=1000.500000" in out
        assert "Fatal Error at line 110 in routine read_data" in out
        assert os.path.exists(oob + ".capowbin")
        assert api.read_sidecar(oob + ".capowbin", oob, los=3).shape == (2, 4)


def test_neither_periodic_nor_survey_is_rejected(tmp_path):
    pf = write_params(str(tmp_path / "p.dat"), Periodic=0, Survey=0)
    rc, out = run(pf, "d", "r", "o")
    assert rc == 1
    assert "neither `Periodic' nor `Survey'" in out


def test_survey_ingest_failures(tmp_path):
    """The survey branch's fatal cases before any GPU work
    (src/read_data.c:113, 102, 155; src/main.c:72)."""
    kw = dict(Periodic=0, Survey=1, XMIN=-1000.0, XMAX=1000.0, YMIN=-1000.0, YMAX=1000.0,
              ZMIN=-1000.0, ZMAX=1000.0, Coord_Type=1, REDMININ=0.0, REDMAXIN=0.5, NOBJ_Max=2)
    pf = write_params(str(tmp_path / "p.dat"), **kw)
    rc, out = run(pf, str(tmp_path / "missing"), "r", "o")
    assert rc == 1 and "Creating redshift-distance lookup table" in out
    assert "cannot open input file" in out and "Fatal Error at line 113 in routine read_data" in out
    row = "10 20 %s 1 1e-4 1e-4 0 0 0 0 0 0\n"
    bad = tmp_path / "bad.txt"
    bad.write_text(row % 0.1 + "1 2 3\n")
    rc, out = run(pf, str(bad), "r", "o")
    assert rc == 1 and "has error reading file: 1 2 3" in out
    assert "Fatal Error at line 102 in routine read_data" in out
    far = tmp_path / "far.txt"
    far.write_text(row % 0.45)     # r(0.45) ~ 1210 Mpc/h > 1000
    rc, out = run(pf, str(far), "r", "o")
    assert rc == 1 and "object out of grid bounds for chosen InterpOrder" in out
    assert "Fatal Error at line 155 in routine read_data" in out
    many = tmp_path / "many.txt"
    many.write_text(row % 0.1 * 3)
    rc, out = run(pf, str(many), "r", "o")
    assert rc == 1 and "more objects than NOBJ_Max" in out
    assert "Fatal Error at line 72 in routine main" in out
