"""On-disk formats of the reference driver (io.c, main.c:452-455) through the
library's C entry points, against the reference's own tests-wall files
(tests/golden/tests-wall.pos, tests-wall.for: data of the reference's tests)."""
import os

import numpy as np
import pytest

from conftest import load_fixture, random_system

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_system_reads_the_reference_file(pkg):
    s = pkg.read_system(os.path.join(GOLD, "tests-wall.pos"))
    ref, _ = load_fixture("tests_wall.npz")
    assert s["L"] == 10.0 and len(s["q"]) == 300
    for k in "xyzq":
        np.testing.assert_array_equal(s[k], np.asarray(ref[k]))


def test_read_exact_forces_reads_all_seven_columns(pkg):
    """tests-wall.for rows are 'i f_x f_y f_z f_kx f_ky f_kz' with 22 digits:
    both force triples are read exactly (correctly rounded decimal parse)."""
    f, fk = pkg.read_exact_forces(os.path.join(GOLD, "tests-wall.for"), 300)
    table = np.loadtxt(os.path.join(GOLD, "tests-wall.for"))
    np.testing.assert_array_equal(table[:, 0], np.arange(300))
    np.testing.assert_array_equal(f, table[:, 1:4].T)
    np.testing.assert_array_equal(fk, table[:, 4:7].T)
    _, fref = load_fixture("tests_wall.npz")
    np.testing.assert_array_equal(f, np.asarray(fref).reshape(3, -1))


def test_system_round_trip_six_decimals(pkg, tmp_path):
    """Write_system prints '%lf' (six decimals), so a round trip returns each
    value rounded as C's printf rounds it (Python's '%f' is the same)."""
    s = random_system(257, 7.5, seed=3)
    path = str(tmp_path / "sys.dat")
    pkg.write_system(path, s)
    text = open(path).read().splitlines()
    assert text[0] == "# Teilchenzahl: 257" and text[1] == "# Len: 7.500000"
    assert text[2] == "\t".join("%f" % s[k][0] for k in "xyzq")
    back = pkg.read_system(path)
    assert back["L"] == 7.5
    for k in "xyzq":
        np.testing.assert_array_equal(back[k], [float("%f" % v) for v in s[k]])


def test_force_file_round_trip_and_four_columns(pkg, tmp_path):
    rng = np.random.default_rng(5)
    f, fk = rng.normal(size=(3, 40)) * 30, rng.normal(size=(3, 40))
    path = str(tmp_path / "forces.dat")
    pkg.write_exact_forces(path, f, fk)
    rows = open(path).read().splitlines()
    assert rows[7] == "\t".join(["7"] + ["%f" % v for v in list(f[:, 7]) + list(fk[:, 7])])
    g, gk = pkg.read_exact_forces(path, 40)
    np.testing.assert_array_equal(g, np.vectorize(lambda v: float("%f" % v))(f))
    np.testing.assert_array_equal(gk, np.vectorize(lambda v: float("%f" % v))(fk))
    # the four-column rows Read_exact_forces itself expects leave f_k at zero
    four = str(tmp_path / "four.dat")
    with open(four, "w") as fh:
        for i in range(40):
            fh.write("%d\t%.17g\t%.17g\t%.17g\n" % (i, f[0, i], f[1, i], f[2, i]))
    g, gk = pkg.read_exact_forces(four, 40)
    np.testing.assert_array_equal(g, f)
    assert not gk.any()


def test_out_row_matches_the_driver_format(pkg):
    v = (0.85, 3.2e-5, 2.9e-5, 1.1e-6, 2.8e-5, 2.7e-5, 4.0e-9, 0.0)
    row = pkg.format_out_row(*v)
    # C's "% lf" and "% e" (space flag) as Python formats them
    assert row == "% f\t% e\t% e\t% e\t% e\t% e\t% e\t% e\n" % v
    assert row.startswith(" 0.850000\t 3.200000e-05")


@pytest.mark.parametrize("content,what", [
    ("", "header"), ("# Teilchenzahl: 3\n# Len: 5.0\n1 2 3 1\n1 2 3\n", "particle 1"),
    ("# Teilchenzahl: 2\n# Len: 5.0\n1 2 3 1\n", "particle 1")])
def test_read_system_reports_bad_files(pkg, tmp_path, content, what):
    path = tmp_path / "bad.dat"
    path.write_text(content)
    with pytest.raises(RuntimeError, match=what):
        pkg.read_system(str(path))
    with pytest.raises(RuntimeError, match="cannot open"):
        pkg.read_system(str(tmp_path / "missing.dat"))
