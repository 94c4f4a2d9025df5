"""Coordinate readers and the command line (host side; the GPU run is in
test_gpu_configs).  Fixtures: upstream's own molecule files
(test/test_singlepoint/mols/{H2O,NO2}/coord, NO2/mol.xyz, AD7en+/.CHRG),
copied to tests/golden/mols."""

from pathlib import Path

import numpy as np
import pytest

from dxtb_amd import io
from dxtb_amd import cli

G = Path(__file__).resolve().parent / "golden" / "mols"


def test_turbomole_coord():
    z, p = io.read(G / "H2O.coord")
    assert list(z) == [8, 1, 1]
    # bohr, as written (the file's $coord block has no unit option)
    assert np.allclose(p[0], [0.0, 0.0, -0.74288549752983], atol=0, rtol=0)
    assert np.allclose(p[1], [-1.43472674945442, 0.0, 0.37144274876492], atol=0, rtol=0)


def test_xyz_and_coord_of_the_same_molecule_agree():
    """NO2's xyz (Angstrom, 14 decimals) and coord (bohr) files describe the
    same geometry: after the unit conversion they agree to the files'
    rounding (1e-9 bohr)."""
    zc, pc = io.read(G / "NO2.coord")
    zx, px = io.read(G / "NO2.xyz")
    assert list(zc) == list(zx) == [7, 8, 8]
    assert np.abs(pc - px).max() < 1e-9


def test_charge_file_and_defaults(tmp_path):
    assert io.read_chrg(G / "AD7en_plus.CHRG") == 1
    d = tmp_path / "mol"
    d.mkdir()
    (d / "coord").write_text((G / "H2O.coord").read_text())
    assert io.attr_next_to(d / "coord", "chrg") == 0
    (d / ".CHRG").write_text("-1\n")
    (d / ".UHF").write_text("2\n")
    assert io.attr_next_to(d / "coord", "chrg") == -1
    assert io.attr_next_to(d / "coord", "spin") == 2


def test_reader_errors(tmp_path):
    bad = tmp_path / "x.xyz"
    bad.write_text("3\ncomment\nO 0 0 0\nH 0 0 1\n")
    with pytest.raises(ValueError, match="announced"):
        io.read(bad)
    bad.write_text("1\n\nQq 0 0 0\n")
    with pytest.raises(ValueError, match="unknown element"):
        io.read(bad)
    nc = tmp_path / "coord"
    nc.write_text("$end\n")
    with pytest.raises(ValueError, match="no \\$coord"):
        io.read(nc)
    with pytest.raises(ValueError, match="file type"):
        io.read(tmp_path / "structure.dat")


def test_angstrom_coord_block(tmp_path):
    f = tmp_path / "coord"
    f.write_text("$coord angs\n 0.0 0.0 0.0 h\n 0.0 0.0 0.74 h\n$end\n")
    z, p = io.read(f)
    assert list(z) == [1, 1] and p[1, 2] == pytest.approx(0.74 / 0.52917721090380, rel=1e-15)


def test_cli_batch_loading_and_table(tmp_path):
    """Several files become one padded batch (atomic number 0 pads); the
    charge comes from .CHRG next to each file; the table has upstream's
    layout (driver.py:45-67)."""
    a = tmp_path / "a"
    a.mkdir()
    (a / "coord").write_text((G / "H2O.coord").read_text())
    (a / ".CHRG").write_text("1\n")
    args = cli.build_parser().parse_args([str(a / "coord"), str(G / "NO2.xyz"), "--grad"])
    z, p, ch, sp = cli.load(args)
    assert z.shape == (2, 3) and p.shape == (2, 3, 3)
    assert list(z[0]) == [8, 1, 1] and list(z[1]) == [7, 8, 8]
    assert ch == [1, 0] and sp == [0, 0]
    t = cli.format_table(np.array([[1.0, -2.0, 0.5]]), np.array([8]), "Forces").splitlines()
    assert t[2:4] == ["Forces", "------"]
    assert t[5] == "  #  Z            dX              dY              dZ"
    assert t[7] == "  1  O       1.000000000    -2.000000000     0.500000000"
    args = cli.build_parser().parse_args([str(G / "H2O.coord"), "--chrg", "-1", "--xtol", "1e-9"])
    z, p, ch, sp = cli.load(args)
    assert z.shape == (3,) and ch == -1 and cli.options(args) == {"x_atol": 1e-9}
