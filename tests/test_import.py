"""Mesh import (NASTRAN, Abaqus) against the reference's own known answers
(test/test_meshing.jl:3426-3432 and :6406-6413) on the reference's fixture
files, copied under tests/golden/meshes as data."""
import os

import numpy as np
import pytest

from feasm import finetools as F

HERE = os.path.dirname(os.path.abspath(__file__))
NAS = os.path.join(HERE, "golden", "meshes", "cylinder.nas")
INP = os.path.join(HERE, "golden", "meshes", "problem_LE10NAFEMS_MST10.inp")


def test_cylinder_nastran_counts():
    out = F.import_NASTRAN(NAS, allocationchunk=13, expectfixedformat=True)
    assert out["fens"].count() == 1406
    assert out["fesets"][0].count() == 829
    assert isinstance(out["fesets"][0], F.FESetT10)
    assert out["property_ids"] == [1]
    conn = out["fesets"][0].conn
    assert conn.min() >= 1 and conn.max() <= 1406
    assert len(np.unique(conn)) == 1406  # every node is used by some element


def test_fixupdecimal():
    assert float(F._fixupdecimal("-1.66618812195+1")) == -16.6618812195
    assert float(F._fixupdecimal("3.85740337853-2")) == 3.85740337853e-2
    assert F._fixupdecimal("1.5E-3") == "1.5E-3"


def test_nastran_free_format_and_errors(tmp_path):
    good = tmp_path / "t.nas"
    good.write_text("BEGIN BULK\nGRID,1,,0.0,0.0,0.0\nGRID,2,,1.0,0.0,0.0\nGRID,3,,0.0,1.0,0.0\n"
                    "GRID,4,,0.0,0.0,1.0\nCTETRA,1,7,1,2,3,4\nENDDATA\n")
    out = F.import_NASTRAN(str(good))
    assert out["fens"].count() == 4 and out["property_ids"] == [7]
    assert isinstance(out["fesets"][0], F.FESetT4)
    assert np.array_equal(out["fesets"][0].conn, [[1, 2, 3, 4]])
    bad = tmp_path / "b.nas"
    bad.write_text("GRID,2,,0.0,0.0,0.0\nGRID,1,,1.0,0.0,0.0\nENDDATA\n")
    with pytest.raises(ValueError, match="serial order"):
        F.import_NASTRAN(str(bad))


def test_abaqus_nafems_counts():
    out = F.import_ABAQUS(INP, allocationchunk=11)
    assert out["fesets"][0].count() == 60
    assert isinstance(out["fesets"][0], F.FESetT10)
    assert len(out["nsets"]["L3"]) == 7
    assert any("SFM3D6" in w for w in out["warnings"])


def test_abaqus_renumbering_generate_and_continuation(tmp_path):
    f = tmp_path / "m.inp"
    f.write_text("*NODE\n10, 0.0, 0.0, 0.0\n20, 1.0, 0.0, 0.0\n30, 0.0, 1.0, 0.0\n40, 0.0, 0.0, 1.0\n"
                 "*ELEMENT, TYPE=C3D4, ELSET=E\n5, 10, 20,\n30, 40\n"
                 "*NSET, NSET=ALL, GENERATE\n10, 40, 10\n*ELSET, ELSET=E1\n5\n")
    out = F.import_ABAQUS(str(f))
    assert out["fens"].count() == 4
    assert np.array_equal(out["fesets"][0].conn, [[1, 2, 3, 4]])
    assert np.array_equal(out["nsets"]["ALL"], [1, 2, 3, 4])
    assert np.array_equal(out["elsets"]["E1"], [5])
