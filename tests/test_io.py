"""File formats of the reference (CPU): displacements.csv, text matrices,
parameters.csv and the recorder output, plus the reference's patch-9 linear
case (tests/golden/patch9)."""
import os

import numpy as np

from altar_amd import io
from altar_amd.distributions import Contiguous, Uniform

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_displacements_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    n = 50
    oid = rng.integers(0, 2, n)
    x, y, d = rng.normal(size=(3, n))
    th, ph = rng.uniform(0, 1, n), rng.uniform(-3, 3, n)
    p = tmp_path / "displacements.csv"
    io.write_displacements(p, oid, x, y, d, th, ph)
    assert open(p).readline().strip() == "oid,x,y,d,theta,phi"
    r = io.read_displacements(p)
    assert np.array_equal(r["oid"], oid) and np.array_equal(r["x"], x) and np.array_equal(r["phi"], ph)
    # LOS of models/mogi/mogi/Mogi.py:112-117
    assert np.allclose(r["los"][:, 2], np.cos(th)) and np.allclose(np.linalg.norm(r["los"], axis=1), 1.0)


def test_columns_by_header_name(tmp_path):
    p = tmp_path / "d.csv"
    p.write_text("x,y,oid,phi,theta,d\n1,2,0,0.5,0.25,0.01\n3,4,1,0.5,0.25,0.02\n")
    r = io.read_displacements(p)
    assert np.array_equal(r["oid"], [0, 1]) and np.array_equal(r["d"], [0.01, 0.02])


def test_patch9_case():
    case = io.load_linear_case(os.path.join(GOLDEN, "patch9"), observations=108, parameters=18)
    assert case["G"].shape == (108, 18) and case["d"].shape == (108,)
    assert np.allclose(case["cd"], 2.5e-5 * np.eye(108))


def test_matrix_and_recorder_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(7, 3))
    io.write_matrix(tmp_path / "m.txt", a)
    assert np.array_equal(io.read_matrix(tmp_path / "m.txt"), a)
    v = rng.normal(size=9)
    io.write_vector(tmp_path / "v.txt", v)
    assert np.array_equal(io.read_vector(tmp_path / "v.txt", 9), v)
    ps = {"location": Contiguous(2, Uniform((-1, 1)), offset=0), "depth": Contiguous(1, Uniform((0, 1)), offset=2)}
    io.write_parameters(tmp_path / "parameters.csv", ps)
    assert (tmp_path / "parameters.csv").read_text().splitlines() == ["name,count,offset", "location,2,0",
                                                                      "depth,1,2"]
