"""Input and output files of the reference's models and recorder.

* ``displacements.csv`` with a header and the columns ``oid, x, y, d, theta,
  phi`` (models/mogi/mogi/Data.py:17-39; the same sheet for CDM).  LOS unit
  vectors are (sin theta cos phi, sin theta sin phi, cos theta)
  (models/mogi/mogi/Mogi.py:104-117).
* ``cd.txt``, ``green.txt``, ``data.txt``: whitespace-separated text
  matrices and vectors, one matrix row per line (models/mogi/mogi/Mogi.py
  :276-324, models/linear/linear/Linear.py:231-296; the example cases
  models/linear/examples/patch-9, patch-144).
* ``parameters.csv``: the parameter-set layout, ``name, count, offset``
  (Mogi.py:359-381).
* Recorder output ``theta.txt``, ``sigma.txt``, ``llk.txt`` (the final
  sample, its covariance and posterior log-likelihood;
  altar/altar/simulations/Recorder.py:46-58), written with 17 significant
  digits so that a reload is exact.
"""
import csv
import os

import numpy as np

DISPLACEMENT_COLUMNS = ("oid", "x", "y", "d", "theta", "phi")


def los_vectors(theta, phi):
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=1)


def read_displacements(path):
    """Columns by header name; returns a dict of numpy arrays plus ``los``."""
    with open(path, newline="") as f:
        rows = list(csv.reader(f))
    if not rows:
        raise ValueError(f"{path}: empty displacement file")
    header = [h.strip() for h in rows[0]]
    missing = [c for c in DISPLACEMENT_COLUMNS if c not in header]
    if missing:
        raise ValueError(f"{path}: missing columns {missing}")
    idx = {c: header.index(c) for c in DISPLACEMENT_COLUMNS}
    body = [r for r in rows[1:] if r and any(x.strip() for x in r)]
    out = {}
    for c in DISPLACEMENT_COLUMNS:
        vals = [r[idx[c]] for r in body]
        out[c] = np.array(vals, dtype=np.int32 if c == "oid" else np.float64)
    out["los"] = los_vectors(out["theta"], out["phi"])
    return out


def write_displacements(path, oid, x, y, d, theta, phi):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(DISPLACEMENT_COLUMNS)
        for row in zip(oid, x, y, d, theta, phi):
            w.writerow([int(row[0])] + [repr(float(v)) for v in row[1:]])


def read_matrix(path, shape=None):
    """A whitespace-separated text matrix (rows on lines)."""
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if shape is not None:
        a = a.reshape(shape)
    return a


def read_vector(path, n=None):
    a = np.loadtxt(path, dtype=np.float64, ndmin=1).reshape(-1)
    if n is not None and a.shape[0] != n:
        raise ValueError(f"{path}: expected {n} values, found {a.shape[0]}")
    return a


def write_matrix(path, a):
    a = np.atleast_2d(np.asarray(a, dtype=np.float64))
    np.savetxt(path, a, fmt="%.17g")


def write_vector(path, v):
    np.savetxt(path, np.asarray(v, dtype=np.float64).reshape(-1, 1), fmt="%.17g")


def write_parameters(path, psets):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["name", "count", "offset"])
        for name, pset in psets.items():
            w.writerow([name, pset.count, pset.offset])


def load_linear_case(directory, observations=None, parameters=None, green="green.txt", data="data.txt",
                     cd="cd.txt"):
    """The reference's linear example layout (models/linear/examples/patch-9)."""
    G = read_matrix(os.path.join(directory, green))
    N, P = G.shape
    if observations is not None and N != observations:
        raise ValueError("green functions: wrong number of observations")
    if parameters is not None and P != parameters:
        raise ValueError("green functions: wrong number of parameters")
    d = read_vector(os.path.join(directory, data), N)
    C = read_matrix(os.path.join(directory, cd), (N, N))
    return dict(G=G, d=d, cd=C)


def load_source_case(directory, displacements="displacements.csv", covariance="cd.txt"):
    """A Mogi / CDM case directory (Mogi.py:276-324)."""
    obs = read_displacements(os.path.join(directory, displacements))
    N = len(obs["x"])
    obs["cd"] = read_matrix(os.path.join(directory, covariance), (N, N))
    return obs


class Recorder:
    """Writes the final state of an annealing run (Recorder.py:19-58)."""

    def __init__(self, theta="theta.txt", sigma="sigma.txt", llk="llk.txt", directory="."):
        self.theta, self.sigma, self.llk, self.directory = theta, sigma, llk, directory

    def record(self, step, **kwds):
        os.makedirs(self.directory, exist_ok=True)
        write_matrix(os.path.join(self.directory, self.theta), step.theta.cpu().numpy())
        write_matrix(os.path.join(self.directory, self.sigma), step.sigma.cpu().numpy())
        write_vector(os.path.join(self.directory, self.llk), step.posterior.cpu().numpy())
        return self
