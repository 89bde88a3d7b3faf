"""Write CDM golden vectors from the reference's numpy CDM (models/cdm/cdm/libcdm.py).

Run in the development container only (the reference tree is not present on
GPU machines):

    python tests/golden/make_cdm_vectors.py /root/reference

The reference's MATLAB known-answer test (models/cdm/tests/libcdm.py:27-46)
pins one geometry at 12 points.  These vectors cover what it leaves open:
random geometries over the cfg2 prior box, the single-zero-axis branches
(libcdm.py:230-235), sides that are exactly vertical (the beta = 0 / pi skip,
libcdm.py:320-324), points on the configuration I/II boundary
(libcdm.py:344) and large rotations / aspect ratios.  Output:
tests/golden/cdm_vectors.npz with arrays ``X``, ``Y`` (points), ``params``
(cases x 11: X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ, opening, nu)
and ``ue``, ``un``, ``uv`` (cases x points).
"""
import importlib.util
import os
import sys

import numpy as np


def load_libcdm(root):
    path = os.path.join(root, "models", "cdm", "cdm", "libcdm.py")
    spec = importlib.util.spec_from_file_location("ref_libcdm", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cases(rng):
    out = []
    # random geometries over the cfg2 prior box (metres, degrees)
    while len(out) < 40:
        p = [rng.uniform(-10e3, 10e3), rng.uniform(-10e3, 10e3), rng.uniform(1e3, 6e3),
             rng.uniform(100, 3000), rng.uniform(100, 3000), rng.uniform(100, 3000),
             rng.uniform(-90, 90), rng.uniform(-90, 90), rng.uniform(-90, 90),
             rng.uniform(1e-3, 1.0), 0.25]
        out.append(p)
    # no rotation: the vertical sides of two RDs hit the beta = 0 / pi skip
    out.append([0.0, 0.0, 4000.0, 1000.0, 1200.0, 800.0, 0.0, 0.0, 0.0, 0.1, 0.25])
    out.append([1500.0, -700.0, 3000.0, 500.0, 500.0, 500.0, 0.0, 0.0, 45.0, 0.3, 0.25])
    # single zero axis branches
    out.append([1000.0, -500.0, 4000.0, 0.0, 1200.0, 800.0, 5.0, -8.0, 30.0, 0.1, 0.25])
    out.append([1000.0, -500.0, 4000.0, 1000.0, 0.0, 800.0, 5.0, -8.0, 30.0, 0.1, 0.25])
    out.append([1000.0, -500.0, 4000.0, 1000.0, 1200.0, 0.0, 5.0, -8.0, 30.0, 0.1, 0.25])
    # large rotations and aspect ratios
    out.append([-3000.0, 2000.0, 5500.0, 3000.0, 100.0, 1500.0, 89.0, -89.0, 89.0, 1.0, 0.25])
    out.append([0.0, 0.0, 6000.0, 100.0, 3000.0, 100.0, -60.0, 75.0, -10.0, 0.5, 0.3])
    # the reference's MATLAB geometry in km, and a shallow source
    out.append([0.5, -0.25, 2.75, 0.4, 0.45, 0.8, 5.0, -8.0, 30.0, 1e-3, 0.25])
    out.append([200.0, 300.0, 1000.0, 300.0, 400.0, 500.0, 10.0, 20.0, 30.0, 0.05, 0.25])
    return np.array(out)


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    libcdm = load_libcdm(root)
    rng = np.random.default_rng(20261203)
    params = cases(rng)
    # 60 random points over the 30 km cfg2 box, plus points on axes through the
    # source centre (where y1A = 0 for sides aligned with the axes)
    X = np.concatenate([rng.uniform(-15e3, 15e3, 60), [0.0, 0.0, 1000.0, -1000.0]])
    Y = np.concatenate([rng.uniform(-15e3, 15e3, 60), [1000.0, -1000.0, 0.0, 0.0]])
    keep, ue, un, uv = [], [], [], []
    for k, p in enumerate(params):
        X0, Y0, depth, ax, ay, az, ox, oy, oz, op, nu = p
        if depth < 100:  # km-scaled MATLAB geometry: use km points
            Xk, Yk = X / 1e3, Y / 1e3
        else:
            Xk, Yk = X, Y
        try:
            e, n, v = libcdm.CDM(Xk, Yk, X0, Y0, depth, ax, ay, az, ox, oy, oz, op, nu)
        except ValueError:
            continue  # pierces the free surface
        keep.append(k)
        ue.append(e)
        un.append(n)
        uv.append(v)
    params = params[keep]
    scale = np.where(params[:, 2] < 100, 1e-3, 1.0)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cdm_vectors.npz")
    np.savez_compressed(out, X=X, Y=Y, point_scale=scale, params=params,
                        ue=np.array(ue), un=np.array(un), uv=np.array(uv))
    print(f"wrote {out}: {len(params)} geometries x {len(X)} points")


if __name__ == "__main__":
    main()
