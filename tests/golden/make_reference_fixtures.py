"""Writes tests/golden/reference_fixtures.npz from the data files the
reference ships for its sampler loop (read as data, nothing of the reference
is run):

  Likelihood/makefile/Sphere.dat           Matrix D2D {1234, 2, 80160, 3, typeId 0}:
                                           80,160 integer points (x, y, z rows)
  Likelihood/makefile/LikelihoodInput.txt  PETsigma2, PETscale, T1uc, T2uc, slope
  makefile/Example/HGG_TumorIC.bin         12 float32 (x, y, z of the tumour centre first)
  makefile/Example/HGG_InputParameters.txt Dw [cm^2/day], rho [1/day], T [days]

Run from the repository root where the reference tree is present:
    python tests/golden/make_reference_fixtures.py
"""
import os

import numpy as np

REF = "/root/reference/source"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.npz")


def main():
    path = os.path.join(REF, "Likelihood", "makefile", "Sphere.dat")
    head = np.fromfile(path, np.int32, 5)
    assert list(head[:2]) == [1234, 2] and head[4] == 0
    nx, ny = int(head[2]), int(head[3])
    pts = np.fromfile(path, np.float64, offset=20)
    assert pts.size == nx * ny and np.all(pts == np.round(pts))
    sphere = pts.reshape(ny, nx).astype(np.int16)  # row c = coordinate c of every point
    lik_in = np.loadtxt(os.path.join(REF, "Likelihood", "makefile", "LikelihoodInput.txt"))
    tumor_ic = np.fromfile(os.path.join(REF, "makefile", "Example", "HGG_TumorIC.bin"), np.float32)
    with open(os.path.join(REF, "makefile", "Example", "HGG_InputParameters.txt")) as f:
        params_text = f.read()
    params = np.array([float(v) for v in params_text.split()])
    np.savez_compressed(OUT, sphere=sphere, sphere_header=head, likelihood_input=lik_in, tumor_ic=tumor_ic,
                        input_parameters=params, input_parameters_text=np.array(params_text))
    print("wrote %s: %d sphere points, likelihood input %s, tumour IC %s, parameters %s"
          % (OUT, nx, lik_in.tolist(), tumor_ic[:3].tolist(), params.tolist()))


if __name__ == "__main__":
    main()
