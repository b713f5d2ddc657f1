"""Write the reference's own known-answer data as small fixtures.

Reads two reference source files *as text* (it never compiles or imports them) and writes:

* ``icls_golden.npz`` — the Matlab ``quadprog`` problem and solutions embedded in
  ``testing/cmd/testing_unit_tests_icls.cpp:39-59`` (column-major arrays, sizes at ``:39-42``),
  reshaped here to row-major ``H[100,50]``, ``A_ineq[290,50]``, ``A_eq[10,50]``, ``b[100]``,
  ``t_ineq[290]``, ``t_eq[10]`` and the four solution vectors ``[50]``.
* ``mrtrix3_amd/data/electrostatic_repulsion_300.txt`` — the built-in 300 constraint directions
  ``src/dwi/directions/predefined.cpp:92-393`` as ``azimuth elevation`` rows (radians), the same
  text format ``dwi2fod -directions`` accepts.  This table is an input of the algorithm, so the
  product ships it; it is MRtrix3 data under MPL-2.0 (``LICENCE.txt`` of the reference).
* ``mrtrix3_amd/data/sh2peaks_seeds_60.txt`` — the 60 built-in search directions of ``sh2peaks``
  (``cmd/sh2peaks.cpp:363-424``, ``default_directions``) as ``azimuth elevation`` rows, the text
  format of ``sh2peaks -seeds``; the reference holds them as float.
* ``mrtrix3_amd/data/tesselation_1281.txt`` — the 1281 directions ``fod2fixel`` segments FODs on
  (``src/dwi/directions/predefined.cpp``, ``tesselation_1281_data``; ``cmd/fod2fixel.cpp:318``).

Usage: ``python tests/golden/extract_fixtures.py [/path/to/reference]``.  Run once; the outputs
are committed, and nothing at test or run time reads the reference tree.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))


def _array(text, name):
    m = re.search(r"\b%s\s*\[\]\s*=\s*\{([^}]*)\}" % re.escape(name), text)
    if not m:
        raise RuntimeError("array %s not found" % name)
    vals = [v for v in re.split(r"[\s,]+", m.group(1)) if v]
    return np.array([float(v) for v in vals], dtype=np.float64)


def main(ref):
    icls = open(os.path.join(ref, "testing/cmd/testing_unit_tests_icls.cpp")).read()
    sizes = {k: int(re.search(r"const size_t %s\s*=\s*(\d+)" % k, icls).group(1))
             for k in ("num_sig", "num_coef", "num_ineq", "num_eq")}
    ns, nc, ni, ne = sizes["num_sig"], sizes["num_coef"], sizes["num_ineq"], sizes["num_eq"]

    def colmajor(name, rows, cols):
        a = _array(icls, name)
        assert a.size == rows * cols, (name, a.size, rows, cols)
        return a.reshape(cols, rows).T.copy()   # Eigen::Map<MatrixXd> is column-major

    out = dict(
        H=colmajor("problem_matrix_data", ns, nc),
        A_ineq=colmajor("inequality_constraint_matrix_data", ni, nc),
        A_eq=colmajor("equality_constraint_matrix_data", ne, nc),
        b=_array(icls, "problem_vector_data"),
        t_ineq=_array(icls, "inequality_constraint_vector_data"),
        t_eq=_array(icls, "equality_constraint_vector_data"),
        x_eq=_array(icls, "solution_data"),
        x_no_eq=_array(icls, "solution_no_eq_data"),
        x_eq_zero=_array(icls, "solution_zero_data"),
        x_no_eq_zero=_array(icls, "solution_no_eq_zero_data"),
    )
    assert out["b"].size == ns and out["t_ineq"].size == ni and out["t_eq"].size == ne
    for k in ("x_eq", "x_no_eq", "x_eq_zero", "x_no_eq_zero"):
        assert out[k].size == nc, k
    np.savez_compressed(os.path.join(HERE, "icls_golden.npz"), **out)

    pre = open(os.path.join(ref, "src/dwi/directions/predefined.cpp")).read()
    d = _array(pre, "electrostatic_repulsion_300_data").reshape(-1, 2)
    assert d.shape == (300, 2)
    path = os.path.join(REPO, "mrtrix3_amd", "data", "electrostatic_repulsion_300.txt")
    with open(path, "w") as f:
        f.write("# 300 constraint directions [ azimuth elevation ] (radians)\n")
        f.write("# MRtrix3 src/dwi/directions/predefined.cpp:92-393, MPL-2.0\n")
        for az, el in d:
            f.write("%s %s\n" % (repr(float(az)), repr(float(el))))
    pk = open(os.path.join(ref, "cmd/sh2peaks.cpp")).read()
    seeds = _array(pk, "default_directions").reshape(-1, 2)
    assert seeds.shape == (60, 2)
    path = os.path.join(REPO, "mrtrix3_amd", "data", "sh2peaks_seeds_60.txt")
    with open(path, "w") as f:
        f.write("# 60 sh2peaks search directions [ azimuth elevation ] (radians)\n")
        f.write("# MRtrix3 cmd/sh2peaks.cpp:363-424 (default_directions, float), MPL-2.0\n")
        for az, el in seeds:
            f.write("%s %s\n" % (repr(float(az)), repr(float(el))))
    d = _array(pre, "tesselation_1281_data").reshape(-1, 2)
    assert d.shape == (1281, 2)
    path = os.path.join(REPO, "mrtrix3_amd", "data", "tesselation_1281.txt")
    with open(path, "w") as f:
        f.write("# 1281 FOD segmentation directions [ azimuth elevation ] (radians)\n")
        f.write("# MRtrix3 src/dwi/directions/predefined.cpp (tesselation_1281_data), MPL-2.0\n")
        for az, el in d:
            f.write("%s %s\n" % (repr(float(az)), repr(float(el))))
    print("wrote icls_golden.npz, electrostatic_repulsion_300.txt, sh2peaks_seeds_60.txt and tesselation_1281.txt")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
