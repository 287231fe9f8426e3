"""Writes the golden fixtures under tests/golden/ and the Sobol direction-number table
shipped with the package.

Sources (read in this container only; nothing here runs on the GPU box):
  * /root/reference/test/inchworm.h5, topology_eval.h5, bethe.h5  -- the reference's own
    golden vectors (test/inchworm.jl:180-212, test/topology_eval.jl:37-155) and DLR results
    (test/bethe.jl, test/bethe_gf.jl), dumped with h5dump;
  * SciPy's scipy.stats.qmc.Sobol -- the library the reference's Sobol tests were written
    against (test/scrambled_sobol.jl:47-93, with the mock RNG of :97-117);
  * SciPy's Joe-Kuo direction numbers (new-joe-kuo-6.21201), the table behind Sobol.jl's
    sobol_a / sobol_minit (src/scrambled_sobol.jl:79-83).

Usage:  python tests/golden/make_golden.py
"""
import json
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
H5DUMP = "/opt/conda/bin/h5dump"


def h5_dataset(path, name):
    """Return (shape, flat list) of a dataset; complex compounds as [re, im] pairs."""
    out = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, path], check=True,
                         capture_output=True, text=True).stdout
    shape = re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", out).group(1)
    shape = [int(s) for s in shape.split(",")]
    body = out[out.index("DATA {") + len("DATA {"):]
    if "H5T_COMPOUND" in out:
        vals = re.findall(r"\{\s*(-?[0-9.eE+-]+|-?nan|-?inf),\s*(-?[0-9.eE+-]+|-?nan|-?inf)\s*\}", body)
        flat = [[float(a), float(b)] for a, b in vals]
    else:
        body = re.sub(r"\([0-9, ]+\):", " ", body)
        flat = [float(v) for v in re.findall(r"-?[0-9][0-9.eE+-]*", body)]
    return shape, flat


def inchworm_fixture():
    path = os.path.join(REF, "test", "inchworm.h5")
    fx = {"source": "test/inchworm.h5 (test/inchworm.jl)"}
    for name in ["/inchworm/P/1", "/inchworm/P/2", "/inchworm/g",
                 "/inchworm_step/value/1", "/inchworm_step/value/2",
                 "/inchworm_step_bare/value/1", "/inchworm_step_bare/value/2",
                 "/diff_inchworm/P/1", "/diff_inchworm/P/2"]:
        shape, flat = h5_dataset(path, name)
        fx[name] = {"shape": shape, "data": flat}
    return fx


def topology_eval_fixture():
    path = os.path.join(REF, "test", "topology_eval.h5")
    fx = {"source": "test/topology_eval.h5 (test/topology_eval.jl)"}
    for name in ["/x1_list", "/xs_list", "/values"]:
        shape, flat = h5_dataset(path, name)
        fx[name] = {"shape": shape, "data": flat}
    return fx


def bethe_fixture():
    """Density matrices and Green's functions of the DLR reference results in
    test/bethe.h5 (test/bethe.jl:48-58, test/bethe_gf.jl:47-58)."""
    path = os.path.join(REF, "test", "bethe.h5")
    fx = {"source": "test/bethe.h5 (test/bethe.jl, test/bethe_gf.jl)"}
    for name in ["/rho/exact", "/rho/NCA", "/rho/OCA", "/rho/TCA", "/g/NCA", "/g/OCA", "/g/TCA"]:
        shape, flat = h5_dataset(path, name)
        fx[name] = {"shape": shape, "data": flat}
    return fx


def sobol_fixture():
    from scipy.stats import qmc
    import scipy._lib._util as util
    import scipy.stats._qmc as qmcmod

    fx = {"source": "scipy.stats.qmc.Sobol(d, scramble=..., bits=32)"}
    for d, m in [(1, 3), (5, 3), (8, 6)]:
        fx[f"unscrambled_d{d}"] = qmc.Sobol(d=d, scramble=False, bits=32).random_base2(m=m).tolist()

    # Mock RNG of test/scrambled_sobol.jl:97-117: the i-th value drawn from range(high)
    # is (popcount(i) % high), arrays filled in column-major order.
    def mock_rng_integers(rng, high, /, size, dtype):
        r = list(range(high))
        seq = np.array([r[bin(i).count("1") % len(r)] for i in range(int(np.prod(size)))], dtype=dtype)
        return np.ascontiguousarray(seq.reshape(size, order="F"))

    saved_util, saved_qmc = util.rng_integers, qmcmod.rng_integers
    util.rng_integers = mock_rng_integers
    qmcmod.rng_integers = mock_rng_integers
    try:
        for d in (1, 5):
            fx[f"mock_scrambled_d{d}"] = qmc.Sobol(d=d, scramble=True, bits=32).random_base2(m=3).tolist()
    finally:
        util.rng_integers, qmcmod.rng_integers = saved_util, saved_qmc
    return fx


def joe_kuo_table(n_dims=64):
    import scipy
    npz = np.load(os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz"))
    poly = npz["poly"][:n_dims].astype(int).tolist()
    vinit = npz["vinit"][:n_dims].astype(int).tolist()
    return {"source": "Joe & Kuo new-joe-kuo-6.21201 (as shipped with SciPy)",
            "poly": poly, "vinit": vinit}


def main():
    def dump(obj, path):
        with open(path, "w") as f:
            json.dump(obj, f, indent=0, separators=(",", ":"))
        print("wrote", os.path.relpath(path, REPO))

    dump(inchworm_fixture(), os.path.join(HERE, "inchworm_h5.json"))
    dump(topology_eval_fixture(), os.path.join(HERE, "topology_eval_h5.json"))
    dump(bethe_fixture(), os.path.join(HERE, "bethe_h5.json"))
    dump(sobol_fixture(), os.path.join(HERE, "sobol_scipy.json"))
    dump(joe_kuo_table(), os.path.join(REPO, "qinchworm.jl_amd", "qiw", "data", "sobol_joe_kuo.json"))


if __name__ == "__main__":
    main()
