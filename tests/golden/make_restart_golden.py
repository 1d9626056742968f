"""Writes the restart-file fixtures with an installed libhdf5 (the conda
HDF5 under /opt/conda in this image): make_restart_h5.c is compiled against
it and run on seeded inputs, the files it writes are kept as

    restart_libhdf5.h5            RestartIO::write layout
    restart_libhdf5_par.h5.rank0  RestartIO::writeParallel fallback (rank 0 of 1)
    restart_libhdf5_big.h5        the same layout at 50 000 elements (not
                                  committed: re-made by the test when
                                  libhdf5 is present)
    restart_expected.npz          the inputs

    python tests/golden/make_restart_golden.py [--hdf5 /opt/conda] [--big OUT]
"""
import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs(n_elem, seed):
    """elem_n_sp alternates P2 quadrilaterals (9) and P2 triangles (6);
    solution values are seeded normals with a few special values."""
    rng = np.random.default_rng(seed)
    nsp = np.where(rng.random(n_elem) < 0.7, 9, 6).astype(np.int32)
    sol = rng.normal(size=int(nsp.sum()) * 4)
    if sol.size >= 4:
        sol[:4] = [0.0, -0.0, 5e-324, 1.7976931348623157e308]
    return dict(n_elem=n_elem, time=0.3125 + seed, iteration=1234 + seed, poly_order=2, gamma=1.4,
                elem_n_sp=nsp, solution=sol)


def write_inputs(path, d):
    with open(path, "wb") as f:
        np.array([d["n_elem"], d["solution"].size], np.int64).tofile(f)
        np.array([d["time"]], np.float64).tofile(f)
        np.array([d["iteration"], d["poly_order"]], np.int32).tofile(f)
        np.array([d["gamma"]], np.float64).tofile(f)
        d["elem_n_sp"].astype(np.int32).tofile(f)
        d["solution"].astype(np.float64).tofile(f)


def build_tool(hdf5, out_dir):
    exe = os.path.join(out_dir, "make_restart_h5")
    subprocess.run(["gcc", "-O1", "-o", exe, os.path.join(HERE, "make_restart_h5.c"), f"-I{hdf5}/include",
                    f"-L{hdf5}/lib", "-lhdf5", f"-Wl,-rpath,{hdf5}/lib"], check=True)
    return exe


def run(exe, d, out, par_prefix, tmp):
    inp = os.path.join(tmp, "inputs.bin")
    write_inputs(inp, d)
    subprocess.run([exe, inp, out, par_prefix, "0", "1"], check=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--hdf5", default="/opt/conda")
    ap.add_argument("--big", default=None, help="write only the 50 000-element file to this path")
    a = ap.parse_args()
    with tempfile.TemporaryDirectory() as tmp:
        exe = build_tool(a.hdf5, tmp)
        if a.big:
            d = inputs(50000, 7)
            run(exe, d, a.big, os.path.join(tmp, "par"), tmp)
            np.savez(a.big + ".npz", **{k: np.asarray(v) for k, v in d.items()})
            return
        d = inputs(7, 1)
        run(exe, d, os.path.join(HERE, "restart_libhdf5.h5"), os.path.join(HERE, "restart_libhdf5_par.h5"), tmp)
        np.savez(os.path.join(HERE, "restart_expected.npz"), **{k: np.asarray(v) for k, v in d.items()})
    print("wrote restart fixtures", file=sys.stderr)


if __name__ == "__main__":
    main()
