"""Convert the reference's own HDF5 test fixtures into small committed .npz files.

Run in the development container only (it needs the HDF5 C library and the reference
tree); the resulting .npz files are the committed golden vectors:

    python tests/golden/make_fixtures.py /root/reference

Each .npz holds the datasets ("D_<name>") and attributes ("A_<name>") of one fixture,
complex compound types as complex128. The fixtures are data written by the reference's
test generators (test/**/test_*.py); no reference code is copied or run.
"""

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FIXTURES = [
    "test/tensor/data/test_block_sparse_tensor_transpose.hdf5",
    "test/tensor/data/test_block_sparse_tensor_reshape.hdf5",
    "test/tensor/data/test_block_sparse_tensor_dot.hdf5",
    "test/tensor/data/test_block_sparse_tensor_cyclic_partial_trace.hdf5",
    "test/tensor/data/test_block_sparse_tensor_serialize.hdf5",
    "test/tensor/data/test_dense_tensor_dot_update.hdf5",
    "test/tensor/data/test_dense_tensor_qr.hdf5",
    "test/tensor/data/test_dense_tensor_rq.hdf5",
    "test/tensor/data/test_dense_tensor_svd.hdf5",
    "test/util/data/test_lanczos_iteration_d.hdf5",
    "test/util/data/test_eigensystem_krylov_symmetric.hdf5",
    "test/algorithm/data/test_retained_bond_indices.hdf5",
    "test/state/data/test_mps_split_tensor_svd.hdf5",
    "test/operator/data/test_heisenberg_xxz_1d_mpo.hdf5",
    "test/operator/data/test_fermi_hubbard_1d_mpo.hdf5",
]


def build_tool(tmp):
    exe = os.path.join(tmp, "h5_extract")
    subprocess.check_call(["gcc", "-O1", "-I/opt/conda/include", os.path.join(HERE, "h5_extract.c"),
                           "-o", exe, "-L/opt/conda/lib", "-lhdf5", "-Wl,-rpath,/opt/conda/lib"])
    return exe


def convert(exe, src, dst):
    with tempfile.TemporaryDirectory() as out:
        subprocess.check_call([exe, src, out])
        arrays = {}
        with open(os.path.join(out, "manifest.txt")) as f:
            for line in f:
                kind, name, tname, nd, *dims = line.split()
                dims = tuple(int(x) for x in dims[:int(nd)])
                raw = np.fromfile(os.path.join(out, f"{kind}_{name}.bin"),
                                  dtype={"f8": "<f8", "i8": "<i8", "c16": "<c16"}[tname])
                arrays[f"{kind}_{name}"] = raw.reshape(dims)
        np.savez_compressed(dst, **arrays)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    with tempfile.TemporaryDirectory() as tmp:
        exe = build_tool(tmp)
        for rel in FIXTURES:
            name = os.path.basename(rel).replace(".hdf5", ".npz")
            convert(exe, os.path.join(ref, rel), os.path.join(HERE, name))
            print("wrote", name, os.path.getsize(os.path.join(HERE, name)))


if __name__ == "__main__":
    main()
