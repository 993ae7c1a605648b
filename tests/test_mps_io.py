"""save_mps / load_mps (src/state/mps.c:1219-1461) on the CPU: the HDF5 layout of the reference
(root attributes nsites, qsite, qbond_i; datasets tensor_i holding the dense site tensors), read back
by an independent HDF5 dump tool, and exact round trips (the reference's test_save_mps,
test/state/test_mps.c:743-785, with a real double MPS)."""

import os
import shutil
import subprocess

import numpy as np
import pytest

import chemtensor_amd as ct
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
H5_INC, H5_LIB = "/opt/conda/include", "/opt/conda/lib"

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(ct.__file__), "libchemtensor_h5.so")),
                                reason="built without HDF5 support")


def reference_test_mps():
    # test_save_mps: nsites 7, d 4, qsite {0, 2, 0, -1}, sector 1, max_vdim 14, seed 43
    return ct.MPS.random(7, 4, [0, 2, 0, -1], 1, 14, 43)


def dump(tmp_path, filename):
    """root attributes and datasets of an HDF5 file through tests/golden/h5_extract.c (its own HDF5
    reader), as {"A_name" | "D_name": array}"""
    if shutil.which("gcc") is None or not os.path.exists(os.path.join(H5_INC, "hdf5.h")):
        pytest.skip("no C compiler or HDF5 headers for the dump tool")
    exe = tmp_path / "h5_extract"
    if not exe.exists():
        subprocess.run(["gcc", "-O1", "-I", H5_INC, os.path.join(HERE, "golden", "h5_extract.c"), "-o", str(exe),
                        "-L", H5_LIB, "-lhdf5", "-Wl,-rpath," + H5_LIB], check=True)
    out = tmp_path / ("dump_" + os.path.basename(filename))
    out.mkdir()
    subprocess.run([str(exe), filename, str(out)], check=True, timeout=60)
    arrays = {}
    with open(out / "manifest.txt") as f:
        for line in f:
            kind, name, tname, nd, *dims = line.split()
            shape = tuple(int(x) for x in dims[:int(nd)])
            raw = np.fromfile(out / f"{kind}_{name}.bin", dtype={"f8": "<f8", "i8": "<i8", "c16": "<c16"}[tname])
            arrays[f"{kind}_{name}"] = (tname, raw.reshape(shape))
    return arrays


def test_save_mps_layout_matches_reference(tmp_path):
    psi = reference_test_mps()
    fn = str(tmp_path / "mps.hdf5")
    ct.save_mps(fn, psi)
    a = dump(tmp_path, fn)
    n = psi.nsites
    bonds = psi.bond_dims()
    assert a["A_nsites"][0] == "i8" and a["A_nsites"][1].shape == () and int(a["A_nsites"][1]) == n
    assert a["A_qsite"][1].tolist() == [0, 2, 0, -1]
    want = {"A_nsites", "A_qsite"} | {f"A_qbond_{i}" for i in range(n + 1)} | {f"D_tensor_{i}" for i in range(n)}
    assert set(a) == want
    for i in range(n + 1):
        site = orc.BST(psi.site_ptr(min(i, n - 1)), owned=False)
        q = site.meta()[2][0 if i < n else 2]
        assert a[f"A_qbond_{i}"][1].tolist() == q.tolist()
        assert len(q) == bonds[i]
    for i in range(n):
        tname, t = a[f"D_tensor_{i}"]
        assert tname == "f8"
        assert t.shape == (bonds[i], 4, bonds[i + 1])
        # the dense site tensor of the oracle's block-sparse to dense conversion, entry for entry
        assert np.array_equal(t, orc.BST(psi.site_ptr(i), owned=False).to_dense())


def test_save_load_round_trip_exact(tmp_path):
    psi = reference_test_mps()
    fn = str(tmp_path / "mps.hdf5")
    ct.save_mps(fn, psi)
    got = ct.load_mps(fn)
    assert got.nsites == psi.nsites and got.bond_dims() == psi.bond_dims()
    assert list(got.s.qsite[i] for i in range(4)) == [0, 2, 0, -1]
    for i in range(psi.nsites):
        p, q = orc.BST(psi.site_ptr(i), owned=False), orc.BST(got.site_ptr(i), owned=False)
        d1, a1, q1 = p.meta()
        d2, a2, q2 = q.meta()
        assert list(d1) == list(d2) and list(a1) == list(a2)
        assert all(np.array_equal(x, y) for x, y in zip(q1, q2))
        assert np.array_equal(p.serialize(), q.serialize())
    # saving the loaded MPS again gives the same datasets and attributes
    fn2 = str(tmp_path / "again.hdf5")
    ct.save_mps(fn2, got)
    a, b = dump(tmp_path, fn), dump(tmp_path, fn2)
    assert set(a) == set(b)
    assert all(np.array_equal(a[k][1], b[k][1]) for k in a)


def test_load_mps_errors(tmp_path):
    with pytest.raises(RuntimeError):
        ct.load_mps(str(tmp_path / "missing.hdf5"))
    bad = tmp_path / "not_hdf5.hdf5"
    bad.write_bytes(b"not an HDF5 file")
    with pytest.raises(RuntimeError):
        ct.load_mps(str(bad))


def test_saved_tensors_contract_to_state_vector(tmp_path):
    """a larger MPS (Hubbard sites, several sectors): the saved dense tensors contract to the oracle's
    state vector of the MPS"""
    mpo = ct.MPO.hubbard(6, 1.0, 4.0, 0.0)
    psi = ct.MPS.random(6, 4, mpo.qsite, ct.encode_pair(6, 0), 24, 5)
    fn = str(tmp_path / "h.hdf5")
    ct.save_mps(fn, psi)
    a = dump(tmp_path, fn)
    t = a["D_tensor_0"][1]
    for i in range(1, 6):
        s = a[f"D_tensor_{i}"][1]
        t = np.tensordot(t, s, axes=(2, 0)).reshape(t.shape[0], t.shape[1] * s.shape[1], s.shape[2])
    vec = np.zeros(4 ** 6)
    assert orc.lib().orc_h_mps_vector(psi.ptr, orc.dptr(vec), vec.size) == vec.size
    assert np.max(np.abs(t.reshape(-1) - vec)) < 1e-14
