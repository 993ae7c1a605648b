"""The reference's Python entry points for the DMRG path, over this library's C ABI.

    import chemtensor_amd.compat as chemtensor
    hamiltonian = chemtensor.construct_fermi_hubbard_1d_mpo(6, 1.0, 4.0, 1.5)
    psi, en_sweeps, entropy = chemtensor.dmrg(hamiltonian, num_sweeps=4, maxiter_lanczos=5, tol_split=1e-8,
                                              max_vdim=32, qnum_sector=chemtensor.encode_quantum_number_pair(7, 1))
    psi.bond_dims; psi.to_statevector(); hamiltonian.to_matrix()

Names, argument meanings, defaults, return values and exceptions follow python/pymodule.c of the
reference (Py_dmrg :3238-3358, the MPS / MPO getters and to_statevector / to_matrix, the quantum
number pair helpers of include/tensor/qnumber.h:47-70). Only real FP64 tensors exist here.
"""

import math
import os
import warnings

import numpy as np

import chemtensor_amd as _ct


def encode_quantum_number_pair(qa, qb):
    """qnumber.h:47-50: (qa << 16) + qb"""
    return (int(qa) << 16) + int(qb)


def decode_quantum_number_pair(qnum):
    """qnumber.h:56-70: the inverse of encode_quantum_number_pair for |qb| < 2^15"""
    q1 = int(math.fmod(int(qnum), 1 << 16))     # C remainder (sign of the dividend)
    if q1 >= (1 << 15):
        q1 -= 1 << 16
    elif q1 < -(1 << 15):
        q1 += 1 << 16
    q0 = (int(qnum) - q1) >> 16
    return q0, q1


def get_max_openmp_threads():
    """host threads the library lends to block factorisations of one split (no OpenMP here)"""
    return min(16, os.cpu_count() or 1)


def _to_dense(t):
    """dense numpy array of a struct block_sparse_tensor (blocks placed at their logical indices)"""
    nd = t.ndim
    dims = [int(t.dim_logical[i]) for i in range(nd)]
    dblk = [int(t.dim_blocks[i]) for i in range(nd)]
    qlog = [np.array([t.qnums_logical[i][j] for j in range(dims[i])], dtype=np.int64) for i in range(nd)]
    qblk = [[int(t.qnums_blocks[i][j]) for j in range(dblk[i])] for i in range(nd)]
    out = np.zeros(dims)
    for k in range(int(np.prod(dblk))):
        if not t.blocks[k]:
            continue
        blk = t.blocks[k].contents
        ib = np.unravel_index(k, dblk)
        idx = [np.nonzero(qlog[i] == qblk[i][ib[i]])[0] for i in range(nd)]
        shape = [int(blk.dim[i]) for i in range(nd)]
        data = np.ctypeslib.as_array(_ct.C.cast(blk.data, _ct.DP), shape=(int(np.prod(shape)),)).reshape(shape)
        out[np.ix_(*idx)] = data
    return out


class MPS:
    """matrix product state returned by dmrg (python/pymodule.c PyMPS)"""

    def __init__(self, mps):
        self._m = mps

    @property
    def nsites(self):
        return self._m.nsites

    @property
    def d(self):
        return int(self._m.s.d)

    @property
    def qsite(self):
        return [int(self._m.s.qsite[i]) for i in range(self.d)]

    @property
    def bond_dims(self):
        return self._m.bond_dims()

    def site(self, i):
        """dense site tensor i (left bond, physical, right bond)"""
        return _to_dense(self._m.s.a[i])

    def to_statevector(self):
        """mps_to_statevector (mps.c:1186-1210): the contraction over all bonds, first site slowest"""
        t = self.site(0)
        for i in range(1, self.nsites):
            a = self.site(i)
            t = np.tensordot(t, a, axes=(2, 0)).reshape(t.shape[0], t.shape[1] * a.shape[1], a.shape[2])
        return t.reshape(-1)

    def norm(self):
        e = np.ones((1, 1))
        for i in range(self.nsites):
            a = self.site(i)
            e = np.einsum("ab,asc,bsd->cd", e, a, a, optimize=True)
        return math.sqrt(max(e.reshape(-1)[0], 0.0))


class MPO:
    """matrix product operator (python/pymodule.c PyMPO)"""

    def __init__(self, mpo):
        self._m = mpo

    @property
    def nsites(self):
        return self._m.nsites

    @property
    def d(self):
        return int(self._m.d)

    @property
    def qsite(self):
        return [int(x) for x in self._m.qsite]

    @property
    def bond_dims(self):
        s = self._m.s
        dims = [int(s.a[0].dim_logical[0])]
        for i in range(s.nsites):
            dims.append(int(s.a[i].dim_logical[3]))
        return dims

    def site(self, i):
        """dense MPO tensor i (left bond, physical out, physical in, right bond)"""
        return _to_dense(self._m.s.a[i])

    def to_matrix(self):
        """mpo_to_matrix (mpo.c:287-311): merged pairwise, physical output axis first"""
        t = self.site(0)
        for i in range(1, self.nsites):
            w = self.site(i)
            t = np.einsum("aoim,mpqb->aopiqb", t, w, optimize=True)
            t = t.reshape(t.shape[0], t.shape[1] * t.shape[2], t.shape[3] * t.shape[4], t.shape[5])
        return t.reshape(t.shape[1], t.shape[2])


def construct_heisenberg_xxz_1d_mpo(nsites, J, D, h):
    """hamiltonian.c:102-155: XXZ chain H = sum J (X X + Y Y) + D Z Z - h Z (open boundaries)"""
    return MPO(_ct.MPO.xxz(nsites, J, D, h))


def construct_fermi_hubbard_1d_mpo(nsites, t, u, mu):
    """hamiltonian.c:240-366: Fermi-Hubbard chain, quantum numbers (particle number, 2 Sz)"""
    return MPO(_ct.MPO.hubbard(nsites, t, u, mu))


def construct_spin_molecular_hamiltonian_mpo(tkin, vint, optimize=False):
    """hamiltonian.c:2755-2990: H = sum t_ij a+_{i s} a_{j s} + 1/2 sum v_ijkl a+_{i s} a+_{j u} a_{l u} a_{k s}
    (physicists' convention); 'optimize' is accepted for the reference's signature (the bond basis
    here is always the finite-state machine over the terms)"""
    tkin = np.asarray(tkin, dtype=np.float64)
    vint = np.asarray(vint, dtype=np.float64)
    n = tkin.shape[0] if tkin.ndim == 2 else -1
    if tkin.ndim != 2 or tkin.shape != (n, n):
        raise ValueError("'tkin' must be a square matrix")
    if vint.shape != (n, n, n, n):
        raise ValueError("'vint' must be a degree-4 tensor with the dimensions of 'tkin'")
    return MPO(_ct.MPO.spin_molecular(tkin, vint))


def dmrg(mpo, num_sweeps=5, maxiter_lanczos=20, tol_split=1e-10, max_vdim=256, qnum_sector=0, rng_seed=42):
    """Two-site DMRG on the GPU: (psi, en_sweeps, entropy), as chemtensor.dmrg (pymodule.c:3238-3358).

    The start state is construct_random_mps seeded with rng_seed in the sector qnum_sector."""
    if not isinstance(mpo, MPO):
        raise SyntaxError("error parsing input; syntax: dmrg(mpo, num_sweeps=5, maxiter_lanczos=20, tol_split=1e-10, "
                          "max_vdim=256, qnum_sector=0, rng_seed=42)")
    if num_sweeps <= 0:
        raise ValueError(f"'num_sweeps' must be a positive integer, received {num_sweeps}")
    if maxiter_lanczos <= 0:
        raise ValueError(f"'maxiter_lanczos' must be a positive integer, received {maxiter_lanczos}")
    if tol_split < 0 or tol_split > 1:
        raise ValueError(f"'tol_split' must be in the interval [0, 1], received {tol_split:g}")
    if max_vdim <= 0:
        raise ValueError(f"'max_vdim' must be a positive integer, received {max_vdim}")
    m = mpo._m
    psi = _ct.MPS.random(m.nsites, m.d, m.qsite, qnum_sector, max_vdim, rng_seed)
    if MPS(psi).norm() == 0:
        warnings.warn("initial random MPS has norm zero (possibly due to mismatching quantum numbers)", RuntimeWarning)
    try:
        en, ent = _ct.dmrg(m, psi, num_sweeps, maxiter_lanczos, tol_split, max_vdim)
    except RuntimeError:
        raise RuntimeError("two-site DMRG failed internally (likely since Krylov subspace method did not converge)")
    return MPS(psi), [float(x) for x in en], [float(x) for x in ent]
