"""ctypes front end of the CPU FP64 oracle (TEST INFRASTRUCTURE ONLY).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker of
the GPU product. The product package (chemtensor_amd) never imports this module.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libct_oracle.so")
REF_RNG_PATH = os.path.join(HERE, "_ref", "libct_ref_rng.so")

ct_long = C.c_int64
P = C.c_void_p
DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_int64)
IP = C.POINTER(C.c_int32)
UP = C.POINTER(C.c_uint64)

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)
    if os.path.isdir("/root/reference/src"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"], stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_h_bst_new": (P, [C.c_int, LP, IP, IP]),
            "orc_h_bst_free": (None, [P]),
            "orc_h_bst_ndim": (C.c_int, [P]),
            "orc_h_bst_dims": (None, [P, LP, IP]),
            "orc_h_bst_qnums": (None, [P, C.c_int, IP]),
            "orc_h_bst_num_blocks_present": (ct_long, [P]),
            "orc_h_bst_set_dense": (None, [P, DP]),
            "orc_h_bst_get_dense": (None, [P, DP]),
            "orc_h_bst_fill_random": (None, [P, C.c_uint64, C.c_double]),
            "orc_h_dot": (P, [P, C.c_int, P, C.c_int, C.c_int]),
            "orc_h_transpose": (P, [IP, P]),
            "orc_h_flatten": (P, [P, C.c_int, C.c_int]),
            "orc_h_split": (P, [P, C.c_int, ct_long, ct_long, C.c_int, C.c_int, IP, IP]),
            "orc_h_trace": (P, [P, C.c_int]),
            "orc_h_apply_local_hamiltonian": (P, [P, P, P, P]),
            "orc_h_copy": (P, [P]),
            "orc_h_mpo_xxz": (P, [C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_h_mpo_hubbard": (P, [C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_h_mpo_free": (None, [P]),
            "orc_h_mpo_site": (P, [P, C.c_int]),
            "orc_h_mpo_matrix": (ct_long, [P, DP, ct_long]),
            "orc_h_mpo_merge": (P, [P, P]),
            "orc_h_mps_random": (P, [C.c_int, ct_long, IP, C.c_int32, ct_long, C.c_uint64]),
            "orc_h_mps_free": (None, [P]),
            "orc_h_mps_site": (P, [P, C.c_int]),
            "orc_h_mps_vector": (ct_long, [P, DP, ct_long]),
            "orc_h_mps_orthonormalize": (C.c_double, [P, C.c_int]),
            "orc_h_dmrg": (C.c_int, [P, P, C.c_int, C.c_int, C.c_double, ct_long, DP, DP]),
            "orc_h_local_problem": (None, [P, P, C.c_int, C.POINTER(P)]),
            "orc_h_random_bond_qnums": (None, [C.c_int, ct_long, IP, C.c_int32, ct_long, C.c_uint64, LP, IP]),
            "orc_h_lanczos_dense": (None, [ct_long, DP, DP, C.c_int, DP, DP, DP, C.POINTER(C.c_int)]),
            "orc_h_eigensystem_dense": (C.c_int, [ct_long, DP, DP, C.c_int, C.c_int, DP, DP]),
            "orc_h_mps_split_svd": (C.c_int, [P, ct_long, ct_long, IP, IP, C.c_double, ct_long, C.c_int, C.c_int, C.POINTER(P), C.POINTER(P), DP]),
            "orc_h_mps_merge": (P, [P, P]),
            "orc_h_svd_dense": (C.c_int, [ct_long, ct_long, DP, DP, DP, DP]),
            "orc_h_qr_dense": (C.c_int, [ct_long, ct_long, DP, DP, DP, C.c_int]),
            "orc_h_tridiag_eigh": (C.c_int, [C.c_int, DP, DP, DP]),
            "orc_h_retained_bond_indices": (ct_long, [DP, ct_long, C.c_double, C.c_int, ct_long, LP, DP]),
            "orc_h_rng_draws": (None, [C.c_uint64, C.c_int, ct_long, C.c_uint64, DP, UP]),
            "orc_block_sparse_tensor_num_elements_blocks": (ct_long, [P]),
            "orc_block_sparse_tensor_serialize_entries": (None, [P, DP]),
            "orc_block_sparse_tensor_deserialize_entries": (None, [P, DP]),
            "orc_block_sparse_tensor_norm2": (C.c_double, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def dptr(a):
    return a.ctypes.data_as(DP)


def lptr(a):
    return a.ctypes.data_as(LP)


def iptr(a):
    return a.ctypes.data_as(IP)


class BST:
    """Owning (or borrowed) handle of an oracle block-sparse tensor."""

    def __init__(self, ptr, owned=True):
        self.ptr = ptr
        self.owned = owned

    def __del__(self):
        if self.owned and self.ptr and _lib is not None:
            _lib.orc_h_bst_free(self.ptr)
            self.ptr = None

    @staticmethod
    def zeros(dims, axis_dir, qnums):
        dims = np.asarray(dims, dtype=np.int64)
        ad = np.asarray(axis_dir, dtype=np.int32)
        qc = np.ascontiguousarray(np.concatenate([np.asarray(q, dtype=np.int32) for q in qnums]))
        return BST(lib().orc_h_bst_new(len(dims), lptr(dims), iptr(ad), iptr(qc)))

    @staticmethod
    def from_dense(dense, axis_dir, qnums):
        dense = np.ascontiguousarray(dense, dtype=np.float64)
        t = BST.zeros(dense.shape, axis_dir, qnums)
        lib().orc_h_bst_set_dense(t.ptr, dptr(dense))
        return t

    @property
    def ndim(self):
        return lib().orc_h_bst_ndim(self.ptr)

    def meta(self):
        nd = self.ndim
        dims = np.zeros(max(nd, 1), dtype=np.int64)
        ad = np.zeros(max(nd, 1), dtype=np.int32)
        lib().orc_h_bst_dims(self.ptr, lptr(dims), iptr(ad))
        qn = []
        for i in range(nd):
            q = np.zeros(dims[i], dtype=np.int32)
            lib().orc_h_bst_qnums(self.ptr, i, iptr(q))
            qn.append(q)
        return dims[:nd].copy(), ad[:nd].copy(), qn

    def to_dense(self):
        dims, _, _ = self.meta()
        out = np.zeros(int(np.prod(dims)) if len(dims) else 1, dtype=np.float64)
        lib().orc_h_bst_get_dense(self.ptr, dptr(out))
        return out.reshape(tuple(dims))

    def nelem(self):
        return lib().orc_block_sparse_tensor_num_elements_blocks(self.ptr)

    def serialize(self):
        out = np.zeros(self.nelem(), dtype=np.float64)
        lib().orc_block_sparse_tensor_serialize_entries(self.ptr, dptr(out))
        return out

    def deserialize(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nelem()
        lib().orc_block_sparse_tensor_deserialize_entries(self.ptr, dptr(v))

    def fill_random(self, seed, alpha=1.0):
        lib().orc_h_bst_fill_random(self.ptr, seed, alpha)
        return self

    def num_blocks(self):
        return lib().orc_h_bst_num_blocks_present(self.ptr)

    def norm(self):
        return lib().orc_block_sparse_tensor_norm2(self.ptr)

    def copy(self):
        return BST(lib().orc_h_copy(self.ptr))


LEADING, TRAILING = 0, 1
AXIS_IN, AXIS_OUT = -1, 1


def dot(s, axs, t, axt, ndim_mult):
    return BST(lib().orc_h_dot(s.ptr, axs, t.ptr, axt, ndim_mult))


def transpose(perm, t):
    p = np.asarray(perm, dtype=np.int32)
    return BST(lib().orc_h_transpose(iptr(p), t.ptr))


def flatten(t, i_ax, new_dir):
    return BST(lib().orc_h_flatten(t.ptr, i_ax, new_dir))


def split(t, i_ax, dims2, dirs2, qnums2):
    q0 = np.asarray(qnums2[0], dtype=np.int32)
    q1 = np.asarray(qnums2[1], dtype=np.int32)
    return BST(lib().orc_h_split(t.ptr, i_ax, int(dims2[0]), int(dims2[1]), int(dirs2[0]), int(dirs2[1]), iptr(q0), iptr(q1)))


def partial_trace(t, ndim_trace):
    return BST(lib().orc_h_trace(t.ptr, ndim_trace))


def apply_local_hamiltonian(a, w, l, r):
    return BST(lib().orc_h_apply_local_hamiltonian(a.ptr, w.ptr, l.ptr, r.ptr))


def encode_pair(qa, qb):
    return int(qa) * 65536 + int(qb)


class MPO:
    def __init__(self, ptr, nsites, d, qsite):
        self.ptr = ptr
        self.nsites = nsites
        self.d = d
        self.qsite = np.asarray(qsite, dtype=np.int32)

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.orc_h_mpo_free(self.ptr)
            self.ptr = None

    @staticmethod
    def xxz(nsites, J=1.0, D=0.5, h=0.0):
        return MPO(lib().orc_h_mpo_xxz(nsites, J, D, h), nsites, 2, [1, -1])

    @staticmethod
    def hubbard(nsites, t=1.0, u=4.0, mu=0.0):
        q = [encode_pair(0, 0), encode_pair(1, -1), encode_pair(1, 1), encode_pair(2, 0)]
        return MPO(lib().orc_h_mpo_hubbard(nsites, t, u, mu), nsites, 4, q)

    def site(self, i):
        return BST(lib().orc_h_mpo_site(self.ptr, i), owned=False)

    def to_matrix(self):
        n = self.d ** self.nsites
        out = np.zeros(n * n)
        m = lib().orc_h_mpo_matrix(self.ptr, dptr(out), n * n)
        assert m == n
        return out.reshape(n, n)


class MPS:
    def __init__(self, ptr, nsites):
        self.ptr = ptr
        self.nsites = nsites

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.orc_h_mps_free(self.ptr)
            self.ptr = None

    @staticmethod
    def random(nsites, d, qsite, sector, max_vdim, seed):
        q = np.asarray(qsite, dtype=np.int32)
        return MPS(lib().orc_h_mps_random(nsites, d, iptr(q), sector, max_vdim, seed), nsites)

    def site(self, i):
        return BST(lib().orc_h_mps_site(self.ptr, i), owned=False)

    def to_vector(self, d):
        n = d ** self.nsites
        out = np.zeros(n)
        m = lib().orc_h_mps_vector(self.ptr, dptr(out), n)
        assert m == n
        return out

    def bond_dims(self):
        dims = [self.site(0).meta()[0][0]]
        for i in range(self.nsites):
            dims.append(self.site(i).meta()[0][2])
        return [int(x) for x in dims]

    def orthonormalize(self, mode=1):
        return lib().orc_h_mps_orthonormalize(self.ptr, mode)


def dmrg(mpo, psi, num_sweeps, maxiter_lanczos, tol_split, max_vdim):
    en = np.zeros(num_sweeps)
    ent = np.zeros(mpo.nsites - 1)
    ret = lib().orc_h_dmrg(mpo.ptr, psi.ptr, num_sweeps, maxiter_lanczos, tol_split, max_vdim, dptr(en), dptr(ent))
    return ret, en, ent


def local_problem(mpo, psi, i):
    """(theta, h2, L, R) at bond pair (i, i+1) of a right-normalised MPS (psi is modified)."""
    arr = (P * 4)()
    lib().orc_h_local_problem(mpo.ptr, psi.ptr, i, arr)
    return tuple(BST(arr[k]) for k in range(4))


def random_bond_qnums(nsites, d, qsite, sector, max_vdim, seed):
    q = np.asarray(qsite, dtype=np.int32)
    dims = np.zeros(nsites + 1, dtype=np.int64)
    qb = np.zeros((nsites + 1) * max(max_vdim, 1) + 8, dtype=np.int32)
    lib().orc_h_random_bond_qnums(nsites, d, iptr(q), sector, max_vdim, seed, lptr(dims), iptr(qb))
    out, off = [], 0
    for n in dims:
        out.append(qb[off:off + n].copy())
        off += n
    return out


def mpo_bond_qnums(mpo, i):
    """Quantum numbers of MPO virtual bond i (0..nsites)."""
    if i < mpo.nsites:
        return mpo.site(i).meta()[2][0]
    return mpo.site(mpo.nsites - 1).meta()[2][3]


def matvec_problem(model, nsites, chi, seed):
    """Synthetic H_eff matvec inputs at the centre bond pair: bond quantum numbers as drawn by
    construct_random_mps(seed), theta / left / right blocks N(0,1) with seeds seed+1..seed+3,
    h2 the merged MPO tensors of the model (XXZ J=1 D=0.5 h=0; Hubbard t=1 U=4 mu=0)."""
    if model == "xxz":
        mpo = MPO.xxz(nsites, 1.0, 0.5, 0.0)
        sector = 0
    else:
        mpo = MPO.hubbard(nsites, 1.0, 4.0, 0.0)
        sector = encode_pair(nsites, 0)
    theta, h2, left, right = matvec_problem_mpo(mpo, sector, chi, seed)
    return theta, h2, left, right, mpo


class MPOView(MPO):
    """An MPO struct owned elsewhere (the product library's, which has the reference's layout)."""

    def __del__(self):
        self.ptr = None


def matvec_problem_mpo(mpo, sector, chi, seed):
    """matvec_problem's inputs for any MPO: centre bond pair i = nsites/2 - 1, bond quantum numbers
    as construct_random_mps(mpo's site quantum numbers, sector, chi, seed) draws them, theta / left /
    right N(0,1) with seeds seed+1..seed+3, h2 the merged MPO pair."""
    nsites = mpo.nsites
    d = mpo.d
    qb = random_bond_qnums(nsites, d, mpo.qsite, sector, chi, seed)
    i = nsites // 2 - 1
    qp = (mpo.qsite[:, None] + mpo.qsite[None, :]).reshape(-1)
    theta = BST.zeros([len(qb[i]), d * d, len(qb[i + 2])], [AXIS_OUT, AXIS_OUT, AXIS_IN], [qb[i], qp, qb[i + 2]]).fill_random(seed + 1)
    h2 = BST(lib().orc_h_mpo_merge(mpo.site(i).ptr, mpo.site(i + 1).ptr))
    ml = mpo_bond_qnums(mpo, i)
    mr = mpo_bond_qnums(mpo, i + 2)
    zero = np.zeros(1, dtype=np.int32)
    left = BST.zeros([1, len(qb[i]), len(ml), len(qb[i])], [AXIS_OUT, AXIS_IN, AXIS_IN, AXIS_OUT], [zero, qb[i], ml, qb[i]]).fill_random(seed + 2)
    right = BST.zeros([len(qb[i + 2]), len(mr), len(qb[i + 2]), 1], [AXIS_OUT, AXIS_OUT, AXIS_IN, AXIS_IN], [qb[i + 2], mr, qb[i + 2], zero]).fill_random(seed + 3)
    return theta, h2, left, right


def ref_rng_lib():
    """The reference's own RNG (src/util/rng.c + pcg_basic.c) compiled into oracle/_ref, if built."""
    if not os.path.exists(REF_RNG_PATH):
        return None
    return C.CDLL(REF_RNG_PATH)
