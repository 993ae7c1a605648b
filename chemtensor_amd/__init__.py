"""chemtensor_amd: MI355X drop-in for chemtensor's two-site DMRG matvec path.

The product is the C-ABI library libchemtensor_gpu.so (include/chemtensor_gpu.h). This module
is a thin ctypes front end for Python callers (tests, bench); it mirrors the names and argument
meaning of the reference's C API. It refuses to run without the compiled HIP library: there is
no CPU fallback.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libchemtensor_gpu.so")
REPO = os.path.dirname(HERE)
HEADER = os.path.join(REPO, "include", "chemtensor_gpu.h")

ct_long = C.c_int64
P = C.c_void_p
DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_int64)
IP = C.POINTER(C.c_int32)
FP = C.POINTER(C.c_float)

LEADING, TRAILING = 0, 1
AXIS_IN, AXIS_OUT = -1, 1
PLAIN_FP64, GROUPED_FP64, GROUPED_SPLIT, FUSED_FP64, PAIR_FP64 = 0, 1, 2, 3, 4

CT_ERR_HIP, CT_ERR_UNSUPPORTED, CT_ERR_NOMEM, CT_ERR_ARG = -100, -101, -102, -103
ERRORS = {-1: "Krylov breakdown / factorisation failure", -2: "eigensolver failure", -100: "HIP runtime error",
          -101: "unsupported input", -102: "out of device memory / budget", -103: "invalid argument"}


class DenseTensor(C.Structure):
    _fields_ = [("data", P), ("dim", LP), ("dtype", C.c_int), ("ndim", C.c_int)]


class BlockSparseTensor(C.Structure):
    _fields_ = [("blocks", C.POINTER(C.POINTER(DenseTensor))), ("dim_blocks", LP), ("dim_logical", LP),
                ("axis_dir", IP), ("qnums_blocks", C.POINTER(IP)), ("qnums_logical", C.POINTER(IP)),
                ("dtype", C.c_int), ("ndim", C.c_int)]


class MPSStruct(C.Structure):
    _fields_ = [("a", C.POINTER(BlockSparseTensor)), ("qsite", IP), ("d", ct_long), ("nsites", C.c_int)]


MPOStruct = MPSStruct

_lib = None


# include/util/krylov.h:10
LINEAR_FUNC_D = C.CFUNCTYPE(None, ct_long, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def lib():
    """The compiled HIP library; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with chemtensor_amd.build.build() (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        bst = C.POINTER(BlockSparseTensor)
        sig = {
            "allocate_block_sparse_tensor": (None, [C.c_int, C.c_int, LP, IP, C.POINTER(IP), bst]),
            "delete_block_sparse_tensor": (None, [bst]),
            "copy_block_sparse_tensor": (None, [P, bst]),
            "block_sparse_tensor_num_elements_blocks": (ct_long, [P]),
            "block_sparse_tensor_serialize_entries": (None, [P, P]),
            "block_sparse_tensor_deserialize_entries": (None, [P, P]),
            "block_sparse_tensor_norm2": (C.c_double, [P]),
            "block_sparse_tensor_dot": (None, [P, C.c_int, P, C.c_int, C.c_int, bst]),
            "block_sparse_tensor_transpose": (None, [IP, P, bst]),
            "block_sparse_tensor_cyclic_partial_trace": (None, [P, C.c_int, bst]),
            "apply_local_hamiltonian": (None, [P, P, P, P, bst]),
            "mpo_merge_tensor_pair": (None, [P, P, bst]),
            "ct_matvec_plan_create": (C.c_int, [P, P, P, P, C.c_int, C.POINTER(P)]),
            "ct_matvec_plan_destroy": (C.c_int, [P]),
            "ct_matvec_plan_check": (C.c_int, [P, P, P, P]),
            "ct_matvec_pair_check": (C.c_int, [P, P, P, P, C.c_int, DP]),
            "ct_matvec_pair_stats": (C.c_int, [P, P, P, P, C.c_int, DP]),
            "ct_matvec_plan_dim": (ct_long, [P]),
            "ct_matvec_plan_flops": (C.c_double, [P]),
            "ct_matvec_plan_bytes": (C.c_double, [P]),
            "ct_matvec_plan_stats": (C.c_int, [P, DP]),
            "ct_matvec_plan_stage_flops": (C.c_int, [P, DP]),
            "ct_matvec_plan_stage_bytes": (C.c_int, [P, DP]),
            "ct_matvec_plan_forms": (C.c_int, [P, IP, FP]),
            "ct_matvec_apply": (C.c_int, [P, P, P]),
            "ct_matvec_apply_noalloc": (C.c_int, [P, P, P]),
            "ct_matvec_plan_stage_time": (C.c_int, [P, C.c_int, P, P, C.c_int, FP]),
            "ct_matvec_apply_timed": (C.c_int, [P, P, P, FP]),
            "ct_matvec_plan_stream": (P, [P]),
            "ct_lanczos_ground_state": (C.c_int, [P, P, C.c_int, DP, P]),
            "ct_matvec_batch_create": (C.c_int, [C.c_int, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
            "ct_matvec_batch_create_mode": (C.c_int, [C.c_int, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.c_int, C.POINTER(P)]),
            "ct_matvec_batch_destroy": (C.c_int, [P]),
            "ct_matvec_batch_dim": (ct_long, [P]),
            "ct_matvec_batch_offsets": (C.c_int, [P, LP]),
            "ct_matvec_batch_flops": (C.c_double, [P]),
            "ct_matvec_batch_bytes": (C.c_double, [P]),
            "ct_matvec_batch_form": (C.c_int, [P, IP, FP]),
            "ct_matvec_batch_stream": (P, [P]),
            "ct_matvec_batch_apply": (C.c_int, [P, P, P]),
            "ct_lanczos_ground_state_batch": (C.c_int, [P, P, C.c_int, DP, IP, P]),
            "ct_serve_create": (C.c_int, [C.c_double, C.c_int, C.POINTER(P)]),
            "ct_serve_submit": (C.c_int, [P, P, C.POINTER(C.c_int64)]),
            "ct_serve_poll": (C.c_int, [P, C.c_int64, C.c_int, P]),
            "ct_serve_energies": (C.c_int, [P, C.c_int64, DP, C.c_int]),
            "ct_serve_stats": (C.c_int, [P, DP]),
            "ct_serve_destroy": (C.c_int, [P]),
            "eigensystem_krylov_symmetric": (C.c_int, [ct_long, LINEAR_FUNC_D, P, DP, C.c_int, C.c_int, DP, DP]),
            "ct_device_alloc": (C.c_int, [C.POINTER(P), ct_long]),
            "ct_device_free": (C.c_int, [P]),
            "ct_memcpy_h2d": (C.c_int, [P, P, ct_long]),
            "ct_memcpy_d2h": (C.c_int, [P, P, ct_long]),
            "ct_device_synchronize": (C.c_int, []),
            "ct_device_count": (C.c_int, []),
            "ct_event_create": (C.c_int, [C.POINTER(P)]),
            "ct_event_record": (C.c_int, [P, P]),
            "ct_event_elapsed_ms": (C.c_int, [P, P, FP]),
            "ct_event_destroy": (C.c_int, [P]),
            "ct_stream_synchronize": (C.c_int, [P]),
            "ct_device_copy": (C.c_int, [P, P, ct_long]),
            "ct_generate_matvec_problem": (C.c_int, [C.c_int, C.c_int, ct_long, C.c_uint64, bst, bst, bst, bst]),
            "ct_generate_matvec_problem_mpo": (C.c_int, [C.POINTER(MPOStruct), C.c_int32, ct_long, C.c_uint64, bst, bst, bst, bst]),
            "ct_construct_heisenberg_xxz_1d_mpo": (None, [C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(MPOStruct)]),
            "ct_construct_fermi_hubbard_1d_mpo": (None, [C.c_int, C.c_double, C.c_double, C.c_double, C.POINTER(MPOStruct)]),
            "ct_construct_spin_molecular_hamiltonian_mpo": (None, [C.c_int, DP, DP, C.POINTER(MPOStruct)]),
            "delete_mpo": (None, [C.POINTER(MPOStruct)]),
            "ct_dense_svd_device": (C.c_int, [ct_long, ct_long, DP, DP, DP, DP]),
            "ct_dense_gram_factor_device": (C.c_int, [ct_long, ct_long, DP, C.c_int, DP, DP, DP]),
            "delete_mps": (None, [C.POINTER(MPSStruct)]),
            "save_mps": (C.c_int, [C.c_char_p, C.POINTER(MPSStruct)]),
            "load_mps": (C.c_int, [C.c_char_p, C.POINTER(MPSStruct)]),
            "seed_rng_state": (None, [C.c_uint64, P]),
            "construct_random_mps": (None, [C.c_int, C.c_int, ct_long, IP, C.c_int32, ct_long, P, C.POINTER(MPSStruct)]),
            "dmrg_twosite": (C.c_int, [C.POINTER(MPOStruct), C.c_int, C.c_int, C.c_double, ct_long, C.POINTER(MPSStruct), DP, DP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(ret, what):
    if ret != 0:
        raise RuntimeError(f"{what} failed: {ERRORS.get(ret, ret)}")


def gpu_available():
    try:
        return lib().ct_device_count() > 0
    except (OSError, RuntimeError):
        return False


def _ptr(t):
    """Address of a block-sparse tensor struct (product HostTensor or any object with .ptr)."""
    if isinstance(t, HostTensor):
        return C.cast(C.pointer(t.s), P)
    return t.ptr


class HostTensor:
    """A struct block_sparse_tensor allocated by the product library (freed with its delete)."""

    def __init__(self):
        self.s = BlockSparseTensor()
        self.live = False

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), P)

    @property
    def sptr(self):
        return C.pointer(self.s)

    def __del__(self):
        if self.live and _lib is not None:
            _lib.delete_block_sparse_tensor(C.pointer(self.s))
            self.live = False

    @staticmethod
    def zeros(dims, axis_dir, qnums):
        t = HostTensor()
        nd = len(dims)
        dims = np.asarray(dims, dtype=np.int64)
        ad = np.asarray(axis_dir, dtype=np.int32)
        qarrs = [np.ascontiguousarray(q, dtype=np.int32) for q in qnums]
        qp = (IP * nd)(*[q.ctypes.data_as(IP) for q in qarrs])
        lib().allocate_block_sparse_tensor(1, nd, dims.ctypes.data_as(LP), ad.ctypes.data_as(IP), qp, t.sptr)
        t.live = True
        return t

    def nelem(self):
        return lib().block_sparse_tensor_num_elements_blocks(self.ptr)

    def serialize(self):
        out = np.zeros(self.nelem())
        lib().block_sparse_tensor_serialize_entries(self.ptr, out.ctypes.data_as(P))
        return out

    def deserialize(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert v.size == self.nelem()
        lib().block_sparse_tensor_deserialize_entries(self.ptr, v.ctypes.data_as(P))

    def meta(self):
        nd = self.s.ndim
        dims = [int(self.s.dim_logical[i]) for i in range(nd)]
        ad = [int(self.s.axis_dir[i]) for i in range(nd)]
        qn = [np.array([self.s.qnums_logical[i][j] for j in range(dims[i])], dtype=np.int32) for i in range(nd)]
        return dims, ad, qn


def dot(s, axs, t, axt, ndim_mult):
    r = HostTensor()
    lib().block_sparse_tensor_dot(_ptr(s), axs, _ptr(t), axt, ndim_mult, r.sptr)
    r.live = True
    return r


def transpose(perm, t):
    r = HostTensor()
    p = np.asarray(perm, dtype=np.int32)
    lib().block_sparse_tensor_transpose(p.ctypes.data_as(IP), _ptr(t), r.sptr)
    r.live = True
    return r


def cyclic_partial_trace(t, ndim_trace):
    """block_sparse_tensor_cyclic_partial_trace (bst.c:1560) on the GPU"""
    r = HostTensor()
    lib().block_sparse_tensor_cyclic_partial_trace(_ptr(t), int(ndim_trace), r.sptr)
    r.live = True
    return r


def apply_local_hamiltonian(a, w, l, r):
    b = HostTensor()
    lib().apply_local_hamiltonian(_ptr(a), _ptr(w), _ptr(l), _ptr(r), b.sptr)
    b.live = True
    return b


def dense_svd_device(a):
    """u, s, vh of a dense matrix through the device block-SVD kernel of the DMRG split"""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m, n = a.shape
    k = min(m, n)
    u = np.zeros((m, k)); s = np.zeros(k); vh = np.zeros((k, n))
    ret = lib().ct_dense_svd_device(m, n, a.ctypes.data_as(DP), u.ctypes.data_as(DP), s.ctypes.data_as(DP), vh.ctypes.data_as(DP))
    check(ret, "ct_dense_svd_device")
    return u, s, vh


def dense_gram_factor_device(a, left):
    """(u, s, vh) of a dense matrix through the device Gram factorisation of the DMRG split: u
    orthonormal and vh = s Vh when left, else vh orthonormal and u = U s"""
    a = np.ascontiguousarray(a, dtype=np.float64)
    m, n = a.shape
    k = min(m, n)
    u = np.zeros((m, k)); s = np.zeros(k); vh = np.zeros((k, n))
    ret = lib().ct_dense_gram_factor_device(m, n, a.ctypes.data_as(DP), 1 if left else 0, u.ctypes.data_as(DP), s.ctypes.data_as(DP),
                                            vh.ctypes.data_as(DP))
    check(ret, "ct_dense_gram_factor_device")
    return u, s, vh


def mpo_merge_tensor_pair(a0, a1):
    r = HostTensor()
    lib().mpo_merge_tensor_pair(_ptr(a0), _ptr(a1), r.sptr)
    r.live = True
    return r


def matvec_pair_check(a, w, l, r, mfma_cap=0):
    """Host-only replay of the sector-pair tables (ct_matvec_pair_check): (status, y), status 0 when
    every index and write-once check holds, y the product computed by the replay."""
    n = lib().block_sparse_tensor_num_elements_blocks(_ptr(a))
    y = np.zeros(n, dtype=np.float64)
    st = lib().ct_matvec_pair_check(_ptr(a), _ptr(w), _ptr(l), _ptr(r), int(mfma_cap), y.ctypes.data_as(DP))
    return st, y


def matvec_pair_stats(a, w, l, r, mfma_cap=0):
    """Host-only statistics of the sector-pair tables (ct_matvec_pair_stats)."""
    st = np.zeros(6, dtype=np.float64)
    check(lib().ct_matvec_pair_stats(_ptr(a), _ptr(w), _ptr(l), _ptr(r), int(mfma_cap), st.ctypes.data_as(DP)), "ct_matvec_pair_stats")
    return dict(zip(["units", "partial_doubles", "cost_sum", "cost_max", "lds_max", "record_words_max"], st.tolist()))


def matvec_plan_check(a, w, l, r):
    """Host-only validation of the grouped matvec tables (ct_matvec_plan_check); 0 when valid."""
    return lib().ct_matvec_plan_check(_ptr(a), _ptr(w), _ptr(l), _ptr(r))


def generate_matvec_problem(model, nsites, chi, seed):
    """theta, h2, left, right of the synthetic H_eff matvec at the centre bond pair (model 'xxz' or 'hubbard')."""
    ts = [HostTensor() for _ in range(4)]
    check(lib().ct_generate_matvec_problem(0 if model == "xxz" else 1, nsites, chi, seed, *[t.sptr for t in ts]), "ct_generate_matvec_problem")
    for t in ts:
        t.live = True
    return tuple(ts)


def generate_matvec_problem_mpo(mpo, sector, chi, seed):
    """theta, h2, left, right of the synthetic H_eff matvec at the centre bond pair of any MPO built by
    the library (the spin molecular Hamiltonian, for instance), quantum-number sector `sector`."""
    ts = [HostTensor() for _ in range(4)]
    check(lib().ct_generate_matvec_problem_mpo(C.pointer(mpo.s), sector, chi, seed, *[t.sptr for t in ts]),
          "ct_generate_matvec_problem_mpo")
    for t in ts:
        t.live = True
    return tuple(ts)


def eigensystem_krylov_symmetric(afunc, vstart, maxiter, numeig):
    """Lanczos eigensystem of the symmetric map afunc(v) -> A v (numpy in, numpy out) through the
    library's eigensystem_krylov_symmetric (src/util/krylov.c:172-251): returns (lambda[numeig],
    u[n, numeig])."""
    vstart = np.ascontiguousarray(vstart, dtype=np.float64)
    n = vstart.size

    def cb(nn, data, v, ret):
        x = np.ctypeslib.as_array(v, shape=(nn,))
        y = np.ctypeslib.as_array(ret, shape=(nn,))
        y[:] = afunc(x.copy())

    fn = LINEAR_FUNC_D(cb)
    lam = np.zeros(numeig)
    u = np.zeros((n, numeig))
    ret = lib().eigensystem_krylov_symmetric(n, fn, None, vstart.ctypes.data_as(DP), maxiter, numeig,
                                             lam.ctypes.data_as(DP), u.ctypes.data_as(DP))
    check(ret, "eigensystem_krylov_symmetric")
    return lam, u


class DeviceArray:
    """A device buffer of float64 values owned by Python (hipMalloc via the library)."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = P()
        check(lib().ct_device_alloc(C.byref(self.ptr), max(self.n, 1) * 8), "ct_device_alloc")

    @staticmethod
    def from_numpy(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = DeviceArray(a.size)
        check(lib().ct_memcpy_h2d(d.ptr, a.ctypes.data_as(P), a.size * 8), "ct_memcpy_h2d")
        return d

    def numpy(self):
        out = np.zeros(self.n)
        check(lib().ct_memcpy_d2h(out.ctypes.data_as(P), self.ptr, self.n * 8), "ct_memcpy_d2h")
        return out

    def __del__(self):
        if self.ptr and _lib is not None:
            _lib.ct_device_free(self.ptr)
            self.ptr = None


class MatvecPlan:
    """Device-resident H_eff plan (ct_matvec_plan_*): w, l, r stay on the GPU."""

    def __init__(self, a, w, l, r, mode=GROUPED_FP64):
        self.h = P()
        check(lib().ct_matvec_plan_create(_ptr(a), _ptr(w), _ptr(l), _ptr(r), mode, C.byref(self.h)), "ct_matvec_plan_create")
        self.mode = mode

    def __del__(self):
        if self.h and _lib is not None:
            _lib.ct_matvec_plan_destroy(self.h)
            self.h = None

    @property
    def dim(self):
        return lib().ct_matvec_plan_dim(self.h)

    @property
    def flops(self):
        return lib().ct_matvec_plan_flops(self.h)

    @property
    def bytes(self):
        return lib().ct_matvec_plan_bytes(self.h)

    def stats(self):
        s = np.zeros(4)
        lib().ct_matvec_plan_stats(self.h, s.ctypes.data_as(DP))
        return dict(terms=int(s[0]), tasks=int(s[1]), items=int(s[2]), intermediate_doubles=int(s[3]))

    def forms(self):
        f = np.zeros(3, dtype=np.int32)
        t = np.zeros(12, dtype=np.float32)
        lib().ct_matvec_plan_forms(self.h, f.ctypes.data_as(IP), t.ctypes.data_as(FP))
        return f.tolist(), t.reshape(3, 4).tolist()

    def stage_bytes(self):
        s = np.zeros(3)
        lib().ct_matvec_plan_stage_bytes(self.h, s.ctypes.data_as(DP))
        return s

    def stage_flops(self):
        s = np.zeros(3)
        lib().ct_matvec_plan_stage_flops(self.h, s.ctypes.data_as(DP))
        return s

    @property
    def stream(self):
        return lib().ct_matvec_plan_stream(self.h)

    def apply(self, x, y):
        check(lib().ct_matvec_apply(self.h, x.ptr, y.ptr), "ct_matvec_apply")

    def apply_timed(self, x, y):
        ms = np.zeros(3, dtype=np.float32)
        check(lib().ct_matvec_apply_timed(self.h, x.ptr, y.ptr, ms.ctypes.data_as(FP)), "ct_matvec_apply_timed")
        return ms

    def synchronize(self):
        check(lib().ct_stream_synchronize(self.stream), "ct_stream_synchronize")

    def stage_time(self, stage, x, y, reps=50):
        """average ms of one stage's launch over reps back-to-back launches"""
        ms = C.c_float(0)
        check(lib().ct_matvec_plan_stage_time(self.h, stage, x.ptr, y.ptr, reps, C.byref(ms)), "ct_matvec_plan_stage_time")
        return ms.value

    def lanczos(self, vstart, maxiter):
        lam = C.c_double(0)
        u = DeviceArray(self.dim)
        ret = lib().ct_lanczos_ground_state(self.h, vstart.ptr, maxiter, C.byref(lam), u.ptr)
        check(ret, "ct_lanczos_ground_state")
        return lam.value, u


class MatvecBatch:
    """Independent H_eff problems applied in one fused launch (ct_matvec_batch_*); problems are
    (theta, w, l, r) tuples of block-sparse tensors."""

    def __init__(self, problems, mode=None):
        n = len(problems)
        arrs = [(P * n)(*[_ptr(p[i]) for p in problems]) for i in range(4)]
        self.h = P()
        if mode is None:
            check(lib().ct_matvec_batch_create(n, *arrs, C.byref(self.h)), "ct_matvec_batch_create")
        else:
            check(lib().ct_matvec_batch_create_mode(n, *arrs, mode, C.byref(self.h)), "ct_matvec_batch_create_mode")
        self.mode = GROUPED_FP64 if mode is None else mode
        self.nprob = n
        self.dim = lib().ct_matvec_batch_dim(self.h)
        off = np.zeros(n + 1, dtype=np.int64)
        lib().ct_matvec_batch_offsets(self.h, off.ctypes.data_as(LP))
        self.offsets = off
        self.flops = lib().ct_matvec_batch_flops(self.h)
        self.bytes = lib().ct_matvec_batch_bytes(self.h)
        self.stream = lib().ct_matvec_batch_stream(self.h)
        f = C.c_int(0)
        ms = (C.c_float * 2)()
        lib().ct_matvec_batch_form(self.h, C.byref(f), ms)
        self.fused = bool(f.value)
        self.form_ms = (ms[0], ms[1])

    def __del__(self):
        if getattr(self, "h", None):
            lib().ct_matvec_batch_destroy(self.h)
            self.h = None

    def apply(self, x, y):
        check(lib().ct_matvec_batch_apply(self.h, x.ptr, y.ptr), "ct_matvec_batch_apply")

    def synchronize(self):
        check(lib().ct_stream_synchronize(self.stream), "ct_stream_synchronize")

    def lanczos(self, vstart, maxiter):
        lam = np.zeros(self.nprob)
        st = np.zeros(self.nprob, dtype=np.int32)
        u = DeviceArray(self.dim)
        lib().ct_lanczos_ground_state_batch(self.h, vstart.ptr, maxiter, lam.ctypes.data_as(DP), st.ctypes.data_as(IP), u.ptr)
        return lam, st, u


class ServeRequest(C.Structure):
    _fields_ = [("model", C.c_int), ("nsites", C.c_int), ("param", C.c_double * 3), ("sector", C.c_int32), ("pad", C.c_int),
                ("max_vdim", ct_long), ("num_sweeps", C.c_int), ("maxiter_lanczos", C.c_int), ("tol_split", C.c_double),
                ("tol_energy", C.c_double), ("seed", C.c_uint64)]


class ServeResult(C.Structure):
    _fields_ = [("status", C.c_int), ("sweeps", C.c_int), ("energy", C.c_double), ("energy_change", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


def serve_request(model, nsites, param, sector, max_vdim, num_sweeps=8, maxiter_lanczos=10, tol_split=1e-10,
                  tol_energy=1e-10, seed=0):
    r = ServeRequest()
    r.model = 0 if model == "xxz" else 1
    r.nsites = nsites
    for i, v in enumerate(param):
        r.param[i] = v
    r.sector = sector
    r.max_vdim = max_vdim
    r.num_sweeps = num_sweeps
    r.maxiter_lanczos = maxiter_lanczos
    r.tol_split = tol_split
    r.tol_energy = tol_energy
    r.seed = seed
    return r


class Server:
    """Continuous-batching server of DMRG ground-state solves (ct_serve_*)."""

    def __init__(self, hbm_budget_bytes=200e9, max_active=64):
        self.h = P()
        check(lib().ct_serve_create(float(hbm_budget_bytes), max_active, C.byref(self.h)), "ct_serve_create")

    def submit(self, req):
        t = C.c_int64(0)
        check(lib().ct_serve_submit(self.h, C.byref(req), C.byref(t)), "ct_serve_submit")
        return t.value

    def poll(self, ticket, wait=False):
        r = ServeResult()
        ret = lib().ct_serve_poll(self.h, ticket, 1 if wait else 0, C.byref(r))
        if ret < 0:
            raise RuntimeError(f"ct_serve_poll failed: {ERRORS.get(ret, ret)}")
        return r

    def energies(self, ticket):
        buf = np.zeros(256)
        n = lib().ct_serve_energies(self.h, ticket, buf.ctypes.data_as(DP), buf.size)
        return buf[:max(n, 0)].copy()

    def stats(self):
        st = np.zeros(14)
        lib().ct_serve_stats(self.h, st.ctypes.data_as(DP))
        keys = ["ticks", "lanczos_batches", "completed", "queued", "reserved_bytes", "max_batch", "busy_s", "uptime_s",
                "start_s", "plan_s", "lanczos_s", "split_device_sum_s", "split_host_sum_s", "splits_wall_s"]
        return dict(zip(keys, st.tolist()))

    def close(self):
        if getattr(self, "h", None):
            lib().ct_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Event:
    def __init__(self):
        self.e = P()
        check(lib().ct_event_create(C.byref(self.e)), "ct_event_create")

    def record(self, stream):
        check(lib().ct_event_record(self.e, stream), "ct_event_record")

    def elapsed_ms(self, later):
        ms = C.c_float(0)
        check(lib().ct_event_elapsed_ms(self.e, later.e, C.byref(ms)), "ct_event_elapsed_ms")
        return ms.value

    def __del__(self):
        if self.e and _lib is not None:
            _lib.ct_event_destroy(self.e)
            self.e = None


def encode_pair(qa, qb):
    return int(qa) * 65536 + int(qb)


class MPO:
    """struct mpo built by the product library (ct_construct_*_mpo)."""

    def __init__(self):
        self.s = MPOStruct()
        self.live = False

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), P)

    @property
    def nsites(self):
        return self.s.nsites

    @property
    def d(self):
        return self.s.d

    @property
    def qsite(self):
        return np.array([self.s.qsite[i] for i in range(self.s.d)], dtype=np.int32)

    def __del__(self):
        if self.live and _lib is not None:
            _lib.delete_mpo(C.pointer(self.s))
            self.live = False

    @staticmethod
    def xxz(nsites, J=1.0, D=0.5, h=0.0):
        m = MPO()
        lib().ct_construct_heisenberg_xxz_1d_mpo(nsites, J, D, h, C.pointer(m.s))
        m.live = True
        return m

    @staticmethod
    def hubbard(nsites, t=1.0, u=4.0, mu=0.0):
        m = MPO()
        lib().ct_construct_fermi_hubbard_1d_mpo(nsites, t, u, mu, C.pointer(m.s))
        m.live = True
        return m

    @staticmethod
    def spin_molecular(tkin, vint):
        """spin molecular Hamiltonian, tkin (n, n), vint (n, n, n, n) in physicists' convention"""
        t = np.ascontiguousarray(tkin, dtype=np.float64)
        v = np.ascontiguousarray(vint, dtype=np.float64)
        n = t.shape[0]
        assert t.shape == (n, n) and v.shape == (n, n, n, n)
        m = MPO()
        lib().ct_construct_spin_molecular_hamiltonian_mpo(n, t.ctypes.data_as(DP), v.ctypes.data_as(DP), C.pointer(m.s))
        m.live = True
        return m


def molecular_coefficients(n, seed):
    """Synthetic molecular integrals (seeded): t_ij symmetric with entries N(0,1)/n; chemists'
    (ij|kl) with 8-fold symmetry and entries N(0, 0.1^2), made positive semidefinite as the pair
    matrix [(ij), (kl)] (negative eigenvalues set to zero, which keeps the symmetries); returned as
    (tkin, vint) with vint in physicists' convention, v_ijkl = (ik|jl)."""
    rng = np.random.default_rng(seed)
    t = rng.standard_normal((n, n)) / n
    t = (t + t.T) / 2
    g = rng.standard_normal((n, n, n, n)) * 0.1
    g = (g + g.transpose(1, 0, 2, 3) + g.transpose(0, 1, 3, 2) + g.transpose(1, 0, 3, 2)
         + g.transpose(2, 3, 0, 1) + g.transpose(3, 2, 0, 1) + g.transpose(2, 3, 1, 0) + g.transpose(3, 2, 1, 0)) / 8
    w, u = np.linalg.eigh(g.reshape(n * n, n * n))
    g = ((u * np.maximum(w, 0)) @ u.T).reshape(n, n, n, n)
    v = np.ascontiguousarray(g.transpose(0, 2, 1, 3))
    return t, v


class MPS:
    """struct mps built by the product library (construct_random_mps with a seeded RNG)."""

    def __init__(self):
        self.s = MPSStruct()
        self.live = False

    @property
    def ptr(self):
        return C.cast(C.pointer(self.s), P)

    @property
    def nsites(self):
        return self.s.nsites

    def site_ptr(self, i):
        return C.cast(C.pointer(self.s.a[i]), P)

    def bond_dims(self):
        dims = [int(self.s.a[0].dim_logical[0])]
        for i in range(self.s.nsites):
            dims.append(int(self.s.a[i].dim_logical[2]))
        return dims

    def __del__(self):
        if self.live and _lib is not None:
            _lib.delete_mps(C.pointer(self.s))
            self.live = False

    @staticmethod
    def random(nsites, d, qsite, sector, max_vdim, seed):
        m = MPS()
        rng = C.create_string_buffer(64)
        lib().seed_rng_state(seed, rng)
        q = np.ascontiguousarray(qsite, dtype=np.int32)
        lib().construct_random_mps(1, nsites, d, q.ctypes.data_as(IP), sector, max_vdim, rng, C.pointer(m.s))
        m.live = True
        return m


def save_mps(filename, mps):
    """save_mps (src/state/mps.c:1219): the MPS in an HDF5 file with the reference's layout."""
    if lib().save_mps(os.fsencode(filename), C.pointer(mps.s)) != 0:
        raise RuntimeError(f"save_mps failed for '{filename}'")


def load_mps(filename):
    """load_mps (src/state/mps.c:1309): an MPS from an HDF5 file written by either library."""
    m = MPS()
    if lib().load_mps(os.fsencode(filename), C.pointer(m.s)) != 0:
        raise RuntimeError(f"load_mps failed for '{filename}'")
    m.live = True
    return m


def dmrg(mpo, psi, num_sweeps=5, maxiter_lanczos=20, tol_split=1e-10, max_vdim=256):
    """dmrg_twosite (dmrg.c:262-399) on the GPU; returns (en_sweeps, entropy); psi is updated in place."""
    en = np.zeros(num_sweeps)
    ent = np.zeros(mpo.nsites - 1)
    ret = lib().dmrg_twosite(C.pointer(mpo.s), num_sweeps, maxiter_lanczos, tol_split, max_vdim, C.pointer(psi.s),
                             en.ctypes.data_as(DP), ent.ctypes.data_as(DP))
    check(ret, "dmrg_twosite")
    return en, ent
