"""ctypes binding of the gfx950 library ``libspinoza_amd.so`` (C ABI in
``include/spinoza_amd.h``).

There is no CPU fallback: if the library is missing, importing the package
raises, and if no GPU is visible every call fails with SpinozaError.
"""
import atexit
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SPINOZA_AMD_LIB") or os.path.join(HERE, "libspinoza_amd.so")

# status codes (enum spz_status)
OK, ERR_QUBIT, ERR_GATE, ERR_OOM, ERR_HIP, ERR_ARG, ERR_PROB, ERR_NODEVICE = range(8)
# precision
F64, F32 = 0, 1
# gate kinds (enum spz_gate_kind)
H, X, Y, Z, P, RX, RY, RZ, U, SWAP, BITFLIP, M, UNITARY = range(13)
# exec modes
EXEC_AUTO, EXEC_PLAIN, EXEC_STRICT, EXEC_DENSE5 = 0, 1, 2, 3
OPF_CLASSICAL = 1
# kernel classes
K_PAIR, K_DIAG, K_TILE, K_REDUCE, K_DENSE, K_ADJOINT, K_OTHER = range(7)
K_NAMES = ["pair", "diag", "tile", "reduce", "dense", "adjoint", "other"]


class SpinozaError(RuntimeError):
    """A non-zero status from the C ABI (the reference panics instead)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Gate_t(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("t1", ctypes.c_int32),
                ("theta", ctypes.c_double), ("phi", ctypes.c_double),
                ("lam", ctypes.c_double)]


class Op_t(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("target", ctypes.c_int32),
                ("theta", ctypes.c_double), ("phi", ctypes.c_double),
                ("lam", ctypes.c_double),
                ("t1", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("ones", ctypes.c_uint64), ("zeros", ctypes.c_uint64),
                ("u", ctypes.c_double),
                ("m", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("u_re", ctypes.c_void_p), ("u_im", ctypes.c_void_p)]


class Tracker_t(ctypes.Structure):
    _fields_ = [("measured", ctypes.c_uint64), ("values", ctypes.c_uint64)]


class ExecOpts_t(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int32), ("tile_qubits", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 6)]


class BatchCircuit_t(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("n_ops", ctypes.c_int32),
                ("op_offset", ctypes.c_uint64), ("grad_offset", ctypes.c_uint64)]


class BatchOpts_t(ctypes.Structure):
    _fields_ = [("wg_max_qubits", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("group_qubits", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 4)]


class PlanInfo_t(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_uint64), ("tile_gates", ctypes.c_uint64),
                ("layouts", ctypes.c_uint64), ("plain_gates", ctypes.c_uint64),
                ("lane_swaps", ctypes.c_uint64), ("frame_gates", ctypes.c_uint64),
                ("reserved", ctypes.c_uint64 * 2)]


class BatchStats_t(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("n_small", ctypes.c_uint64),
                ("n_large", ctypes.c_uint64), ("wg_max_qubits", ctypes.c_int32),
                ("n_group", ctypes.c_int32)]


# name -> (restype, argtypes) for every exported entry point
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i = ctypes.c_int
_d = ctypes.c_double
_u64 = ctypes.c_uint64
SIGNATURES = {
    "spz_init": (_i, []),
    "spz_last_error": (ctypes.c_char_p, []),
    "spz_version": (ctypes.c_char_p, []),
    "spz_device_info": (_i, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(_i), _u64p]),
    "spz_stream": (_vp, []),
    "spz_sync": (_i, []),
    "spz_state_create": (_i, [_i, _i, ctypes.POINTER(_vp)]),
    "spz_state_destroy": (_i, [_vp]),
    "spz_state_qubits": (_i, [_vp, ctypes.POINTER(_i)]),
    "spz_state_precision": (_i, [_vp, ctypes.POINTER(_i)]),
    "spz_state_set_basis": (_i, [_vp, _u64]),
    "spz_state_upload": (_i, [_vp, _dp, _dp, _u64, _u64]),
    "spz_state_download": (_i, [_vp, _dp, _dp, _u64, _u64]),
    "spz_state_copy": (_i, [_vp, _vp]),
    "spz_state_device_ptrs": (_i, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "spz_apply": (_i, [_vp, ctypes.POINTER(Gate_t), _i, _d]),
    "spz_c_apply": (_i, [_vp, ctypes.POINTER(Gate_t), _i, _i]),
    "spz_cc_apply": (_i, [_vp, ctypes.POINTER(Gate_t), _i, _i, _i]),
    "spz_mc_apply": (_i, [_vp, ctypes.POINTER(Gate_t), _u64, _u64, _i]),
    "spz_transform_u": (_i, [_vp, _i, _dp, _dp, _i, _i]),
    "spz_execute": (_i, [_vp, ctypes.POINTER(Op_t), _u64, ctypes.POINTER(Tracker_t),
                         ctypes.POINTER(ExecOpts_t), ctypes.POINTER(ctypes.c_uint8)]),
    "spz_norm": (_i, [_vp, _dp]),
    "spz_prob0": (_i, [_vp, _i, _dp]),
    "spz_measure_qubit": (_i, [_vp, _i, _i, _i, _d, ctypes.POINTER(ctypes.c_uint8)]),
    "spz_qubit_expectation_value": (_i, [_vp, _i, _dp]),
    "spz_xyz_expectation_value": (_i, [_vp, ctypes.c_char, ctypes.POINTER(ctypes.c_int32),
                                       _u64, _dp]),
    "spz_expect_zz": (_i, [_vp, _i, _i, _dp]),
    "spz_ising_energy": (_i, [_vp, _d, _dp]),
    "spz_ising_energy_noalloc": (_i, [_vp, _d, _dp]),
    "spz_inner": (_i, [_vp, _vp, _dp, _dp]),
    "spz_sample": (_i, [_vp, _dp, _u64, _u64p]),
    "spz_adjoint_grad": (_i, [_vp, ctypes.POINTER(Op_t), _u64, _d, _i, _dp, _dp, _dp]),
    "spz_batch_adjoint": (_i, [ctypes.POINTER(BatchCircuit_t), _u64, ctypes.POINTER(Op_t), _d,
                               _dp, _dp, ctypes.POINTER(BatchOpts_t),
                               ctypes.POINTER(BatchStats_t)]),
    "spz_batch_autotune": (_i, [ctypes.POINTER(BatchCircuit_t), _u64, ctypes.POINTER(Op_t),
                                ctypes.POINTER(_i)]),
    "spz_plan_stats": (_i, [_i, _i, ctypes.POINTER(Op_t), _u64, ctypes.POINTER(ExecOpts_t),
                           ctypes.POINTER(PlanInfo_t)]),
    "spz_qasm_parse": (_i, [ctypes.c_char_p, _u64, ctypes.POINTER(Op_t), _u64, _u64p,
                            ctypes.POINTER(_i)]),
    "spz_profile_enable": (_i, [_i]),
    "spz_profile_reset": (_i, []),
    "spz_profile_get": (_i, [_i, _u64p, _dp, _dp]),
}

_lib = None
_exiting = False


def _mark_exit():
    global _exiting
    _exiting = True


# States still alive when the interpreter shuts down (held by a module global
# or a stored traceback) are not freed one by one: the HIP runtime may already
# be tearing down while Python finalises objects, and the process's device
# memory goes with the process anyway.
atexit.register(_mark_exit)


def exiting():
    return _exiting


def load():
    """Load the library (no compute call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the gfx950 library first "
            "(python -c 'import __graft_entry__ as g; g.build()'). There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return _lib if _lib is not None else load()


def check(rc):
    if rc != OK:
        msg = lib().spz_last_error()
        raise SpinozaError(rc, msg.decode() if msg else "error")
    return rc


def dptr(a):
    """float64 numpy array -> double*."""
    return a.ctypes.data_as(_dp)
