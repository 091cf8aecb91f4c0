"""Python mirror of the reference's hot-path API over the libflygpu C-ABI (include/flygpu.h).

Names and argument meanings follow the reference classes:
  system::Box (box.hpp:103-120)            -> Box
  neigh::List (list.hpp:71-210)            -> List
  potential::DataEAM + EAM (data.hpp, eam.hpp) -> EAM
  minimise::StepLBFGS / LBFGS (lbfgs_core.hpp, lbfgs.hpp) -> StepLBFGS / LBFGS
  saddle::Rotor + Dimer (rotor.hpp, dimer.hpp) -> Dimer
  saddle::Master::find_batch (find.cpp:684-755) on the device -> Batch
Errors map to FlyError (fly::RuntimeError) and SplineOutOfBounds (fly::Spline::OutOfBounds).

Arrays: numpy arrays are host data; torch tensors on a GPU are passed as device pointers (the caller
orders torch's stream against the context's stream, e.g. with torch.cuda.synchronize()).
There is no CPU fallback: without the built library every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflygpu.so")

FLYG_OK, FLYG_INVALID_ARG, FLYG_SPLINE_OOB, FLYG_DEVICE_ERROR, FLYG_BUSY, FLYG_STOPPED = 0, 1, 2, 3, 4, 5
HOST, DEVICE = 0, 1
FP64, MIXED = 0, 1
EXIT_NAMES = {0: "success", 1: "convex", 2: "iter_max", 3: "collision", 4: "uninit", 5: "region", 6: "spline OOB"}


class FlyError(RuntimeError):
    """fly::RuntimeError."""


class SplineOutOfBounds(FlyError):
    """fly::Spline::OutOfBounds."""


class DeviceError(FlyError):
    pass


class Busy(FlyError):
    """Serving: the device memory budget cannot take this now (retry after searches finish)."""


class Stopped(FlyError):
    """A step observer asked the run to stop."""


_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_vp = C.c_void_p
# int fn(void* user, int iteration, const double* xyz, int64_t n) (flyg_step_observer)
STEP_OBSERVER = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_int64)
_lib = None


class LBFGSOptions(C.Structure):
    _fields_ = [("n", C.c_int), ("iter_max", C.c_int), ("f2norm", C.c_double), ("skin_frac", C.c_double),
                ("proj_tol", C.c_double), ("max_trust", C.c_double), ("min_trust", C.c_double),
                ("grow_trust", C.c_double), ("shrink_trust", C.c_double)]


class DimerOptions(C.Structure):
    _fields_ = [("n", C.c_int), ("max_steps", C.c_int), ("hist_check_freq", C.c_int), ("f2norm", C.c_double),
                ("skin_frac", C.c_double), ("proj_tol", C.c_double), ("max_trust", C.c_double),
                ("min_trust", C.c_double), ("grow_trust", C.c_double), ("shrink_trust", C.c_double),
                ("convex_max", C.c_int), ("use_history", C.c_int)]


class RotorOptions(C.Structure):
    _fields_ = [("n", C.c_int), ("iter_max_rot", C.c_int), ("delta_r", C.c_double), ("theta_tol", C.c_double),
                ("relax_in_convex", C.c_int)]


class SearchStats(C.Structure):
    _fields_ = [("exit_code", C.c_int), ("steps", C.c_int), ("grads", C.c_int64), ("rebuilds", C.c_int),
                ("curv", C.c_double), ("energy", C.c_double), ("max_disp", C.c_double), ("launches", C.c_int),
                ("rot_iters", C.c_int), ("device_ms", C.c_double), ("device_grad_ms", C.c_double),
                ("retries", C.c_int), ("resumes", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("block_threads", C.c_int), ("lanes_per_atom", C.c_int),
                ("list_skin", C.c_double), ("max_concurrent", C.c_int)]


class MechOptions(C.Structure):
    _fields_ = [("enabled", C.c_int), ("nudge_frac", C.c_double), ("basin_tol", C.c_double),
                ("stationary_tol", C.c_double), ("minimiser", LBFGSOptions)]


class Mechanism(C.Structure):
    _fields_ = [("code", C.c_int), ("min_iters", C.c_int), ("barrier", C.c_double), ("delta", C.c_double),
                ("drift", C.c_double * 3)]

    def as_dict(self):
        return {"code": self.code, "min_iters": self.min_iters, "barrier": self.barrier, "delta": self.delta,
                "drift": np.array(self.drift[:])}


MECH_NAMES = {0: "pathway", 1: "minimisation failed", 2: "unlinked", 3: "back to initial", 4: "min-sp stationary",
              5: "sp-min stationary", 6: "centred elsewhere", 7: "saddle below minimum", -1: "not checked"}


class ServerOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("block_threads", C.c_int), ("lanes_per_atom", C.c_int),
                ("list_skin", C.c_double), ("max_concurrent", C.c_int), ("streams", C.c_int), ("max_batch", C.c_int),
                ("mem_budget", C.c_int64)]


class ServerState(C.Structure):
    _fields_ = [("queued", C.c_int64), ("in_flight", C.c_int64), ("systems", C.c_int64), ("device_bytes", C.c_int64),
                ("budget_bytes", C.c_int64), ("launches", C.c_int64), ("searches_done", C.c_int64),
                ("kernel_ms", C.c_double), ("busy_streams", C.c_int), ("block_threads", C.c_int),
                ("lanes_per_atom", C.c_int), ("host_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_SIGS = {
    "flyg_ctx_create": [C.c_int, C.POINTER(_vp)],
    "flyg_ctx_destroy": [_vp],
    "flyg_ctx_last_error": [_vp],
    "flyg_ctx_synchronize": [_vp],
    "flyg_ctx_stream": [_vp],
    "flyg_ctx_device_info": [_vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_int)],
    "flyg_eam_load": [_vp, C.c_char_p, C.POINTER(_vp)],
    "flyg_eam_load_text": [_vp, C.c_char_p, C.c_size_t, C.POINTER(_vp)],
    "flyg_eam_destroy": [_vp],
    "flyg_eam_info": [_vp, C.POINTER(C.c_int), _d, _d, C.c_char_p],
    "flyg_eam_set_precision": [_vp, C.c_int],
    "flyg_eam_set_lanes": [_vp, C.c_int],
    "flyg_eam_set_timing": [_vp, C.c_int],
    "flyg_eam_timing": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
    "flyg_eam_table_eval": [_vp, C.c_int, C.c_int, C.c_int, _d, C.c_int64, C.c_int, _d],
    "flyg_list_create": [_vp, _d, C.POINTER(C.c_int), C.c_double, C.POINTER(_vp)],
    "flyg_list_destroy": [_vp],
    "flyg_list_rebuild": [_vp, _vp, C.c_int64, C.c_int],
    "flyg_list_update": [_vp, _vp, C.c_int64, C.c_int],
    "flyg_list_size": [_vp, _i64],
    "flyg_list_num_ghosts": [_vp, _i64],
    "flyg_list_count": [_vp, C.c_double, _i64],
    "flyg_list_export": [_vp, C.c_double, _i64, _i32, _i32, _d],
    "flyg_list_positions": [_vp, _vp, C.c_int],
    "flyg_eam_energy": [_vp, _vp, _vp, _vp, C.c_int64, C.c_int, _d],
    "flyg_eam_gradient": [_vp, _vp, _vp, _vp, C.c_int64, C.c_int, _vp],
    "flyg_eam_gradient_noalloc": [_vp, _vp, _vp, _vp, C.c_int64, _vp],
    "flyg_eam_hvp": [_vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, C.c_int],
    "flyg_eam_hessian_eigenvalues": [_vp, _vp, _vp, _vp, C.c_int64, C.c_int, _d, _i64],
    "flyg_log_prod_eigen": [_d, C.c_int64, C.c_double, _d, _i64, _i64],
    "flyg_prefactor": [C.c_double, C.c_double, _d],
    "flyg_steplbfgs_create": [_vp, C.c_int, C.POINTER(_vp)],
    "flyg_steplbfgs_destroy": [_vp],
    "flyg_steplbfgs_clear": [_vp],
    "flyg_steplbfgs_newton_step": [_vp, _vp, _vp, C.c_int64, _vp, C.c_int],
    "flyg_lbfgs_default_options": [C.POINTER(LBFGSOptions)],
    "flyg_lbfgs_create": [_vp, C.POINTER(LBFGSOptions), _d, C.POINTER(C.c_int), C.POINTER(_vp)],
    "flyg_lbfgs_destroy": [_vp],
    "flyg_lbfgs_minimise": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int64, C.c_int, C.POINTER(C.c_int),
                            C.POINTER(C.c_int)],
    "flyg_lbfgs_set_observer": [_vp, STEP_OBSERVER, _vp],
    "flyg_eam_stage_atoms": [_vp, _vp, _vp, C.c_int64, C.c_int],
    "flyg_dimer_set_observer": [_vp, STEP_OBSERVER, _vp],
    "flyg_dimer_default_options": [C.POINTER(DimerOptions)],
    "flyg_rotor_default_options": [C.POINTER(RotorOptions)],
    "flyg_dimer_create": [_vp, C.POINTER(DimerOptions), C.POINTER(RotorOptions), _d, C.POINTER(C.c_int),
                          C.POINTER(_vp)],
    "flyg_dimer_destroy": [_vp],
    "flyg_rotor_create": [_vp, C.POINTER(RotorOptions), C.POINTER(_vp)],
    "flyg_rotor_destroy": [_vp],
    "flyg_rotor_r_cut": [_vp, _vp, _d],
    "flyg_rotor_eff_gradient": [_vp, _vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, C.c_int, C.c_int, _d],
    "flyg_dimer_find_sp": [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int64, _vp, _vp, C.c_int, C.c_double, C.c_int,
                           C.c_int, C.POINTER(SearchStats)],
    "flyg_batch_default_options": [C.POINTER(BatchOptions)],
    "flyg_batch_create": [_vp, _vp, C.POINTER(DimerOptions), C.POINTER(RotorOptions), C.POINTER(BatchOptions),
                          _d, C.POINTER(C.c_int), _d, _u32, C.c_int64, C.POINTER(_vp)],
    "flyg_batch_destroy": [_vp],
    "flyg_batch_submit": [_vp, C.c_int, _i64, _d, _d, _d, _i64],
    "flyg_batch_submit_ex": [_vp, C.c_int, _i64, _d, _d, _d, C.c_double, _d, C.c_int, _i64],
    "flyg_batch_submit_perturbed": [_vp, C.c_int, _i64, _d, C.POINTER(C.c_uint64), C.c_double, C.c_double,
                                    C.c_double, _d, C.c_int, _i64],
    "flyg_perturb": [_d, C.POINTER(C.c_int), _d, _u8, C.c_int64, C.c_int64, C.c_double, C.c_double,
                     C.POINTER(C.c_uint64), _d, _d, _d],
    "flyg_theta_tol": [C.c_int, _d],
    "flyg_batch_run": [_vp],
    "flyg_batch_result": [_vp, C.c_int64, C.POINTER(SearchStats), _d, _d],
    "flyg_batch_region": [_vp, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "flyg_batch_shape": [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "flyg_mech_default_options": [C.POINTER(MechOptions)],
    "flyg_batch_set_mechanisms": [_vp, C.POINTER(MechOptions)],
    "flyg_batch_mechanism": [_vp, C.c_int64, C.POINTER(Mechanism), _d, _d],
    "flyg_server_set_mechanisms": [_vp, C.POINTER(MechOptions)],
    "flyg_server_mechanism": [_vp, C.c_int64, C.POINTER(Mechanism), _d, _d],
    "flyg_server_default_options": [C.POINTER(ServerOptions)],
    "flyg_server_create": [_vp, _vp, C.POINTER(DimerOptions), C.POINTER(RotorOptions), C.POINTER(ServerOptions),
                           C.POINTER(_vp)],
    "flyg_server_destroy": [_vp],
    "flyg_server_add_system": [_vp, _d, C.POINTER(C.c_int), _d, _u32, C.c_int64, _i64],
    "flyg_server_remove_system": [_vp, C.c_int64],
    "flyg_server_submit": [_vp, C.c_int64, C.c_int, _i64, _d, _i64, _i64, _d, _d, _i64],
    "flyg_server_submit_perturbed": [_vp, C.c_int64, C.c_int, _i64, _d, C.POINTER(C.c_uint64), C.c_double,
                                     C.c_double, _i64],
    "flyg_server_progress": [_vp, C.c_double, _i64, C.c_int, C.POINTER(C.c_int)],
    "flyg_server_result": [_vp, C.c_int64, C.POINTER(SearchStats), C.POINTER(C.c_int), _i64, _d, _d],
    "flyg_server_release": [_vp, C.c_int64],
    "flyg_server_status": [_vp, C.POINTER(ServerState)],
}
_VOID = {"flyg_lbfgs_default_options", "flyg_dimer_default_options", "flyg_rotor_default_options",
         "flyg_batch_default_options", "flyg_server_default_options", "flyg_mech_default_options"}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlyError(f"{LIB_PATH} is not built (run openfly_amd/build.py); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            if name in _VOID:
                f.restype = None
            elif name == "flyg_ctx_last_error":
                f.restype = C.c_char_p
            elif name == "flyg_ctx_stream":
                f.restype = _vp
            else:
                f.restype = C.c_int
        _lib = L
    return _lib


def declared_symbols():
    return list(_SIGS)


def _raise(ctx_h, rc):
    msg = (lib().flyg_ctx_last_error(ctx_h) or b"").decode()
    if rc == FLYG_SPLINE_OOB:
        raise SplineOutOfBounds(msg)
    if rc == FLYG_DEVICE_ERROR:
        raise DeviceError(msg)
    if rc == FLYG_BUSY:
        raise Busy(msg)
    if rc == FLYG_STOPPED:
        raise Stopped(msg)
    raise FlyError(msg)


class _Observer:
    """A Python callable fn(iteration, positions (n, 3) copy) as a step observer (the reference's
    Options::fout hook): a true return value stops the run; an exception stops it and is re-raised by
    the call that ran it."""

    def __init__(self, fn):
        self.fn = fn
        self.error = None

        def cb(_user, it, xyz, n):
            try:
                return 1 if fn(int(it), np.ctypeslib.as_array(xyz, shape=(int(n), 3)).copy()) else 0
            except BaseException as e:   # noqa: BLE001 - carried to the caller, never across C frames
                self.error = e
                return 1

        self.c = STEP_OBSERVER(cb)

    def reraise(self):
        if self.error is not None:
            e, self.error = self.error, None
            raise e


def _is_torch(a):
    return type(a).__module__.startswith("torch")


_TORCH_DTYPES = {np.uint32: "uint32", np.uint8: "uint8", np.float64: "float64"}


def _ptr(a, dtype, numel=None, device=None):
    """(pointer, where, keepalive) of a numpy array or a torch tensor.

    A device tensor is passed as is, so it must already have the element type, be contiguous, hold
    `numel` elements and live on `device`; anything else raises FlyError (the library would read it
    as raw memory).  Host arrays are converted (a copy when needed) and must hold `numel` elements.
    """
    if _is_torch(a):
        if a.is_cuda:
            want = _TORCH_DTYPES[dtype]
            got = str(a.dtype).replace("torch.", "")
            if got != want:
                raise FlyError(f"device tensor has dtype {got}, expected {want}")
            if not a.is_contiguous():
                raise FlyError("device tensor must be contiguous")
            if numel is not None and a.numel() != numel:
                raise FlyError(f"device tensor has {a.numel()} elements, expected {numel}")
            if device is not None and a.device.index != device:
                raise FlyError(f"device tensor is on cuda:{a.device.index}, the context is on cuda:{device}")
            return C.c_void_p(a.data_ptr()), DEVICE, a
        a = a.numpy()
    arr = np.ascontiguousarray(a, dtype=dtype)
    if numel is not None and arr.size != numel:
        raise FlyError(f"array has {arr.size} elements, expected {numel}")
    return C.c_void_p(arr.ctypes.data), HOST, arr


class Context:
    """A device, one HIP stream and the last error (one per host thread)."""

    def __init__(self, device=0):
        h = _vp()
        rc = lib().flyg_ctx_create(device, C.byref(h))
        if rc:
            _raise(None, rc)
        self.h = h
        self.device = device

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_ctx_destroy(self.h)
            self.h = None

    def check(self, rc):
        if rc:
            _raise(self.h, rc)

    def synchronize(self):
        self.check(lib().flyg_ctx_synchronize(self.h))

    @property
    def stream(self):
        return lib().flyg_ctx_stream(self.h)

    def device_info(self):
        buf = C.create_string_buffer(64)
        cu = C.c_int()
        self.check(lib().flyg_ctx_device_info(self.h, buf, 64, C.byref(cu)))
        return buf.value.decode(), cu.value


class Box:
    """system::Box(basis, periodic): basis columns are the box vectors."""

    def __init__(self, basis, periodic=(True, True, True)):
        self.basis = np.ascontiguousarray(basis, dtype=np.float64).reshape(3, 3)
        self.periodic = np.ascontiguousarray(np.asarray(periodic, dtype=np.int32).reshape(3))

    def args(self):
        return self.basis.ctypes.data_as(_d), self.periodic.ctypes.data_as(C.POINTER(C.c_int))


class EAM:
    """potential::DataEAM(file) + potential::EAM: setfl tables resident on the device."""

    def __init__(self, ctx: Context, path):
        self.ctx = ctx
        h = _vp()
        ctx.check(lib().flyg_eam_load(ctx.h, os.fsencode(str(path)), C.byref(h)))
        self.h = h
        nt = C.c_int()
        rc = C.c_double()
        ctx.check(lib().flyg_eam_info(h, C.byref(nt), C.byref(rc), None, None))
        self.ntypes = nt.value
        self._r_cut = rc.value
        self.masses = np.zeros(self.ntypes)
        names = C.create_string_buffer(16 * self.ntypes)
        ctx.check(lib().flyg_eam_info(h, None, None, self.masses.ctypes.data_as(_d), names))
        self.names = [names.raw[16 * i:16 * (i + 1)].split(b"\0")[0].decode() for i in range(self.ntypes)]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_eam_destroy(self.h)
            self.h = None

    def r_cut(self):
        return self._r_cut

    def set_precision(self, precision):
        self.ctx.check(lib().flyg_eam_set_precision(self.h, precision))

    def set_lanes(self, lanes):
        self.ctx.check(lib().flyg_eam_set_lanes(self.h, lanes))

    def set_timing(self, on=True):
        """Record HIP events around each gradient call's kernels (clears the record)."""
        self.ctx.check(lib().flyg_eam_set_timing(self.h, 1 if on else 0))

    def timing(self):
        """(summed kernel ms, timed gradient calls) since set_timing(True); synchronises."""
        ms, n = C.c_double(), C.c_int64()
        self.ctx.check(lib().flyg_eam_timing(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def table(self, kind, a, b, x, deriv=0):
        x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        out = np.zeros_like(x)
        k = {"F": 0, "phi": 1, "V": 2}[kind]
        self.ctx.check(lib().flyg_eam_table_eval(self.h, k, a, b, x.ctypes.data_as(_d), x.size, deriv,
                                                 out.ctypes.data_as(_d)))
        return out

    def energy(self, types, frozen, nl: "List"):
        """types = frozen = None: the copy kept by stage()."""
        dev = self.ctx.device
        if types is None and frozen is None:
            tp = fp = None
            w1 = w2 = HOST
        else:
            tp, w1, k1 = _ptr(types, np.uint32, nl.n, dev)
            fp, w2, k2 = _ptr(frozen, np.uint8, nl.n, dev)
        if w1 != w2:
            raise FlyError("types and frozen must both be host or both be device arrays")
        E = C.c_double()
        self.ctx.check(lib().flyg_eam_energy(self.h, nl.h, tp, fp, nl.n, w1, C.byref(E)))
        return E.value

    def stage(self, types, frozen):
        """flyg_eam_stage_atoms: keep species and frozen flags (host or device arrays) in the handle;
        later calls given types=None, frozen=None use them."""
        dev = self.ctx.device
        n = len(types)
        tp, w1, k1 = _ptr(types, np.uint32, n, dev)
        fp, w2, k2 = _ptr(frozen, np.uint8, n, dev)
        if w1 != w2:
            raise FlyError("types and frozen must both be host or both device")
        self.ctx.check(lib().flyg_eam_stage_atoms(self.h, tp, fp, n, w1))

    def gradient(self, types, frozen, nl: "List", out=None):
        dev = self.ctx.device
        if out is None:
            out = np.zeros((nl.n, 3))
        op, w3, k3 = _ptr(out, np.float64, 3 * nl.n, dev)
        if types is None and frozen is None:   # the staged copy
            self.ctx.check(lib().flyg_eam_gradient(self.h, nl.h, None, None, nl.n, w3, op))
            return out if w3 == DEVICE else k3.reshape(-1, 3)
        tp, w1, k1 = _ptr(types, np.uint32, nl.n, dev)
        fp, w2, k2 = _ptr(frozen, np.uint8, nl.n, dev)
        if not (w1 == w2 == w3):
            raise FlyError("all arrays must be host or all device")
        self.ctx.check(lib().flyg_eam_gradient(self.h, nl.h, tp, fp, nl.n, w1, op))
        return out if w3 == DEVICE else k3.reshape(-1, 3)

    def hvp(self, types, frozen, nl: "List", v, out=None):
        """types = frozen = None: the copy kept by stage()."""
        dev = self.ctx.device
        vp, w3, k3 = _ptr(v, np.float64, 3 * nl.n, dev)
        if out is None:
            out = np.zeros((nl.n, 3))
        op, w4, k4 = _ptr(out, np.float64, 3 * nl.n, dev)
        if types is None and frozen is None:
            tp = fp = None
            w1 = w2 = w3
        else:
            tp, w1, k1 = _ptr(types, np.uint32, nl.n, dev)
            fp, w2, k2 = _ptr(frozen, np.uint8, nl.n, dev)
        if not (w1 == w2 == w3 == w4):
            raise FlyError("all arrays must be host or all device")
        self.ctx.check(lib().flyg_eam_hvp(self.h, nl.h, tp, fp, nl.n, vp, op, w1))
        return out if w4 == DEVICE else k4.reshape(-1, 3)


    def hessian_eigenvalues(self, types, frozen, nl: "List"):
        """Sorted eigenvalues of the mass-weighted Hessian of the non-frozen atoms (host arrays)."""
        t = np.ascontiguousarray(types, dtype=np.uint32).reshape(nl.n)
        f = np.ascontiguousarray(frozen, dtype=np.uint8).reshape(nl.n)
        w = np.zeros(3 * nl.n)
        m = C.c_int64()
        self.ctx.check(lib().flyg_eam_hessian_eigenvalues(self.h, nl.h, C.c_void_p(t.ctypes.data),
                                                          C.c_void_p(f.ctypes.data), nl.n, HOST,
                                                          w.ctypes.data_as(_d), C.byref(m)))
        return w[:m.value]


def log_prod_eigen(eigenvalues, zero_tol=1e-6):
    """(sum of log lambda over lambda > zero_tol, zero modes, negative modes), as Master reads a spectrum."""
    w = np.ascontiguousarray(eigenvalues, dtype=np.float64)
    lp = C.c_double()
    z, neg = C.c_int64(), C.c_int64()
    rc = lib().flyg_log_prod_eigen(w.ctypes.data_as(_d), w.size, float(zero_tol), C.byref(lp), C.byref(z),
                                   C.byref(neg))
    if rc:
        _raise(None, rc)
    return lp.value, z.value, neg.value


def prefactor(log_prod_min, log_prod_sp):
    """Mechanism::kinetic_pre (find.cpp:287) from the two log-products (Hz)."""
    a = C.c_double()
    rc = lib().flyg_prefactor(float(log_prod_min), float(log_prod_sp), C.byref(a))
    if rc:
        _raise(None, rc)
    return a.value


DataEAM = EAM


class List:
    """neigh::List(box, r_cut)."""

    def __init__(self, ctx: Context, box: Box, r_cut: float):
        self.ctx = ctx
        self.box = box
        self.r_cut = float(r_cut)
        h = _vp()
        b, p = box.args()
        ctx.check(lib().flyg_list_create(ctx.h, b, p, self.r_cut, C.byref(h)))
        self.h = h
        self.n = 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_list_destroy(self.h)
            self.h = None

    def rebuild(self, positions, num_threads=1):
        size = positions.numel() if _is_torch(positions) else np.size(positions)
        if size % 3:
            raise FlyError("positions must hold 3 values per atom")
        p, w, keep = _ptr(positions, np.float64, size, self.ctx.device)
        n = size // 3
        self.ctx.check(lib().flyg_list_rebuild(self.h, p, n, w))
        self.n = n

    def update(self, delta):
        p, w, keep = _ptr(delta, np.float64, 3 * self.n, self.ctx.device)
        self.ctx.check(lib().flyg_list_update(self.h, p, self.n, w))

    def size(self):
        return self.n

    def num_ghosts(self):
        g = C.c_int64()
        self.ctx.check(lib().flyg_list_num_ghosts(self.h, C.byref(g)))
        return g.value

    def pairs(self, r_cut=None):
        """(offsets, j, shift, dr) of every neighbour with r < r_cut (List::for_neighbours)."""
        rc = self.r_cut if r_cut is None else float(r_cut)
        m = C.c_int64()
        self.ctx.check(lib().flyg_list_count(self.h, rc, C.byref(m)))
        off = np.zeros(self.n + 1, np.int64)
        j = np.zeros(max(m.value, 1), np.int32)
        s = np.zeros(max(m.value, 1), np.int32)
        dr = np.zeros((max(m.value, 1), 3))
        self.ctx.check(lib().flyg_list_export(self.h, rc, off.ctypes.data_as(_i64), j.ctypes.data_as(_i32),
                                              s.ctypes.data_as(_i32), dr.ctypes.data_as(_d)))
        return off, j[:m.value], s[:m.value], dr[:m.value]

    def for_neighbours(self, i, r_cut, f):
        off, j, s, dr = self.pairs(r_cut)
        for k in range(off[i], off[i + 1]):
            f(int(j[k]), float(np.sqrt(dr[k] @ dr[k])), dr[k])

    def positions(self):
        out = np.zeros((self.n, 3))
        self.ctx.check(lib().flyg_list_positions(self.h, C.c_void_p(out.ctypes.data), HOST))
        return out


class StepLBFGS:
    """minimise::StepLBFGS(n)."""

    def __init__(self, ctx: Context, n: int):
        self.ctx = ctx
        h = _vp()
        ctx.check(lib().flyg_steplbfgs_create(ctx.h, n, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_steplbfgs_destroy(self.h)
            self.h = None

    def clear(self):
        self.ctx.check(lib().flyg_steplbfgs_clear(self.h))

    def newton_step(self, x, g):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        g = np.ascontiguousarray(g, dtype=np.float64).ravel()
        q = np.zeros_like(x)
        self.ctx.check(lib().flyg_steplbfgs_newton_step(self.h, C.c_void_p(x.ctypes.data), C.c_void_p(g.ctypes.data),
                                                        x.size, C.c_void_p(q.ctypes.data), HOST))
        return q


def _fill(struct, defaults_fn, over):
    o = struct()
    defaults_fn(C.byref(o))
    for k, v in (over or {}).items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


def lbfgs_options(**kw):
    return _fill(LBFGSOptions, lib().flyg_lbfgs_default_options, kw)


def dimer_options(**kw):
    return _fill(DimerOptions, lib().flyg_dimer_default_options, kw)


def rotor_options(**kw):
    return _fill(RotorOptions, lib().flyg_rotor_default_options, kw)


def batch_options(**kw):
    return _fill(BatchOptions, lib().flyg_batch_default_options, kw)


class LBFGS:
    """minimise::LBFGS(Options, Box)."""

    def __init__(self, ctx: Context, box: Box, **opts):
        self.ctx = ctx
        self.opt = lbfgs_options(**opts)
        h = _vp()
        b, p = box.args()
        ctx.check(lib().flyg_lbfgs_create(ctx.h, C.byref(self.opt), b, p, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_lbfgs_destroy(self.h)
            self.h = None

    def minimise(self, positions, types, frozen, pot: EAM):
        """Returns (positions, gradient, failed, iterations); failed mirrors the reference's bool.
        types = frozen = None: the copy kept by pot.stage()."""
        pos = np.array(positions, dtype=np.float64, copy=True).reshape(-1, 3)
        n = len(pos)
        grad = np.zeros_like(pos)
        staged = types is None and frozen is None
        t = None if staged else np.ascontiguousarray(types, dtype=np.uint32)
        f = None if staged else np.ascontiguousarray(frozen, dtype=np.uint8)
        failed = C.c_int()
        it = C.c_int()
        rc = lib().flyg_lbfgs_minimise(self.h, pot.h, C.c_void_p(pos.ctypes.data), C.c_void_p(grad.ctypes.data),
                                       None if staged else C.c_void_p(t.ctypes.data),
                                       None if staged else C.c_void_p(f.ctypes.data), n, HOST,
                                       C.byref(failed), C.byref(it))
        if self._obs is not None:
            self._obs.reraise()
        self.ctx.check(rc)
        return pos, grad, bool(failed.value), it.value

    _obs = None

    def set_observer(self, fn):
        """fn(iteration, positions) at the start of every iteration (LBFGS::Options::fout, lbfgs.cpp:75-77);
        None removes it."""
        self._obs = _Observer(fn) if fn is not None else None
        self.ctx.check(lib().flyg_lbfgs_set_observer(self.h, self._obs.c if self._obs else STEP_OBSERVER(), None))


class Rotor:
    """saddle::Rotor(options): eff_gradient on a ready neighbour list (rotor.hpp:105-123)."""

    def __init__(self, ctx: Context, **rotor):
        self.ctx = ctx
        self.ropt = rotor_options(**rotor)
        h = _vp()
        ctx.check(lib().flyg_rotor_create(ctx.h, C.byref(self.ropt), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_rotor_destroy(self.h)
            self.h = None

    def r_cut(self, pot: EAM):
        r = C.c_double()
        self.ctx.check(lib().flyg_rotor_r_cut(self.h, pot.h, C.byref(r)))
        return r.value

    def eff_gradient(self, axis, types, frozen, pot: EAM, nl: "List", count_frozen=None):
        """(curvature, rotated axis, effective gradient); host arrays."""
        n = nl.n
        ax = np.ascontiguousarray(axis, dtype=np.float64).reshape(n, 3).copy()
        t = np.ascontiguousarray(types, dtype=np.uint32).reshape(n)
        f = np.ascontiguousarray(frozen, dtype=np.uint8).reshape(n)
        if count_frozen is None:
            count_frozen = int(f.sum())
        eff = np.zeros((n, 3))
        c = C.c_double()
        self.ctx.check(lib().flyg_rotor_eff_gradient(self.h, pot.h, nl.h, C.c_void_p(t.ctypes.data),
                                                     C.c_void_p(f.ctypes.data), n, C.c_void_p(ax.ctypes.data),
                                                     C.c_void_p(eff.ctypes.data), int(count_frozen), HOST,
                                                     C.byref(c)))
        return c.value, ax, eff


class Dimer:
    """saddle::Dimer(Options, Rotor::Options, Box)."""

    def __init__(self, ctx: Context, box: Box, dimer=None, rotor=None):
        self.ctx = ctx
        self.opt = dimer_options(**(dimer or {}))
        self.ropt = rotor_options(**(rotor or {}))
        h = _vp()
        b, p = box.args()
        ctx.check(lib().flyg_dimer_create(ctx.h, C.byref(self.opt), C.byref(self.ropt), b, p, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_dimer_destroy(self.h)
            self.h = None

    def find_sp(self, positions, axis, types, frozen, in_min, pot: EAM, hist_sp=(), theta_tol=0.0,
                count_frozen=None):
        pos = np.array(positions, dtype=np.float64, copy=True).reshape(-1, 3)
        ax = np.array(axis, dtype=np.float64, copy=True).reshape(-1, 3)
        n = len(pos)
        staged = types is None and frozen is None   # the copy kept by pot.stage()
        t = None if staged else np.ascontiguousarray(types, dtype=np.uint32)
        f = None if staged else np.ascontiguousarray(frozen, dtype=np.uint8)
        mn = np.ascontiguousarray(in_min, dtype=np.float64)
        hs = np.ascontiguousarray(np.stack([np.ravel(h) for h in hist_sp]) if len(hist_sp) else np.zeros(1))
        if count_frozen is None:
            if staged:
                raise FlyError("count_frozen must be given with staged frozen flags")
            count_frozen = int(f.sum())
        st = SearchStats()
        rc = lib().flyg_dimer_find_sp(self.h, pot.h, C.c_void_p(pos.ctypes.data), C.c_void_p(ax.ctypes.data),
                                      None if staged else C.c_void_p(t.ctypes.data),
                                      None if staged else C.c_void_p(f.ctypes.data), n,
                                      C.c_void_p(mn.ctypes.data), C.c_void_p(hs.ctypes.data), len(hist_sp),
                                      float(theta_tol), int(count_frozen), HOST, C.byref(st))
        if self._obs is not None:
            self._obs.reraise()
        self.ctx.check(rc)
        out = st.as_dict()
        out.update(pos=pos, axis=ax, exit=st.exit_code)
        return out

    _obs = None

    def set_observer(self, fn):
        """fn(iteration, centre positions) at the start of every iteration (Dimer::Options::fout,
        dimer.cpp:88-90); None removes it."""
        self._obs = _Observer(fn) if fn is not None else None
        self.ctx.check(lib().flyg_dimer_set_observer(self.h, self._obs.c if self._obs else STEP_OBSERVER(), None))


class Batch:
    """Batched dimer searches on one resident base system (Master::find_batch on the device)."""

    def __init__(self, ctx: Context, pot: EAM, box: Box, min_pos, types, dimer=None, rotor=None, **batch):
        self.ctx = ctx
        self.pot = pot
        self.opt = dimer_options(**(dimer or {}))
        self.ropt = rotor_options(**(rotor or {}))
        self.bopt = batch_options(**batch)
        self.base = np.ascontiguousarray(min_pos, dtype=np.float64).reshape(-1, 3)
        self.n = len(self.base)
        self.types = np.ascontiguousarray(types, dtype=np.uint32)
        h = _vp()
        b, p = box.args()
        ctx.check(lib().flyg_batch_create(ctx.h, pot.h, C.byref(self.opt), C.byref(self.ropt), C.byref(self.bopt), b, p,
                                          self.base.ctypes.data_as(_d), self.types.ctypes.data_as(_u32), self.n,
                                          C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_batch_destroy(self.h)
            self.h = None

    def _hist(self, hist):
        if hist is None or len(hist) == 0:
            return np.zeros(1), 0
        h = np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.float64).reshape(self.n, 3) for x in hist]))
        return h, len(hist)

    def submit(self, centres, radii, start_pos, start_axis, theta_tol=0.0, hist=None):
        """Queue searches from explicit starts; `hist` (known saddles, full positions) with theta_tol adds
        Dimer::find_sp's collision test."""
        c = np.ascontiguousarray(np.atleast_1d(centres), dtype=np.int64)
        r = np.ascontiguousarray(np.atleast_1d(radii), dtype=np.float64)
        k = len(c)
        sp = np.ascontiguousarray(start_pos, dtype=np.float64).reshape(k, self.n, 3)
        sa = np.ascontiguousarray(start_axis, dtype=np.float64).reshape(k, self.n, 3)
        ids = np.zeros(k, np.int64)
        h, nh = self._hist(hist)
        self.ctx.check(lib().flyg_batch_submit_ex(self.h, k, c.ctypes.data_as(_i64), r.ctypes.data_as(_d),
                                                  sp.ctypes.data_as(_d), sa.ctypes.data_as(_d), float(theta_tol),
                                                  h.ctypes.data_as(_d), nh, ids.ctypes.data_as(_i64)))
        return ids

    def submit_perturbed(self, centres, radii, prng, stddev=0.6, r_pert=4.0, theta_tol=0.0, hist=None):
        """Master::find_batch's searches: starts from Master::perturb with per-search Xoshiro256** states
        prng (k x 4 uint64, advanced in place), re-perturbed up to three times after a spline error."""
        c = np.ascontiguousarray(np.atleast_1d(centres), dtype=np.int64)
        r = np.ascontiguousarray(np.atleast_1d(radii), dtype=np.float64)
        k = len(c)
        if not (isinstance(prng, np.ndarray) and prng.dtype == np.uint64 and prng.flags.c_contiguous and
                prng.shape == (k, 4)):
            raise FlyError("prng must be a C-contiguous (k, 4) uint64 array")
        ids = np.zeros(k, np.int64)
        h, nh = self._hist(hist)
        self.ctx.check(lib().flyg_batch_submit_perturbed(self.h, k, c.ctypes.data_as(_i64), r.ctypes.data_as(_d),
                                                         prng.ctypes.data_as(C.POINTER(C.c_uint64)), float(stddev),
                                                         float(r_pert), float(theta_tol), h.ctypes.data_as(_d), nh,
                                                         ids.ctypes.data_as(_i64)))
        return ids

    def run(self):
        self.ctx.check(lib().flyg_batch_run(self.h))

    def shape(self):
        """(block_threads, lanes_per_atom) the batch ran with; (0, 0) before its first run."""
        t, l = C.c_int(), C.c_int()
        self.ctx.check(lib().flyg_batch_shape(self.h, C.byref(t), C.byref(l)))
        return t.value, l.value

    def set_mechanisms(self, **kw):
        """Check every converged saddle as a min->sp->min mechanism (Master::do_adj_min / saddle_2_mech).
        The mixed form needs minimiser=dict(f2norm=1e-4) or looser (FlyError otherwise)."""
        self.mopt = mech_options(**kw)
        self.ctx.check(lib().flyg_batch_set_mechanisms(self.h, C.byref(self.mopt)))

    def mechanism(self, sid, positions=False):
        m = Mechanism()
        rev = np.zeros((self.n, 3)) if positions else None
        fwd = np.zeros((self.n, 3)) if positions else None
        self.ctx.check(lib().flyg_batch_mechanism(self.h, int(sid), C.byref(m),
                                                  rev.ctypes.data_as(_d) if positions else None,
                                                  fwd.ctypes.data_as(_d) if positions else None))
        out = m.as_dict()
        if positions:
            out.update(rev=rev, fwd=fwd)
        return out

    def result(self, sid, positions=False):
        st = SearchStats()
        pos = np.zeros((self.n, 3)) if positions else None
        ax = np.zeros((self.n, 3)) if positions else None
        self.ctx.check(lib().flyg_batch_result(self.h, int(sid), C.byref(st),
                                               pos.ctypes.data_as(_d) if positions else None,
                                               ax.ctypes.data_as(_d) if positions else None))
        out = st.as_dict()
        out["barrier"] = st.energy
        out["exit"] = st.exit_code
        if positions:
            out["pos"] = pos
            out["axis"] = ax
        nm = C.c_int()
        nr = C.c_int()
        self.ctx.check(lib().flyg_batch_region(self.h, int(sid), C.byref(nm), C.byref(nr)))
        out["n_mobile"] = nm.value
        out["n_region"] = nr.value
        return out


def mech_options(minimiser=None, **kw):
    """flyg_mech_options with enabled = 1; minimiser: dict of LBFGS option overrides."""
    o = MechOptions()
    lib().flyg_mech_default_options(C.byref(o))
    o.enabled = 1
    for k, v in kw.items():
        setattr(o, k, v)
    for k, v in (minimiser or {}).items():
        setattr(o.minimiser, k, v)
    return o


def server_options(**kw):
    return _fill(ServerOptions, lib().flyg_server_default_options, kw)


class Server:
    """Continuous-batching saddle-search server over many resident systems (serving form of
    Master::find_batch).  Systems are admitted under a device memory budget; searches are queued with
    sparse start states and packed into launches on several streams as they arrive."""

    def __init__(self, ctx: Context, pot: EAM, dimer=None, rotor=None, **opts):
        self.ctx = ctx
        self.pot = pot
        self.opt = dimer_options(**(dimer or {}))
        self.ropt = rotor_options(**(rotor or {}))
        self.sopt = server_options(**opts)
        h = _vp()
        ctx.check(lib().flyg_server_create(ctx.h, pot.h, C.byref(self.opt), C.byref(self.ropt), C.byref(self.sopt),
                                           C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.flyg_server_destroy(self.h)
            self.h = None

    def add_system(self, box: Box, min_pos, types):
        """System id, or raises Busy when the memory budget cannot hold it now."""
        pos = np.ascontiguousarray(min_pos, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(types, dtype=np.uint32)
        sid = C.c_int64()
        b, p = box.args()
        self.ctx.check(lib().flyg_server_add_system(self.h, b, p, pos.ctypes.data_as(_d), t.ctypes.data_as(_u32),
                                                    len(pos), C.byref(sid)))
        return sid.value

    def remove_system(self, sys_id):
        self.ctx.check(lib().flyg_server_remove_system(self.h, int(sys_id)))

    def submit(self, sys_id, centres, radii, kicks):
        """kicks[k] = (atom indices, positions (m, 3), axis (m, 3)) of search k's perturbed atoms."""
        c = np.ascontiguousarray(np.atleast_1d(centres), dtype=np.int64)
        r = np.ascontiguousarray(np.atleast_1d(radii), dtype=np.float64)
        k = len(c)
        offs = np.zeros(k + 1, np.int64)
        offs[1:] = np.cumsum([len(kk[0]) for kk in kicks])
        atoms = np.ascontiguousarray(np.concatenate([np.asarray(kk[0], np.int64) for kk in kicks]) if k else
                                     np.zeros(0, np.int64), dtype=np.int64)
        kp = np.ascontiguousarray(np.concatenate([np.asarray(kk[1], np.float64).reshape(-1, 3) for kk in kicks])
                                  if k else np.zeros((0, 3)), dtype=np.float64)
        ka = np.ascontiguousarray(np.concatenate([np.asarray(kk[2], np.float64).reshape(-1, 3) for kk in kicks])
                                  if k else np.zeros((0, 3)), dtype=np.float64)
        ids = np.zeros(k, np.int64)
        self.ctx.check(lib().flyg_server_submit(self.h, int(sys_id), k, c.ctypes.data_as(_i64), r.ctypes.data_as(_d),
                                                offs.ctypes.data_as(_i64), atoms.ctypes.data_as(_i64),
                                                kp.ctypes.data_as(_d), ka.ctypes.data_as(_d), ids.ctypes.data_as(_i64)))
        return ids

    def submit_perturbed(self, sys_id, centres, radii, prng, stddev=0.6, r_pert=4.0):
        """Searches started by Master::perturb; prng: C-contiguous (k, 4) uint64 states, advanced in place."""
        c = np.ascontiguousarray(np.atleast_1d(centres), dtype=np.int64)
        r = np.ascontiguousarray(np.atleast_1d(radii), dtype=np.float64)
        k = len(c)
        if not (isinstance(prng, np.ndarray) and prng.dtype == np.uint64 and prng.flags.c_contiguous and
                prng.shape == (k, 4)):
            raise FlyError("prng must be a C-contiguous (k, 4) uint64 array")
        ids = np.zeros(k, np.int64)
        self.ctx.check(lib().flyg_server_submit_perturbed(self.h, int(sys_id), k, c.ctypes.data_as(_i64),
                                                          r.ctypes.data_as(_d),
                                                          prng.ctypes.data_as(C.POINTER(C.c_uint64)),
                                                          float(stddev), float(r_pert), ids.ctypes.data_as(_i64)))
        return ids

    def progress(self, wait_ms=0.0, cap=4096):
        """Ids of searches that finished since the last call (waits up to wait_ms for one)."""
        buf = np.zeros(cap, np.int64)
        n = C.c_int()
        self.ctx.check(lib().flyg_server_progress(self.h, float(wait_ms), buf.ctypes.data_as(_i64), cap, C.byref(n)))
        return buf[: n.value].copy()

    def result(self, sid, positions=False, release=False):
        st = SearchStats()
        nm = C.c_int()
        self.ctx.check(lib().flyg_server_result(self.h, int(sid), C.byref(st), C.byref(nm), None, None, None))
        out = st.as_dict()
        out["barrier"] = st.energy
        out["exit"] = st.exit_code
        out["n_mobile"] = nm.value
        if positions:
            atoms = np.zeros(nm.value, np.int64)
            pos = np.zeros((nm.value, 3))
            ax = np.zeros((nm.value, 3))
            self.ctx.check(lib().flyg_server_result(self.h, int(sid), None, None, atoms.ctypes.data_as(_i64),
                                                    pos.ctypes.data_as(_d), ax.ctypes.data_as(_d)))
            out.update(atoms=atoms, pos=pos, axis=ax)
        if release:
            self.ctx.check(lib().flyg_server_release(self.h, int(sid)))
        return out

    def set_mechanisms(self, **kw):
        self.mopt = mech_options(**kw)
        self.ctx.check(lib().flyg_server_set_mechanisms(self.h, C.byref(self.mopt)))

    def mechanism(self, sid, n_mobile=None, positions=False):
        """Mechanism of a finished search (call before release); rev/fwd per mobile atom."""
        m = Mechanism()
        rev = fwd = None
        if positions:
            rev = np.zeros((n_mobile, 3))
            fwd = np.zeros((n_mobile, 3))
        self.ctx.check(lib().flyg_server_mechanism(self.h, int(sid), C.byref(m),
                                                   rev.ctypes.data_as(_d) if positions else None,
                                                   fwd.ctypes.data_as(_d) if positions else None))
        out = m.as_dict()
        if positions:
            out.update(rev=rev, fwd=fwd)
        return out

    def status(self):
        st = ServerState()
        self.ctx.check(lib().flyg_server_status(self.h, C.byref(st)))
        return st.as_dict()


def perturb(box: Box, pos, frozen, centre, prng, stddev=0.6, r_pert=4.0):
    """Master::perturb (flyg_perturb): (positions, axis, sd); prng (4 uint64) advances in place."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = len(pos)
    f = np.ascontiguousarray(frozen, dtype=np.uint8).reshape(n)
    if not (isinstance(prng, np.ndarray) and prng.dtype == np.uint64 and prng.shape == (4,)):
        raise FlyError("prng must be a (4,) uint64 array")
    out = np.zeros((n, 3))
    ax = np.zeros((n, 3))
    sd = C.c_double()
    b, p = box.args()
    rc = lib().flyg_perturb(b, p, pos.ctypes.data_as(_d), f.ctypes.data_as(_u8), n, int(centre), float(stddev),
                            float(r_pert), prng.ctypes.data_as(C.POINTER(C.c_uint64)), out.ctypes.data_as(_d),
                            ax.ctypes.data_as(_d), C.byref(sd))
    if rc:
        _raise(None, rc)
    return out, ax, sd.value


def theta_tol(tot):
    """Master::find_batch's collision angle (radians) after `tot` searches at one site."""
    t = C.c_double()
    rc = lib().flyg_theta_tol(int(tot), C.byref(t))
    if rc:
        _raise(None, rc)
    return t.value
