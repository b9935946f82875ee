"""Loader for libnbg_hip.so (the HIP kernels and the C ABI of include/nbg.h).

The library is built in-tree (`make -C nbodygradient.jl_amd/csrc` or
`__graft_entry__.build()`).  There is no CPU fallback: every integration
entry point runs on the GPU, and a missing library or device raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbg_hip.so")

D = C.POINTER(C.c_double)
I64 = C.POINTER(C.c_int64)
I32 = C.POINTER(C.c_int32)
U8 = C.POINTER(C.c_uint8)
VP = C.c_void_p


class NbgError(RuntimeError):
    pass


class BatchStats(C.Structure):
    _fields_ = [("system_steps", C.c_int64), ("transits", C.c_int64), ("newton_iters", C.c_int64),
                ("kernel_ms", C.c_double), ("launch_ms", C.c_double * 13), ("launch_units", (C.c_int64 * 13) * 13),
                ("launch_grid", C.c_int32 * 13), ("paired_systems", C.c_int32)]


_lib = None


def lib():
    """The loaded C library (raises NbgError when it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NbgError("libnbg_hip.so is not built: run `make -C nbodygradient.jl_amd/csrc`")
    L = C.CDLL(LIB_PATH)
    L.nbg_ctx_create.argtypes = [C.POINTER(VP), C.c_int]
    L.nbg_ctx_destroy.argtypes = [VP]
    L.nbg_last_error.argtypes = [VP]
    L.nbg_last_error.restype = C.c_char_p
    L.nbg_device_info.argtypes = [VP, C.c_char_p, C.c_int, C.POINTER(C.c_int)]
    L.nbg_lds_bytes_per_system.argtypes = [C.c_int]
    L.nbg_lds_bytes_per_system.restype = C.c_size_t
    st = [D, D, D, D, D, U8, D, D, D]  # x v xe ve m pair jac jacerr dqdt
    L.nbg_ahl21_step.argtypes = [VP, C.c_int32, C.c_double, C.c_int32] + st
    L.nbg_integrate_steps.argtypes = [VP, C.c_int32, C.c_double, C.c_int64, C.c_int32, D] + st
    L.nbg_integrate_to.argtypes = [VP, C.c_int32, C.c_double, C.c_double, C.c_int32, D] + st
    L.nbg_integrate_transits.argtypes = [VP, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32, D] + st + \
        [D, C.c_int64, D, D, D, I64]
    L.nbg_integrate_cartesian.argtypes = [VP, C.c_int32, C.c_double, C.c_double, C.c_int64, C.c_int32, D] + st + \
        [D, D, D, D, D, D, D]
    L.nbg_integrate_snapshot.argtypes = [VP, C.c_int32, C.c_double, C.c_int64, D, C.c_int32, D] + st + [D, D, D, D, D]
    L.nbg_batch_snapshot.argtypes = [VP, C.c_int32, I32, D, D, I64, D, C.c_int32, D, D, D, U8, D, D, D, D, D]
    L.nbg_batch_cartesian.argtypes = [VP, C.c_int32, I32, D, D, D, I64, C.c_int32, C.c_int32, D, D, D, U8,
                                      D, D, D, D, D, D, D]
    L.nbg_batch_create.argtypes = [VP, C.c_int32, I32, D, D, D, I32, I64, C.c_int32, C.c_int32, D, D, D, U8, D,
                                   C.POINTER(VP)]
    L.nbg_batch_create_elements.argtypes = [VP, C.c_int32, I32, D, D, D, I32, I64, C.c_int32, C.c_int32, D, D, U8,
                                            C.POINTER(VP)]
    L.nbg_batch_initial.argtypes = [VP, D, D, D]
    L.nbg_batch_autotune.argtypes = [VP, C.c_double, I32, D, C.c_int32]
    L.nbg_batch_autotune_candidate.argtypes = [C.c_int32, I32, I32, I32]
    L.nbg_batch_reset.argtypes = [VP]
    L.nbg_batch_run.argtypes = [VP, C.POINTER(BatchStats)]
    L.nbg_batch_round.argtypes = [VP, C.c_int64, C.c_int32, C.POINTER(BatchStats)]
    L.nbg_batch_autotune_round.argtypes = [VP, C.c_int64, I32, D, C.c_int32]
    L.nbg_batch_download.argtypes = [VP, D, D, D, I64, D, D]
    L.nbg_batch_destroy.argtypes = [VP]
    L.nbg_set_launch.argtypes = [VP, C.c_int32, C.c_int32]
    L.nbg_set_grid_cap.argtypes = [VP, C.c_int32]
    L.nbg_set_coop.argtypes = [VP, C.c_int32, C.c_int32]
    L.nbg_set_coop_wait.argtypes = [VP, C.c_int64]
    L.nbg_server_create.argtypes = [VP, C.c_int64, C.c_int64, C.POINTER(VP)]
    L.nbg_server_destroy.argtypes = [VP]
    L.nbg_server_submit.argtypes = [VP, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int64, C.c_int32,
                                    C.c_int32, D, D, D, U8, D, I64]
    L.nbg_server_step.argtypes = [VP, I32, D]
    L.nbg_server_finished.argtypes = [VP, I64, C.c_int32, I32]
    L.nbg_server_fetch.argtypes = [VP, C.c_int64, D, D, D, I64, D, D]
    L.nbg_server_status.argtypes = [VP, I64, I64, I64, I64]
    L.nbg_server_work.argtypes = [VP, I64, I64, D]
    L.nbg_server_footprint.argtypes = [C.c_int32, C.c_int32, C.c_int64]
    L.nbg_server_footprint.restype = C.c_int64
    L.nbg_copy_calibrate.argtypes = [VP, C.c_int64, C.c_int32, D]
    L.nbg_set_systems_per_wave.argtypes = [VP, C.c_int32]
    L.nbg_set_row_update.argtypes = [VP, C.c_int32]
    L.nbg_batch_autotune_systems_per_wave.argtypes = [C.c_int32]
    L.nbg_batch_autotune_systems_per_wave.restype = C.c_int32
    L.nbg_probe_row_update.argtypes = [VP, C.c_int32, C.c_int32, C.c_int32, D, D]
    L.nbg_libm_eval_device.argtypes = [VP, C.c_int32, C.c_int64, D, D]
    L.nbg_selftest_pair_host.argtypes = [D, D, C.c_double, C.c_double, C.c_int32, D]
    _lib = L
    return L


def dptr(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim <= 1):
        raise NbgError("expected a Fortran-ordered float64 array")
    return a.ctypes.data_as(D)


def i64ptr(a):
    if a.dtype != np.int64:
        raise NbgError("expected int64")
    return a.ctypes.data_as(I64)


def i32ptr(a):
    if a.dtype != np.int32:
        raise NbgError("expected int32")
    return a.ctypes.data_as(I32)


def u8ptr(a):
    if a is None:
        return None
    if a.dtype != np.uint8:
        raise NbgError("expected uint8")
    return a.ctypes.data_as(U8)


class Context:
    """A libnbg context on one GPU (C ABI: nbg_ctx_create/destroy)."""

    def __init__(self, device=0):
        L = lib()
        h = VP()
        rc = L.nbg_ctx_create(C.byref(h), int(device))
        if rc != 0:
            raise NbgError("nbg_ctx_create(device=%d) failed with %d: no usable HIP device" % (device, rc))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def check(self, rc, what):
        if rc != 0:
            msg = lib().nbg_last_error(self._h).decode()
            raise NbgError("%s failed (%d): %s" % (what, rc, msg))

    def info(self):
        name = C.create_string_buffer(128)
        cu = C.c_int(0)
        lib().nbg_device_info(self._h, name, 128, C.byref(cu))
        return name.value.decode(), cu.value

    def set_launch(self, max_groups_per_cu=0, streams=0):
        self.check(lib().nbg_set_launch(self._h, int(max_groups_per_cu), int(streams)), "nbg_set_launch")

    def set_grid_cap(self, grid_cap=0):
        self.check(lib().nbg_set_grid_cap(self._h, int(grid_cap)), "nbg_set_grid_cap")

    def set_systems_per_wave(self, k=1):
        """Batch runs and rounds: 1 system per wave, or 2 (transit runs of
        equal body count, n <= 9) in lockstep with their Kepler pairs sharing
        the wave's lanes; results are the same bits either way (C ABI:
        nbg_set_systems_per_wave)."""
        self.check(lib().nbg_set_systems_per_wave(self._h, int(k)), "nbg_set_systems_per_wave")

    def set_row_update(self, mode=0):
        """Form of the Jacobian updates (pair sweeps and phisalpha!'s
        jac_phi): 0 the reference's order (bitwise), 1 factored, without the
        zero-addend compensated passes (Jacobians and their gradients to
        ~1e-11 relative, dq/dt and transit times within 1e-12, states
        bitwise).  C ABI: nbg_set_row_update."""
        self.check(lib().nbg_set_row_update(self._h, int(mode)), "nbg_set_row_update")

    def set_coop(self, mode=1, max_systems=128):
        """Two-wave workgroups for the longest transit-search systems of batch
        runs: 0 off, 1 automatic (default), 2 the longest max_systems."""
        self.check(lib().nbg_set_coop(self._h, int(mode), int(max_systems)), "nbg_set_coop")

    def set_coop_wait(self, polls=0):
        """Diagnostic: bound each wait between the two waves of a cooperative
        workgroup (0: defaults)."""
        self.check(lib().nbg_set_coop_wait(self._h, int(polls)), "nbg_set_coop_wait")

    def close(self):
        if getattr(self, "_h", None):
            lib().nbg_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = {}


def default_context(device=0):
    ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        _default[device] = ctx
    return ctx
