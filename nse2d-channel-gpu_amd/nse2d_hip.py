"""ctypes binding of libnse2d_hip.so (the C-ABI declared in include/nse2d_hip.h).

This module is plumbing for Python callers (tests, bench.py): it loads the in-tree HIP library,
declares the argument types of every entry point and offers small device-buffer helpers built on
the library's own runtime calls. There is no CPU fallback: if the library is missing or no GPU is
visible, the calls that need one raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnse2d_hip.so")
EXE_PATH = os.path.join(HERE, "nse2d")
EXE_FREESLIP_PATH = os.path.join(HERE, "nse2d_freeslip")

_d, _i, _sz = ctypes.c_double, ctypes.c_int, ctypes.c_size_t
_p, _vp = ctypes.POINTER(ctypes.c_double), ctypes.c_void_p
_stream = ctypes.c_void_p

# name: (restype, [argtypes]); pointers to device memory are passed as c_void_p
SIGNATURES = {
    # test_api.cu wrappers
    "testapi_uv_from_psi_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "testapi_omega_from_psi_interior_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _stream]),
    "testapi_thom_wall_vorticity_mapped": (None, [_vp, _vp, _i, _i, _vp, _d, _d, _stream]),
    "testapi_arakawa_J_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "testapi_compute_adv_from_uv_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "testapi_compute_adv_flux_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "testapi_arakawa_J_mapped_split": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    # operator launchers
    "launch_uv_from_psi_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "launch_omega_from_psi_interior_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _stream]),
    "launch_thom_wall_vorticity_mapped": (None, [_vp, _vp, _i, _i, _vp, _d, _d, _stream]),
    "launch_arakawa_J_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp, _stream]),
    "launch_laplacian_fd_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _stream]),
    "launch_rhs_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _vp, _stream]),
    "launch_mms_add_forcing": (None, [_vp, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _i, _vp, _stream]),
    "launch_axpy": (None, [_vp, _vp, _vp, _d, _sz, _stream]),
    "launch_ssprk2": (None, [_vp, _vp, _vp, _vp, _d, _sz, _stream]),
    "launch_ssprk3": (None, [_vp, _vp, _vp, _vp, _d, _sz, _stream]),
    "nse2d_set_wall_type": (None, [_i]),
    "nse2d_get_wall_type": (_i, []),
    # Poisson
    "poisson2d_create": (_vp, [_i, _i, _d, _stream]),
    "poisson2d_enable_mapped": (_i, [_vp, _vp, _vp, _d, _stream]),
    "poisson2d_solve": (None, [_vp, _vp, _vp, _d, _d, _d, _stream]),
    "poisson2d_set_profile": (None, [_vp, _i]),
    "poisson2d_profile_reset": (None, [_vp]),
    "poisson2d_profile_get": (None, [_vp, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    "poisson2d_destroy": (None, [_vp]),
    # diagnostics
    "launch_enstrophy_diss_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    "launch_enstrophy_diss_mapped_full": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    "launch_enstrophy_sbp_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    "launch_kinetic_energy_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _stream]),
    "launch_energy_dissipation_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    "launch_power_input_mapped": (None, [_vp, _i, _i, _d, _d, _d, _i, _vp, _vp, _vp, _vp, _stream]),
    "launch_enstrophy_input_mapped_full": (None, [_vp, _i, _i, _d, _d, _i, _vp, _vp, _vp, _stream]),
    "launch_enstrophy_input_mapped": (None, [_vp, _i, _i, _d, _d, _i, _vp, _vp, _vp, _stream]),
    "launch_row_moments_from_psi_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp, _stream]),
    "launch_mean_wall_omega": (None, [_vp, _i, _i, _vp, _vp, _stream]),
    "launch_mean_wall_omega_omegay_mapped": (None, [_vp, _i, _i, _d, _vp, _vp, _vp, _stream]),
    "launch_enstrophy_diss_balance_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _stream]),
    "launch_omega_rhs_balance_mapped": (None, [_vp, _vp, _i, _i, _vp, _vp, _stream]),
    "launch_max_uv_from_psi_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    "nse2d_diag_row": (_i, [_vp, _vp, _vp, _i, _i, _d, _d, _d, _d, _i, _vp, _vp, _vp, _vp, _vp, _vp, _stream]),
    "launch_max_uv_from_psi_mapped_noalloc": (_i, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _stream]),
    # whole step
    "nse2d_channel_create": (_vp, [_vp, _stream]),
    "nse2d_channel_destroy": (None, [_vp]),
    "nse2d_channel_field": (_vp, [_vp, _i]),
    "nse2d_channel_init": (None, [_vp]),
    "nse2d_channel_step": (None, [_vp, _d, _d, _i]),
    "nse2d_channel_step_auto": (None, [_vp, _i]),
    "nse2d_channel_time_ptr": (_vp, [_vp]),
    "nse2d_channel_dt_ptr": (_vp, [_vp]),
    "nse2d_channel_wall_bookkeeping": (None, [_vp, _p, _d]),
    "nse2d_channel_profile": (None, [_vp, _i]),
    "nse2d_channel_profile_every": (None, [_vp, _i]),
    "nse2d_channel_profile_read": (None, [_vp, _p, ctypes.POINTER(ctypes.c_longlong)]),
    # spectra
    "nse2d_spectra1d_create": (_vp, [_i, _i, _d, _stream]),
    "nse2d_spectra1d_accumulate": (None, [_vp, _vp, _vp, _stream]),
    "nse2d_spectra1d_read": (_i, [_vp, _vp, _vp, _vp, _stream]),
    "nse2d_spectra1d_destroy": (None, [_vp]),
    # learned closure network
    "nse2d_cnn_create": (_vp, [_i, _i, _i, ctypes.c_ulonglong, _stream]),
    "nse2d_cnn_destroy": (None, [_vp]),
    "nse2d_cnn_num_params": (_sz, [_vp]),
    "nse2d_cnn_get_params": (None, [_vp, _vp]),
    "nse2d_cnn_set_params": (None, [_vp, _vp]),
    "nse2d_cnn_get_grads": (None, [_vp, _vp]),
    "nse2d_cnn_forward": (_i, [_vp, _vp, _vp, _i]),
    "nse2d_cnn_backward": (_i, [_vp, _vp, _i]),
    "nse2d_cnn_infer": (_i, [_vp, _vp, _vp, _i]),
    "nse2d_cnn_set_direct": (_i, [_vp, _i]),
    "nse2d_cnn_train_step": (_d, [_vp, _vp, _vp, _i, ctypes.c_float]),
    "nse2d_cnn_adam_step": (_i, [_vp, ctypes.c_float]),
    "nse2d_cnn_eval_loss": (_d, [_vp, _vp, _vp, _i]),
    # learned closure of a coarse channel (LES)
    "nse2d_channel_set_closure": (_i, [_vp, _vp, _vp]),
    "nse2d_channel_stream": (_vp, [_vp]),
    "nse2d_les_create": (_vp, [_vp, _vp, _i, _i, ctypes.c_ulonglong, _stream]),
    "nse2d_les_destroy": (None, [_vp]),
    "nse2d_les_cnn": (_vp, [_vp]),
    "nse2d_les_filter": (_i, [_vp, _vp, _vp]),
    "nse2d_les_pair": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "nse2d_les_num_pairs": (_i, [_vp]),
    "nse2d_les_clear_pairs": (None, [_vp]),
    "nse2d_les_fit_scales": (_i, [_vp, _i]),
    "nse2d_les_get_scales": (None, [_vp, _p, _p]),
    "nse2d_les_set_scales": (_i, [_vp, _p, _p]),
    "nse2d_les_train": (_d, [_vp, _i, _i, _i, ctypes.c_float, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float)]),
    "nse2d_les_eval": (_i, [_vp, _i, _i, _p, _p]),
    "nse2d_les_closure": (_i, [_vp, _vp, _vp, _vp]),
    "nse2d_les_set_constraints": (_i, [_vp, _i]),
    "nse2d_les_attach": (_i, [_vp, _vp, _i]),
    # ensembles
    "nse2d_ensemble_create": (_vp, [_vp, _i, _stream]),
    "nse2d_ensemble_destroy": (None, [_vp]),
    "nse2d_ensemble_size": (_i, [_vp]),
    "nse2d_ensemble_field": (_vp, [_vp, _i, _i]),
    "nse2d_ensemble_scalars": (_vp, [_vp]),
    "nse2d_ensemble_stream": (_vp, [_vp]),
    "nse2d_ensemble_init": (None, [_vp]),
    "nse2d_ensemble_step_auto": (None, [_vp, _i]),
    "nse2d_ensemble_set_graph": (None, [_vp, _i]),
    "nse2d_ensemble_sync": (None, [_vp]),
    "nse2d_ensemble_autotune": (_i, [_vp, _i, ctypes.c_char_p, _sz]),
    "nse2d_host_metrics": (_i, [_i, _d, ctypes.c_char_p, _d, _vp, _vp, _vp, _vp, _vp]),
    "nse2d_host_init_omega": (_i, [_i, ctypes.POINTER(ctypes.c_char_p), _vp]),
    # runtime
    "nse2d_dev_alloc": (_vp, [_sz]),
    "nse2d_dev_free": (None, [_vp]),
    "nse2d_memcpy_h2d": (None, [_vp, _vp, _sz]),
    "nse2d_memcpy_d2h": (None, [_vp, _vp, _sz]),
    "nse2d_memcpy_d2d": (None, [_vp, _vp, _sz]),
    "nse2d_device_sync": (None, []),
    "nse2d_stream_create": (_vp, []),
    "nse2d_stream_destroy": (None, [_vp]),
    "nse2d_stream_sync": (None, [_vp]),
    "nse2d_device_count": (_i, []),
    "nse2d_last_error": (ctypes.c_char_p, []),
    "nse2d_version": (ctypes.c_char_p, []),
    "nse2d_main": (_i, [_i, ctypes.POINTER(ctypes.c_char_p)]),
}

FIELD_OMEGA, FIELD_PSI, FIELD_RHS, FIELD_W1, FIELD_W2 = 0, 1, 2, 3, 4
PROF_CLASSES = ["stage", "fft_forward", "thomas", "fft_inverse", "thom", "cfl"]
STEP_STORE_RHS, STEP_WALL_BOOKKEEPING = 1, 2
CLOSURE_DISSIPATIVE, CLOSURE_MEAN_FREE = 1, 2
# outputs of nse2d_diag_row (NSE2D_DIAG_* of include/nse2d_hip.h)
(DIAG_UMAX, DIAG_VMAX, DIAG_VMAXM, DIAG_K, DIAG_OMEGA_W2, DIAG_OMEGA_GRAD2, DIAG_OMEGA_RHS, DIAG_PIN, DIAG_EPS,
 DIAG_POMEGA, DIAG_WALL_OMEGA_OMEGAY_BOT, DIAG_WALL_OMEGA_OMEGAY_TOP, DIAG_WALL_OMEGA_BOT, DIAG_WALL_OMEGA_TOP,
 DIAG_NOUT) = range(15)


class ChannelDesc(ctypes.Structure):
    """nse2d_channel_desc of include/nse2d_hip.h (field order must match)."""
    _fields_ = [("Nx", _i), ("Ny", _i), ("Lx", _d), ("h", _d), ("nu", _d), ("F0", _d), ("lin_drag", _d),
                ("nforce", _i), ("psi_bot", _d), ("psi_top", _d), ("mms", _i), ("mms_A0", _d),
                ("mms_eps", _d), ("mms_Om", _d), ("mms_kx", _i), ("cfl", _d), ("cvisc", _d),
                ("dt_max", _d), ("dy_min", _d), ("d_a_node", _vp), ("d_a_edge", _vp), ("d_y_node", _vp),
                ("d_dy_edge", _vp), ("d_w_node", _vp), ("wall_type", _i)]


_LIB = None


def lib():
    """Load libnse2d_hip.so once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C nse2d-channel-gpu_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def have_gpu():
    try:
        return lib().nse2d_device_count() > 0
    except (OSError, RuntimeError):
        return False


def require_gpu():
    if lib().nse2d_device_count() <= 0:
        raise RuntimeError("no HIP device is visible; the nse2d HIP path needs a GPU")


class DeviceBuffer:
    """A device allocation of float64 values made with nse2d_dev_alloc."""

    def __init__(self, n, init=None):
        self.n = int(n)
        self.ptr = lib().nse2d_dev_alloc(max(1, self.n) * 8)
        if self.ptr is None:
            raise MemoryError("nse2d_dev_alloc failed")
        if init is not None:
            self.upload(init)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return cls(a.size, a)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch: buffer {self.n}, array {a.size}")
        lib().nse2d_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes)

    def download(self, shape=None):
        out = np.empty(self.n, dtype=np.float64)
        lib().nse2d_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out.reshape(shape) if shape is not None else out

    def free(self):
        if self.ptr:
            lib().nse2d_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceArray:
    """A device copy of a numpy array of any dtype (nse2d_dev_alloc), with download()."""

    def __init__(self, a=None, shape=None, dtype=np.float32):
        if a is not None:
            a = np.ascontiguousarray(a)
            shape, dtype = a.shape, a.dtype
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().nse2d_dev_alloc(max(1, self.nbytes))
        if self.ptr is None:
            raise MemoryError("nse2d_dev_alloc failed")
        if a is not None:
            lib().nse2d_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes)

    def download(self):
        out = np.empty(self.shape, self.dtype)
        lib().nse2d_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes)
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        lib().nse2d_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes)

    def free(self):
        if self.ptr:
            lib().nse2d_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ptr, n, shape=None):
    out = np.empty(int(n), dtype=np.float64)
    lib().nse2d_memcpy_d2h(out.ctypes.data, ptr, out.nbytes)
    return out.reshape(shape) if shape is not None else out


def upload_ptr(ptr, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    lib().nse2d_memcpy_h2d(ptr, a.ctypes.data, a.nbytes)


def run_main(args):
    """nse2d_main with the reference command line (list of strings, without the program name)."""
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    return lib().nse2d_main(len(args), argv)


def last_error():
    return lib().nse2d_last_error().decode()
