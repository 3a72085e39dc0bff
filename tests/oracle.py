"""ctypes binding of the oracle (oracle/liboracle_nse2d.so): the FP64 CPU restatement of the
reference path used as the checker in these tests. Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle_nse2d.so")
EXE_PATH = os.path.join(ORACLE_DIR, "oracle_nse2d")

_d, _i, _sz, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p
SIG = {
    "orc_build_y_metrics": (_i, [_i, _d, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    "orc_table_y_metrics": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_uv_from_psi_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "orc_omega_from_psi_interior_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp]),
    "orc_thom_wall_vorticity_mapped": (None, [_vp, _vp, _i, _i, _vp, _d, _d, _i]),
    "orc_arakawa_J_mapped": (None, [_vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "orc_arakawa_J_mapped_split": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "orc_compute_adv_from_uv_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "orc_compute_adv_flux_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _vp]),
    "orc_laplacian_fd_mapped": (None, [_vp, _vp, _i, _i, _d, _d, _vp, _vp]),
    "orc_rhs_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _vp]),
    "orc_mms_add_forcing": (None, [_vp, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _i, _vp]),
    "orc_axpy": (None, [_vp, _vp, _vp, _d, _sz]),
    "orc_ssprk2": (None, [_vp, _vp, _vp, _vp, _d, _sz]),
    "orc_ssprk3": (None, [_vp, _vp, _vp, _vp, _d, _sz]),
    "orc_fft_z2z": (None, [_vp, _i, _i, _i]),
    "orc_dft_naive": (None, [_vp, _vp, _i, _i]),
    "orc_set_eval_order": (None, [_i]),
    "orc_get_eval_order": (_i, []),
    "orc_fill_kx2": (None, [_vp, _i, _d]),
    "orc_precompute_cp_mapped": (None, [_vp, _vp, _vp, _vp, _i, _i, _d]),
    "orc_poisson_tridiag_mapped": (None, [_vp, _vp, _vp, _vp, _vp, _i, _i, _d, _d, _d]),
    "orc_poisson_solve": (None, [_vp, _vp, _i, _i, _d, _vp, _vp, _d, _d, _d]),
    "orc_poisson_solve_ld": (None, [_vp, _vp, _i, _i, _d, _vp, _vp, _d, _d, _d]),
    "orc_spectra1d_accumulate": (None, [_vp, _vp, _i, _i, _vp]),
    "orc_enstrophy_diss_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp]),
    "orc_enstrophy_diss_mapped_full": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp]),
    "orc_enstrophy_diss_balance_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp]),
    "orc_enstrophy_sbp_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp]),
    "orc_kinetic_energy_mapped": (_d, [_vp, _i, _i, _d, _d, _vp, _vp]),
    "orc_energy_dissipation_mapped": (_d, [_vp, _vp, _i, _i, _d, _d, _vp, _vp, _vp]),
    "orc_power_input_mapped": (_d, [_vp, _i, _i, _d, _d, _d, _i, _vp, _vp, _vp]),
    "orc_enstrophy_input_mapped_full": (_d, [_vp, _i, _i, _d, _d, _i, _vp, _vp]),
    "orc_enstrophy_input_mapped": (_d, [_vp, _i, _i, _d, _d, _i, _vp, _vp]),
    "orc_row_moments_from_psi_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_mean_wall_omega": (None, [_vp, _i, _i, _vp]),
    "orc_mean_wall_omega_omegay_mapped": (None, [_vp, _i, _i, _d, _vp, _vp]),
    "orc_omega_rhs_balance_mapped": (_d, [_vp, _vp, _i, _i, _vp]),
    "orc_max_uv_from_psi_mapped": (None, [_vp, _i, _i, _d, _d, _vp, _vp]),
    "orc_init_omega": (_i, [_i, ctypes.POINTER(ctypes.c_char_p), _vp]),
    "orc_run_main": (_i, [_i, ctypes.POINTER(ctypes.c_char_p)]),
    "orc_time_steps": (_d, [_i, ctypes.POINTER(ctypes.c_char_p), _i, _i]),
    "orc_num_threads": (_i, []),
}

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.isfile(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in SIG.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def P(a):
    """Data pointer of a contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def argv(args):
    return (ctypes.c_char_p * len(args))(*[a.encode() for a in args])


class Metrics:
    """Grid metrics of grid.cu:7-94 computed by the oracle."""

    def __init__(self, Ny, h=1.0, stretch=1, beta=2.0):
        self.Ny, self.h = Ny, h
        self.y = np.zeros(Ny)
        self.a_node = np.zeros(Ny)
        self.a_edge = np.zeros(Ny - 1)
        self.dy_edge = np.zeros(Ny - 1)
        self.w_node = np.zeros(Ny)
        rc = lib().orc_build_y_metrics(Ny, h, stretch, beta, P(self.y), P(self.a_node), P(self.a_edge),
                                       P(self.dy_edge), P(self.w_node))
        assert rc == 0
        self.wt_node = np.zeros(Ny)
        self.wt_node[1:-1] = (2.0 / (Ny - 1)) / self.a_node[1:-1]


def table_metrics(y):
    """Metrics of a tabulated grid (grid.cu:46-78) from its Ny values of y, by the oracle:
    (y, a_node, a_edge, dy_edge, w_node)."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    Ny = len(y)
    out = [np.zeros(Ny), np.zeros(Ny), np.zeros(Ny - 1), np.zeros(Ny - 1), np.zeros(Ny)]
    rc = lib().orc_table_y_metrics(Ny, P(y), *[P(a) for a in out])
    assert rc == 0, rc
    return out


def init_omega(args, Nx, Ny):
    w = np.zeros(Nx * Ny)
    rc = lib().orc_init_omega(len(args), argv(args), P(w))
    assert rc == 0, rc
    return w.reshape(Ny, Nx)


def run_main(args):
    return lib().orc_run_main(len(args), argv(args))


def poisson_solve(omega, Lx, m, psi_bot, psi_top, extended=False):
    """The oracle's FP64 solve (poisson.cu restated), or with extended=True the same operator
    solved in 80-bit arithmetic (a measurement of FP64 rounding, no reference counterpart)."""
    Ny, Nx = omega.shape
    psi = np.zeros_like(omega)
    f = lib().orc_poisson_solve_ld if extended else lib().orc_poisson_solve
    f(P(np.ascontiguousarray(omega)), P(psi), Nx, Ny, Lx, P(m.a_node), P(m.a_edge), (Ny - 1) / 2.0, psi_bot, psi_top)
    return psi
