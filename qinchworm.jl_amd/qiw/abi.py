"""ctypes mirror of include/qiw.h and loader of the native library libqiw.so.

The structures here are the drop-in boundary: the same layout a Julia `ccall` shim fills
from an `Expansion` (see qinchworm.jl_amd/julia/QInchwormAMD.jl).
"""
import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(PKG_DIR), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libqiw.so")

QIW_OK = 0
QIW_E_ARG, QIW_E_NOT_POW2, QIW_E_OFFDIAG, QIW_E_NOT_SBM = -1, -2, -3, -4
QIW_E_DIM, QIW_E_STATE, QIW_E_DEVICE, QIW_E_LIMIT, QIW_E_NO_GPU = -5, -6, -7, -8, -9

QIW_GF_GRID_LINEAR, QIW_GF_CUBIC_SPLINE, QIW_GF_INC_SPLINE, QIW_GF_DLR = 0, 1, 2, 3
QIW_PREC_FP64, QIW_PREC_FP32 = 0, 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)


class qiw_sbm(C.Structure):
    _fields_ = [("n_blocks", C.c_int32), ("s_i", _i32p), ("s_f", _i32p), ("data", _f64p)]


class qiw_model(C.Structure):
    _fields_ = [("beta", C.c_double), ("n_tau", C.c_int32), ("n_sectors", C.c_int32),
                ("sector_dim", _i32p), ("energies", _f64p), ("ppgf_kind", C.c_int32),
                ("n_pairs", C.c_int32), ("pair_op_i", C.POINTER(qiw_sbm)),
                ("pair_op_f", C.POINTER(qiw_sbm)), ("pair_prop", _f64p),
                ("pair_prop_kind", _i32p), ("pair_prop_xi", _i32p), ("n_corr", C.c_int32),
                ("corr_A", C.POINTER(qiw_sbm)), ("corr_B", C.POINTER(qiw_sbm)),
                ("dlr_rank", C.c_int32), ("pair_dlr_omega", _f64p), ("pair_dlr_coef", _f64p)]


class qiw_rand(C.Structure):
    _fields_ = [("N_seqs", C.c_int32), ("scramble", C.c_int32), ("seed", C.c_uint64), ("target_std", C.c_double)]


class qiw_run(C.Structure):
    _fields_ = [("n_orders", C.c_int32), ("orders", _i32p), ("n_orders_bare", C.c_int32),
                ("orders_bare", _i32p), ("N_samples", C.c_int64), ("n_pts_after_max", C.c_int32),
                ("precision", C.c_int32), ("rand_params", C.POINTER(qiw_rand))]


class qiw_warp_config(C.Structure):
    _fields_ = [("n_layers", C.c_int32), ("n_bins", C.c_int32), ("hidden", C.c_int32),
                ("iters", C.c_int32), ("batch", C.c_int32), ("n_train_steps", C.c_int32),
                ("n_train", C.c_int64), ("lr", C.c_double), ("seed", C.c_uint64)]


class qiw_tuning(C.Structure):
    _fields_ = [("chunk_words", C.c_int32), ("samples_per_task", C.c_int32), ("chunk_cost", C.c_int64),
                ("n_tried", C.c_int32), ("pad_", C.c_int32), ("best_ms", C.c_double)]


class qiw_stats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("kernel_ms", C.c_double),
                ("n_kernel_launches", C.c_int64), ("diagram_evals", C.c_int64),
                ("leaves", C.c_int64), ("flops", C.c_double), ("graph_replays", C.c_int64)]


def ptr(a, ctype):
    """Pointer to the data of a contiguous numpy array (None for None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


class ModelStruct:
    """A qiw_model plus the numpy buffers it points into (kept alive together)."""

    def __init__(self, beta, n_tau, dims, energies, ppgf_kind, pairs, props, prop_kinds,
                 prop_xis, corr_ops, dlr=None):
        """dlr: None, or (omega [n_pairs][rank], coef [n_pairs][rank] complex) for the
        QIW_GF_DLR pairs (rows of other pairs are ignored)."""
        self._keep = []
        self.dims = np.ascontiguousarray(dims, dtype=np.int32)
        self.energies = np.ascontiguousarray(np.concatenate([np.asarray(e, float) for e in energies]))
        self.props = np.ascontiguousarray(np.asarray(props, dtype=np.complex128).reshape(len(pairs), n_tau))
        self.prop_kinds = np.ascontiguousarray(prop_kinds, dtype=np.int32)
        self.prop_xis = np.ascontiguousarray(prop_xis, dtype=np.int32)
        op_i = (qiw_sbm * max(len(pairs), 1))()
        op_f = (qiw_sbm * max(len(pairs), 1))()
        for p, (oi, of) in enumerate(pairs):
            op_i[p] = self._sbm(oi)
            op_f[p] = self._sbm(of)
        cA = (qiw_sbm * max(len(corr_ops), 1))()
        cB = (qiw_sbm * max(len(corr_ops), 1))()
        for c, (a, b) in enumerate(corr_ops):
            cA[c] = self._sbm(a)
            cB[c] = self._sbm(b)
        self._keep += [op_i, op_f, cA, cB]
        self.struct = qiw_model(
            beta=float(beta), n_tau=int(n_tau), n_sectors=len(self.dims),
            sector_dim=ptr(self.dims, C.c_int32), energies=ptr(self.energies, C.c_double),
            ppgf_kind=int(ppgf_kind), n_pairs=len(pairs), pair_op_i=op_i, pair_op_f=op_f,
            pair_prop=self.props.ctypes.data_as(_f64p),
            pair_prop_kind=ptr(self.prop_kinds, C.c_int32),
            pair_prop_xi=ptr(self.prop_xis, C.c_int32), n_corr=len(corr_ops), corr_A=cA, corr_B=cB)
        if dlr is not None:
            self.dlr_w = np.ascontiguousarray(dlr[0], dtype=np.float64)
            self.dlr_c = np.ascontiguousarray(dlr[1], dtype=np.complex128)
            self.struct.dlr_rank = int(self.dlr_w.shape[1])
            self.struct.pair_dlr_omega = self.dlr_w.ctypes.data_as(_f64p)
            self.struct.pair_dlr_coef = self.dlr_c.ctypes.data_as(_f64p)

    def _sbm(self, blocks):
        """blocks: dict s_i -> (s_f, complex matrix)."""
        keys = sorted(blocks)
        s_i = np.array(keys, dtype=np.int32)
        s_f = np.array([blocks[k][0] for k in keys], dtype=np.int32)
        data = [np.asarray(blocks[k][1], dtype=np.complex128).ravel(order="F") for k in keys]
        data = np.ascontiguousarray(np.concatenate(data) if data else np.zeros(1, np.complex128))
        self._keep += [s_i, s_f, data]
        return qiw_sbm(n_blocks=len(keys), s_i=ptr(s_i, C.c_int32), s_f=ptr(s_f, C.c_int32),
                       data=data.ctypes.data_as(_f64p))

    def ref(self):
        return C.byref(self.struct)


class RunStruct:
    """qiw_run; rand: None (RandomizationParams()) or (N_seqs, scramble, seed, target_std)."""

    def __init__(self, orders, orders_bare, N_samples, n_pts_after_max=0, precision=QIW_PREC_FP64, rand=None):
        self.orders = np.ascontiguousarray(list(orders), dtype=np.int32)
        self.orders_bare = np.ascontiguousarray(list(orders_bare), dtype=np.int32)
        self.rand = None
        if rand is not None:
            n_seqs, scramble, seed, target_std = rand
            self.rand = qiw_rand(N_seqs=int(n_seqs), scramble=int(scramble), seed=int(seed) & (2 ** 64 - 1),
                                 target_std=float(target_std))
        self.struct = qiw_run(n_orders=len(self.orders), orders=ptr(self.orders, C.c_int32),
                              n_orders_bare=len(self.orders_bare),
                              orders_bare=ptr(self.orders_bare, C.c_int32),
                              N_samples=int(N_samples), n_pts_after_max=int(n_pts_after_max or 0),
                              precision=int(precision),
                              rand_params=C.pointer(self.rand) if self.rand is not None else None)

    def ref(self):
        return C.byref(self.struct)


# Exported symbols of libqiw (every declaration of include/qiw.h)
EXPORTS = ["qiw_create", "qiw_destroy", "qiw_last_error", "qiw_abi_version", "qiw_set_model",
           "qiw_set_sobol", "qiw_set_scrambles", "qiw_set_points", "qiw_scramble_dims", "qiw_set_ppgf", "qiw_get_ppgf", "qiw_inchworm", "qiw_inchworm_step",
           "qiw_inchworm_step_bare", "qiw_inchworm_step_noalloc", "qiw_correlator_2p", "qiw_correlator_2p_std", "qiw_diff_inchworm",
           "qiw_partition_function",
           "qiw_density_matrix", "qiw_sobol_points", "qiw_simplex_points", "qiw_get_stats",
           "qiw_warp_init", "qiw_warp_params", "qiw_warp_apply", "qiw_warp_loss_grad", "qiw_warp_train",
           "qiw_warp_enable", "qiw_set_tuning", "qiw_get_tuning", "qiw_autotune", "qiw_set_graphs"]

_lib = None


def load_library(path=None):
    """Load libqiw.so (RTLD_GLOBAL so a Julia-style single copy is shared).  Raises if
    the native library has not been built: there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"libqiw native library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    vp = C.c_void_p
    lib.qiw_create.argtypes = [C.POINTER(vp), C.c_int]
    lib.qiw_destroy.argtypes = [vp]
    lib.qiw_destroy.restype = None
    lib.qiw_last_error.argtypes = [vp]
    lib.qiw_last_error.restype = C.c_char_p
    lib.qiw_set_model.argtypes = [vp, C.POINTER(qiw_model)]
    lib.qiw_set_sobol.argtypes = [vp, C.c_int32, _u32p, _u32p]
    lib.qiw_set_scrambles.argtypes = [vp, C.c_int64, C.c_int32, _u32p, _u32p]
    lib.qiw_set_points.argtypes = [vp, C.c_int64, C.c_int64, C.c_int32, _f64p]
    lib.qiw_scramble_dims.argtypes = [vp, C.POINTER(qiw_run), C.c_int32, C.c_int64, _i32p]
    lib.qiw_scramble_dims.restype = C.c_int64
    lib.qiw_set_ppgf.argtypes = [vp, C.c_int32, _f64p, _f64p]
    lib.qiw_get_ppgf.argtypes = [vp, _i32p, _f64p, _f64p]
    lib.qiw_inchworm.argtypes = [vp, C.POINTER(qiw_run), _f64p, _f64p, _f64p]
    lib.qiw_inchworm_step.argtypes = [vp, C.POINTER(qiw_run), C.c_int32, _f64p, _f64p]
    lib.qiw_inchworm_step_bare.argtypes = [vp, C.POINTER(qiw_run), _f64p, _f64p]
    lib.qiw_inchworm_step_noalloc.argtypes = [vp, C.POINTER(qiw_run), C.c_int32, _f64p, _f64p]
    lib.qiw_correlator_2p.argtypes = [vp, C.POINTER(qiw_run), _f64p]
    lib.qiw_correlator_2p_std.argtypes = [vp, C.POINTER(qiw_run), _f64p, _f64p]
    lib.qiw_diff_inchworm.argtypes = [vp, C.POINTER(qiw_run), _f64p, _f64p, _f64p]
    lib.qiw_partition_function.argtypes = [vp, _f64p]
    lib.qiw_density_matrix.argtypes = [vp, _f64p]
    lib.qiw_sobol_points.argtypes = [vp, C.c_int32, C.c_int64, C.c_int64, _f64p]
    lib.qiw_simplex_points.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                       C.c_double, C.c_int64, _f64p]
    lib.qiw_get_stats.argtypes = [vp, C.POINTER(qiw_stats)]
    lib.qiw_warp_init.argtypes = [vp, C.POINTER(qiw_warp_config)]
    lib.qiw_warp_params.argtypes = [vp, C.POINTER(C.c_int64), _f32p, _f32p]
    lib.qiw_warp_apply.argtypes = [vp, C.c_int32, C.c_int64, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p]
    lib.qiw_warp_loss_grad.argtypes = [vp, C.c_int64, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f32p]
    lib.qiw_warp_train.argtypes = [vp, C.POINTER(qiw_run), _f64p]
    lib.qiw_warp_enable.argtypes = [vp, C.c_int32]
    lib.qiw_set_tuning.argtypes = [vp, C.POINTER(qiw_tuning)]
    lib.qiw_get_tuning.argtypes = [vp, C.POINTER(qiw_tuning)]
    lib.qiw_autotune.argtypes = [vp, C.POINTER(qiw_run), C.POINTER(qiw_tuning)]
    lib.qiw_set_graphs.argtypes = [vp, C.c_int32]
    _lib = lib
    return lib


class QiwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libqiw error {code}: {msg}")
        self.code = code


class Context:
    """Owns one qiw_ctx (one GPU)."""

    def __init__(self, device=0):
        self.lib = load_library()
        self.h = C.c_void_p()
        rc = self.lib.qiw_create(C.byref(self.h), int(device))
        if rc != QIW_OK:
            msg = self.lib.qiw_last_error(self.h).decode() if self.h else "qiw_create failed"
            raise QiwError(rc, msg)
        self._model = None

    def check(self, rc):
        if rc != QIW_OK:
            raise QiwError(rc, self.lib.qiw_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.qiw_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
