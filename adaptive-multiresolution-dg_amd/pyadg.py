"""ctypes binding of libadg_hip.so (C-ABI in include/adg.h).

Arrays are ordered by element hash key, per-element tensors row-major.  The
library must be built in-tree (`make -C adaptive-multiresolution-dg_amd`);
loading fails loudly when it is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadg_hip.so")
_LIB = None

FIELD_UCOE, FIELD_RHS, FIELD_UP, FIELD_TN, FIELD_UINTP, FIELD_FUCOE, FIELD_FP = 0, 1, 2, 3, 4, 10, 20
TAB_ALPT_AT_LAGR, TAB_LA_UV, TAB_LA_UVX, TAB_LA_FLX = 0, 1, 2, 3
TAB_AA_UV, TAB_AA_UVX, TAB_AA_ULFT, TAB_AA_URGT, TAB_AA_UJP = 4, 5, 6, 7, 8
LU_FULL, LU_L, LU_U = 0, 1, 2
FLUX_BURGERS, FLUX_SINCOS, FLUX_VARCOEF, FLUX_VARCOEF_SCALED = 0, 1, 2, 3
INTERP_LAGRANGE, INTERP_HERMITE = 0, 1
REFINE_PREDICT, REFINE_LEARNED, REFINE_CURRENT = 0, 1, 2
PROPOSER_INPUTS, PROPOSER_PARAMS = 12, 5057
PREC_FP64, PREC_FP32_SPLIT = 0, 1
RK_EULER, RK2, RK3 = 0, 2, 3

POINT_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)


class RunStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("dof_stages", ctypes.c_int64), ("max_elem", ctypes.c_int64),
                ("seconds", ctypes.c_double), ("final_time", ctypes.c_double), ("redone", ctypes.c_int64),
                ("proposed", ctypes.c_int64)]


class Problem(ctypes.Structure):
    _fields_ = [("flux_kind", ctypes.c_int), ("flux_param", ctypes.c_double), ("alpha", ctypes.c_double * 4),
                ("wave_speed", ctypes.c_double * 4), ("cfl", ctypes.c_double), ("t_final", ctypes.c_double),
                ("adaptive", ctypes.c_int), ("eps", ctypes.c_double), ("eta", ctypes.c_double),
                ("max_steps", ctypes.c_int64), ("t_start", ctypes.c_double), ("form", ctypes.c_int)]


class ICParams(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("shift", ctypes.c_double), ("base", ctypes.c_double),
                ("ampl", ctypes.c_double), ("nbump", ctypes.c_int), ("bump_amp", ctypes.c_double * 8),
                ("bump_width", ctypes.c_double * 8), ("bump_center", (ctypes.c_double * 4) * 8),
                ("inside", ctypes.c_double), ("outside", ctypes.c_double), ("radius", ctypes.c_double),
                ("center", ctypes.c_double * 4), ("scale", ctypes.c_double * 4), ("offset", ctypes.c_double * 4)]


SEP_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
COEF_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C adaptive-multiresolution-dg_amd`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        P = ctypes.POINTER
        I32P, DP = P(ctypes.c_int32), P(dbl)
        FP, U8P = P(ctypes.c_float), P(ctypes.c_uint8)
        sig = {
            "adg_create": (i32, [i32, i32, i32, i32, P(vp)]),
            "adg_destroy": (i32, [vp]),
            "adg_last_error": (ctypes.c_char_p, [vp]),
            "adg_sync": (i32, [vp]),
            "adg_grid_init": (i32, [vp, i32, i32]),
            "adg_grid_set": (i32, [vp, I32P, I32P, i64]),
            "adg_num_elem": (i64, [vp]),
            "adg_grid_get": (i32, [vp, I32P, I32P, I32P]),
            "adg_max_level": (i32, [vp, I32P]),
            "adg_hash_key": (ctypes.c_int32, [I32P, I32P, i32]),
            "adg_field_set": (i32, [vp, i32, DP]),
            "adg_field_get": (i32, [vp, i32, DP]),
            "adg_sweep": (i32, [vp, i32, i32, i32, i32, I32P, I32P, dbl, DP, DP]),
            "adg_sweep_noalloc": (i32, [vp, i32, i32, i32, i32, I32P, I32P, dbl, DP, DP]),
            "adg_eval_up": (i32, [vp]),
            "adg_set_level_sum_cap": (i32, [vp, i32]),
            "adg_set_interpolation": (i32, [vp, i32]),
            "adg_set_precision": (i32, [vp, i32]),
            "adg_set_refine_mode": (i32, [vp, i32, dbl]),
            "adg_set_judge_defer": (i32, [vp, i32]),
            "adg_proposer_stats": (i32, [vp, P(i64), P(i64), P(i64)]),
            "adg_viscosity_elements": (i32, [vp, dbl, P(i64), I32P]),
            "adg_viscosity_flux": (i32, [vp, dbl, i32]),
            "adg_viscosity_operator": (i32, [vp, dbl, i32, ctypes.POINTER(i64)]),
            "adg_rhs_viscosity_operator": (i32, [vp]),
            "adg_imex43_init": (i32, [vp, dbl]),
            "adg_imex43_stage": (i32, [vp, i32, ctypes.POINTER(i32)]),
            "adg_proposer_init": (i32, [vp, ctypes.c_uint64]),
            "adg_proposer_set_params": (i32, [vp, FP]),
            "adg_proposer_get_params": (i32, [vp, FP]),
            "adg_proposer_gradient": (i32, [vp, FP, U8P, FP, i64, FP, DP]),
            "adg_proposer_train": (i32, [vp, FP, U8P, i64, i32, i32, ctypes.c_float, ctypes.c_uint64, DP]),
            "adg_proposer_predict": (i32, [vp, FP, i64, FP]),
            "adg_proposer_features": (i32, [vp, DP, FP, I32P]),
            "adg_decision_log": (i32, [vp, i32]),
            "adg_decision_log_size": (i64, [vp]),
            "adg_decision_log_get": (i32, [vp, FP, U8P]),
            "adg_decision_log_write": (i32, [vp, ctypes.c_char_p]),
            "adg_state_write": (i32, [vp, ctypes.c_char_p]),
            "adg_state_read": (i32, [vp, ctypes.c_char_p]),
            "adg_lagr_points": (i32, [vp, DP]),
            "adg_flux": (i32, [vp, i32, DP]),
            "adg_flux_set": (i32, [vp, i32, DP]),
            "adg_fp_to_coe": (i32, [vp]),
            "adg_rhs_zero": (i32, [vp]),
            "adg_rhs_lagr_vol": (i32, [vp]),
            "adg_rhs_lagr_flx": (i32, [vp]),
            "adg_rhs_penalty": (i32, [vp, DP]),
            "adg_rhs_linear": (i32, [vp, DP]),
            "adg_rk_init": (i32, [vp]),
            "adg_rk_stage": (i32, [vp, i32, i32, dbl]),
            "adg_copy_to_predict": (i32, [vp]),
            "adg_copy_predict_to_ucoe": (i32, [vp]),
            "adg_indicator_sumsq": (i32, [vp, DP]),
            "adg_refine": (i32, [vp, dbl, P(i64)]),
            "adg_coarsen": (i32, [vp, dbl, P(i64)]),
            "adg_set_device_adapt": (i32, [vp, i32]),
            "adg_device_adapt_stats": (i32, [vp, P(i64), P(i64), P(i32)]),
            "adg_grid_check": (i32, [vp, P(i64)]),
            "adg_init_adaptive_intp": (i32, [vp, POINT_FN, vp, dbl]),
            "adg_ic_eval": (dbl, [DP, i32, vp]),
            "adg_run_linear": (i32, [vp, DP, dbl, i64, P(RunStats)]),
            "adg_run_hyperbolic": (i32, [vp, P(Problem), P(RunStats)]),
            "adg_stage": (i32, [vp, P(Problem), i32, i32, dbl]),
            "adg_set_kernel_variant": (i32, [vp, i32]),
            "adg_state_save": (i32, [vp, i32]),
            "adg_state_restore": (i32, [vp, i32]),
            "adg_profile": (i32, [vp, i32]),
            "adg_profile_get": (i32, [vp, DP, P(i64), DP]),
            "adg_profile_host": (i32, [vp, DP]),
            "adg_init_separable": (i32, [vp, SEP_FN, vp, i32, dbl]),
            "adg_sep_eval": (dbl, [dbl, i32, i32, vp]),
            "adg_filter": (i32, [vp, dbl, DP, dbl, i32]),
            "adg_filter_local": (i32, [vp, dbl, COEF_FN, vp, dbl, i32]),
            "adg_eval_tensor": (i32, [vp, i32, DP, DP]),
            "adg_error_no_separable": (i32, [vp, POINT_FN, vp, i32, DP]),
            "adg_grid_init_batch": (i32, [vp, i32, i32, i32]),
            "adg_init_adaptive_intp_batch": (i32, [vp, POINT_FN, P(vp), dbl]),
            "adg_num_instances": (i32, [vp]),
            "adg_instance_range": (i32, [vp, i32, P(i64), P(i64)]),
            "adg_set_instance_thresholds": (i32, [vp, DP, DP]),
            "adg_set_instance_params": (i32, [vp, DP, DP, DP]),
            "adg_init_from_points": (i32, [vp, DP]),
            "adg_set_stage_launch": (i32, [vp, i32]),
            "adg_stage_launch_count": (i32, [vp, P(i64), P(i64)]),
            "adg_set_graph_step": (i32, [vp, i32]),
            "adg_kernel_tuning": (i32, [vp, I32P, i32, P(i32)]),
            "adg_set_class_variants": (i32, [vp, P(i32), i32]),
            "adg_graph_step_count": (i32, [vp, P(i64), P(i64)]),
            "adg_rk_step": (i32, [vp, P(Problem), i32, dbl]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def ic_eval_ptr():
    """Address of the library's built-in point function adg_ic_eval."""
    return ctypes.cast(lib().adg_ic_eval, ctypes.c_void_p).value


def sep_eval_ptr():
    """Address of the library's built-in separable function adg_sep_eval."""
    return ctypes.cast(lib().adg_sep_eval, ctypes.c_void_p).value


class AdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"adg error {code}: {msg}")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _u8p(a):
    assert a.dtype == np.uint8 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class Solution:
    """One device-resident DG solution (a context of the C-ABI)."""

    def __init__(self, dim, k, m, nmax, sparse=True, level_init=None):
        self.dim, self.k, self.m, self.nmax = dim, k, m, nmax
        self.K = (k + 1) ** dim
        self.P = (m + 1) ** dim
        h = ctypes.c_void_p()
        rc = lib().adg_create(dim, k, m, nmax, ctypes.byref(h))
        self._h = h
        if rc != 0:
            msg = lib().adg_last_error(h).decode() if h.value else "create failed"
            if h.value:
                lib().adg_destroy(h)
            self._h = None
            raise AdgError(rc, msg)
        self._keep = []
        self._check(lib().adg_grid_init(self._h, int(sparse), nmax if level_init is None else level_init))

    def _check(self, rc):
        if rc != 0:
            raise AdgError(rc, lib().adg_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().adg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_elem(self):
        return int(lib().adg_num_elem(self._h))

    def grid(self):
        E = self.num_elem
        lev = np.zeros((E, self.dim), np.int32)
        sup = np.zeros((E, self.dim), np.int32)
        key = np.zeros(E, np.int32)
        self._check(lib().adg_grid_get(self._h, _ip(lev), _ip(sup), _ip(key)))
        return lev, sup, key

    def set_grid(self, level, suppt):
        lev = np.ascontiguousarray(level, np.int32)
        sup = np.ascontiguousarray(suppt, np.int32)
        self._check(lib().adg_grid_set(self._h, _ip(lev), _ip(sup), lev.shape[0]))

    def max_level(self):
        out = np.zeros(self.dim, np.int32)
        self._check(lib().adg_max_level(self._h, _ip(out)))
        return out

    def _fsize(self, field):
        return self.K if field in (FIELD_UCOE, FIELD_RHS, FIELD_TN) else self.P

    def get(self, field=FIELD_UCOE):
        out = np.zeros(self.num_elem * self._fsize(field))
        self._check(lib().adg_field_get(self._h, field, _dp(out)))
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.num_elem * self._fsize(field)
        self._check(lib().adg_field_set(self._h, field, _dp(v)))

    def sweep(self, table, flx, lu, dim, size_in, size_out, coef, x):
        si = np.asarray(size_in, np.int32)
        so = np.asarray(size_out, np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.num_elem * int(np.prod(so)))
        self._check(lib().adg_sweep(self._h, table, int(flx), lu, dim, _ip(si), _ip(so), coef, _dp(x), _dp(out)))
        return out

    def sweep_noalloc(self, table, flx, lu, dim, size_in, size_out, coef, x):
        """adg_sweep through the allocation-free entry point (per-GPU scratch staging)."""
        si = np.asarray(size_in, np.int32)
        so = np.asarray(size_out, np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.num_elem * int(np.prod(so)))
        self._check(lib().adg_sweep_noalloc(self._h, table, int(flx), lu, dim, _ip(si), _ip(so), coef, _dp(x),
                                            _dp(out)))
        return out

    def lagr_points(self):
        """Coordinates of every element's Lagrange points, shape (E, P, dim)."""
        x = np.zeros(self.num_elem * self.P * self.dim)
        self._check(lib().adg_lagr_points(self._h, _dp(x)))
        return x.reshape(self.num_elem, self.P, self.dim)

    def set_interpolation(self, family):
        """INTERP_LAGRANGE or INTERP_HERMITE (P3, needs m = 3) for eval_up, flux,
        fp_to_coe, rhs_lagr_*, stage and run_hyperbolic."""
        self._check(lib().adg_set_interpolation(self._h, family))

    # ---- artificial viscosity (matrix-free) ----
    def viscosity_elements(self, kappa):
        """count and per-element flags of DGAdapt::update_viscosity_element"""
        n = ctypes.c_int64()
        flags = np.zeros(self.num_elem, np.int32)
        self._check(lib().adg_viscosity_elements(self._h, kappa, ctypes.byref(n), _ip(flags)))
        return n.value, flags

    def viscosity_flux(self, coef, overlap=False):
        self._check(lib().adg_viscosity_flux(self._h, coef, 1 if overlap else 0))

    def viscosity_operator(self, coef, option):
        """ArtificialViscosity::assemble_matrix; returns the number of coupled element blocks"""
        n = ctypes.c_int64()
        self._check(lib().adg_viscosity_operator(self._h, coef, option, ctypes.byref(n)))
        return n.value

    def rhs_viscosity_operator(self):
        self._check(lib().adg_rhs_viscosity_operator(self._h))

    def imex43_init(self, dt):
        self._check(lib().adg_imex43_init(self._h, dt))

    def imex43_stage(self, stage):
        """IMEX43::step_stage + final; returns the largest CG iteration count"""
        it = ctypes.c_int()
        self._check(lib().adg_imex43_stage(self._h, stage, ctypes.byref(it)))
        return it.value

    # ---- learned refinement proposer ----
    def set_judge_defer(self, defer):
        """0: the judge redoes the instances it refines; 1: it refines the missed leaves after the step."""
        self._check(lib().adg_set_judge_defer(self._h, int(defer)))

    def proposer_stats(self):
        """(leaves proposed, instance steps redone, leaves refined by a deferring judge) so far."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().adg_proposer_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_refine_mode(self, mode, threshold=0.5):
        """REFINE_PREDICT (forward-Euler prediction, the reference), REFINE_LEARNED
        (network proposer + threshold judge) or REFINE_CURRENT (current norms + judge)."""
        self._check(lib().adg_set_refine_mode(self._h, mode, threshold))

    def proposer_init(self, seed):
        self._check(lib().adg_proposer_init(self._h, seed))

    def proposer_params(self):
        out = np.zeros(PROPOSER_PARAMS, np.float32)
        self._check(lib().adg_proposer_get_params(self._h, _fp(out)))
        return out

    def set_proposer_params(self, params):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == PROPOSER_PARAMS
        self._check(lib().adg_proposer_set_params(self._h, _fp(p)))

    def proposer_gradient(self, feats, labels, weights=None):
        x = np.ascontiguousarray(feats, np.float32).reshape(-1, PROPOSER_INPUTS)
        y = np.ascontiguousarray(labels, np.uint8)
        w = None if weights is None else np.ascontiguousarray(weights, np.float32)
        g = np.zeros(PROPOSER_PARAMS, np.float32)
        loss = ctypes.c_double()
        self._check(lib().adg_proposer_gradient(self._h, _fp(x), _u8p(y), None if w is None else _fp(w), len(y),
                                                _fp(g), ctypes.byref(loss)))
        return g, loss.value

    def proposer_train(self, feats, labels, epochs, batch, lr, seed):
        x = np.ascontiguousarray(feats, np.float32).reshape(-1, PROPOSER_INPUTS)
        y = np.ascontiguousarray(labels, np.uint8)
        loss = ctypes.c_double()
        self._check(lib().adg_proposer_train(self._h, _fp(x), _u8p(y), len(y), epochs, batch, lr, seed,
                                             ctypes.byref(loss)))
        return loss.value

    def proposer_predict(self, feats):
        x = np.ascontiguousarray(feats, np.float32).reshape(-1, PROPOSER_INPUTS)
        out = np.zeros(len(x), np.float32)
        self._check(lib().adg_proposer_predict(self._h, _fp(x), len(x), _fp(out)))
        return out

    def proposer_features(self, eps):
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, np.float64), (self.num_instances,)))
        feats = np.zeros((self.num_elem, PROPOSER_INPUTS), np.float32)
        leaf = np.zeros(self.num_elem, np.int32)
        self._check(lib().adg_proposer_features(self._h, _dp(e), _fp(feats), _ip(leaf)))
        return feats, leaf

    def decision_log(self, enable):
        self._check(lib().adg_decision_log(self._h, 1 if enable else 0))

    def decision_log_get(self):
        n = int(lib().adg_decision_log_size(self._h))
        feats = np.zeros((n, PROPOSER_INPUTS), np.float32)
        labels = np.zeros(n, np.uint8)
        self._check(lib().adg_decision_log_get(self._h, _fp(feats), _u8p(labels)))
        return feats, labels

    def decision_log_write(self, path):
        self._check(lib().adg_decision_log_write(self._h, os.fsencode(path)))

    def set_precision(self, precision):
        """PREC_FP64 or PREC_FP32_SPLIT for stage() and the fused form of run_hyperbolic."""
        self._check(lib().adg_set_precision(self._h, precision))

    def state_write(self, path):
        """Binary state file: element sets and Alpert coefficients of every instance."""
        self._check(lib().adg_state_write(self._h, os.fsencode(path)))

    def state_read(self, path):
        """Replace the grids and coefficients by those of a state file."""
        self._check(lib().adg_state_read(self._h, os.fsencode(path)))

    def set_level_sum_cap(self, cap):
        """Coarse grid: sweeps ignore elements with level sum above cap (cap < 0: none)."""
        self._check(lib().adg_set_level_sum_cap(self._h, cap))

    def eval_up(self):
        self._check(lib().adg_eval_up(self._h))

    def flux(self, kind, params=(0.0,)):
        p = np.asarray(params, np.float64)
        self._check(lib().adg_flux(self._h, kind, _dp(p)))

    def flux_set(self, values, ncomp):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.num_elem * self.P * ncomp
        self._check(lib().adg_flux_set(self._h, ncomp, _dp(v)))

    def fp_to_coe(self):
        self._check(lib().adg_fp_to_coe(self._h))

    def rhs_zero(self):
        self._check(lib().adg_rhs_zero(self._h))

    def rhs_lagr_vol(self):
        self._check(lib().adg_rhs_lagr_vol(self._h))

    def rhs_lagr_flx(self):
        self._check(lib().adg_rhs_lagr_flx(self._h))

    def rhs_penalty(self, alpha):
        a = np.asarray(alpha, np.float64)
        self._check(lib().adg_rhs_penalty(self._h, _dp(a)))

    def rhs_linear(self, coef):
        c = np.asarray(coef, np.float64)
        self._check(lib().adg_rhs_linear(self._h, _dp(c)))

    def rk_init(self):
        self._check(lib().adg_rk_init(self._h))

    def rk_stage(self, scheme, stage, dt):
        self._check(lib().adg_rk_stage(self._h, scheme, stage, dt))

    def copy_to_predict(self):
        self._check(lib().adg_copy_to_predict(self._h))

    def copy_predict_to_ucoe(self):
        self._check(lib().adg_copy_predict_to_ucoe(self._h))

    def indicator_sumsq(self):
        out = np.zeros(self.num_elem)
        self._check(lib().adg_indicator_sumsq(self._h, _dp(out)))
        return out

    def refine(self, eps):
        n = ctypes.c_int64()
        self._check(lib().adg_refine(self._h, eps, ctypes.byref(n)))
        return n.value

    def coarsen(self, eta):
        n = ctypes.c_int64()
        self._check(lib().adg_coarsen(self._h, eta, ctypes.byref(n)))
        return n.value

    def set_device_adapt(self, on):
        """1: refine / coarsen and the element-table rebuild run on the GPU; 0: on the host."""
        self._check(lib().adg_set_device_adapt(self._h, int(on)))

    def device_adapt_stats(self, reasons=False):
        """(refine / coarsen calls done on the device, calls handed to the host[, bits of why:
        1 a norm at a threshold, 2 new elements beyond the table])."""
        a, b, r = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        self._check(lib().adg_device_adapt_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(r)))
        return (a.value, b.value, r.value) if reasons else (a.value, b.value)

    def grid_check(self):
        """Entries of the device element and fiber tables that differ from a host build."""
        n = ctypes.c_int64()
        self._check(lib().adg_grid_check(self._h, ctypes.byref(n)))
        return n.value

    def init_adaptive_intp(self, func, eps):
        """func(x: ndarray[dim]) -> float; evaluated on the host at Lagrange points."""
        dim = self.dim

        def cb(xp, d, user):
            return float(func(np.ctypeslib.as_array(xp, shape=(dim,))))

        cfun = POINT_FN(cb)
        self._keep.append(cfun)
        self._check(lib().adg_init_adaptive_intp(self._h, cfun, None, eps))

    def init_from_points(self, values):
        """Values at the interpolation points of the current family -> Alpert coefficients."""
        v = np.ascontiguousarray(values, np.float64)
        assert v.size == self.num_elem * self.P
        self._check(lib().adg_init_from_points(self._h, _dp(v)))

    def init_adaptive_intp_builtin(self, params, eps):
        """Initialise with the library's adg_ic_eval and an ICParams."""
        self._keep.append(params)
        fn = POINT_FN(ic_eval_ptr())
        self._keep.append(fn)
        self._check(lib().adg_init_adaptive_intp(self._h, fn, ctypes.cast(ctypes.pointer(params), ctypes.c_void_p), eps))

    def init_separable(self, funcs, eps):
        """funcs[i](x, d) -> float: u0 = sum_i prod_d funcs[i](x_d, d) (host callbacks)."""
        def cb(x, d, i, user):
            return float(funcs[i](x, d))

        cfun = SEP_FN(cb)
        self._keep.append(cfun)
        self._check(lib().adg_init_separable(self._h, cfun, None, len(funcs), eps))

    def init_separable_builtin(self, params, nfunc, eps):
        """Initialise with the library's adg_sep_eval and an ICParams."""
        self._keep.append(params)
        fn = SEP_FN(sep_eval_ptr())
        self._keep.append(fn)
        self._check(lib().adg_init_separable(self._h, fn, ctypes.cast(ctypes.pointer(params), ctypes.c_void_p), nfunc,
                                             eps))

    def filter(self, damp, wave_speed, dt, start_level_sum):
        w = np.ascontiguousarray(wave_speed, np.float64)
        self._check(lib().adg_filter(self._h, damp, _dp(w), dt, start_level_sum))

    def filter_local(self, damp, speed, dt, start_level_sum):
        """speed(x: ndarray[dim], d) -> float, evaluated on the host at element corners."""
        dim = self.dim

        def cb(xp, d, user):
            return float(speed(np.ctypeslib.as_array(xp, shape=(dim,)), d))

        cfun = COEF_FN(cb)
        self._keep.append(cfun)
        self._check(lib().adg_filter_local(self._h, damp, cfun, None, dt, start_level_sum))

    def eval_tensor(self, x1d):
        """u on the tensor grid x1d^dim, shape (len(x1d),) * dim."""
        x = np.ascontiguousarray(x1d, np.float64)
        out = np.zeros(x.size ** self.dim)
        self._check(lib().adg_eval_tensor(self._h, x.size, _dp(x), _dp(out)))
        return out.reshape((x.size,) * self.dim)

    def error_no_separable_builtin(self, params, gauss_points):
        """[L1, L2, Linf] of u - f, f = adg_ic_eval with params."""
        self._keep.append(params)
        fn = POINT_FN(ic_eval_ptr())
        self._keep.append(fn)
        err = np.zeros(3)
        self._check(lib().adg_error_no_separable(self._h, fn, ctypes.cast(ctypes.pointer(params), ctypes.c_void_p),
                                                 gauss_points, _dp(err)))
        return err

    # ---- batches of independent solves of one discretisation ----
    def init_batch(self, ninst, sparse=True, level_init=None):
        """Replace the grid by ninst copies of the initial grid."""
        self._check(lib().adg_grid_init_batch(self._h, ninst, int(sparse), self.nmax if level_init is None else level_init))

    @property
    def num_instances(self):
        return int(lib().adg_num_instances(self._h))

    def instance_range(self, i):
        a, n = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().adg_instance_range(self._h, i, ctypes.byref(a), ctypes.byref(n)))
        return a.value, n.value

    def init_adaptive_intp_builtin_batch(self, params_list, eps):
        """Initialise instance i with adg_ic_eval and params_list[i]."""
        if len(params_list) != self.num_instances:
            raise ValueError("one ICParams per instance expected")
        self._keep.extend(params_list)
        fn = POINT_FN(ic_eval_ptr())
        self._keep.append(fn)
        users = (ctypes.c_void_p * len(params_list))(
            *[ctypes.cast(ctypes.pointer(p), ctypes.c_void_p).value for p in params_list])
        self._keep.append(users)
        self._check(lib().adg_init_adaptive_intp_batch(self._h, fn, users, eps))

    def set_instance_thresholds(self, eps, eta):
        e = np.ascontiguousarray(eps, np.float64)
        t = np.ascontiguousarray(eta, np.float64)
        if e.size != self.num_instances or t.size != self.num_instances:
            raise ValueError("one threshold per instance expected")
        self._check(lib().adg_set_instance_thresholds(self._h, _dp(e), _dp(t)))

    def set_instance_params(self, coef_scale=None, alpha=None, wave_speed=None):
        """Per-instance (ninst x dim) variable-coefficient scales, LF constants
        and time-step wave speeds; None keeps the problem's own."""
        arrs = []
        for a in (coef_scale, alpha, wave_speed):
            if a is None:
                arrs.append(None)
                continue
            a = np.ascontiguousarray(a, np.float64)
            if a.size != self.num_instances * self.dim:
                raise ValueError("ninst x dim values expected")
            arrs.append(a)
        self._keep_params = arrs
        self._check(lib().adg_set_instance_params(self._h, *[_dp(a) if a is not None else None for a in arrs]))

    def instance_grid(self, i):
        """(level, suppt, hash) of instance i."""
        a, n = self.instance_range(i)
        lv, sp, hk = self.grid()
        return lv[a:a + n], sp[a:a + n], hk[a:a + n]

    def instance_get(self, i, field=FIELD_UCOE):
        a, n = self.instance_range(i)
        return self.get(field)[a * self.K:(a + n) * self.K]

    def run_linear(self, coef, dt, nsteps):
        c = np.asarray(coef, np.float64)
        st = RunStats()
        self._check(lib().adg_run_linear(self._h, _dp(c), dt, nsteps, ctypes.byref(st)))
        return st

    def run_hyperbolic(self, problem):
        st = RunStats()
        self._check(lib().adg_run_hyperbolic(self._h, ctypes.byref(problem), ctypes.byref(st)))
        return st

    def set_kernel_variant(self, variant):
        """Thread mapping of the fused kernels (-1: autotuned)."""
        self._check(lib().adg_set_kernel_variant(self._h, variant))

    def set_stage_launch(self, mode):
        """0: one launch per phase of a fused stage; 1: one persistent cooperative launch per stage."""
        self._check(lib().adg_set_stage_launch(self._h, mode))

    def stage_launch_count(self):
        """(kernel launches, stages) of the fused form so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().adg_stage_launch_count(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def stage(self, problem, scheme, stage, dt):
        self._check(lib().adg_stage(self._h, ctypes.byref(problem), scheme, stage, dt))

    def set_class_variants(self, variants):
        """Fixed mapping per kernel class (classes 0..3 FP64, 4..7 FP32; -1: autotuner)."""
        arr = (ctypes.c_int * max(len(variants), 1))(*variants)
        self._check(lib().adg_set_class_variants(self._h, arr, len(variants)))

    def kernel_tuning(self):
        """{(kernel class, grid-size class): variant} chosen by the autotuner so far."""
        n = ctypes.c_int()
        self._check(lib().adg_kernel_tuning(self._h, None, 0, ctypes.byref(n)))
        buf = np.zeros(2 * max(n.value, 1), dtype=np.int32)
        self._check(lib().adg_kernel_tuning(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n.value,
                                            ctypes.byref(n)))
        return {(int(buf[2 * i]) // 64, int(buf[2 * i]) % 64): int(buf[2 * i + 1]) for i in range(n.value)}

    def set_graph_step(self, on):
        """Record whole fused RK steps into HIP graphs and replay them (1) or not (0)."""
        self._check(lib().adg_set_graph_step(self._h, int(on)))

    def graph_step_count(self):
        """(steps recorded into graphs, steps replayed from graphs) so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(lib().adg_graph_step_count(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def rk_step(self, problem, scheme, dt):
        """One whole fused RK step (u^n <- u, then every stage of the scheme)."""
        self._check(lib().adg_rk_step(self._h, ctypes.byref(problem), scheme, dt))

    def save_state(self, slot=0):
        self._check(lib().adg_state_save(self._h, slot))

    def restore_state(self, slot=0):
        self._check(lib().adg_state_restore(self._h, slot))

    def profile(self, enable=True):
        self._check(lib().adg_profile(self._h, int(enable)))

    def profile_get(self):
        ms = np.zeros(6)
        cnt = np.zeros(6, np.int64)
        by = np.zeros(6)
        self._check(lib().adg_profile_get(self._h, _dp(ms), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(by)))
        names = ["sweep", "surplus", "flux", "rk", "other", "rhs_phase"]
        return {n: {"ms": float(ms[i]), "count": int(cnt[i]), "bytes": float(by[i])} for i, n in enumerate(names)}

    def profile_host(self):
        ms = np.zeros(6)
        self._check(lib().adg_profile_host(self._h, _dp(ms)))
        return dict(zip(["adapt", "grid", "plan", "step", "build", "upload"], (float(v) for v in ms)))

    def sync(self):
        self._check(lib().adg_sync(self._h))
