"""ctypes binding of the CPU oracle (oracle.h).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  Arrays are ordered by element hash key.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FIELD_UCOE, FIELD_RHS, FIELD_UP, FIELD_TN, FIELD_UINTP = 0, 1, 2, 3, 4
FIELD_FUCOE, FIELD_FP = 10, 20  # + flux component
FUNC_T = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)
SEP_T = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_void_p)
COEF_T = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_void_p)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        P = ctypes.POINTER
        sig = {
            "or_create": (vp, [i32, i32, i32, i32, i32, i32]),
            "or_destroy": (None, [vp]),
            "or_num_elem": (i64, [vp]),
            "or_get_grid": (None, [vp, P(ctypes.c_int32), P(ctypes.c_int32), P(ctypes.c_int32)]),
            "or_get_field": (None, [vp, i32, P(dbl)]),
            "or_set_field": (None, [vp, i32, P(dbl)]),
            "or_sweep": (None, [vp, i32, i32, i32, i32, P(ctypes.c_int32), P(ctypes.c_int32), dbl, P(dbl), P(dbl)]),
            "or_eval_up": (None, [vp]),
            "or_set_interpolation": (None, [vp, i32]),
            "or_viscosity_elements": (i64, [vp, dbl, P(ctypes.c_int32)]),
            "or_viscosity_flux": (None, [vp, dbl, i32]),
            "or_viscosity_operator": (i64, [vp, dbl, i32]),
            "or_viscosity_operator_coo": (None, [vp, P(i64), P(i64), P(dbl)]),
            "or_rhs_viscosity_operator": (None, [vp]),
            "or_product_volume_interv": (dbl, [vp, i32, i32, i32, i32, i32, i32]),
            "or_imex43_init": (None, [vp, dbl]),
            "or_imex43_stage": (i32, [vp, i32]),
            "or_set_level_sum_cap": (None, [vp, ctypes.c_int]),
            "or_flux": (None, [vp, i32, P(dbl)]),
            "or_fp_to_coe": (None, [vp]),
            "or_rhs_zero": (None, [vp]),
            "or_rhs_lagr_vol": (None, [vp]),
            "or_rhs_lagr_flx": (None, [vp]),
            "or_rhs_penalty": (None, [vp, P(dbl)]),
            "or_rhs_linear": (None, [vp, P(dbl)]),
            "or_rhs_linear_assembled": (None, [vp, P(dbl)]),
            "or_rk_init": (None, [vp]),
            "or_rk_stage": (None, [vp, i32, i32, dbl]),
            "or_copy_to_predict": (None, [vp]),
            "or_copy_predict_to_ucoe": (None, [vp]),
            "or_indicator_sumsq": (None, [vp, P(dbl)]),
            "or_refine": (i64, [vp, dbl]),
            "or_coarsen": (i64, [vp, dbl]),
            "or_max_level": (None, [vp, P(ctypes.c_int32)]),
            "or_init_adaptive_intp": (None, [vp, FUNC_T, vp, dbl]),
            "or_init_separable": (None, [vp, SEP_T, vp, i32, dbl]),
            "or_eval_tensor": (None, [vp, i32, P(dbl), P(dbl)]),
            "or_error_no_separable": (None, [vp, FUNC_T, vp, i32, P(dbl)]),
            "or_filter": (None, [vp, dbl, P(dbl), dbl, i32]),
            "or_filter_local": (None, [vp, dbl, COEF_T, vp, dbl, i32]),
            "or_l2_error_split": (dbl, [vp, vp]),
            "or_set_threads": (None, [i32]),
            "or_table_size": (i64, [vp, i32, P(ctypes.c_int32), P(ctypes.c_int32)]),
            "or_table": (None, [vp, i32, P(dbl)]),
            "or_neighbors": (i64, [vp, i64, i32, i32, P(ctypes.c_int32), i64]),
            "or_basis_val": (dbl, [vp, i32, i32, dbl, i32, i32]),
            "or_lagr_point": (dbl, [vp, i32]),
            "or_product_volume": (dbl, [vp, i32, i32, i32, i32, i32, i32]),
            "or_hash_key": (ctypes.c_int32, [P(ctypes.c_int32), P(ctypes.c_int32), i32]),
            "or_surplus_tables": (None, [vp, P(ctypes.c_int32), P(ctypes.c_int32), P(dbl)]),
            "or_lagr_mass": (None, [vp, P(dbl)]),
            "or_gl": (None, [i32, P(dbl), P(dbl)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class Oracle:
    """One DGAdapt-like solution held by the oracle."""

    def __init__(self, dim, k, m, nmax, sparse=True, level_init=None):
        self.dim, self.k, self.m, self.nmax = dim, k, m, nmax
        self.K = (k + 1) ** dim
        self.P = (m + 1) ** dim
        self._h = lib().or_create(dim, k, m, nmax, int(sparse), nmax if level_init is None else level_init)
        if not self._h:
            raise RuntimeError("oracle: or_create failed")
        self._keep = []

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_elem(self):
        return int(lib().or_num_elem(self._h))

    def grid(self):
        E = self.num_elem
        lev = np.zeros((E, self.dim), np.int32)
        sup = np.zeros((E, self.dim), np.int32)
        key = np.zeros(E, np.int32)
        lib().or_get_grid(self._h, _ip(lev), _ip(sup), _ip(key))
        return lev, sup, key

    def _fsize(self, field):
        return self.P if field in (FIELD_UP, FIELD_UINTP) or field >= 10 else self.K

    def get(self, field=FIELD_UCOE):
        out = np.zeros(self.num_elem * self._fsize(field))
        lib().or_get_field(self._h, field, _dp(out))
        return out

    def set(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        assert v.size == self.num_elem * self._fsize(field)
        lib().or_set_field(self._h, field, _dp(v))

    def sweep(self, table, flx, lu, dim, size_in, size_out, coef, x):
        si = np.asarray(size_in, np.int32)
        so = np.asarray(size_out, np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.num_elem * int(np.prod(so)))
        lib().or_sweep(self._h, table, int(flx), lu, dim, _ip(si), _ip(so), coef, _dp(x), _dp(out))
        return out

    def viscosity_elements(self, kappa):
        """count and per-element flags (hash order) of DGAdapt::update_viscosity_element"""
        flags = np.zeros(self.num_elem, np.int32)
        n = lib().or_viscosity_elements(self._h, kappa, flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return int(n), flags

    def viscosity_flux(self, coef, overlap=False):
        lib().or_viscosity_flux(self._h, coef, 1 if overlap else 0)

    def viscosity_operator(self, coef, option):
        """ArtificialViscosity::assemble_matrix; returns the number of entries"""
        self._visc_nnz = int(lib().or_viscosity_operator(self._h, coef, option))
        return self._visc_nnz

    def viscosity_operator_coo(self):
        n = self._visc_nnz
        rows, cols, vals = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n)
        P = ctypes.POINTER
        lib().or_viscosity_operator_coo(self._h, rows.ctypes.data_as(P(ctypes.c_int64)),
                                        cols.ctypes.data_as(P(ctypes.c_int64)), _dp(vals))
        return rows, cols, vals

    def rhs_viscosity_operator(self):
        lib().or_rhs_viscosity_operator(self._h)

    def product_volume_interv(self, ru, rv, rw, du, dv, nq):
        return float(lib().or_product_volume_interv(self._h, ru, rv, rw, du, dv, nq))

    def imex43_init(self, dt):
        lib().or_imex43_init(self._h, dt)

    def imex43_stage(self, stage):
        """IMEX43::step_stage + final; returns the CG iteration count (-1: no convergence)"""
        return int(lib().or_imex43_stage(self._h, stage))

    def set_interpolation(self, family):
        """0 Lagrange, 1 Hermite P3 (needs m = 3)."""
        lib().or_set_interpolation(self._h, family)

    def eval_up(self):
        lib().or_eval_up(self._h)

    def set_level_sum_cap(self, cap):
        lib().or_set_level_sum_cap(self._h, cap)

    def flux(self, kind, params=(0.0,)):
        p = np.asarray(params, np.float64)
        lib().or_flux(self._h, kind, _dp(p))

    def fp_to_coe(self):
        lib().or_fp_to_coe(self._h)

    def rhs_zero(self):
        lib().or_rhs_zero(self._h)

    def rhs_lagr_vol(self):
        lib().or_rhs_lagr_vol(self._h)

    def rhs_lagr_flx(self):
        lib().or_rhs_lagr_flx(self._h)

    def rhs_penalty(self, alpha):
        a = np.asarray(alpha, np.float64)
        lib().or_rhs_penalty(self._h, _dp(a))

    def rhs_linear(self, coef, assembled=False):
        c = np.asarray(coef, np.float64)
        (lib().or_rhs_linear_assembled if assembled else lib().or_rhs_linear)(self._h, _dp(c))

    def rk_init(self):
        lib().or_rk_init(self._h)

    def rk_stage(self, scheme, stage, dt):
        lib().or_rk_stage(self._h, scheme, stage, dt)

    def copy_to_predict(self):
        lib().or_copy_to_predict(self._h)

    def copy_predict_to_ucoe(self):
        lib().or_copy_predict_to_ucoe(self._h)

    def indicator_sumsq(self):
        out = np.zeros(self.num_elem)
        lib().or_indicator_sumsq(self._h, _dp(out))
        return out

    def refine(self, eps):
        return int(lib().or_refine(self._h, eps))

    def coarsen(self, eta):
        return int(lib().or_coarsen(self._h, eta))

    def max_level(self):
        out = np.zeros(self.dim, np.int32)
        lib().or_max_level(self._h, _ip(out))
        return out

    def init_adaptive_intp(self, func, eps):
        """func(x: ndarray[dim]) -> float, evaluated at Lagrange points."""
        dim = self.dim

        def cb(xp, d, user):
            return float(func(np.ctypeslib.as_array(xp, shape=(dim,))))

        cfun = FUNC_T(cb)
        self._keep.append(cfun)
        lib().or_init_adaptive_intp(self._h, cfun, None, eps)

    def init_adaptive_intp_c(self, cfunc_ptr, user_ptr, eps):
        lib().or_init_adaptive_intp(self._h, ctypes.cast(cfunc_ptr, FUNC_T), user_ptr, eps)

    def init_separable(self, funcs, eps):
        """funcs[i](x, d) -> float: u0 = sum_i prod_d funcs[i](x_d, d)."""
        def cb(x, d, i, user):
            return float(funcs[i](x, d))

        cfun = SEP_T(cb)
        self._keep.append(cfun)
        lib().or_init_separable(self._h, cfun, None, len(funcs), eps)

    def init_separable_c(self, cfunc_ptr, user_ptr, nfunc, eps):
        lib().or_init_separable(self._h, ctypes.cast(cfunc_ptr, SEP_T), user_ptr, nfunc, eps)

    def eval_tensor(self, x1d):
        x = np.ascontiguousarray(x1d, np.float64)
        out = np.zeros(x.size ** self.dim)
        lib().or_eval_tensor(self._h, x.size, _dp(x), _dp(out))
        return out.reshape((x.size,) * self.dim)

    def error_no_separable(self, func, points):
        dim = self.dim

        def cb(xp, d, user):
            return float(func(np.ctypeslib.as_array(xp, shape=(dim,))))

        cfun = FUNC_T(cb)
        err = np.zeros(3)
        lib().or_error_no_separable(self._h, cfun, None, points, _dp(err))
        return err

    def error_no_separable_c(self, cfunc_ptr, user_ptr, points):
        err = np.zeros(3)
        lib().or_error_no_separable(self._h, ctypes.cast(cfunc_ptr, FUNC_T), user_ptr, points, _dp(err))
        return err

    def filter(self, damp, wave_speed, dt, start_level_sum):
        w = np.ascontiguousarray(wave_speed, np.float64)
        lib().or_filter(self._h, damp, _dp(w), dt, start_level_sum)

    def filter_local(self, damp, speed, dt, start_level_sum):
        dim = self.dim

        def cb(xp, d, user):
            return float(speed(np.ctypeslib.as_array(xp, shape=(dim,)), d))

        cfun = COEF_T(cb)
        self._keep.append(cfun)
        lib().or_filter_local(self._h, damp, cfun, None, dt, start_level_sum)

    def table(self, table_id):
        r = ctypes.c_int32()
        c = ctypes.c_int32()
        lib().or_table_size(self._h, table_id, ctypes.byref(r), ctypes.byref(c))
        out = np.zeros(r.value * c.value)
        lib().or_table(self._h, table_id, _dp(out))
        return out.reshape(r.value, c.value)

    def neighbors(self, elem, dim, flx):
        cap = self.num_elem
        out = np.zeros(cap, np.int32)
        n = lib().or_neighbors(self._h, elem, dim, int(flx), _ip(out), cap)
        return out[:n].copy()

    def basis_val(self, kind, order, x, deriv=0, sgn=0):
        return float(lib().or_basis_val(self._h, kind, order, x, deriv, sgn))

    def lagr_point(self, order):
        return float(lib().or_lagr_point(self._h, order))

    def product_volume(self, ku, ru, kv, cv, du=0, dv=0):
        return float(lib().or_product_volume(self._h, ku, ru, kv, cv, du, dv))

    def surplus_tables(self):
        p1 = self.m + 1
        ns = 1 << self.nmax
        a = np.zeros(ns * p1, np.int32)
        b = np.zeros(ns * p1, np.int32)
        w = np.zeros(ns * p1 * p1)
        lib().or_surplus_tables(self._h, _ip(a), _ip(b), _dp(w))
        return a.reshape(ns, p1), b.reshape(ns, p1), w.reshape(ns, p1, p1)

    def lagr_mass(self):
        out = np.zeros((1 << self.nmax) * (self.m + 1))
        lib().or_lagr_mass(self._h, _dp(out))
        return out

    def l2_error_split(self, exact):
        return float(lib().or_l2_error_split(self._h, exact._h))


def hash_key(level, suppt):
    lv = np.ascontiguousarray(level, np.int32)
    sp = np.ascontiguousarray(suppt, np.int32)
    return int(lib().or_hash_key(_ip(lv), _ip(sp), lv.size))


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_gl(n, _dp(x), _dp(w))
    return x, w


def set_threads(n):
    lib().or_set_threads(int(n))
