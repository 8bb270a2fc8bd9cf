"""Problem definitions of the benchmark configurations (synthetic, seeded).

Coordinates are the reference's unit cube [0,1]^dim; config 5's physical
domain [-2,2]x[-2.5,1.5] is mapped affinely onto it, which scales both flux
components by 1/4.  Random parameters come from splitmix64 (the project's own
counter-based generator), so a seed fixes a problem on any machine.
"""
import math
from dataclasses import dataclass, field

import pyadg

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, a, b):
        return a + (b - a) * ((self.next() >> 11) * 2.0 ** -53)


@dataclass
class Workload:
    name: str
    description: str
    dim: int
    k: int
    m: int
    nmax: int
    sparse: bool
    level_init: int
    ic: pyadg.ICParams
    init_eps: float
    linear: bool = False
    coef: list = field(default_factory=list)   # linear advection speeds
    dt: float = 0.0                             # linear runs: fixed step
    nsteps: int = 0
    problem: pyadg.Problem = None
    # per-instance parameters of a batch (dim values each, or None): scale of
    # the variable coefficient, LF constants, time-step wave speeds
    coef_scale: list = None
    alpha: list = None
    wave_speed: list = None

    def oracle_problem(self):
        """The problem as the oracle's run loop reads it, with this instance's
        own parameters in place of the batch's shared ones."""
        from types import SimpleNamespace
        p = self.problem
        d = self.dim
        ns = SimpleNamespace(**{f: getattr(p, f) for f, _ in p._fields_})
        ns.alpha = list(self.alpha) if self.alpha is not None else [p.alpha[i] for i in range(d)]
        ns.wave_speed = list(self.wave_speed) if self.wave_speed is not None else [p.wave_speed[i] for i in range(d)]
        if self.coef_scale is not None:
            ns.flux_params = (p.flux_param,) + tuple(self.coef_scale)
        return ns

    def config(self):
        return {"workload": self.name, "description": self.description, "dim": self.dim, "k": self.k,
                "m": self.m, "nmax": self.nmax}


def _ic(kind, **kw):
    p = pyadg.ICParams()
    p.kind = kind
    for name, val in kw.items():
        if name == "bumps":
            p.nbump = len(val)
            for b, (amp, width, center) in enumerate(val):
                p.bump_amp[b], p.bump_width[b] = amp, width
                for d, c in enumerate(center):
                    p.bump_center[b][d] = c
        elif name in ("center", "scale", "offset"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def _problem(**kw):
    p = pyadg.Problem()
    for name, val in kw.items():
        if name in ("alpha", "wave_speed"):
            arr = getattr(p, name)
            for d, c in enumerate(val):
                arr[d] = c
        else:
            setattr(p, name, val)
    return p


def c1_linear(nmax=7, t_final=1.0):
    """2D linear advection u_t+u_x+u_y=0, u0=sin(2 pi (x+y)), k=2, sparse grid N,
    RK3, dt = dx*0.1/2 (example/02_hyperbolic_01_scalar_const_coefficient.cpp:148-152)."""
    dx = 1.0 / 2.0 ** nmax
    dt = dx * 0.1 / 2
    n = math.ceil(t_final / dt) + 1
    dt = t_final / n
    return Workload("c1_linear_advection_2d_k2_sparseN%d" % nmax,
                    "2D linear advection, sparse grid N=%d, k=2, RK3, t=%g" % (nmax, t_final),
                    2, 2, 3, nmax, True, nmax, _ic(0, shift=0.0), 1e10, linear=True, coef=[1.0, 1.0],
                    dt=dt, nsteps=n)


def c2_burgers(seed=1, nmax=8, level_init=6, t_final=0.1, eps=1e-4, eta=1e-5):
    """2D Burgers, u0 = 0.5 + 0.25 sin(2 pi (x+y)) + 4 seeded Gaussian bumps,
    k=2, Lagrange P3 flux interpolation, adaptive, RK3, LF alpha 1.2, cfl 0.2
    (example/02_hyperbolic_05_burgers_adapt.cpp:214-244)."""
    g = SplitMix64(seed)
    bumps = []
    for _ in range(4):
        amp = g.uniform(0.05, 0.2)
        width = g.uniform(0.05, 0.15)
        center = (g.uniform(0.0, 1.0), g.uniform(0.0, 1.0))
        bumps.append((amp, width, center))
    ic = _ic(2, base=0.5, ampl=0.25, bumps=bumps)
    prob = _problem(flux_kind=pyadg.FLUX_BURGERS, flux_param=0.0, alpha=[1.2, 1.2], wave_speed=[1.0, 1.0], cfl=0.2,
                    t_final=t_final, adaptive=1, eps=eps, eta=eta, max_steps=0, t_start=0.0)
    return Workload("c2_burgers_adaptive_2d_k2_m3_nmax%d" % nmax,
                    "2D Burgers, 4 seeded bumps (seed %d), adaptive eps=%g eta=%g, NMAX=%d, sparse init N0=%d, t=%g"
                    % (seed, eps, eta, nmax, level_init, t_final),
                    2, 2, 3, nmax, True, level_init, ic, eps, problem=prob)


def c3_varcoef_4d(nmax=8, t_final=0.5):
    """4D u_t + sum_i a_i(x) u_{x_i} = 0, a_i = 1 + 0.5 sin(2 pi x_{i+1 mod 4}),
    u0 = prod sin(2 pi x_i), k=1, Lagrange P2, sparse N=8, RK3, LF alpha 1.5
    (example/02_hyperbolic_02_var_coeff_coarse_grid.cpp:242-250)."""
    ic = _ic(1, shift=0.0)
    prob = _problem(flux_kind=pyadg.FLUX_VARCOEF, flux_param=0.5, alpha=[1.5] * 4, wave_speed=[1.5] * 4, cfl=0.2,
                    t_final=t_final, adaptive=0, eps=0.0, eta=0.0, max_steps=0, t_start=0.0)
    return Workload("c3_varcoef_4d_k1_m2_sparseN%d" % nmax, "4D variable-coefficient advection, sparse N=%d, k=1" % nmax,
                    4, 1, 2, nmax, True, nmax, ic, 1e10, problem=prob)


def c5_rotating(seed=1, nmax=9, level_init=6, t_final=1.0, eps=1e-3):
    """2D rotating wave u_t + (sin u)_x + (cos u)_y = 0 on [-2,2]x[-2.5,1.5]
    mapped to [0,1]^2, u0 = 7 pi/4 inside a circle of seeded radius, pi/4
    outside, k=2, Lagrange P3, adaptive eps, eta = eps/10."""
    g = SplitMix64(seed)
    radius = g.uniform(0.8, 1.2)
    ic = _ic(3, inside=7 * math.pi / 4, outside=math.pi / 4, radius=radius, center=[0.0, 0.0], scale=[4.0, 4.0],
             offset=[-2.0, -2.5])
    prob = _problem(flux_kind=pyadg.FLUX_SINCOS, flux_param=0.25, alpha=[0.25, 0.25], wave_speed=[0.25, 0.25],
                    cfl=0.2, t_final=t_final, adaptive=1, eps=eps, eta=eps / 10, max_steps=0, t_start=0.0)
    return Workload("c5_rotating_wave_2d_k2_m3_nmax%d" % nmax,
                    "2D rotating wave (KPP), circle radius %.4f, adaptive eps=%g, NMAX=%d" % (radius, eps, nmax),
                    2, 2, 3, nmax, True, level_init, ic, eps, problem=prob)


def c4_advection_3d(seed=1, nmax=7, level_init=3, t_final=0.25, amp=0.5):
    """One of the 256 independent solves of the batching configuration: 3D
    u_t + sum_d a_d(x) u_{x_d} = 0 on [0,1]^3 periodic with
    a_d(x) = c_d (1 + amp sin(2 pi x_{(d+1) mod 3})), seeded speeds
    c ~ U[0.5,2]^3, u0 a Gaussian bump exp(-|x - c0|^2 / w^2) (periodic
    distance) with centre U[0.2,0.8]^3 and width U[0.05,0.15], adaptive with
    eps ~ U[1e-4,1e-3], eta = eps/10, k=2, Lagrange P3 interpolation of a u
    (LagrInterpolation::var_coeff_u_Lagr_fast), LF constants and time-step
    speeds max|a_d| = 1.5 c_d, cfl 0.2 (the rule of
    example/02_hyperbolic_02_var_coeff_coarse_grid.cpp:242-250)."""
    g = SplitMix64(seed)
    speed = [g.uniform(0.5, 2.0) for _ in range(3)]
    center = [g.uniform(0.2, 0.8) for _ in range(3)]
    width = g.uniform(0.05, 0.15)
    eps = g.uniform(1e-4, 1e-3)
    ic = _ic(4, bumps=[(1.0, width, center)])
    amax = [c * (1.0 + amp) for c in speed]
    prob = _problem(flux_kind=pyadg.FLUX_VARCOEF_SCALED, flux_param=amp, alpha=amax, wave_speed=amax, cfl=0.2,
                    t_final=t_final, adaptive=1, eps=eps, eta=eps / 10, max_steps=0, t_start=0.0)
    return Workload("c4_advection_3d_k2_m3_nmax%d" % nmax,
                    "3D variable-coefficient advection (seed %d), speeds %s, bump width %.4f, adaptive eps=%.3g, "
                    "NMAX=%d, sparse init N0=%d, t=%g" % (seed, ["%.3f" % c for c in speed], width, eps, nmax,
                                                          level_init, t_final),
                    3, 2, 3, nmax, True, level_init, ic, eps, problem=prob, coef_scale=speed, alpha=amax,
                    wave_speed=amax)


WORKLOADS = {"c1": c1_linear, "c2": c2_burgers, "c3": c3_varcoef_4d, "c4": c4_advection_3d, "c5": c5_rotating}


def _instance_params(s, ws):
    """Per-instance parameters of the workloads that carry them."""
    if all(w.coef_scale is None and w.alpha is None and w.wave_speed is None for w in ws):
        return
    def arr(name):
        vals = [getattr(w, name) for w in ws]
        if all(v is None for v in vals):
            return None
        return [x for w, v in zip(ws, vals) for x in (v if v is not None else
                                                          [getattr(w.problem, name)[d] for d in range(w.dim)])]
    s.set_instance_params(arr("coef_scale"), arr("alpha"), arr("wave_speed"))


def make_solution(w):
    s = pyadg.Solution(w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init)
    s.init_adaptive_intp_builtin(w.ic, w.init_eps)
    _instance_params(s, [w])
    return s


def make_batch(workloads):
    """One context holding every workload of the list as an instance.  All must
    share (dim, k, m, nmax, sparse, level_init); different initialisation
    thresholds become per-instance thresholds (also used by the run)."""
    w0 = workloads[0]
    for w in workloads[1:]:
        if (w.dim, w.k, w.m, w.nmax, w.sparse, w.level_init) != (
                w0.dim, w0.k, w0.m, w0.nmax, w0.sparse, w0.level_init):
            raise ValueError("instances of one batch must share the discretisation")
    s = pyadg.Solution(w0.dim, w0.k, w0.m, w0.nmax, w0.sparse, w0.level_init)
    s.init_batch(len(workloads), w0.sparse, w0.level_init)
    per_eps = len({w.init_eps for w in workloads}) > 1 or len({w.problem.eps for w in workloads if w.problem}) > 1
    if per_eps:
        if any(w.problem is None or w.init_eps != w.problem.eps for w in workloads):
            raise ValueError("per-instance thresholds: the initialisation threshold must be the run's eps")
        s.set_instance_thresholds([w.problem.eps for w in workloads], [w.problem.eta for w in workloads])
    s.init_adaptive_intp_builtin_batch([w.ic for w in workloads], w0.init_eps)
    _instance_params(s, workloads)
    return s
