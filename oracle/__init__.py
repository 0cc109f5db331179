"""CPU oracle for the hom3 finite-volume hot path — TEST INFRASTRUCTURE ONLY.

A ctypes front-end over ``oracle/_build/liboracle.so``, the faithful C++17
restatement of gnzlbg/hom3's grid, solver, physics, boundary conditions and
time integration (see oracle/oracle.hpp for the per-function citations).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``hom3_amd``) never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

INVALID = np.iinfo(np.int64).max
INVALID_SIND = np.iinfo(np.int32).max

# enums (oracle.hpp; identical numbering to include/hom3g.h)
PHYS_EULER, PHYS_CNS, PHYS_HEAT = 0, 1, 2
FLUX_AUSM, FLUX_LLF, FLUX_THREE_POINT = 0, 1, 2
TI_EULER_FWD, TI_HEUN_RK2, TI_SSP_RK3 = 0, 1, 2
RECON_NONE, RECON_MUSCL = 0, 1
LIM_ZERO, LIM_MINMOD, LIM_VANLEER = 0, 1, 2
GEOM_NONE, GEOM_EDGE, GEOM_SQUARE, GEOM_SPHERE = 0, 1, 2, 3
(BC_EULER_NEUMANN, BC_EULER_DIRICHLET, BC_ISENTROPIC_VORTEX, BC_CNS_INFLOW,
 BC_CNS_OUTFLOW, BC_ADIABATIC_NOSLIP, BC_ADIABATIC_SLIP, BC_ISOTHERMAL_NOSLIP,
 BC_HEAT_DIRICHLET, BC_HEAT_NEUMANN, BC_PERIODIC, BC_CHT_FLUID,
 BC_CHT_SOLID, BC_EULER_COUPLING) = range(14)
(IC_CONSTANT_PV, IC_SHOCK_TUBE, IC_ISENTROPIC_VORTEX, IC_SPHERE, IC_TGV,
 IC_FOUR_QUADRANT, IC_HEAT_CONSTANT, IC_HEAT_SINE, IC_LINEAR_PV,
 IC_CALLBACK) = range(10)


class PhysDesc(C.Structure):
    _fields_ = [("physics", C.c_int32), ("nd", C.c_int32), ("flux", C.c_int32),
                ("ti", C.c_int32), ("recon", C.c_int32), ("limiter", C.c_int32),
                ("gamma", C.c_double), ("cfl", C.c_double), ("cfl_viscous", C.c_double),
                ("Rspecific", C.c_double), ("T0", C.c_double), ("rho0", C.c_double),
                ("Minfty", C.c_double), ("ReInfty", C.c_double), ("Pr0", C.c_double),
                ("Sutherland0", C.c_double), ("diffusivity", C.c_double),
                ("T_ref", C.c_double), ("time_end", C.c_double), ("max_cells", C.c_int64)]


class GeomDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("invert", C.c_int32),
                ("a", C.c_double * 3), ("b", C.c_double * 3)]


IC_CB = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


def build(force=False):
    """Compile the oracle with its Makefile (strict IEEE flags)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def long_double():
    """The same front-end over ``liboracle_ld.so``, the oracle built with
    long double (x87, 64-bit significand) arithmetic: a checker for central
    differences at small steps, where the FP64 rounding of the two runs is
    of the size of their difference.  Values cross the C API as double."""
    import importlib.util
    import sys
    name = __name__ + "_ld"
    mod = sys.modules.get(name)
    if mod is None:
        spec = importlib.util.spec_from_file_location(name, __file__)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod._LIB_PATH = os.path.join(_HERE, "_build", "liboracle_ld.so")
        sys.modules[name] = mod
    return mod


def fast(path=None):
    """The same front-end over the timing build (``liboracle_fast.so``, or
    the library at ``path``): the reference's Release flags (-O3
    -ffast-math) with OpenMP over the cell loops.  Used only to time the CPU
    baseline, never as the checker (fast-math reorders the arithmetic)."""
    import importlib.util
    import sys
    lp = path or os.path.join(_HERE, "_build", "liboracle_fast.so")
    name = __name__ + "_fast_" + str(abs(hash(lp)))
    mod = sys.modules.get(name)
    if mod is None:
        spec = importlib.util.spec_from_file_location(name, __file__)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod._LIB_PATH = lp
        sys.modules[name] = mod
    return mod


def set_threads(n):
    """OpenMP threads of the timing build (the checker builds run serially
    and return 1)."""
    return lib().orc_set_threads(n)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i64, dbl = C.c_void_p, C.c_int64, C.c_double
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_set_threads": (C.c_int, [C.c_int]),
            "orc_grid_create": (vp, [C.c_int, C.POINTER(dbl), C.POINTER(dbl), C.c_int]),
            "orc_grid_destroy": (None, [vp]),
            "orc_grid_min_level": (C.c_int, [vp, C.c_int]),
            "orc_grid_size": (i64, [vp]),
            "orc_grid_refine": (i64, [vp, i64]),
            "orc_grid_refine_region": (C.c_int, [vp, C.c_int, C.POINTER(GeomDesc), dbl]),
            "orc_grid_balance": (C.c_int, [vp]),
            "orc_grid_tree": (None, [vp, vp, vp]),
            "orc_grid_find_samelvl_neighbor": (i64, [vp, i64, C.c_int]),
            "orc_grid_level": (C.c_int, [vp, i64]),
            "orc_grid_position_in_parent": (C.c_int, [vp, i64]),
            "orc_grid_is_leaf": (C.c_int, [vp, i64]),
            "orc_grid_cell_coordinates": (None, [vp, i64, C.POINTER(dbl)]),
            "orc_grid_cell_length": (dbl, [vp, i64]),
            "orc_grid_is_cut_by": (C.c_int, [vp, i64, C.POINTER(GeomDesc)]),
            "orc_signed_distance": (dbl, [C.POINTER(GeomDesc), C.POINTER(dbl), C.c_int]),
            "orc_grid_all_neighbors": (None, [vp, vp]),
            "orc_solver_create": (vp, [vp, C.c_int, C.POINTER(PhysDesc)]),
            "orc_solver_destroy": (None, [vp]),
            "orc_solver_append_bc": (C.c_int, [vp, C.POINTER(GeomDesc), C.c_int, C.POINTER(dbl), C.c_int, vp, dbl]),
            "orc_solver_set_bc_other": (C.c_int, [vp, C.c_int, vp]),
            "orc_solver_set_domain": (None, [vp, C.POINTER(GeomDesc), C.c_int]),
            "orc_solver_set_domain_mask": (None, [vp, vp, i64]),
            "orc_solver_set_ic": (None, [vp, C.c_int, C.POINTER(dbl), C.c_int, IC_CB, vp]),
            "orc_solver_set_ghost_order": (None, [vp, C.c_int]),
            "orc_solver_initialize": (C.c_int, [vp]),
            "orc_solver_solve": (C.c_int, [vp, C.c_int]),
            "orc_solver_apply_bcs_lhs": (C.c_int, [vp]),
            "orc_solver_evolve": (C.c_int, [vp]),
            "orc_solver_set_dt": (C.c_int, [vp]),
            "orc_solver_advance": (None, [vp]),
            "orc_solver_compute_dt": (dbl, [vp]),
            "orc_solver_force_dt": (None, [vp, dbl]),
            "orc_solver_dt": (dbl, [vp]),
            "orc_solver_time": (dbl, [vp]),
            "orc_solver_step": (i64, [vp]),
            "orc_solver_set_time": (None, [vp, dbl, dbl, i64]),
            "orc_solver_ncells": (i64, [vp]),
            "orc_solver_first_gc": (i64, [vp]),
            "orc_solver_nvars": (C.c_int, [vp]),
            "orc_solver_cap": (i64, [vp]),
            "orc_solver_quantities": (None, [vp, C.POINTER(dbl)]),
            "orc_solver_get": (None, [vp, C.c_char_p, vp]),
            "orc_solver_nfine": (C.c_int64, [vp]),
            "orc_solver_fine": (None, [vp, vp]),
            "orc_solver_set_perturbation": (None, [vp, C.c_uint64, dbl]),
            "orc_node_morton": (C.c_uint64, [vp, C.c_int64]),
            "orc_solver_set_lhs": (None, [vp, C.POINTER(dbl)]),
            "orc_solver_set_source": (None, [vp, C.POINTER(dbl), i64]),
            "orc_solver_conservation_sums": (None, [vp, C.POINTER(dbl)]),
            "orc_solver_periodic_image": (i64, [vp, i64]),
            "orc_solver_cht_partner": (i64, [vp, i64]),
            "orc_solver_face_flux": (None, [vp, i64, i64, C.c_int, C.POINTER(dbl)]),
            "orc_solver_num_flux": (None, [vp, i64, C.POINTER(dbl)]),
            "orc_solver_slope": (dbl, [vp, i64, C.c_int, C.c_int]),
            "orc_solver_cell_dt": (dbl, [vp, i64]),
            "orc_ausm_split": (None, [C.c_int, dbl, C.POINTER(dbl), C.POINTER(dbl)]),
            "orc_solver_ausm_states": (None, [vp, C.POINTER(dbl), C.POINTER(dbl), C.c_int, C.POINTER(dbl)]),
            "orc_solver_ic_point": (C.c_int, [vp, C.POINTER(dbl), C.POINTER(dbl)]),
            "orc_couple_cns_heat": (None, [dbl, vp, vp]),
            "orc_couple_cns_heat_fixed": (C.c_int, [dbl, dbl, vp, vp]),
            "orc_cht_wall_heat_flux": (dbl, [vp]),
            "orc_couple_same": (None, [vp, vp]),
            "orc_time_solve": (dbl, [vp, C.c_int]),
            "orc_prng_uniform": (dbl, [C.c_uint64, C.c_uint64, C.c_uint64]),
            "orc_morton_encode": (C.c_uint64, [C.c_int, C.POINTER(C.c_uint32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def geom(kind=GEOM_NONE, a=(0, 0, 0), b=(0, 0, 0), invert=False):
    g = GeomDesc()
    g.kind = kind
    g.invert = 1 if invert else 0
    for i in range(3):
        g.a[i] = a[i] if i < len(a) else 0.0
        g.b[i] = b[i] if i < len(b) else 0.0
    return g


class Grid:
    """grid::CartesianHSP<nd> with the MinLevel generator (grid/grid.hpp,
    grid/generation.hpp, containers/hierarchical/implementation.hpp)."""

    def __init__(self, nd, xmin, xmax, max_solvers=1):
        self.nd = nd
        xmin = np.ascontiguousarray(xmin, dtype=np.float64)
        xmax = np.ascontiguousarray(xmax, dtype=np.float64)
        self.xmin, self.xmax = xmin, xmax
        self.h = lib().orc_grid_create(nd, _dptr(xmin), _dptr(xmax), max_solvers)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_grid_destroy(self.h)
            self.h = None

    def min_level(self, level):
        _check(lib().orc_grid_min_level(self.h, level))
        return self

    def size(self):
        return lib().orc_grid_size(self.h)

    def refine(self, n):
        return lib().orc_grid_refine(self.h, n)

    def refine_region(self, max_level, geom, band):
        """Refine leaves within `band` of the geometry's surface up to max_level."""
        _check(lib().orc_grid_refine_region(self.h, max_level, C.byref(geom), band))
        return self

    def balance(self):
        """2:1 balance across faces."""
        _check(lib().orc_grid_balance(self.h))
        return self

    def node_morton(self, n):
        return lib().orc_node_morton(self.h, n)

    def tree(self):
        n = self.size()
        par = np.empty(n, np.int64)
        fc = np.empty(n, np.int64)
        lib().orc_grid_tree(self.h, par.ctypes.data, fc.ctypes.data)
        return par, fc

    def neighbor(self, n, pos):
        return lib().orc_grid_find_samelvl_neighbor(self.h, n, pos)

    def all_neighbors(self):
        out = np.empty((self.size(), 2 * self.nd), np.int64)
        lib().orc_grid_all_neighbors(self.h, out.ctypes.data)
        return out

    def level(self, n):
        return lib().orc_grid_level(self.h, n)

    def position_in_parent(self, n):
        return lib().orc_grid_position_in_parent(self.h, n)

    def is_leaf(self, n):
        return bool(lib().orc_grid_is_leaf(self.h, n))

    def cell_coordinates(self, n):
        x = np.zeros(3)
        lib().orc_grid_cell_coordinates(self.h, n, _dptr(x))
        return x[: self.nd]

    def cell_length(self, n):
        return lib().orc_grid_cell_length(self.h, n)

    def is_cut_by(self, n, g):
        return bool(lib().orc_grid_is_cut_by(self.h, n, C.byref(g)))


def signed_distance(g, x):
    x = np.ascontiguousarray(np.pad(np.asarray(x, float), (0, 3 - len(x))))
    return lib().orc_signed_distance(C.byref(g), _dptr(x), len(x))


def phys(physics=PHYS_EULER, nd=2, flux=FLUX_AUSM, ti=TI_EULER_FWD, recon=RECON_NONE,
         limiter=LIM_ZERO, gamma=1.4, cfl=0.2, cfl_viscous=0.4, Rspecific=287.058,
         T0=273.15, rho0=1.225, Minfty=0.1, ReInfty=100.0, Pr0=0.72, Sutherland0=110.4,
         diffusivity=1.0, T_ref=273.15, time_end=1e30, max_cells=0):
    p = PhysDesc()
    for k, v in dict(physics=physics, nd=nd, flux=flux, ti=ti, recon=recon, limiter=limiter,
                     gamma=gamma, cfl=cfl, cfl_viscous=cfl_viscous, Rspecific=Rspecific, T0=T0,
                     rho0=rho0, Minfty=Minfty, ReInfty=ReInfty, Pr0=Pr0,
                     Sutherland0=Sutherland0, diffusivity=diffusivity, T_ref=T_ref,
                     time_end=time_end, max_cells=max_cells).items():
        setattr(p, k, v)
    return p


def cube_boundaries(nd, xmin, xmax):
    """grid/helpers.hpp:112-160 make_boundaries: left, right, bottom, top,
    front, back Edge geometries inset by eps = 1e-15 * length."""
    xmin = np.asarray(xmin, float)
    xmax = np.asarray(xmax, float)
    length = max(np.finfo(float).eps, float(np.max(xmax - xmin)))
    centre = xmin + 0.5 * (xmax - xmin)
    rmin = centre - 0.5 * length
    rmax = centre + 0.5 * length
    eps = length * 1e-15
    x_max = rmax[0] - eps
    x_min = rmin[0] + eps
    off = x_min + 0.5 * length
    out = []
    for d in range(nd):
        for side, val, nrm in (("neg", x_min, 1.0), ("pos", x_max, -1.0)):
            a = [off] * nd
            a[d] = val
            b = [0.0] * nd
            b[d] = nrm
            out.append(geom(GEOM_EDGE, a, b))
    return out


class Solver:
    """fv::Solver<Physics, TimeIntegration> (solver/fv/solver.hpp)."""

    def __init__(self, grid, phys_desc, solver_idx=0):
        self.grid = grid
        self.p = phys_desc
        self.nd = phys_desc.nd
        self.h = lib().orc_solver_create(grid.h, solver_idx, C.byref(phys_desc))
        self._cb = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_solver_destroy(self.h)
            self.h = None

    def append_bc(self, g, kind, params=(), other=None, kratio=1.0):
        prm = np.ascontiguousarray(params, dtype=np.float64)
        if prm.size == 0:
            prm = np.zeros(1)
            n = 0
        else:
            n = prm.size
        _check(lib().orc_solver_append_bc(self.h, C.byref(g), kind, _dptr(prm), n,
                                          other.h if other is not None else None, kratio))

    def set_bc_other(self, k, other):
        _check(lib().orc_solver_set_bc_other(self.h, k, other.h))

    def append_cube_bcs(self, kinds_params):
        """utilities.hpp:27-33 append_bcs with make_boundaries."""
        g = self.grid
        for geo, kp in zip(cube_boundaries(self.nd, g.xmin, g.xmax), kinds_params):
            kind, params = kp if isinstance(kp, tuple) else (kp, ())
            self.append_bc(geo, kind, params)

    def set_domain(self, g=None, sign=0):
        lib().orc_solver_set_domain(self.h, C.byref(g) if g is not None else None, sign)

    def set_domain_mask(self, mask):
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        self._mask = m
        lib().orc_solver_set_domain_mask(self.h, m.ctypes.data, m.size)

    def set_ic(self, kind, params=(), callback=None):
        prm = np.ascontiguousarray(params if len(params) else [0.0], dtype=np.float64)
        self._ic_params = prm
        cb = IC_CB(0)
        if callback is not None:
            nd = self.nd

            def _f(x, out, ctx):
                xs = np.ctypeslib.as_array(x, shape=(nd,)).copy()
                res = callback(xs)
                for i, v in enumerate(res):
                    out[i] = float(v)
            cb = IC_CB(_f)
        self._cb = cb
        lib().orc_solver_set_ic(self.h, kind, _dptr(prm), len(params), cb, None)

    def set_ghost_order(self, order):
        lib().orc_solver_set_ghost_order(self.h, order)

    def initialize(self):
        _check(lib().orc_solver_initialize(self.h))
        return self

    def solve(self, nsteps=1):
        _check(lib().orc_solver_solve(self.h, nsteps))

    def apply_bcs_lhs(self):
        _check(lib().orc_solver_apply_bcs_lhs(self.h))

    def evolve(self):
        _check(lib().orc_solver_evolve(self.h))

    def compute_dt(self):
        return lib().orc_solver_compute_dt(self.h)

    min_dt = compute_dt

    def force_dt(self, dt):
        lib().orc_solver_force_dt(self.h, dt)

    def dt(self):
        return lib().orc_solver_dt(self.h)

    def time(self):
        return lib().orc_solver_time(self.h)

    def step(self):
        return lib().orc_solver_step(self.h)

    def set_time(self, t, dt, step):
        lib().orc_solver_set_time(self.h, t, dt, step)

    @property
    def ncells(self):
        return lib().orc_solver_ncells(self.h)

    @property
    def first_gc(self):
        return lib().orc_solver_first_gc(self.h)

    @property
    def nvars(self):
        return lib().orc_solver_nvars(self.h)

    def quantities(self):
        out = np.zeros(7)
        lib().orc_solver_quantities(self.h, _dptr(out))
        return dict(zip(["T_inf", "rho_inf", "u_inf", "p_inf", "Re0", "sOverTref", "sOverTrefP1"], out))

    def get(self, field):
        n = self.ncells
        cols = {"lhs": self.nvars, "rhs": self.nvars, "neighbors": 2 * self.nd,
                "x_center": self.nd, "distances": 2 * self.nd, "length": 1, "node_idx": 1,
                "bc_idx": 1}[field]
        dtype = {"neighbors": np.int64, "node_idx": np.int64, "bc_idx": np.int32}.get(field, np.float64)
        out = np.empty((cols, n), dtype)
        lib().orc_solver_get(self.h, field.encode(), out.ctypes.data)
        return out

    def fine(self):
        """Fine-face table of refined grids: (nfine, 2^(nd-1)) cell ids."""
        nf = lib().orc_solver_nfine(self.h)
        out = np.empty((nf, 1 << (self.nd - 1)), np.int64)
        if nf:
            lib().orc_solver_fine(self.h, out.ctypes.data)
        return out

    def set_perturbation(self, seed, amp):
        lib().orc_solver_set_perturbation(self.h, seed, amp)

    def set_lhs(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        lib().orc_solver_set_lhs(self.h, _dptr(q))

    def set_source(self, src):
        s = np.ascontiguousarray(src, dtype=np.float64)
        self._src = s
        lib().orc_solver_set_source(self.h, _dptr(s), s.size)

    def conservation_sums(self):
        out = np.zeros(5)
        lib().orc_solver_conservation_sums(self.h, _dptr(out))
        return out[: self.nvars]

    def periodic_image(self, g):
        return lib().orc_solver_periodic_image(self.h, g)

    def cht_partner(self, g):
        return lib().orc_solver_cht_partner(self.h, g)

    def face_flux(self, l, r, d):
        out = np.zeros(5)
        lib().orc_solver_face_flux(self.h, l, r, d, _dptr(out))
        return out[: self.nvars]

    def num_flux(self, c):
        out = np.zeros(5)
        lib().orc_solver_num_flux(self.h, c, _dptr(out))
        return out[: self.nvars]

    def slope(self, c, v, d):
        return lib().orc_solver_slope(self.h, c, v, d)

    def cell_dt(self, c):
        return lib().orc_solver_cell_dt(self.h, c)

    def ausm_states(self, ql, qr, d):
        """AUSM flux of two explicit conserved states along direction d."""
        a = np.zeros(5)
        b = np.zeros(5)
        a[: self.nvars] = ql
        b[: self.nvars] = qr
        out = np.zeros(5)
        lib().orc_solver_ausm_states(self.h, _dptr(a), _dptr(b), d, _dptr(out))
        return out[: self.nvars]

    def ic_point(self, x):
        """The initial condition's conserved state at the point x."""
        xx = np.zeros(3)
        xx[: self.nd] = x
        out = np.zeros(5)
        _check(lib().orc_solver_ic_point(self.h, _dptr(xx), _dptr(out)))
        return out[: self.nvars]

    def time_solve(self, nsteps):
        return lib().orc_time_solve(self.h, nsteps)


def couple_cns_heat(Tr0, cns, heat):
    lib().orc_couple_cns_heat(Tr0, cns.h, heat.h)


def cht_wall_heat_flux(heat):
    """Interface heat flux out of the solid: sum of h^(nd-2) 2 (T_b - T_s)."""
    return lib().orc_cht_wall_heat_flux(heat.h)


def cht_objective(Tr0, cns, heat, nsteps, dts=None):
    """Time-integrated wall heat flux J = sum_n dt_heat^n Q(state^{n+1}) over
    nsteps coupled steps (coupling.hpp:47-63, then cns.solve(), heat.solve());
    with ``dts`` the fluid dt of every step is prescribed (heat dt = dt Tr0
    Pr0 Re0), otherwise it comes from the CFL bounds.  Returns (J, dts)."""
    used = []
    J = 0.0
    for n in range(nsteps):
        if dts is None:
            couple_cns_heat(Tr0, cns, heat)
        else:
            _check(lib().orc_couple_cns_heat_fixed(Tr0, float(dts[n]), cns.h, heat.h))
        cns.solve()
        heat.solve()
        used.append(cns.dt())
        J += heat.dt() * cht_wall_heat_flux(heat)
    return J, np.array(used)


def couple_same(a, b):
    lib().orc_couple_same(a.h, b.h)


def ausm_split(sign, M):
    """(m_int_<sign>(M), p_int_<sign>(M)) of euler/physics.hpp:297-310."""
    m, p = C.c_double(), C.c_double()
    lib().orc_ausm_split(sign, M, C.byref(m), C.byref(p))
    return m.value, p.value


def prng_uniform(seed, stream, counter):
    return lib().orc_prng_uniform(seed, stream, counter)


def morton_encode(nd, ijk):
    a = (C.c_uint32 * 3)(*([int(v) for v in ijk] + [0] * (3 - len(ijk))))
    return lib().orc_morton_encode(nd, a)
