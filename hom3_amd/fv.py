"""Host-side mirror of hom3's finite-volume solver interface over libhom3g.

Names follow the reference so that a hom3 driver reads the same:

* ``Grid``          — grid::CartesianHSP<nd> + grid::helpers::cube::properties
                      + generation::MinLevel (grid/grid.hpp:147-221,
                      grid/helpers.hpp:163-179, grid/generation.hpp:39-60)
* ``Solver``        — solver::fv::Solver<Physics, TimeIntegration>
                      (solver/fv/solver.hpp:60-145)
* ``Quantities``    — euler/cns quantities (euler/quantities.hpp:14-208,
                      cns/quantities.hpp:15-135)
* ``euler.bc`` / ``cns.bc`` / ``heat.bc`` conditions and ``Periodic``
* ``append_bcs``    — solver/fv/utilities.hpp:27-33 with make_boundaries
* ``coupling.apply`` and ``run_solver`` — solver/fv/coupling.hpp:47-73,
                      solver/fv/utilities.hpp:98-113

Every compute call goes to the HIP library; there is no host fallback.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from ._lib import check, lib

# ---------------------------------------------------------------------------
# context and geometry
# ---------------------------------------------------------------------------


class Context:
    """One HIP device + stream (hom3g_create)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().hom3g_create(C.byref(h), device))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            lib().hom3g_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().hom3g_synchronize(self.h))


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _geom(kind, a=(0, 0, 0), b=(0, 0, 0), invert=False):
    g = L.Geometry()
    g.kind = kind
    g.invert = 1 if invert else 0
    for i in range(3):
        g.a[i] = float(a[i]) if i < len(a) else 0.0
        g.b[i] = float(b[i]) if i < len(b) else 0.0
    return g


def Edge(point, normal):
    """geometry::implicit::Edge (implicit.hpp:91-103)."""
    return _geom(L.GEOM_EDGE, point, normal)


def Square(center, lengths):
    """geometry::implicit::Square (implicit.hpp:52-85)."""
    return _geom(L.GEOM_SQUARE, center, lengths)


def Sphere(center, radius):
    """geometry::implicit::Sphere (implicit.hpp:39-46)."""
    return _geom(L.GEOM_SPHERE, center, (radius, 0, 0))


def invert(g):
    """geometry::implicit::adaptors::invert (implicit.hpp:119-132)."""
    h = L.Geometry()
    C.pointer(h)[0] = g
    h.invert = 0 if g.invert else 1
    return h


def make_cube(center, dims, cell_length, tol=0.6):
    """geometry::make_cube (implicit.hpp:181-204): (cube0, cube1 = dims + h,
    inverted cube2 = dims - tol*h)."""
    center = list(center)
    dims = np.asarray(dims, float)
    return (Square(center, dims), Square(center, dims + cell_length),
            invert(Square(center, dims - tol * cell_length)))


# ---------------------------------------------------------------------------
# grid
# ---------------------------------------------------------------------------


class Grid:
    """grid::CartesianHSP<nd> generated with MinLevel on the device."""

    def __init__(self, nd, xmin, xmax, min_level, max_solvers=1, ctx=None):
        self.ctx = ctx or default_context()
        self.nd = nd
        self.xmin = np.asarray(xmin, float)
        self.xmax = np.asarray(xmax, float)
        self.min_level = min_level
        gid = C.c_int()
        xa = np.ascontiguousarray(self.xmin)
        xb = np.ascontiguousarray(self.xmax)
        check(lib().hom3g_grid_create(self.ctx.h, nd, xa.ctypes.data_as(C.POINTER(C.c_double)),
                                      xb.ctypes.data_as(C.POINTER(C.c_double)), min_level,
                                      max_solvers, C.byref(gid)))
        self.id = gid.value
        # RootCell (grid.hpp:108-132)
        lengths = self.xmax - self.xmin
        self.root_length = max(np.finfo(float).eps, float(np.max(lengths)))
        self.root_center = self.xmin + 0.5 * lengths

    def cell_length_at_level(self, level):
        """grid.hpp:344-346"""
        return self.root_length / math.pow(2, level)

    def root_cell(self):
        return (self.root_center - 0.5 * self.root_length, self.root_center + 0.5 * self.root_length)

    def size(self):
        n = C.c_int64()
        check(lib().hom3g_grid_size(self.ctx.h, self.id, C.byref(n)))
        return n.value

    def refine_region(self, max_level, geometry, band):
        """Refine leaves near the geometry's surface up to max_level (2:1
        refinement extension; children appended at the end of the tree)."""
        n = C.c_int64()
        check(lib().hom3g_grid_refine_region(self.ctx.h, self.id, max_level, C.byref(geometry), band,
                                             C.byref(n)))
        return n.value

    def balance(self):
        """2:1 face balance; returns the number of refined nodes."""
        n = C.c_int64()
        check(lib().hom3g_grid_balance(self.ctx.h, self.id, C.byref(n)))
        return n.value

    def set_tree(self, parent, first_child):
        p = np.ascontiguousarray(parent, np.int64)
        f = np.ascontiguousarray(first_child, np.int64)
        check(lib().hom3g_grid_set_tree(self.ctx.h, self.id, p.size, p.ctypes.data_as(C.POINTER(C.c_int64)),
                                        f.ctypes.data_as(C.POINTER(C.c_int64))))

    def tree(self):
        n = self.size()
        p = np.empty(n, np.int64)
        f = np.empty(n, np.int64)
        check(lib().hom3g_grid_tree(self.ctx.h, self.id, p.ctypes.data_as(C.POINTER(C.c_int64)),
                                    f.ctypes.data_as(C.POINTER(C.c_int64))))
        return p, f

    def neighbors(self):
        """find_samelvl_neighbor for every node, (nnodes, 2nd)."""
        n = self.size()
        out = np.empty((n, 2 * self.nd), np.int64)
        check(lib().hom3g_grid_neighbors(self.ctx.h, self.id, out.ctypes.data_as(C.POINTER(C.c_int64))))
        return out

    def geometry(self):
        n = self.size()
        x = np.empty((n, self.nd))
        ln = np.empty(n)
        check(lib().hom3g_grid_geometry(self.ctx.h, self.id, x.ctypes.data_as(C.POINTER(C.c_double)),
                                        ln.ctypes.data_as(C.POINTER(C.c_double))))
        return x, ln


def cube_boundaries(grid):
    """grid/helpers.hpp:112-160 make_boundaries: left, right, bottom, top,
    front, back Edge geometries inset by 1e-15 * length."""
    length = grid.root_length
    rmin = grid.root_center - 0.5 * length
    rmax = grid.root_center + 0.5 * length
    eps = length * 1e-15
    x_max = rmax[0] - eps
    x_min = rmin[0] + eps
    off = x_min + 0.5 * length
    out = []
    for d in range(grid.nd):
        for val, nrm in ((x_min, 1.0), (x_max, -1.0)):
            a = [off] * grid.nd
            a[d] = val
            b = [0.0] * grid.nd
            b[d] = nrm
            out.append(Edge(a, b))
    return out


# ---------------------------------------------------------------------------
# quantities (euler/quantities.hpp, cns/quantities.hpp)
# ---------------------------------------------------------------------------


class Quantities:
    def __init__(self, gamma=1.4, Minfty=0.1, ReInfty=None, Pr0=None, Sutherland0=None, T0=273.15,
                 Rspecific=287.058, rho0=1.225):
        self.gamma_ = gamma
        self.Minfty = Minfty
        self.ReInfty = ReInfty
        self.Pr0_ = Pr0
        self.T0 = T0
        self.Rspecific = Rspecific
        self.rho0 = rho0
        if Sutherland0 is not None:
            self.sOverTref = Sutherland0 / T0
            self.sOverTrefP1 = self.sOverTref + 1.0
        else:
            self.sOverTref = self.sOverTrefP1 = None

    def gamma(self):
        return self.gamma_

    def gammaM1(self):
        return self.gamma_ - 1.0

    def T_infinity(self):
        return 1.0 / (1 + 0.5 * self.gammaM1() * (self.Minfty * self.Minfty))

    def rho_infinity(self):
        return math.pow(self.T_infinity(), 1.0 / self.gammaM1())

    def u_infinity(self):
        return self.Minfty * math.sqrt(self.T_infinity())

    def p_infinity(self):
        return math.pow(self.T_infinity(), self.gamma_ / self.gammaM1()) / self.gamma_

    def sutherlands_law(self, T):
        return math.pow(T, 1.5) * self.sOverTrefP1 / (T + self.sOverTref)

    def mu_infinity(self):
        return self.sutherlands_law(self.T_infinity())

    def Re0(self):
        return self.ReInfty * self.mu_infinity() / (self.rho_infinity() * self.u_infinity())

    def Pr0(self):
        return self.Pr0_

    def a0(self):
        return math.sqrt(self.gamma_ * self.Rspecific * self.T0)

    def cv(self, pv):
        """euler/quantities.hpp:153-166"""
        pv = np.asarray(pv, float)
        nd = len(pv) - 2
        out = np.empty_like(pv)
        out[nd] = pv[nd]
        vm2 = 0.0
        for d in range(nd):
            out[d] = pv[nd] * pv[d]
            vm2 += pv[d] ** 2
        out[nd + 1] = pv[nd + 1] / (self.gamma_ - 1.0) + 0.5 * pv[nd] * vm2
        return out


# ---------------------------------------------------------------------------
# boundary conditions (one class per reference Condition)
# ---------------------------------------------------------------------------


class Condition:
    kind = None

    def __init__(self, params=(), other=None, kratio=1.0):
        self.params = list(params)
        self.other = other
        self.kratio = kratio


class Periodic(Condition):
    """Extension: ghost = periodic image; slope hook = image slope."""
    kind = L.BC_PERIODIC


class _EulerBC:
    class Neumann(Condition):
        """euler/boundary_conditions.hpp:17-38 (constant surface flux)."""
        kind = L.BC_NEUMANN

        def __init__(self, flux=()):
            super().__init__(flux)

    class Dirichlet(Condition):
        """euler/boundary_conditions.hpp:42-60 (constant surface values)."""
        kind = L.BC_DIRICHLET

        def __init__(self, values):
            super().__init__(values)

    class IsentropicVortex(Condition):
        """euler/boundary_conditions.hpp:65-89 (centre (-5,-5) unless given)."""
        kind = L.BC_ISENTROPIC_VORTEX

        def __init__(self, center=None):
            super().__init__(() if center is None else center)


class _EulerCoupling:
    class Euler(Condition):
        """euler/coupling_conditions.hpp:17-35: the ghost takes the other
        solver's state at the partner cell (same buffer role)."""
        kind = L.BC_EULER_COUPLING

        def __init__(self, other_solver):
            super().__init__((), other_solver)


_EulerBC.coupling = _EulerCoupling


class _CNSExternal:
    class Inflow(Condition):
        """cns/boundary_conditions.hpp:28-65; pv = (u.., rho, p) surface values."""
        kind = L.BC_INFLOW

        def __init__(self, pv):
            super().__init__(pv)

    class Outflow(Condition):
        """cns/boundary_conditions.hpp:72-105; pv[p] is the surface pressure."""
        kind = L.BC_OUTFLOW

        def __init__(self, pv):
            super().__init__(pv)


class _CNSWall:
    class AdiabaticNoSlip(Condition):
        kind = L.BC_ADIABATIC_NOSLIP

    class AdiabaticSlip(Condition):
        kind = L.BC_ADIABATIC_SLIP

    class IsothermalNoSlip(Condition):
        kind = L.BC_ISOTHERMAL_NOSLIP

        def __init__(self, T):
            super().__init__((T,))


class _CNSCoupling:
    class Heat(Condition):
        """cns/coupling_conditions.hpp:21-43 (T_srfc = (T_f + T_s)/2)."""
        kind = L.BC_CHT_FLUID

        def __init__(self, heat_solver, kratio=1.0):
            super().__init__((), heat_solver, kratio)


class _CNSBC:
    Neumann = _EulerBC.Neumann
    external = _CNSExternal
    wall = _CNSWall
    coupling = _CNSCoupling


class _HeatCoupling:
    class CNS(Condition):
        """heat/coupling_conditions.hpp:18-38."""
        kind = L.BC_CHT_SOLID

        def __init__(self, cns_solver, kratio=1.0):
            super().__init__((), cns_solver, kratio)


class _HeatBC:
    class Dirichlet(Condition):
        kind = L.BC_HEAT_DIRICHLET

        def __init__(self, T):
            super().__init__((T,))

    class Neumann(Condition):
        kind = L.BC_HEAT_NEUMANN

        def __init__(self, flux=0.0):
            super().__init__((flux,))

    coupling = _HeatCoupling


class GhostSet:
    """What a host-evaluated condition sees (hom3g_bc_callback): the buffer
    (0 lhs, 1 rhs, 2 SSP-RK3 stage), the solver's time, the ghosts it fills
    and their boundary cells (ids and centres, (nd, n)), the boundary cells'
    states and the ghost states to fill ((nvars, n), writable; boundary
    cells may be overwritten too)."""

    def __init__(self, buffer, time, n, ghosts, bndry, xg, xb, qg, qb, nd, nvars):
        self.buffer, self.time, self.n = buffer, time, n
        self.ghosts = np.ctypeslib.as_array(ghosts, shape=(n,)) if n else np.empty(0, np.int64)
        self.bndry = np.ctypeslib.as_array(bndry, shape=(n,)) if n else np.empty(0, np.int64)
        self.x_ghost = np.ctypeslib.as_array(xg, shape=(nd, n)) if n else np.empty((nd, 0))
        self.x_bndry = np.ctypeslib.as_array(xb, shape=(nd, n)) if n else np.empty((nd, 0))
        self.q_ghost = np.ctypeslib.as_array(qg, shape=(nvars, n)) if n else np.empty((nvars, 0))
        self.q_bndry = np.ctypeslib.as_array(qb, shape=(nvars, n)) if n else np.empty((nvars, 0))


class HostCondition(Condition):
    """Any other Condition (bc::Interface, boundary_condition.hpp:22-67),
    evaluated on the host: ``fn(ghost_set)`` fills ``ghost_set.q_ghost``
    from ``ghost_set.q_bndry`` (and may overwrite the boundary cells, as
    IsentropicVortex does).  ``hook``: the Condition's slope<> hook,
    "default" (the boundary cell's slope) or "wall" (velocity slopes
    negated, cns/boundary_conditions.hpp:143-147)."""
    kind = L.BC_HOST_CALLBACK

    def __init__(self, fn, hook="default"):
        super().__init__(())
        self.fn = fn
        self.hook = {"default": L.SLOPE_DEFAULT, "wall": L.SLOPE_WALL}[hook]


class _NS:
    def __init__(self, **kw):
        self.__dict__.update(kw)


euler = _NS(bc=_EulerBC)
cns = _NS(bc=_CNSBC)
heat = _NS(bc=_HeatBC)


# ---------------------------------------------------------------------------
# solver
# ---------------------------------------------------------------------------

_PHYSICS = {"euler": L.EULER, "cns": L.CNS, "heat": L.HEAT}
_FLUX = {"ausm": L.FLUX_AUSM, "lax_friedrichs": L.FLUX_LLF, "llf": L.FLUX_LLF,
         "three_point": L.FLUX_THREE_POINT, "standard": L.FLUX_AUSM}
_TI = {"euler_forward": L.TI_EULER_FWD, "runge_kutta_2": L.TI_HEUN_RK2, "ssp_rk3": L.TI_SSP_RK3}


class Solver:
    """solver::fv::Solver<Physics, TimeIntegration> on the device.

    ``properties`` mirrors the io::Properties keys the reference reads:
    timeEnd, CFL, CFL_viscous, gamma, Rspecific, T0, rho0, Minfty, ReInfty,
    Pr0, Sutherland0, diffusivity, T_ref, maxNoCells.
    """

    def __init__(self, grid, physics="euler", flux="ausm", time_integration="euler_forward",
                 solver_idx=0, recon="none", limiter="minmod", precision="fp64", kernel="auto",
                 **properties):
        self.grid = grid
        self.ctx = grid.ctx
        self.nd = grid.nd
        self.physics = physics
        p = L.Physics()
        p.physics = _PHYSICS[physics]
        p.nd = grid.nd
        p.flux = _FLUX[flux]
        p.ti = _TI[time_integration]
        p.recon = {"none": L.RECON_NONE, "muscl": L.RECON_MUSCL}[recon]
        p.limiter = {"zero": L.LIM_ZERO, "minmod": L.LIM_MINMOD, "vanleer": L.LIM_VANLEER}[limiter]
        prop = dict(gamma=1.4, CFL=0.2, CFL_viscous=0.4, Rspecific=287.058, T0=273.15, rho0=1.225,
                    Minfty=0.1, ReInfty=100.0, Pr0=0.72, Sutherland0=110.4, diffusivity=1.0,
                    T_ref=273.15, timeEnd=1e30, maxNoCells=0)
        prop.update(properties)
        self.properties = prop
        p.gamma = prop["gamma"]
        p.cfl = prop["CFL"]
        p.cfl_viscous = prop["CFL_viscous"]
        p.Rspecific = prop["Rspecific"]
        p.T0 = prop["T0"]
        p.rho0 = prop["rho0"]
        p.Minfty = prop["Minfty"]
        p.ReInfty = prop["ReInfty"]
        p.Pr0 = prop["Pr0"]
        p.Sutherland0 = prop["Sutherland0"]
        p.diffusivity = prop["diffusivity"]
        p.T_ref = prop["T_ref"]
        p.time_end = prop["timeEnd"]
        p.max_cells = int(prop["maxNoCells"])
        p.precision = {"fp64": L.FP64, "fp32": L.FP32_STATE, "fp32_state": L.FP32_STATE}[precision]
        p.kernel = {"auto": L.KERNEL_AUTO, "generic": L.KERNEL_GENERIC, "brick": L.KERNEL_BRICK}[kernel]
        self.desc = p
        self.quantities = Quantities(gamma=p.gamma, Minfty=p.Minfty, ReInfty=p.ReInfty, Pr0=p.Pr0,
                                     Sutherland0=p.Sutherland0, T0=p.T0)
        self.nvars = 1 if physics == "heat" else self.nd + 2
        sid = C.c_int()
        check(lib().hom3g_solver_create(self.ctx.h, grid.id, solver_idx, C.byref(p), C.byref(sid)))
        self.id = sid.value
        self.solver_idx = solver_idx
        self._ic = None
        self._bcs = []
        self._pending = []
        self._built = False

    # -- input (solver.hpp:150-169) ---------------------------------------------
    def set_initial_condition(self, kind, params=()):
        """Built-in IC kind (hom3g IC enum) and its parameters, or a callable
        x -> conserved variables (the "initialCondition" property,
        solver.hpp:153-156), evaluated on the host at every cell's Gauss
        points by the library (hom3g_solver_impose_ic_host)."""
        self._ic = (kind, list(params)) if not callable(kind) else (kind, None)

    def set_perturbation(self, seed, amplitude):
        """Seeded density perturbation applied after the initial condition:
        rho *= 1 + amplitude (2U - 1), U keyed by (seed, level, Morton index)."""
        check(lib().hom3g_solver_set_perturbation(self.ctx.h, self.id, int(seed), float(amplitude)))

    def set_initial_domain(self, geometry=None, sign=0, mask=None):
        if mask is not None:
            m = np.ascontiguousarray(mask, np.uint8)
            self._mask = m
            check(lib().hom3g_solver_set_domain_mask(self.ctx.h, self.id, m.ctypes.data, m.size))
        else:
            check(lib().hom3g_solver_set_domain(self.ctx.h, self.id,
                                                C.byref(geometry) if geometry is not None else None, sign))

    def append_bc(self, geometry, condition):
        if condition.kind == L.BC_HOST_CALLBACK:
            nd, nv, fn = self.nd, self.nvars, condition.fn

            def _cb(user, buffer, time, n, ghosts, bndry, xg, xb, qg, qb):
                fn(GhostSet(buffer, time, n, ghosts, bndry, xg, xb, qg, qb, nd, nv))
            cb = L.BC_CALLBACK(_cb)
            self._callbacks = getattr(self, "_callbacks", []) + [cb]
            check(lib().hom3g_solver_append_bc_host(self.ctx.h, self.id, C.byref(geometry), cb, None,
                                                    condition.hook))
            self._bcs.append((geometry, condition))
            return
        other = -1
        if condition.other is not None:
            other = condition.other.id
        prm = np.ascontiguousarray(condition.params if condition.params else [0.0], np.float64)
        check(lib().hom3g_solver_append_bc(self.ctx.h, self.id, C.byref(geometry), condition.kind,
                                           prm.ctypes.data_as(C.POINTER(C.c_double)),
                                           len(condition.params), other, float(condition.kratio)))
        self._bcs.append((geometry, condition))

    def append_bcs(self, pairs):
        for g, c in pairs:
            self.append_bc(g, c)

    # -- lifecycle (solver.hpp:133-145) -----------------------------------------
    def build(self):
        if not self._built:
            check(lib().hom3g_solver_build(self.ctx.h, self.id))
            self._built = True
        return self

    def initialize(self):
        """create_local_cells + impose_initial_condition + set_dt."""
        self.build()
        if self._ic is None:
            raise RuntimeError("initialCondition not set")
        kind, params = self._ic
        if callable(kind):
            nd, nv = self.nd, self.nvars

            def _ic(user, x, q):
                out = kind(np.ctypeslib.as_array(x, shape=(nd,)).copy())
                for v in range(nv):
                    q[v] = float(out[v])
            cb = L.IC_CALLBACK(_ic)
            self._ic_cb = cb
            check(lib().hom3g_solver_impose_ic_host(self.ctx.h, self.id, cb, None))
            return self
        prm = np.ascontiguousarray(params if params else [0.0], np.float64)
        check(lib().hom3g_solver_impose_ic(self.ctx.h, self.id, kind,
                                           prm.ctypes.data_as(C.POINTER(C.c_double)), len(params)))
        return self

    def solve(self, nsteps=1):
        rc = lib().hom3g_solver_solve(self.ctx.h, self.id, nsteps)
        if rc == L.ERR_DIVERGED:
            raise SolutionDiverged(rc, lib().hom3g_last_error().decode())
        check(rc)

    def apply_bcs(self):
        """apply_bcs(lhs) (solver.hpp:561-586) on the current state."""
        check(lib().hom3g_solver_apply_bcs(self.ctx.h, self.id))

    def step(self):
        return self._time()[2]

    def time(self):
        return self._time()[0]

    def dt(self, value=None):
        if value is not None:
            check(lib().hom3g_solver_force_dt(self.ctx.h, self.id, float(value)))
            return None
        return self._time()[1]

    def min_dt(self):
        out = C.c_double()
        check(lib().hom3g_solver_min_dt(self.ctx.h, self.id, C.byref(out)))
        return out.value

    def final_time(self):
        return self.desc.time_end

    def _time(self):
        t, dt, st = C.c_double(), C.c_double(), C.c_int64()
        check(lib().hom3g_solver_time(self.ctx.h, self.id, C.byref(t), C.byref(dt), C.byref(st)))
        return t.value, dt.value, st.value

    def reset_time(self, time, dt, step):
        check(lib().hom3g_solver_reset_time(self.ctx.h, self.id, time, dt, step))

    # -- data -----------------------------------------------------------------
    def counts(self):
        n, f = C.c_int64(), C.c_int64()
        check(lib().hom3g_solver_counts(self.ctx.h, self.id, C.byref(n), C.byref(f)))
        return n.value, f.value

    def kernel(self):
        """'generic' or 'brick' (the kernel family picked at build)."""
        k = C.c_int()
        check(lib().hom3g_solver_kernel(self.ctx.h, self.id, C.byref(k)))
        return {L.KERNEL_GENERIC: "generic", L.KERNEL_BRICK: "brick"}[k.value]

    @property
    def ncells(self):
        return self.counts()[0]

    @property
    def first_gc(self):
        return self.counts()[1]

    def topology(self):
        n, _ = self.counts()
        nd = self.nd
        out = dict(neighbors=np.empty((2 * nd, n), np.int64), node_idx=np.empty((1, n), np.int64),
                   bc_idx=np.empty((1, n), np.int32), x_center=np.empty((nd, n)),
                   length=np.empty((1, n)), distances=np.empty((2 * nd, n)))
        check(lib().hom3g_solver_topology(self.ctx.h, self.id, *(out[k].ctypes.data for k in
                                                                 ("neighbors", "node_idx", "bc_idx",
                                                                  "x_center", "length", "distances"))))
        return out

    def fine_faces(self):
        """Faces towards finer cells of a 2:1 refined grid: (nfine, 2^(nd-1))
        cell ids; the neighbour table holds -(2 + f) for face f."""
        nf = C.c_int64()
        check(lib().hom3g_solver_fine_faces(self.ctx.h, self.id, None, C.byref(nf)))
        out = np.empty((nf.value, 1 << (self.nd - 1)), np.int32)
        if nf.value:
            check(lib().hom3g_solver_fine_faces(self.ctx.h, self.id, out.ctypes.data, C.byref(nf)))
        return out

    @property
    def lhs(self):
        n, _ = self.counts()
        out = np.empty((self.nvars, n))
        check(lib().hom3g_solver_download_state(self.ctx.h, self.id, out.ctypes.data_as(C.POINTER(C.c_double)), n))
        return out

    def set_lhs(self, q):
        q = np.ascontiguousarray(q, np.float64)
        check(lib().hom3g_solver_upload_state(self.ctx.h, self.id, q.ctypes.data_as(C.POINTER(C.c_double)),
                                              q.shape[1]))

    def set_source(self, src):
        s = np.ascontiguousarray(src, np.float64)
        check(lib().hom3g_solver_set_source(self.ctx.h, self.id, s.ctypes.data_as(C.POINTER(C.c_double)), s.size))

    def conservation_sums(self):
        out = np.zeros(5)
        check(lib().hom3g_solver_conservation_sums(self.ctx.h, self.id, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out[: self.nvars]

    def check_state(self):
        n = C.c_int64()
        check(lib().hom3g_solver_check_state(self.ctx.h, self.id, C.byref(n)))
        return n.value

    def time_solve(self, nsteps):
        ms = C.c_float()
        check(lib().hom3g_solver_time_solve(self.ctx.h, self.id, nsteps, C.byref(ms)))
        return ms.value

    def profile_solve(self, nsteps):
        """(total ms, mean ms per stage launch, number of stage launches)."""
        tot, st, n = C.c_float(), C.c_float(), C.c_int()
        check(lib().hom3g_solver_profile_solve(self.ctx.h, self.id, nsteps, C.byref(tot), C.byref(st), C.byref(n)))
        return tot.value, st.value, n.value

    def close(self):
        if getattr(self, "id", None) is not None and self.ctx.h:
            lib().hom3g_solver_destroy(self.ctx.h, self.id)
            self.id = None


def append_bcs(solver, conditions):
    """utilities.hpp:27-33: unit-cube boundaries (left, right, bottom, top,
    front, back) with one condition each."""
    geos = cube_boundaries(solver.grid)
    if not isinstance(conditions, (list, tuple)):
        conditions = [conditions] * len(geos)
    for g, c in zip(geos, conditions):
        solver.append_bc(g, c)


class coupling:
    @staticmethod
    def apply_and_step(Tr0, cns_solver, heat_solver, nsteps=1):
        """nsteps of coupling::apply(Tr0, cns, heat); cns.solve(); heat.solve()."""
        check(lib().hom3g_couple_cht_step(cns_solver.ctx.h, cns_solver.id, heat_solver.id, float(Tr0), nsteps))

    @staticmethod
    def same_and_step(s1, s2, nsteps=1):
        check(lib().hom3g_couple_same_step(s1.ctx.h, s1.id, s2.id, nsteps))


def solver_finished(max_steps, *solvers):
    """utilities.hpp:73-82"""
    return any(s.time() >= s.final_time() or s.step() >= max_steps for s in solvers)


def solution_diverged(*solvers):
    """utilities.hpp:85-92"""
    return any(s.dt() < 1e-300 or s.dt() > 1e300 for s in solvers)


class SolutionDiverged(L.Hom3gError):
    """Solution diverged! (utilities.hpp:107-110): a step's dt left
    [1e-300, 1e300]; hom3g_solver_solve returns HOM3G_ERR_DIVERGED."""


def run_solver(grid, solver, max_steps, output_interval=None, on_output=None):
    """utilities.hpp:98-113: initialize, then solve until solver_finished,
    with on_output(solver) every output_interval steps (write_output; e.g.
    vtk.write_domain) and SolutionDiverged in place of TERMINATE."""
    solver.initialize()
    while not solver_finished(max_steps, solver):
        solver.solve()
        if output_interval and on_output and solver.step() % output_interval == 0:
            on_output(solver)
        if solution_diverged(solver):
            raise SolutionDiverged(L.ERR_DIVERGED, "Solution diverged!")
    return solver
