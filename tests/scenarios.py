"""Build the same hom3 scenario on the oracle (CPU) and on libhom3g (GPU).

A scenario is a plain dict; ``build_oracle`` and ``build_gpu`` turn it into
an initialized oracle Solver and an initialized hom3_amd Solver.  The oracle
is the checker only.
"""
import numpy as np

import oracle as O

PROP_DEFAULTS = dict(gamma=1.4, CFL=0.2, CFL_viscous=0.4, Rspecific=287.058, T0=273.15, rho0=1.225,
                     Minfty=0.1, ReInfty=100.0, Pr0=0.72, Sutherland0=110.4, diffusivity=1.0,
                     T_ref=273.15, timeEnd=1e30, maxNoCells=0)

PHYS = {"euler": O.PHYS_EULER, "cns": O.PHYS_CNS, "heat": O.PHYS_HEAT}
FLUX = {"ausm": O.FLUX_AUSM, "llf": O.FLUX_LLF, "three_point": O.FLUX_THREE_POINT}
TI = {"euler_forward": O.TI_EULER_FWD, "runge_kutta_2": O.TI_HEUN_RK2, "ssp_rk3": O.TI_SSP_RK3}
RECON = {"none": O.RECON_NONE, "muscl": O.RECON_MUSCL}
LIM = {"zero": O.LIM_ZERO, "minmod": O.LIM_MINMOD, "vanleer": O.LIM_VANLEER}
GKIND = {"edge": O.GEOM_EDGE, "square": O.GEOM_SQUARE, "sphere": O.GEOM_SPHERE}


def scenario(nd, level, physics="euler", xmin=None, xmax=None, flux=None, ti="euler_forward",
             recon="none", limiter="minmod", bcs=None, domain=None, ic=None, **props):
    xmin = [0.0] * nd if xmin is None else list(xmin)
    xmax = [1.0] * nd if xmax is None else list(xmax)
    flux = flux or ("three_point" if physics == "heat" else "ausm")
    p = dict(PROP_DEFAULTS)
    p.update(props)
    return dict(nd=nd, level=level, physics=physics, xmin=xmin, xmax=xmax, flux=flux, ti=ti,
                recon=recon, limiter=limiter, bcs=bcs or [], domain=domain, ic=ic, props=p)


def cube_bcs(sc, kind, params=()):
    """make_boundaries on the scenario's root cell with one condition kind."""
    return [("cube", i, kind, list(params)) for i in range(2 * sc["nd"])]


def _oracle_geom(sc, spec):
    if spec[0] == "cube":
        return O.cube_boundaries(sc["nd"], sc["xmin"], sc["xmax"])[spec[1]]
    kind, a, b, inv = spec[0], spec[1], spec[2], spec[3] if len(spec) > 3 else False
    return O.geom(GKIND[kind], a, b, inv)


def phys_desc(sc):
    p = sc["props"]
    return O.phys(physics=PHYS[sc["physics"]], nd=sc["nd"], flux=FLUX[sc["flux"]], ti=TI[sc["ti"]],
                  recon=RECON[sc["recon"]], limiter=LIM[sc["limiter"]], gamma=p["gamma"], cfl=p["CFL"],
                  cfl_viscous=p["CFL_viscous"], Rspecific=p["Rspecific"], T0=p["T0"], rho0=p["rho0"],
                  Minfty=p["Minfty"], ReInfty=p["ReInfty"], Pr0=p["Pr0"], Sutherland0=p["Sutherland0"],
                  diffusivity=p["diffusivity"], T_ref=p["T_ref"], time_end=p["timeEnd"],
                  max_cells=int(p["maxNoCells"]))


def build_oracle(sc, grid=None, solver_idx=0, init=True, others=None):
    g = grid or O.Grid(sc["nd"], sc["xmin"], sc["xmax"], max_solvers=2).min_level(sc["level"])
    s = O.Solver(g, phys_desc(sc), solver_idx)
    for spec in sc["bcs"]:
        geo_spec = spec[:2] if spec[0] == "cube" else spec[0]
        if spec[0] == "cube":
            kind, params = spec[2], spec[3]
            other = None
            kr = 1.0
        else:
            kind, params = spec[1], spec[2]
            other = spec[3] if len(spec) > 3 else None
            kr = spec[4] if len(spec) > 4 else 1.0
        geo = _oracle_geom(sc, geo_spec)
        s.append_bc(geo, kind, params, other=(others or {}).get(other) if isinstance(other, str) else None,
                    kratio=kr)
    if sc["domain"] is not None:
        geo_spec, sign = sc["domain"]
        s.set_domain(_oracle_geom(sc, geo_spec), sign)
    if sc["ic"] is not None:
        s.set_ic(*sc["ic"])
    if init:
        s.initialize()
    return g, s


def build_gpu(sc, grid=None, solver_idx=0, kernel="auto", precision="fp64", build_only=False, others=None):
    import hom3_amd as H
    from hom3_amd import _lib as L
    g = grid or H.Grid(sc["nd"], sc["xmin"], sc["xmax"], sc["level"], max_solvers=2)
    p = dict(sc["props"])
    s = H.Solver(g, physics=sc["physics"], flux=sc["flux"], time_integration=sc["ti"], solver_idx=solver_idx,
                 recon=sc["recon"], limiter=sc["limiter"], precision=precision, kernel=kernel, **p)
    for spec in sc["bcs"]:
        if spec[0] == "cube":
            geo = H.cube_boundaries(g)[spec[1]]
            kind, params, other, kr = spec[2], spec[3], None, 1.0
        else:
            gs = spec[0]
            geo = {"edge": H.Edge, "square": H.Square}.get(gs[0])
            if gs[0] == "sphere":
                geo = H.Sphere(gs[1], gs[2][0])
            else:
                geo = geo(gs[1], gs[2])
            if len(gs) > 3 and gs[3]:
                geo = H.invert(geo)
            kind, params = spec[1], spec[2]
            other = spec[3] if len(spec) > 3 else None
            kr = spec[4] if len(spec) > 4 else 1.0
        cond = H.Condition(params, (others or {}).get(other) if isinstance(other, str) else None, kr)
        cond.kind = kind
        s.append_bc(geo, cond)
    if sc["domain"] is not None:
        gs, sign = sc["domain"]
        geo = H.Square(gs[1], gs[2]) if gs[0] == "square" else H.Sphere(gs[1], gs[2][0])
        if len(gs) > 3 and gs[3]:
            geo = H.invert(geo)
        s.set_initial_domain(geo, sign)
    if build_only:
        s.build()
        return g, s
    if sc["ic"] is not None:
        s.set_initial_condition(sc["ic"][0], sc["ic"][1])
        s.initialize()
    return g, s


def rel_l2(a, b, nd=None):
    """Relative L2 error per conserved variable (rows).

    With ``nd`` given and nd+2 rows (Euler/CNS), the momentum components are
    normalised by the norm of the whole momentum field: a component that is
    zero in exact arithmetic (e.g. transverse momentum of a 1D shock tube)
    carries only round-off, and its own norm is no scale for it."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    num = np.sqrt(np.sum((a - b) ** 2, axis=-1))
    den = np.sqrt(np.sum(b ** 2, axis=-1))
    if nd is not None and a.shape[0] == nd + 2:
        mom = np.sqrt(np.sum(b[:nd] ** 2))
        den = den.copy()
        den[:nd] = mom
    return num / np.where(den > 0, den, 1.0)
