"""Pins for the oracle's physics functions that the reference's own tests
leave open (CPU only).

Each test checks the oracle against something the mathematics fixes
independently of the oracle's code: partition-of-unity and continuity
properties of the AUSM splittings, consistency F(U, U) = F_Euler(U) of the
numerical fluxes, the Liou-Steffen flux written in its average-minus-
dissipation form, closed-form ghost states read as surface statements (the
face mean equals the imposed value, the one-sided difference the imposed
gradient), flux continuity at the conjugate interface, the amplification
factors of the Runge-Kutta schemes on a discrete eigenmode, exact radial
equilibrium and isentropy of the vortex, the incompressible pressure
Poisson equation of the Taylor-Green field, and the full viscous stress
tensor and Fourier heat flux of linear fields.  A dropped term, a wrong
sign, coefficient or index in the oracle fails one of them.
"""
import math

import numpy as np
import pytest

import oracle as O

G = 1.4
E = O.INVALID


# --------------------------------------------------------------------------- helpers
def pv_of(q, nd, gamma=G):
    """Conserved (rows) -> primitive (u_0..u_nd-1, rho, p)."""
    q = np.asarray(q, float)
    rho = q[nd]
    u = q[:nd] / rho
    p = (gamma - 1) * (q[nd + 1] - 0.5 * rho * np.sum(u * u, axis=0))
    return np.concatenate([u, rho[None] if np.ndim(rho) else [rho], p[None] if np.ndim(p) else [p]])


def cv_of(pv, nd, gamma=G):
    pv = np.asarray(pv, float)
    u = pv[:nd]
    rho, p = pv[nd], pv[nd + 1]
    return np.concatenate([rho * u, [rho] if np.ndim(rho) == 0 else rho[None],
                           [p / (gamma - 1) + 0.5 * rho * np.sum(u * u, axis=0)] if np.ndim(rho) == 0
                           else (p / (gamma - 1) + 0.5 * rho * np.sum(u * u, axis=0))[None]])


def euler_flux(q, d, nd, gamma=G):
    """Physical Euler flux along direction d of one conserved state."""
    u, rho, p = (lambda w: (w[:nd], w[nd], w[nd + 1]))(pv_of(q, nd, gamma))
    f = np.empty(nd + 2)
    f[:nd] = q[:nd] * u[d]
    f[d] += p
    f[nd] = q[d]
    f[nd + 1] = u[d] * (q[nd + 1] + p)
    return f


def euler_solver(nd, level=2, **kw):
    g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(level)
    s = O.Solver(g, O.phys(nd=nd, **kw))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * (2 * nd))
    s.set_ic(O.IC_CONSTANT_PV, [0.0] * nd + [1.0, 1.0])
    s.initialize()
    return g, s


def random_state(rng, nd, mach):
    """A conserved state with rho, p ~ U[0.5, 2] and |u| = mach * a."""
    rho = rng.uniform(0.5, 2.0)
    p = rng.uniform(0.5, 2.0)
    a = math.sqrt(G * p / rho)
    n = rng.standard_normal(nd)
    u = mach * a * n / np.linalg.norm(n)
    return cv_of(np.concatenate([u, [rho, p]]), nd)


def close(a, b, rtol, scale=None):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.max(np.abs(b)) if scale is None else scale
    assert np.all(np.abs(a - b) <= rtol * max(s, 1e-300)), (a, b)


# --------------------------------------------------------------------------- AUSM splittings
# Liou & Steffen, J. Comput. Phys. 107 (1993) 23-39: for |M| <= 1
#   M+-(M) = +-(M +- 1)^2 / 4,   P+-(M) = (M +- 1)^2 (2 -+ M) / 4,
# for |M| > 1 the upwind values M+- = (M +- |M|)/2, P+- = (1 +- sign M)/2.
# The oracle follows euler/physics.hpp:297-310 (m_int_, p_int_).
MACHS = np.concatenate([np.linspace(-3, 3, 121), [-1.0, 1.0, -1 + 1e-9, 1 - 1e-9, 1 + 1e-9, -1 - 1e-9, 0.0]])


def test_ausm_split_mach_partition():
    """M+(M) + M-(M) = M for every M (consistency of the mass flux)."""
    for M in MACHS:
        mp, _ = O.ausm_split(+1, M)
        mm, _ = O.ausm_split(-1, M)
        assert abs((mp + mm) - M) <= 2e-16 * max(1.0, abs(M)), M


def test_ausm_split_pressure_partition():
    """P+(M) + P-(M) = 1 for every M (consistency of the pressure flux)."""
    for M in MACHS:
        _, pp = O.ausm_split(+1, M)
        _, pm = O.ausm_split(-1, M)
        assert abs((pp + pm) - 1.0) <= 4e-16, M


def test_ausm_split_expanded_polynomials():
    """Subsonic branch against the expanded polynomials
    M+ = M^2/4 + M/2 + 1/4 and P+ = 1/2 + 3M/4 - M^3/4 (and mirror images)."""
    for M in np.linspace(-1, 1, 41):
        mp, pp = O.ausm_split(+1, M)
        mm, pm = O.ausm_split(-1, M)
        assert mp == pytest.approx(0.25 * M * M + 0.5 * M + 0.25, abs=5e-16)
        assert mm == pytest.approx(-(0.25 * M * M - 0.5 * M + 0.25), abs=5e-16)
        assert pp == pytest.approx(0.5 + 0.75 * M - 0.25 * M ** 3, abs=5e-16)
        assert pm == pytest.approx(0.5 - 0.75 * M + 0.25 * M ** 3, abs=5e-16)


def test_ausm_split_supersonic_is_upwind():
    for M in (1.5, 2.0, 7.0):
        assert O.ausm_split(+1, M) == (M, 1.0)
        assert O.ausm_split(-1, M) == (0.0, 0.0)
        assert O.ausm_split(+1, -M) == (0.0, 0.0)
        assert O.ausm_split(-1, -M) == (-M, 1.0)


def test_ausm_split_c1_at_sonic_points():
    """Both splittings are continuously differentiable at |M| = 1."""
    h = 1e-6
    for s in (+1, -1):
        for M0 in (-1.0, 1.0):
            for k in (0, 1):
                f = lambda M: O.ausm_split(s, M)[k]
                left = (f(M0) - f(M0 - h)) / h
                right = (f(M0 + h) - f(M0)) / h
                assert abs(f(M0 + 1e-12) - f(M0 - 1e-12)) < 1e-11
                assert abs(left - right) < 4e-6, (s, M0, k, left, right)


def test_ausm_split_mirror_symmetry():
    for M in MACHS:
        mp, pp = O.ausm_split(+1, M)
        mm, pm = O.ausm_split(-1, -M)
        assert mp == -mm and pp == pm


# --------------------------------------------------------------------------- AUSM face flux
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("mach", [0.0, 0.3, 0.95, 1.7])
def test_ausm_consistency(nd, mach):
    """F(U, U) = F_Euler(U) on subsonic and supersonic states."""
    rng = np.random.default_rng(int(10 * mach) + nd)
    _, s = euler_solver(nd)
    for _ in range(6):
        q = random_state(rng, nd, mach)
        for d in range(nd):
            close(s.ausm_states(q, q, d), euler_flux(q, d, nd), 4e-15)


@pytest.mark.parametrize("nd", [2, 3])
def test_ausm_supersonic_states_upwind(nd):
    rng = np.random.default_rng(5)
    _, s = euler_solver(nd)
    for sign in (+1, -1):
        for _ in range(4):
            ql = random_state(rng, nd, 0.1)
            qr = random_state(rng, nd, 0.1)
            for q in (ql, qr):       # normal velocity of 1.5-3 sound speeds along x
                pv = pv_of(q, nd)
                pv[0] = sign * rng.uniform(1.5, 3.0) * math.sqrt(G * pv[nd + 1] / pv[nd])
                q[:] = cv_of(pv, nd)
            up = ql if sign > 0 else qr
            close(s.ausm_states(ql, qr, 0), euler_flux(up, 0, nd), 4e-15)


def _liou_steffen(ql, qr, d, nd):
    """Average-minus-dissipation form of the AUSM flux with the expanded
    split polynomials: F = M_h/2 (Phi_L + Phi_R) - |M_h|/2 (Phi_R - Phi_L)
    + P_h e_d, Phi = rho a (u, 1, H)."""
    def split(M, s):
        if abs(M) > 1:
            return 0.5 * (M + s * abs(M)), 0.5 * (1 + s * math.copysign(1.0, M))
        m = s * (0.25 * M * M + s * 0.5 * M + 0.25)
        return m, 0.5 + s * (0.75 * M - 0.25 * M ** 3)

    def phi(q):
        pv = pv_of(q, nd)
        rho, p = pv[nd], pv[nd + 1]
        a = math.sqrt(G * p / rho)
        H = (q[nd + 1] + p) / rho
        return rho * a * np.concatenate([pv[:nd], [1.0, H]]), pv[d] / a, p
    fl, ML, pL = phi(ql)
    fr, MR, pR = phi(qr)
    mp, pp = split(ML, +1)
    mm, pm = split(MR, -1)
    Mh = mp + mm
    f = 0.5 * Mh * (fl + fr) - 0.5 * abs(Mh) * (fr - fl)
    f[d] += pp * pL + pm * pR
    return f


@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("machs", [(0.2, 0.4), (0.9, 1.3), (1.5, 0.5), (0.0, 0.7)])
def test_ausm_matches_liou_steffen(nd, machs):
    rng = np.random.default_rng(17 + nd)
    _, s = euler_solver(nd)
    for _ in range(8):
        ql = random_state(rng, nd, machs[0])
        qr = random_state(rng, nd, machs[1])
        for d in range(nd):
            close(s.ausm_states(ql, qr, d), _liou_steffen(ql, qr, d, nd), 1e-14)


@pytest.mark.parametrize("nd", [2, 3])
def test_ausm_mirror_symmetry(nd):
    """Reflecting the normal velocity and swapping the sides reverses mass,
    energy and tangential-momentum fluxes and keeps the normal-momentum flux."""
    rng = np.random.default_rng(23)
    _, s = euler_solver(nd)
    for _ in range(8):
        ql = random_state(rng, nd, 0.6)
        qr = random_state(rng, nd, 0.8)
        for d in range(nd):
            ml, mr = ql.copy(), qr.copy()
            ml[d], mr[d] = -ml[d], -mr[d]
            f = s.ausm_states(ql, qr, d)
            g = s.ausm_states(mr, ml, d)
            sign = -np.ones(nd + 2)
            sign[d] = 1.0
            close(g, sign * f, 1e-15)


def test_ausm_stationary_contact_is_pressure_only():
    """u = 0, equal p, different rho: the AUSM flux is the pressure alone."""
    _, s = euler_solver(2)
    ql = cv_of([0.0, 0.0, 1.0, 0.7], 2)
    qr = cv_of([0.0, 0.0, 0.125, 0.7], 2)
    assert np.array_equal(s.ausm_states(ql, qr, 0), [0.7, 0.0, 0.0, 0.0])


# --------------------------------------------------------------------------- LLF
@pytest.mark.parametrize("nd", [2, 3])
def test_llf_is_average_plus_dissipation(nd):
    """euler/physics.hpp:262-267: F(L,R) + F(R,L) = F_E(L) + F_E(R) and
    F(L,R) - F(R,L) = (dx/dt)(U_L - U_R); F(U,U) = F_E(U)."""
    rng = np.random.default_rng(31)
    g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(3)
    s = O.Solver(g, O.phys(nd=nd, flux=O.FLUX_LLF, cfl=0.3))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * (2 * nd))
    s.set_ic(O.IC_CONSTANT_PV, [0.1] * nd + [1.0, 1.0])
    s.initialize()
    q = s.get("lhs")
    n = s.first_gc
    for c in range(n):
        q[:, c] = random_state(rng, nd, 0.5)
    s.set_lhs(q)
    nb = s.get("neighbors")
    dt = s.dt()
    h = s.get("length")[0, 0]
    c = O.morton_encode(nd, [3] * nd)
    for d in range(nd):
        r = nb[2 * d + 1, c]
        f = s.face_flux(c, r, d)
        b = s.face_flux(r, c, d)
        fe = euler_flux(q[:, c], d, nd) + euler_flux(q[:, r], d, nd)
        close(f + b, fe, 2e-15, scale=np.max(np.abs(f)))
        close(f - b, (h / dt) * (q[:, c] - q[:, r]), 1e-12, scale=np.max(np.abs(f)))
    # consistency on equal states
    q[:, :n] = q[:, [c]]
    s.set_lhs(q)
    for d in range(nd):
        close(s.face_flux(c, nb[2 * d + 1, c], d), euler_flux(q[:, c], d, nd), 4e-15)


# --------------------------------------------------------------------------- boundary conditions
def _bc_solver(nd, physics, kinds_params, level=3, xmin=None, xmax=None):
    xmin = [0.0] * nd if xmin is None else xmin
    xmax = [1.0] * nd if xmax is None else xmax
    g = O.Grid(nd, xmin, xmax).min_level(level)
    kw = dict(flux=O.FLUX_THREE_POINT) if physics == O.PHYS_HEAT else {}
    s = O.Solver(g, O.phys(physics=physics, nd=nd, **kw))
    s.append_cube_bcs(kinds_params)
    if physics == O.PHYS_HEAT:
        s.set_ic(O.IC_HEAT_CONSTANT, [1.0])
    else:
        s.set_ic(O.IC_CONSTANT_PV, [0.1] * nd + [1.0, 1.0])
    s.initialize()
    return g, s


def _randomize(s, nd, rng, heat=False):
    q = s.get("lhs")
    n = s.first_gc
    if heat:
        q[0, :n] = rng.uniform(0.5, 2.0, n)
    else:
        for c in range(n):
            q[:, c] = random_state(rng, nd, 0.4)
    s.set_lhs(q)
    return q


def _ghosts(s):
    """(ghost, boundary cell, position of the ghost w.r.t. the boundary cell, bc index)."""
    nb = s.get("neighbors")
    bc = s.get("bc_idx")[0]
    out = []
    for gc in range(s.first_gc, s.ncells):
        p = [k for k in range(nb.shape[0]) if nb[k, gc] != E]
        assert len(p) == 1
        out.append((gc, int(nb[p[0], gc]), p[0] ^ 1, int(bc[gc])))
    return out


def _dist(xc, a, b):
    return math.sqrt(np.sum((xc[:, a] - xc[:, b]) ** 2))


@pytest.mark.parametrize("nd", [2, 3])
def test_euler_dirichlet_and_neumann_ghosts(nd):
    """boundary_condition.hpp:70-77: Dirichlet ghosts make the face mean the
    surface value; Neumann ghosts make (Q_b - Q_g)/dist the surface flux."""
    rng = np.random.default_rng(41)
    nv = nd + 2
    vals = [list(rng.uniform(0.5, 1.5, nv)) for _ in range(2 * nd)]
    kp = [(O.BC_EULER_DIRICHLET, vals[i]) if i % 2 == 0 else (O.BC_EULER_NEUMANN, vals[i])
          for i in range(2 * nd)]
    _, s = _bc_solver(nd, O.PHYS_EULER, kp)
    _randomize(s, nd, rng)
    s.apply_bcs_lhs()
    q = s.get("lhs")
    xc = s.get("x_center")
    for gc, b, pos, k in _ghosts(s):
        v = np.array(vals[k])
        if k % 2 == 0:
            close(0.5 * (q[:, gc] + q[:, b]), v, 1e-15)
        else:
            close((q[:, b] - q[:, gc]) / _dist(xc, gc, b), v, 1e-13)


@pytest.mark.parametrize("nd", [2, 3])
def test_cns_external_and_wall_ghosts(nd):
    """cns/boundary_conditions.hpp:17-238 in primitive variables: inflow
    (face-mean u, rho = imposed; p mirrored), outflow (u, rho mirrored; face
    mean p = imposed), adiabatic no-slip (face mean u = 0), slip (normal u
    reflected), isothermal (face mean u = 0, face-mean rho = gamma p_b / T_w)."""
    rng = np.random.default_rng(43)
    pvD = list(rng.uniform(-0.3, 0.3, nd)) + [1.1, 0.9]
    Tw = 1.3
    kinds = [(O.BC_CNS_INFLOW, pvD), (O.BC_CNS_OUTFLOW, pvD), (O.BC_ADIABATIC_NOSLIP, []),
             (O.BC_ADIABATIC_SLIP, []), (O.BC_ISOTHERMAL_NOSLIP, [Tw]), (O.BC_ADIABATIC_NOSLIP, [])][: 2 * nd]
    if nd == 2:
        kinds = [(O.BC_CNS_INFLOW, pvD), (O.BC_CNS_OUTFLOW, pvD), (O.BC_ISOTHERMAL_NOSLIP, [Tw]),
                 (O.BC_ADIABATIC_SLIP, [])]
    _, s = _bc_solver(nd, O.PHYS_CNS, kinds)
    _randomize(s, nd, rng)
    s.apply_bcs_lhs()
    q = s.get("lhs")
    seen = set()
    for gc, b, pos, k in _ghosts(s):
        kind = kinds[k][0]
        seen.add(kind)
        pg, pb = pv_of(q[:, gc], nd), pv_of(q[:, b], nd)
        mean = 0.5 * (pg + pb)
        R, P = nd, nd + 1
        normal = pos // 2
        if kind == O.BC_CNS_INFLOW:
            close(mean[:nd], pvD[:nd], 1e-14, scale=1.0)
            assert mean[R] == pytest.approx(pvD[R], rel=1e-14)
            assert pg[P] == pytest.approx(pb[P], rel=1e-13)
        elif kind == O.BC_CNS_OUTFLOW:
            close(pg[:nd], pb[:nd], 1e-13, scale=1.0)
            assert pg[R] == pytest.approx(pb[R], rel=1e-14)
            assert mean[P] == pytest.approx(pvD[P], rel=1e-13)
        elif kind == O.BC_ADIABATIC_NOSLIP:
            close(mean[:nd], np.zeros(nd), 1e-14, scale=1.0)
            assert pg[R] == pytest.approx(pb[R], rel=1e-14)
            assert pg[P] == pytest.approx(pb[P], rel=1e-13)
        elif kind == O.BC_ADIABATIC_SLIP:
            assert mean[normal] == pytest.approx(0.0, abs=1e-14)
            for d in range(nd):
                if d != normal:
                    assert pg[d] == pytest.approx(pb[d], rel=1e-13, abs=1e-15)
            assert pg[R] == pytest.approx(pb[R], rel=1e-14)
            assert pg[P] == pytest.approx(pb[P], rel=1e-13)
        elif kind == O.BC_ISOTHERMAL_NOSLIP:
            close(mean[:nd], np.zeros(nd), 1e-14, scale=1.0)
            assert pg[P] == pytest.approx(pb[P], rel=1e-13)
            assert mean[R] == pytest.approx(G * pb[P] / Tw, rel=1e-14)
    assert len(seen) == len({k for k, _ in kinds})


@pytest.mark.parametrize("nd", [2, 3])
def test_heat_dirichlet_and_neumann_ghosts(nd):
    """heat/boundary_conditions.hpp:15-63."""
    rng = np.random.default_rng(47)
    vals = list(rng.uniform(0.5, 1.5, 2 * nd))
    kp = [(O.BC_HEAT_DIRICHLET, [vals[i]]) if i % 2 else (O.BC_HEAT_NEUMANN, [vals[i]]) for i in range(2 * nd)]
    _, s = _bc_solver(nd, O.PHYS_HEAT, kp)
    _randomize(s, nd, rng, heat=True)
    s.apply_bcs_lhs()
    q = s.get("lhs")[0]
    xc = s.get("x_center")
    for gc, b, pos, k in _ghosts(s):
        if k % 2:
            assert 0.5 * (q[gc] + q[b]) == pytest.approx(vals[k], rel=1e-15)
        else:
            assert (q[b] - q[gc]) / _dist(xc, gc, b) == pytest.approx(vals[k], rel=1e-13)


@pytest.mark.parametrize("nd", [2, 3])
def test_periodic_ghost_is_wrapped_image(nd):
    """The periodic condition's ghost holds the state of the internal cell at
    the ghost's position shifted by one root length."""
    rng = np.random.default_rng(53)
    _, s = _bc_solver(nd, O.PHYS_EULER, [O.BC_PERIODIC] * (2 * nd), level=3)
    _randomize(s, nd, rng)
    s.apply_bcs_lhs()
    q = s.get("lhs")
    xc = s.get("x_center")
    n = s.first_gc
    index = {tuple(np.round(xc[:, c] * 64).astype(int)): c for c in range(n)}
    for gc, b, pos, k in _ghosts(s):
        x = xc[:, gc].copy()
        d = pos // 2
        x[d] += -1.0 if pos % 2 else 1.0
        im = index[tuple(np.round(x * 64).astype(int))]
        assert s.periodic_image(gc) == im
        assert np.array_equal(q[:, gc], q[:, im])


@pytest.mark.parametrize("nd", [2, 3])
def test_wall_slope_hooks(nd):
    """boundary_condition.hpp:63-67 and cns/boundary_conditions.hpp:143-147,
    230-234: at a ghost the tangential slope is the boundary cell's central
    slope, with the velocity (momentum) slopes negated at no-slip walls."""
    rng = np.random.default_rng(59)
    kinds = [O.BC_ADIABATIC_NOSLIP, O.BC_ADIABATIC_SLIP, O.BC_ISOTHERMAL_NOSLIP, O.BC_EULER_NEUMANN,
             O.BC_ADIABATIC_NOSLIP, O.BC_ADIABATIC_SLIP][: 2 * nd]
    params = [[1.2] if k == O.BC_ISOTHERMAL_NOSLIP else [] for k in kinds]
    _, s = _bc_solver(nd, O.PHYS_CNS, list(zip(kinds, params)), level=3)
    _randomize(s, nd, rng)
    s.apply_bcs_lhs()
    q = s.get("lhs")
    nb = s.get("neighbors")
    h = s.get("length")[0, 0]
    checked = 0
    for gc, b, pos, k in _ghosts(s):
        for d in range(nd):
            if d == pos // 2:
                continue
            m, p = nb[2 * d, b], nb[2 * d + 1, b]
            if m == E or p == E or m >= s.first_gc or p >= s.first_gc:
                continue
            for v in range(nd + 2):
                central = (q[v, p] - q[v, m]) / (2 * h)
                flip = kinds[k] in (O.BC_ADIABATIC_NOSLIP, O.BC_ISOTHERMAL_NOSLIP) and v < nd
                assert s.slope(gc, v, d) == (-central if flip else central)
                checked += 1
    assert checked > 0


def test_vortex_boundary_is_analytic_state_at_solver_time():
    """euler/boundary_conditions.hpp:65-89: ghost and boundary cell take the
    vortex at their centres at the solver's time (the vortex moves with
    (1, 1): its value at (x, t) is the t = 0 value at x - t (1, 1))."""
    g = O.Grid(2, [-8.0, -8.0], [8.0, 8.0]).min_level(4)
    s = O.Solver(g, O.phys(nd=2))
    s.append_cube_bcs([O.BC_ISENTROPIC_VORTEX] * 4)
    s.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0])
    s.initialize()
    t = 0.37
    s.set_time(t, s.dt(), 3)
    s.apply_bcs_lhs()
    q = s.get("lhs")
    xc = s.get("x_center")
    for gc, b, pos, k in _ghosts(s):
        for c in (gc, b):
            assert np.array_equal(q[:, c], s.ic_point(xc[:, c] - t))


# --------------------------------------------------------------------------- conjugate interface
def _cht(nd=2, level=4, kratio=1.0):
    half = 4.0
    xmin, xmax = [-half] * nd, [half] * nd
    h = 2 * half / 2 ** level
    centre, dims = [-1.0] + [0.0] * (nd - 1), [2.0] * nd
    og = O.Grid(nd, xmin, xmax, max_solvers=2).min_level(level)
    cns = O.Solver(og, O.phys(physics=O.PHYS_CNS, nd=nd, cfl=0.4, cfl_viscous=0.4, Minfty=0.2, ReInfty=100.0,
                              Pr0=0.72), 0)
    heat = O.Solver(og, O.phys(physics=O.PHYS_HEAT, nd=nd, flux=O.FLUX_THREE_POINT, cfl=0.5), 1)
    for gg in O.cube_boundaries(nd, xmin, xmax):
        cns.append_bc(gg, O.BC_EULER_NEUMANN, [])
    cns.append_bc(O.geom(O.GEOM_SQUARE, centre, [d + h for d in dims]), O.BC_CHT_FLUID, [], other=heat,
                  kratio=kratio)
    heat.append_bc(O.geom(O.GEOM_SQUARE, centre, [d - 0.6 * h for d in dims], invert=True), O.BC_CHT_SOLID, [],
                   other=cns, kratio=kratio)
    cns.set_domain(O.geom(O.GEOM_SQUARE, centre, dims), +1)
    heat.set_domain(O.geom(O.GEOM_SQUARE, centre, dims), -1)
    cns.set_ic(O.IC_CONSTANT_PV, [0.1] + [0.0] * (nd - 1) + [1.0, 0.7])
    heat.set_ic(O.IC_HEAT_CONSTANT, [1.0])
    heat.initialize()
    cns.initialize()
    return og, cns, heat


@pytest.mark.parametrize("kratio", [1.0, 100.0])
def test_cht_interface_closed_forms(kratio):
    """cns/coupling_conditions.hpp:28-42, heat/coupling_conditions.hpp:24-37:
    the partner of a ghost is the other solver's cell at the ghost's centre;
    the surface temperature satisfies flux continuity
    (T_f - T_w) = kratio (T_w - T_s) (kratio 1: the reference's mean); the
    fluid ghost is an isothermal no-slip wall at T_w, the solid ghost a
    Dirichlet value T_w."""
    rng = np.random.default_rng(61)
    nd = 2
    _, cns, heat = _cht(nd, kratio=kratio)
    qf = _randomize(cns, nd, rng)
    qs = _randomize(heat, nd, rng, heat=True)
    cns.apply_bcs_lhs()
    heat.apply_bcs_lhs()
    qf2, qs2 = cns.get("lhs"), heat.get("lhs")[0]
    xf, xs = cns.get("x_center"), heat.get("x_center")
    nf, ns = cns.first_gc, heat.first_gc
    idx_s = {tuple(xs[:, c]): c for c in range(ns)}
    idx_f = {tuple(xf[:, c]): c for c in range(nf)}

    def wall_T(Tf, Ts):
        Tw = (Tf + kratio * Ts) / (1 + kratio)
        assert (Tf - Tw) == pytest.approx(kratio * (Tw - Ts), rel=1e-12, abs=1e-14)
        return Tw
    nchecked = 0
    for gc, b, pos, k in _ghosts(cns):
        if k != 2 * nd:
            continue
        partner = idx_s[tuple(xf[:, gc])]
        assert cns.cht_partner(gc) == partner
        pb = pv_of(qf[:, b], nd)
        Tw = wall_T(G * pb[nd + 1] / pb[nd], qs[0, partner])
        pg = pv_of(qf2[:, gc], nd)
        close(0.5 * (pg[:nd] + pb[:nd]), np.zeros(nd), 1e-14, scale=1.0)
        assert pg[nd + 1] == pytest.approx(pb[nd + 1], rel=1e-13)
        assert 0.5 * (pg[nd] + pb[nd]) == pytest.approx(G * pb[nd + 1] / Tw, rel=1e-14)
        nchecked += 1
    for gc, b, pos, k in _ghosts(heat):
        partner = idx_f[tuple(xs[:, gc])]
        pf = pv_of(qf[:, partner], nd)
        Tw = wall_T(G * pf[nd + 1] / pf[nd], qs[0, b])
        assert 0.5 * (qs2[gc] + qs2[b]) == pytest.approx(Tw, rel=1e-15)
        nchecked += 1
    assert nchecked > 16


def test_cht_time_scaling():
    """coupling.hpp:47-63: dt_heat = dt_cns Tr0 Pr0 Re0 with
    dt_cns = min(cns.min_dt, heat.min_dt / (Tr0 Pr0 Re0))."""
    for Tr0 in (0.01, 10.0):
        _, cns, heat = _cht()
        scale = Tr0 * 0.72 * cns.quantities()["Re0"]
        a, b = cns.min_dt(), heat.min_dt()
        O.couple_cns_heat(Tr0, cns, heat)
        cns.solve()
        heat.solve()
        assert cns.dt() == min(a, b / scale)
        assert heat.dt() == pytest.approx(cns.dt() * scale, rel=1e-15)


def test_euler_coupling_ghost_takes_partner_state():
    """euler/coupling_conditions.hpp:24-31 and coupling.hpp:66-73: a
    coupling ghost holds the other solver's state at the ghost's centre; both
    solvers step with the smaller of their time steps."""
    rng = np.random.default_rng(67)
    nd, level = 3, 3
    h = 1.0 / 2 ** level
    c, L = [0.5] * 3, 0.25
    og = O.Grid(nd, [0.0] * 3, [1.0] * 3, max_solvers=2).min_level(level)
    p = O.phys(physics=O.PHYS_EULER, nd=nd, cfl=0.2)
    s0, s1 = O.Solver(og, p, 0), O.Solver(og, p, 1)
    s0.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    s0.append_bc(O.geom(O.GEOM_SQUARE, c, [L + h] * 3), O.BC_EULER_COUPLING, [], other=s1)
    s1.append_bc(O.geom(O.GEOM_SQUARE, c, [L - 0.5 * h] * 3, invert=True), O.BC_EULER_COUPLING, [], other=s0)
    s0.set_domain(O.geom(O.GEOM_SQUARE, c, [L] * 3), +1)
    s1.set_domain(O.geom(O.GEOM_SQUARE, c, [L] * 3), -1)
    for s in (s0, s1):
        s.set_ic(O.IC_CONSTANT_PV, [0.1, 0.0, 0.0, 1.0, 1.0])
        s.initialize()
    q0 = _randomize(s0, nd, rng)
    q1 = _randomize(s1, nd, rng)
    a, b = s0.min_dt(), s1.min_dt()
    s0.apply_bcs_lhs()
    s1.apply_bcs_lhs()
    for s, q, o, qo in ((s0, s0.get("lhs"), s1, q1), (s1, s1.get("lhs"), s0, q0)):
        xs, xo = s.get("x_center"), o.get("x_center")
        idx = {tuple(xo[:, k]): k for k in range(o.first_gc)}
        n = 0
        for gc, bb, pos, k in _ghosts(s):
            if s is s0 and k < 6:
                continue
            assert np.array_equal(q[:, gc], qo[:, idx[tuple(xs[:, gc])]])
            n += 1
        assert n > 0
    O.couple_same(s0, s1)
    s0.solve()
    s1.solve()
    assert s0.dt() == s1.dt() == min(a, b)


# --------------------------------------------------------------------------- time integration
@pytest.mark.parametrize("nd", [2, 3])
@pytest.mark.parametrize("ti,order", [(O.TI_EULER_FWD, 1), (O.TI_HEUN_RK2, 2), (O.TI_SSP_RK3, 3)])
def test_rk_amplification_factor(nd, ti, order):
    """solver.hpp:469-494 (+ SSP-RK3): on a discrete eigenmode of the
    periodic heat operator with eigenvalue z, one step multiplies the mode
    by the scheme's stability polynomial sum_{k<=order} z^k / k!."""
    level = 4 if nd == 2 else 3
    g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(level)
    s = O.Solver(g, O.phys(physics=O.PHYS_HEAT, nd=nd, flux=O.FLUX_THREE_POINT, ti=ti, cfl=0.3))
    s.append_cube_bcs([O.BC_PERIODIC] * (2 * nd))
    k = 2 * math.pi * 2
    s.set_ic(O.IC_HEAT_SINE, [0.0, 1.0, k, 0.0])
    s.initialize()
    T0 = s.get("lhs")[0, : s.first_gc].copy()
    dt = s.dt()
    s.solve(1)
    h = 1.0 / 2 ** level
    z = -4 * nd * (dt / h ** 2) * math.sin(k * h / 2) ** 2
    R = sum(z ** j / math.factorial(j) for j in range(order + 1))
    np.testing.assert_allclose(s.get("lhs")[0, : s.first_gc], R * T0, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("ti", [O.TI_EULER_FWD, O.TI_SSP_RK3])
def test_source_term_integrates_exactly(ti):
    """A uniform temperature with a uniform volumetric source S gains dt S
    per step (the heat physics returns dt S; every stage adds it)."""
    g = O.Grid(2, [0.0, 0.0], [1.0, 1.0]).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_HEAT, nd=2, flux=O.FLUX_THREE_POINT, ti=ti, cfl=0.3))
    s.append_cube_bcs([O.BC_PERIODIC] * 4)
    s.set_ic(O.IC_HEAT_CONSTANT, [1.25])
    s.initialize()
    s.set_source(np.full(s.first_gc, 0.5))
    s.solve(1)
    np.testing.assert_allclose(s.get("lhs")[0, : s.first_gc], 1.25 + 0.5 * s.dt(), rtol=1e-15)


# --------------------------------------------------------------------------- time step
@pytest.mark.parametrize("nd", [2, 3])
def test_euler_cfl_number(nd):
    """euler/physics.hpp:182-196: dt max_d(|u_d| + a) / h = CFL in every cell;
    the solver's dt is the minimum (solver.hpp:511-517)."""
    rng = np.random.default_rng(71)
    _, s = _bc_solver(nd, O.PHYS_EULER, [O.BC_EULER_NEUMANN] * (2 * nd))
    q = _randomize(s, nd, rng)
    h = s.get("length")[0, 0]
    dts = []
    for c in range(0, s.first_gc, 3):
        pv = pv_of(q[:, c], nd)
        a = math.sqrt(G * pv[nd + 1] / pv[nd])
        dt = s.cell_dt(c)
        assert dt * np.max(np.abs(pv[:nd]) + a) / h == pytest.approx(0.2, rel=1e-15)
    assert s.compute_dt() == min(s.cell_dt(c) for c in range(s.first_gc))


@pytest.mark.parametrize("reinf", [0.2, 1e4])
def test_cns_cfl_viscous_bound(reinf):
    """cns/physics.hpp:167-187: dt = min(dt_inviscid, CFL_v h / (2 S_v)),
    S_v = max(mu/(rho h Re0), mu/(rho h Re0 Pr0 (gamma-1))), mu by Sutherland's
    law at T = gamma p / rho; both branches (viscous- and inviscid-limited)."""
    rng = np.random.default_rng(73)
    nd = 3
    g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(2)
    s = O.Solver(g, O.phys(physics=O.PHYS_CNS, nd=nd, cfl=0.5, cfl_viscous=0.3, ReInfty=reinf, Pr0=0.71))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    s.set_ic(O.IC_CONSTANT_PV, [0.1, 0.0, 0.0, 1.0, 1.0])
    s.initialize()
    q = _randomize(s, nd, rng)
    qt = s.quantities()
    h = s.get("length")[0, 0]
    S0 = 110.4 / 273.15
    branches = set()
    for c in range(s.first_gc):
        pv = pv_of(q[:, c], nd)
        rho, p = pv[nd], pv[nd + 1]
        a = math.sqrt(G * p / rho)
        T = G * p / rho
        mu = T ** 1.5 * (1 + S0) / (T + S0)
        dt_i = 0.5 * h / np.max(np.abs(pv[:nd]) + a)
        Sc = mu / (rho * h)
        Sv = max(Sc / qt["Re0"], Sc / (qt["Re0"] * 0.71 * 0.4))
        dt_v = 0.3 * h / (2 * Sv)
        assert s.cell_dt(c) == pytest.approx(min(dt_i, dt_v), rel=1e-13)
        branches.add(dt_v < dt_i)
    assert branches == {reinf < 100}


def test_heat_cfl():
    """heat/physics.hpp:82-85: dt = CFL h^2 / 2 (unit diffusivity)."""
    g = O.Grid(3, [0.0] * 3, [2.0] * 3).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_HEAT, nd=3, flux=O.FLUX_THREE_POINT, cfl=0.4))
    s.append_cube_bcs([O.BC_HEAT_NEUMANN] * 6)
    s.set_ic(O.IC_HEAT_CONSTANT, [1.0])
    s.initialize()
    assert s.cell_dt(5) == 0.4 * 0.5 * 0.25 ** 2


# --------------------------------------------------------------------------- viscous flux
@pytest.mark.parametrize("nd", [2, 3])
def test_viscous_flux_of_linear_fields(nd):
    """cns/physics.hpp:101-123, 221-338 on u = u0 + G x, uniform rho, linear
    p: the viscous part of the face flux (CNS flux minus the Euler flux of
    the same states) is tau_{., dir}/Re0 in momentum and
    (u_s . tau_{., dir} + q)/Re0 in energy, with the Newtonian stress
    tau = -mu (G + G^T - 2/3 tr(G) I) and Fourier's law
    q = -mu / ((gamma - 1) Pr0) dT/dx_dir, mu by Sutherland's law at the face."""
    rng = np.random.default_rng(79 + nd)
    level = 3
    Gm = rng.uniform(-0.4, 0.4, (nd, nd))
    u0 = rng.uniform(-0.2, 0.2, nd)
    rho = 1.1
    p0, gp = 0.9, rng.uniform(-0.3, 0.3, nd)
    pr = 0.71

    def solver(physics):
        g = O.Grid(nd, [0.0] * nd, [1.0] * nd).min_level(level)
        s = O.Solver(g, O.phys(physics=physics, nd=nd, Minfty=0.2, ReInfty=50.0, Pr0=pr))
        s.append_cube_bcs([O.BC_EULER_NEUMANN] * (2 * nd))
        s.set_ic(O.IC_CONSTANT_PV, [0.0] * nd + [1.0, 1.0])
        s.initialize()
        xc = s.get("x_center")
        q = s.get("lhs")
        n = s.first_gc
        x = xc[:, :n]
        u = u0[:, None] + Gm @ x
        p = p0 + gp @ x
        q[:, :n] = cv_of(np.vstack([u, np.full(n, rho), p]), nd)
        s.set_lhs(q)
        return s
    cns, eul = solver(O.PHYS_CNS), solver(O.PHYS_EULER)
    re0 = cns.quantities()["Re0"]
    nb = cns.get("neighbors")
    xc = cns.get("x_center")
    S0 = 110.4 / 273.15
    c = O.morton_encode(nd, [3] * nd)
    for d in range(nd):
        r = nb[2 * d + 1, c]
        fv = cns.face_flux(c, r, d) - eul.face_flux(c, r, d)
        xf = 0.5 * (xc[:, c] + xc[:, r])
        us = u0 + Gm @ xf
        T = G * (p0 + gp @ xf) / rho
        mu = T ** 1.5 * (1 + S0) / (T + S0)
        tau = -mu * (Gm + Gm.T - 2.0 / 3.0 * np.trace(Gm) * np.eye(nd))
        qh = -mu / ((G - 1) * pr) * (G * gp[d] / rho)
        np.testing.assert_allclose(fv[:nd], tau[:, d] / re0, rtol=1e-10, atol=1e-14)
        assert fv[nd] == pytest.approx(0.0, abs=1e-15)
        assert fv[nd + 1] == pytest.approx((us @ tau[:, d] + qh) / re0, rel=1e-10, abs=1e-14)


# --------------------------------------------------------------------------- MUSCL extension
@pytest.mark.parametrize("lim", [O.LIM_MINMOD, O.LIM_VANLEER])
def test_muscl_linear_data_reconstructs_face_value(lim):
    """On linear conserved data both limiters keep the slope, so both face
    states equal the value at the face and the flux is F_Euler(Q_face)."""
    nd = 2
    g = O.Grid(nd, [0.0, 0.0], [1.0, 1.0]).min_level(4)
    s = O.Solver(g, O.phys(nd=nd, recon=O.RECON_MUSCL, limiter=lim))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 4)
    s.set_ic(O.IC_CONSTANT_PV, [0.1, 0.0, 1.0, 1.0])
    s.initialize()
    q = s.get("lhs")
    xc = s.get("x_center")
    n = s.first_gc
    A = np.array([[0.2, 0.1], [-0.1, 0.05], [0.3, 0.2], [0.5, -0.4]])
    b = np.array([0.1, -0.05, 1.0, 3.0])
    q[:, :n] = b[:, None] + A @ xc[:, :n]
    s.set_lhs(q)
    nb = s.get("neighbors")
    c = O.morton_encode(2, [6, 9])
    for d in range(2):
        r = nb[2 * d + 1, c]
        qf = b + A @ (0.5 * (xc[:, c] + xc[:, r]))
        close(s.face_flux(c, r, d), euler_flux(qf, d, 2), 1e-14)


@pytest.mark.parametrize("lim", [O.LIM_MINMOD, O.LIM_VANLEER])
def test_muscl_limited_slopes(lim):
    """Face states q_l + s_l/2 and q_r - s_r/2 with the textbook limited
    slopes: minmod(a, b), or van Leer's phi(r) b, phi(r) = (r + |r|)/(1 + |r|),
    r = a / b; zero slope at an extremum (a b <= 0)."""
    nd = 2
    g = O.Grid(nd, [0.0, 0.0], [1.0, 1.0]).min_level(4)
    s = O.Solver(g, O.phys(nd=nd, recon=O.RECON_MUSCL, limiter=lim))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 4)
    s.set_ic(O.IC_CONSTANT_PV, [0.1, 0.0, 1.0, 1.0])
    s.initialize()
    rng = np.random.default_rng(83)
    q = _randomize(s, nd, rng)
    nb = s.get("neighbors")

    def slope(a, b):
        if a * b <= 0:
            return 0.0
        if lim == O.LIM_MINMOD:
            return a if abs(a) < abs(b) else b
        r = a / b
        return (r + abs(r)) / (1 + abs(r)) * b
    for c in (O.morton_encode(2, [5, 7]), O.morton_encode(2, [9, 4])):
        for d in range(2):
            r = nb[2 * d + 1, c]
            ll, rr = nb[2 * d, c], nb[2 * d + 1, r]
            ql = np.array([q[v, c] + 0.5 * slope(q[v, c] - q[v, ll], q[v, r] - q[v, c]) for v in range(4)])
            qr = np.array([q[v, r] - 0.5 * slope(q[v, r] - q[v, c], q[v, rr] - q[v, r]) for v in range(4)])
            close(s.face_flux(c, r, d), s.ausm_states(ql, qr, d), 1e-13)


# --------------------------------------------------------------------------- initial conditions
def _vortex_pv(s, x):
    return pv_of(s.ic_point(x), 2)


def test_isentropic_vortex_point_values():
    """euler/initial_conditions.hpp:14-46 (beta = 5, centre (-5, -5), free
    stream rho = p = u = v = 1): far field is the free stream; p = rho^gamma
    (isentropic); the velocity perturbation is azimuthal with
    u_theta(1) = beta / (2 pi) and u_theta(r) / r ~ exp(1 - r^2); the
    pressure is in radial equilibrium dp/dr = rho u_theta^2 / r."""
    g = O.Grid(2, [-10.0, -10.0], [0.0, 0.0]).min_level(2)
    s = O.Solver(g, O.phys(nd=2))
    s.append_cube_bcs([O.BC_PERIODIC] * 4)
    s.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0])
    s.initialize()
    x0 = np.array([-5.0, -5.0])
    far = _vortex_pv(s, x0 + [30.0, -20.0])
    np.testing.assert_allclose(far, [1, 1, 1, 1], rtol=0, atol=1e-12)
    for r in (0.3, 0.7, 1.0, 1.6, 2.4):
        for th in (0.3, 2.1, 4.4):
            e = np.array([math.cos(th), math.sin(th)])
            pv = _vortex_pv(s, x0 + r * e)
            assert pv[3] == pytest.approx(pv[2] ** G, rel=1e-14)
            du = pv[:2] - 1.0
            assert du @ e == pytest.approx(0.0, abs=1e-14)
            ut = du @ np.array([-e[1], e[0]])
            assert ut == pytest.approx(5 / (2 * math.pi) * r * math.exp(1 - r * r), rel=1e-13)
            hh = 1e-5
            dp = (_vortex_pv(s, x0 + (r + hh) * e)[3] - _vortex_pv(s, x0 + (r - hh) * e)[3]) / (2 * hh)
            assert dp == pytest.approx(pv[2] * ut * ut / r, rel=1e-7)


def test_isentropic_vortex_moves_with_the_free_stream():
    g = O.Grid(2, [-10.0, -10.0], [0.0, 0.0]).min_level(2)
    s0 = O.Solver(g, O.phys(nd=2))
    s1 = O.Solver(g, O.phys(nd=2))
    for s, t in ((s0, 0.0), (s1, 1.25)):
        s.append_cube_bcs([O.BC_PERIODIC] * 4)
        s.set_ic(O.IC_ISENTROPIC_VORTEX, [t])
    for x in ([-4.0, -6.5], [-5.2, -3.0], [-1.0, -9.0]):
        x = np.array(x)
        assert np.array_equal(s1.ic_point(x), s0.ic_point(x - 1.25))


def test_vortex_centre_parameter_translates():
    """Config form of the vortex: centre (5, 5) on [0, 10]^2 is the reference
    vortex translated by (10, 10)."""
    g = O.Grid(2, [0.0, 0.0], [10.0, 10.0]).min_level(2)
    a = O.Solver(g, O.phys(nd=2))
    b = O.Solver(g, O.phys(nd=2))
    a.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0, 5.0, 5.0])
    b.set_ic(O.IC_ISENTROPIC_VORTEX, [0.0])
    for x in ([4.0, 6.5], [5.5, 3.0], [1.0, 9.0]):
        np.testing.assert_allclose(a.ic_point(np.array(x)), b.ic_point(np.array(x) - 10.0), rtol=1e-14)


def test_taylor_green_point_values():
    """Taylor-Green initial field: u = V0 sin x cos y cos z,
    v = -V0 cos x sin y cos z, w = 0, rho = rho_inf,
    p = p_inf + rho V0^2/16 (cos 2x + cos 2y)(cos 2z + 2).  The velocity is
    divergence-free and p solves the incompressible pressure Poisson
    equation  lap p = -rho d_i u_j d_j u_i."""
    V0, rinf, pinf = 0.0995, 0.995, 0.71
    g = O.Grid(3, [0.0] * 3, [2 * math.pi] * 3).min_level(1)
    s = O.Solver(g, O.phys(physics=O.PHYS_CNS, nd=3))
    s.set_ic(O.IC_TGV, [V0, rinf, pinf])
    pv = lambda x: pv_of(s.ic_point(np.asarray(x, float)), 3)
    assert np.allclose(pv([math.pi / 2, 0, 0]), [V0, 0, 0, rinf, pinf], rtol=1e-14, atol=1e-17)
    assert pv([0, 0, 0])[4] == pytest.approx(pinf + rinf * V0 ** 2 / 16 * 2 * 3, rel=1e-14)
    h = 1e-3
    rng = np.random.default_rng(89)
    for _ in range(5):
        x = rng.uniform(0, 2 * math.pi, 3)
        J = np.zeros((3, 3))
        lap = 0.0
        for i in range(3):
            e = np.zeros(3)
            e[i] = h
            J[:, i] = (pv(x + e)[:3] - pv(x - e)[:3]) / (2 * h)
            lap += (pv(x + e)[4] - 2 * pv(x)[4] + pv(x - e)[4]) / h ** 2
        assert abs(np.trace(J)) < 1e-9
        rhs = -rinf * np.sum(J * J.T)
        assert lap == pytest.approx(rhs, rel=1e-4, abs=1e-9)


@pytest.mark.parametrize("nd", [2, 3])
def test_gauss_average_exact_for_cubics(nd):
    """quadrature.hpp:14-74: the 2^nd-point Gauss rule (+-1/sqrt 3) integrates
    per-axis cubics exactly: the cell average of x^3 + x y^2 (+ z^3) equals
    its closed form x_c^3 + x_c h^2/4 + x_c y_c^2 + x_c h^2/12 (+ z_c^3 + z_c h^2/4)."""
    g = O.Grid(nd, [-1.0] * nd, [2.0] * nd).min_level(3)
    s = O.Solver(g, O.phys(physics=O.PHYS_HEAT, nd=nd, flux=O.FLUX_THREE_POINT))
    s.append_cube_bcs([O.BC_HEAT_NEUMANN] * (2 * nd))
    f = (lambda x: [x[0] ** 3 + x[0] * x[1] ** 2]) if nd == 2 else \
        (lambda x: [x[0] ** 3 + x[0] * x[1] ** 2 + x[2] ** 3])
    s.set_ic(O.IC_CALLBACK, [], callback=f)
    s.initialize()
    xc = s.get("x_center")[:, : s.first_gc]
    h = s.get("length")[0, 0]
    T = s.get("lhs")[0, : s.first_gc]
    x, y = xc[0], xc[1]
    exact = x ** 3 + x * h * h / 4 + x * (y * y + h * h / 12)
    if nd == 3:
        z = xc[2]
        exact = exact + z ** 3 + z * h * h / 4
    np.testing.assert_allclose(T, exact, rtol=1e-13, atol=1e-14)


# --------------------------------------------------------------------------- sums
def test_conservation_sums_closed_form():
    """Sum of cell volume x conserved variable of a uniform state is the
    state times the domain volume."""
    g = O.Grid(3, [0.0] * 3, [2.0] * 3).min_level(3)
    s = O.Solver(g, O.phys(nd=3))
    s.append_cube_bcs([O.BC_EULER_NEUMANN] * 6)
    pv = [0.3, -0.2, 0.1, 1.2, 0.8]
    s.set_ic(O.IC_CONSTANT_PV, pv)
    s.initialize()
    np.testing.assert_allclose(s.conservation_sums(), 8.0 * cv_of(np.array(pv), 3), rtol=1e-13)
