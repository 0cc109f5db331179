"""Mutation check of the oracle pins (CPU only): each entry plants one
plausible mistake in a private copy of the oracle's solver (a coefficient,
sign, index or term), rebuilds it, and requires the pin tests
(test_oracle_pins.py, test_oracle_kats.py) to fail on it.  A pin set that
let any of these through would not pin the oracle."""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MUTS = [
 ("ausm m_int coefficient", "s * 0.25 * std::pow(M + s, 2.)", "s * 0.3 * std::pow(M + s, 2.)"),
 ("ausm p_int sign", "(2. - s * M)", "(2. + s * M)"),
 ("ausm swapped MR/ML", "const Num m_i = m_int(+1, ML) + m_int(-1, MR);", "const Num m_i = m_int(+1, MR) + m_int(-1, ML);"),
 ("ausm upwind choice", "const Num* q = m_i >= 0 ? QL : QR;", "const Num* q = m_i >= 0 ? QR : QL;"),
 ("ausm energy theta", "f[nd + 1] = (q[nd + 1] + pc) * ac;", "f[nd + 1] = (q[nd + 1]) * ac;"),
 ("llf dissipation", "dx / dt * (Q(Qb, l, v) - Q(Qb, r, v))", "dx / dt * (Q(Qb, r, v) - Q(Qb, l, v))"),
 ("dirichlet", "Q(buf, g, v) = 2.0 * pp[v] - Q(buf, b, v);", "Q(buf, g, v) = pp[v] - Q(buf, b, v);"),
 ("neumann sign", "Q(buf, g, v) = Q(buf, b, v) - dx * gv;", "Q(buf, g, v) = Q(buf, b, v) + dx * gv;"),
 ("inflow p", "pvg[PP] = pvb[PP] - 0 * 0;\n          } else if (bc.kind == BC_CNS_OUTFLOW)", "pvg[PP] = 2.0 * pp[PP] - pvb[PP];\n          } else if (bc.kind == BC_CNS_OUTFLOW)"),
 ("isothermal rho", "const Num rho_srfc = p_srfc * qty.gamma / Ts;", "const Num rho_srfc = p_srfc / Ts;"),
 ("slip normal", "pvg[normal] = 2.0 * 0 - pvb[normal];", "pvg[normal] = pvb[normal];"),
 ("wall hook", "return v < nd ? -tmp : tmp;", "return tmp;"),
 ("cht kratio", "return (T_fluid + bc.kratio * T_solid) / (1.0 + bc.kratio);", "return (bc.kratio * T_fluid + T_solid) / (1.0 + bc.kratio);"),
 ("cht solid ghost", "Q(buf, g, 0) = 2.0 * cht_surface_T(g, bc) - Q(buf, b, 0);", "Q(buf, g, 0) = cht_surface_T(g, bc);"),
 ("cht scaling", "const Num scaling = Tr0 * cns.qty.Pr0 * cns.qty.Re0;\n  const Num min_dt_cns", "const Num scaling = Tr0 * cns.qty.Re0;\n  const Num min_dt_cns"),
 ("heun", "Q(L, c, v) = 0.5 * (Q(L, c, v) + Q(R, c, v) + nf[v]);", "Q(L, c, v) = 0.5 * (Q(L, c, v) + Q(R, c, v)) + nf[v];"),
 ("rk3 stage2", "0.75 * Q(L, c, v) + 0.25 * (Q(R, c, v) + nf[v] + src[v]);", "0.5 * Q(L, c, v) + 0.5 * (Q(R, c, v) + nf[v] + src[v]);"),
 ("rk3 stage3", "(1. / 3.) * Q(L, c, v) + (2. / 3.) * (Q(S, c, v) + nf[v] + src[v]);", "(1. / 3.) * Q(L, c, v) + (2. / 3.) * (Q(S, c, v) + nf[v]);"),
 ("vortex exponent", "std::exp(2 * (1 - std::pow(r, 2)))", "std::exp(1 - std::pow(r, 2))"),
 ("vortex sign", "pv[1] = v + beta * eT1mr2 * (x_t - x0[0]) / (2 * kPi);", "pv[1] = v - beta * eT1mr2 * (x_t - x0[0]) / (2 * kPi);"),
 ("tgv p coeff", "rinf * V0 * V0 / 16.0", "rinf * V0 * V0 / 8.0"),
 ("viscous 2/3", "const Num div_u = esum(diag, nd) * 2. / 3.;", "const Num div_u = esum(diag, nd) * 1. / 3.;"),
 ("viscous transpose", "nab[i][j] = grad_u_srfc(j, i);", "nab[i][j] = grad_u_srfc(i, i);"),
 ("heat flux Pr", "(qty.gammaM1() * qty.Pr0 * rho_s)", "(qty.gammaM1() * rho_s)"),
 ("cns dt", "S_coeff / (qty.Re0 * qty.Pr0 * qty.gammaM1())", "S_coeff / (qty.Re0 * qty.gammaM1())"),
 ("euler dt", "std::abs(u(Qb, c, d)) + a_", "std::abs(u(Qb, c, d))"),
 ("heat dt", "return P.cfl * 0.5 * std::pow(h, 2.);", "return P.cfl * std::pow(h, 2.);"),
 ("muscl half", "QL[v] = ql + 0.5 * sl;", "QL[v] = ql + sl;"),
 ("vanleer", "return 2.0 * a_ * b_ / (a_ + b_);", "return a_ * b_ / (a_ + b_);"),
 ("source", "if (P.physics == PHYS_HEAT && !source.empty()) out[0] = dt_ * source[c];", "if (P.physics == PHYS_HEAT && !source.empty()) out[0] = source[c];"),
 ("gauss point", "const Num gpv = 1 / std::sqrt(3);", "const Num gpv = 0.5;"),
 ("periodic image", "if (!valid(n) || is_fine_ref(n) || is_ghost(n)) return cur;", "if (!valid(n) || is_fine_ref(n) || is_ghost(n)) return nghbr(cur, dir ^ 1) >= 0 && valid(nghbr(cur, dir^1)) && !is_ghost(nghbr(cur, dir^1)) ? nghbr(cur, dir ^ 1) : cur;"),
 ("euler coupling", "for (int v = 0; v < nvars; ++v) Q(buf, g, v) = o.Q(ob, b1, v);", "for (int v = 0; v < nvars; ++v) Q(buf, g, v) = Q(buf, b, v);"),
 ("vortex bc time", "isentropic_vortex(xg, time_, cx, cy, qg);", "isentropic_vortex(xg, 0, cx, cy, qg);"),
]


def _run(mut):
    name, old, new = mut
    d = tempfile.mkdtemp(prefix="orcmut_")
    try:
        shutil.copytree(os.path.join(ROOT, "oracle"), os.path.join(d, "oracle"),
                        ignore=shutil.ignore_patterns("_build", "__pycache__"))
        os.makedirs(os.path.join(d, "tests"))
        for f in ("conftest.py", "test_oracle_pins.py", "test_oracle_kats.py"):
            shutil.copy(os.path.join(ROOT, "tests", f), os.path.join(d, "tests", f))
        shutil.copy(os.path.join(ROOT, "pytest.ini"), d)
        p = os.path.join(d, "oracle", "solver.cpp")
        s = open(p).read()
        if old not in s:
            return name, "pattern not found"
        open(p, "w").write(s.replace(old, new, 1))
        r = subprocess.run(["make", "-s", "-C", os.path.join(d, "oracle"), "_build/liboracle.so"],
                           capture_output=True, text=True)
        if r.returncode:
            return name, "build failed: " + r.stderr[-300:]
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "tests/test_oracle_pins.py",
                            "tests/test_oracle_kats.py"], capture_output=True, text=True, cwd=d)
        return name, ("caught" if r.returncode == 1 else "missed (rc %d)" % r.returncode)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_every_planted_mistake_fails_a_pin():
    with ThreadPoolExecutor(max_workers=6) as ex:
        res = dict(ex.map(_run, MUTS))
    bad = {k: v for k, v in res.items() if v != "caught"}
    assert not bad, bad
