"""Mutation tests of the oracle pins.

Each entry plants one plausible mistake -- a dropped term, a wrong sign, factor, argument or
time -- in a copy of oracle/tmo.c, builds it on the CPU with the oracle's own flags, and runs
the pins of tests/test_oracle_pins.py that cover that part of the path against it (the copy
is loaded through TMO_LIB).  Every mutant must fail at least one of them; an unmodified copy
built the same way must pass every group.  Nothing here touches a GPU.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
PINS = os.path.join(ROOT, "tests", "test_oracle_pins.py")
CFLAGS = ["-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fopenmp", "-shared"]

# (name, text in tmo.c, replacement, pins expected to catch it (pytest -k expression), reference)
MUTANTS = [
    ("no_transverse_correction",
     "const double scale = -(1.0 / 2.0) * dt / p->dx;", "const double scale = 0.0 * dt / p->dx;",
     "oblique", "LinElastPatchGodunov.cpp:397-452"),
    ("transverse_correction_sign",
     "const double scale = -(1.0 / 2.0) * dt / p->dx;", "const double scale = (1.0 / 2.0) * dt / p->dx;",
     "oblique", "LinElastPatchGodunov.cpp:410-416"),
    ("slip_law_rate_V0_over_L",
     "dPsi = psi - dt * (V / L) * (f - fss);", "dPsi = psi - dt * (p->rs_V0 / L) * (f - fss);",
     "uniform_fault or closed_form", "RSIBCF.ChF:126-132"),
    ("aging_law_without_b",
     "dPsi = psi + dt * (b * p->rs_V0 / L)", "dPsi = psi + dt * (p->rs_V0 / L)",
     "uniform_fault or closed_form", "aging-law extension"),
    ("predictor_fault_update_over_dt",
     "update_boundary(&ibc, &WHalf1[d2], d2, (1.0 / 2.0) * dt, 0.0, 0);",
     "update_boundary(&ibc, &WHalf1[d2], d2, dt, 0.0, 0);",
     "time_centred", "LinElastPatchGodunov.cpp:419-420, LinElastPhysics.cpp:386-395"),
    ("rupture_stamp_at_step_start",
     "update_boundary(&ibc, &WHalf[d], d, dt, time + dt, 1);",
     "update_boundary(&ibc, &WHalf[d], d, dt, time, 1);",
     "rupture_time_is", "LinElastPhysics.cpp:174, RSIBCF.ChF:116-122"),
    ("vanleer_limiter_minmod",
     "double lim = fmin(2.0 * fabs(dWl), 2.0 * fabs(dWr));", "double lim = fmin(fabs(dWl), fabs(dWr));",
     "vanleer or monotone", "Chombo GodunovUtilities::vanLeerSlopes"),
    ("ppm_limiter_factor_3",
     "copysign(fmin(2.0 * fabs(pp), fabs(m)), m);", "copysign(fmin(3.0 * fabs(pp), fabs(m)), m);",
     "ppm_limiter", "Chombo GodunovUtilities::PPMLimiter"),
    ("fourth_order_slope_coefficient",
     "(2.0 / 3.0) * (hi - lo);", "(1.0 / 2.0) * (hi - lo);",
     "fourth_order", "Chombo GodunovUtilities::fourthOrderSlopes"),
    ("ppm_tracing_courant_term",
     "*AT(dWm, i, j, c) = m + sm * ((pp - m) - (pp + m) * (3.0 - 2.0 * sm)) / 2.0;",
     "*AT(dWm, i, j, c) = m + sm * ((pp - m) - (pp + m) * (3.0 - 1.0 * sm)) / 2.0;",
     "ppm_tracing", "Chombo GodunovUtilities::PPMNormalPred"),
    ("boundary_velocity_impedance",
     "*AT(Wg, i, j, o + q->cx) = (wmx - wpx) / (2.0 * Zp);",
     "*AT(Wg, i, j, o + q->cx) = (wmx - wpx) / (2.0 * Zs);",
     "free_surface or absorbing", "LinElastPhysicsF.ChF:968-1014, :1101-1140"),
    ("drucker_prager_pressure_sign",
     "const double Y = tau - fmax(-p->p_mu * sigma + bb, 0.0);",
     "const double Y = tau - fmax(p->p_mu * sigma + bb, 0.0);",
     "plastic", "PlasticPhysicsF.ChF:157-181"),
    ("slip_weakening_strength_max",
     "const double tau = fmin(tau_lock, -sn_syy * fric) - tau_0;",
     "const double tau = fmax(tau_lock, -sn_syy * fric) - tau_0;",
     "sw_fault", "SWIBCF.ChF:211"),
]


def _build(d, name, text):
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "tmo.c"), "w") as f:
        f.write(text)
    shutil.copy(os.path.join(ORACLE, "tmo.h"), d)
    so = os.path.join(d, "libtmo_%s.so" % name)
    r = subprocess.run(["gcc"] + CFLAGS + ["-o", so, os.path.join(d, "tmo.c"), "-lm"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    return so


def _pins(so, kexpr):
    env = dict(os.environ, TMO_LIB=so, OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", PINS, "-q", "-x", "-p", "no:cacheprovider", "-k", kexpr],
                       env=env, capture_output=True, text=True, cwd=ROOT, timeout=900)
    return r.returncode, r.stdout[-2000:]


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    base = str(tmp_path_factory.mktemp("oracle_mutants"))
    src = open(os.path.join(ORACLE, "tmo.c")).read()
    jobs = {}
    for name, old, new, kexpr, _ in MUTANTS:
        assert src.count(old) == 1, (name, old)
        jobs[name] = (src.replace(old, new), kexpr)
    jobs["control"] = (src, " or ".join("(%s)" % m[3] for m in MUTANTS))
    with ThreadPoolExecutor(max_workers=4) as ex:
        sos = dict(zip(jobs, ex.map(lambda n: _build(os.path.join(base, n), n, jobs[n][0]), list(jobs))))
        out = dict(zip(jobs, ex.map(lambda n: _pins(sos[n], jobs[n][1]), list(jobs))))
    return out


def test_unmodified_copy_passes_every_pin_group(results):
    rc, tail = results["control"]
    assert rc == 0, tail


@pytest.mark.parametrize("name", [m[0] for m in MUTANTS])
def test_planted_mistake_fails_a_pin(results, name):
    rc, tail = results[name]
    assert rc == 1, "mutant %s: pins exit %d (1 = a pin failed)\n%s" % (name, rc, tail)
