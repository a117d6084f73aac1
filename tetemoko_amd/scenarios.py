"""Seeded rupture scenarios and level arrays.

Units follow the reference decks in metres, MPa and seconds
(input/labvalues_Psi50_m.input): rho = 2670 kg/m^3 = 2.67e-3 MPa s^2/m^2.

Per-face values are drawn from a counter-based generator keyed by
(seed, stream, level-0 face index), so they do not depend on the box
decomposition, on batching or on the order scenarios are served in.

Array layout (Chombo FArrayBox, Fortran order): a level's state is a numpy
array of shape (22, ny + 8, nx + 8) over cells [-4, nx+3] x [-4, ny+3]; its
fault data has shape (nbdry, nx + 8) over faces x in [-4, nx+3] at j = 0.
"""
import math

import numpy as np

GHOST = 4
NCOMP = 22
NB = {"rs": 11, "sw": 9}

# ---------------------------------------------------------------------------
# counter-based random numbers (SplitMix64)
# ---------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _splitmix64(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return z ^ (z >> np.uint64(31))


def uniform(seed, stream, idx, lo=0.0, hi=1.0):
    """U[lo,hi) values for integer indices idx, keyed by (seed, stream, idx)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream))
        z = _splitmix64(key ^ (idx * np.uint64(0x9E3779B97F4A7C15)))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


def normal(seed, stream, idx, mu=0.0, sigma=1.0):
    u1 = uniform(seed, 2 * stream + 1000, idx)
    u2 = uniform(seed, 2 * stream + 1001, idx)
    u1 = np.maximum(u1, 1e-300)
    return mu + sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------------------
# material (homogeneous, every workload here): cp = 6000, cs = 3464 m/s, rho = 2670 kg/m^3
# ---------------------------------------------------------------------------
RHO = 2.67e-3
CS = 3464.0
CP = 6000.0
G_MPA = RHO * CS * CS


def _base(nx, ny, dx, fault):
    return {
        "nx": nx, "ny": ny, "dx": dx,
        "cs": CS, "cp": CP, "G": G_MPA,
        "fault": fault,
        "bc": (0, 0, 0, 0),
        "numerics": {"normal_predictor": "PPM", "fourth_order": 1, "prim_limiting": 1, "char_limiting": 0},
        "cfl": 0.9,
    }


def dt_of(cfg):
    """dt = cfl dx / max(cs, cp) (LinElastLevelGodunov.cpp:377, AMRLevelLinElast.cpp:479)."""
    return cfg.get("cfl", 0.9) * cfg["dx"] / max(cfg["cs"], cfg["cp"])


def config_c0(seed=0, nx=400, ny=200, dx=100.0):
    """Slip-weakening, uniform grid, seeded per-face parameters, one 3 km nucleation patch."""
    cfg = _base(nx, ny, dx, "sw")
    i = np.arange(nx)
    mu_s = uniform(seed, 1, i, 0.62, 0.70)
    mu_d = mu_s - uniform(seed, 2, i, 0.10, 0.18)
    dc = uniform(seed, 3, i, 0.3, 0.5)
    sn = uniform(seed, 4, i, 110.0, 130.0)         # |normal stress|
    syy0 = -120.0
    strength = mu_s * sn
    tau = uniform(seed, 5, i, 0.55, 0.60) * strength
    x = (i + 0.5) * dx
    xc = 0.5 * nx * dx
    patch = np.abs(x - xc) < 1500.0
    tau = np.where(patch, 1.01 * strength, tau)
    # the bulk carries a uniform shear prestress (SWFAULTBCF splits the traction along
    # the wave-field shear, SWIBCF.ChF:225-230); the per-face rest is a fault-local offset
    sxy0 = 0.5 * 0.62 * 110.0
    cfg["background"] = (syy0, syy0, syy0, sxy0, 0.0, 0.0)
    cfg["sw"] = {"f_static": 0.66, "f_dynamic": 0.52, "d_weak": 0.4, "smooth": 0.0, "rupture_thr": 0.001,
                 "patches": [], "fric_center": (0.0, 0.0), "fric_width": (1e40, 1e40), "outside_f": 1e4}
    cfg["faces"] = {"lo": 0, "fs": mu_s, "fd": mu_d, "dc": dc, "syy": -sn - syy0, "tau": tau - sxy0}
    cfg["steps"] = 1500
    return cfg


def config_c1(seed=0, nx=4096, ny=2048, dx=25.0, V_init=1e-12):
    """Regularized rate-and-state with aging law and seeded per-face a, L and normal stress."""
    cfg = _base(nx, ny, dx, "rs")
    i = np.arange(nx)
    a = uniform(seed, 11, i, 0.008, 0.02)
    L = uniform(seed, 12, i, 0.015, 0.025)
    b = np.full(nx, 0.014)
    V0, f0 = 1e-6, 0.6
    dsyy = normal(seed, 13, i, 0.0, 2.0)
    syy0 = -120.0
    sn = -(syy0 + dsyy)
    psi = f0 - b * np.log(V_init / V0)           # aging-law steady state at V_init
    tau_ss = sn * a * np.arcsinh(V_init / (2.0 * V0) * np.exp(psi / a))
    cfg["background"] = (syy0, syy0, syy0, 0.0, 0.0, 0.0)
    cfg["rs"] = {"a": 0.014, "b": 0.014, "L": 0.02, "V0": V0, "f0": f0, "psi": float(psi[0]),
                 "nuc_R": 1500.0, "nuc_S": 25.0, "nuc_x": 0.5 * nx * dx, "rupture_thr": 0.001,
                 "state_law": 1, "fric_exp": -1.0}
    cfg["faces"] = {"lo": 0, "a": a, "b": b, "L": L, "syy": dsyy, "tau": tau_ss, "psi": psi}
    return cfg


def config_c3(seed=0, nx=8192, ny=4096, dx=12.5):
    """Rate-and-state as config_c1 plus Drucker-Prager viscoplasticity."""
    cfg = config_c1(seed, nx, ny, dx)
    cfg["plasticity"] = 1
    cfg["p_mu"] = 0.6
    cfg["p_beta"] = 0.3
    cfg["p_eta"] = 0.03 * G_MPA     # relaxation time 0.03 s: eta = G * T_relax
    cfg["p_b"] = 1.0                # cohesion 1 MPa
    return cfg


def refine_cfg(cfg, r):
    """The configuration of the next finer level: r times the cells, dx / r.  Per-face
    parameter arrays are injected (each coarse face's value on its r fine faces), so the
    seeded values do not depend on the level; positions are in metres and carry over."""
    out = dict(cfg)
    out["nx"], out["ny"], out["dx"] = cfg["nx"] * r, cfg["ny"] * r, cfg["dx"] / r
    faces = cfg.get("faces")
    if faces:
        out["faces"] = {k: (v * r if k == "lo" else np.repeat(np.asarray(v), r)) for k, v in faces.items()}
    return out


# ---------------------------------------------------------------------------
# level arrays
# ---------------------------------------------------------------------------
def background_vector(cfg):
    """LINELASTSETF background(0:21) (LinElastPhysicsF.ChF:1183-1207)."""
    sxx0, syy0, szz0, sxy0, sxz0, syz0 = cfg.get("background", (0.0,) * 6)
    side = [0.0, 0.0, 0.0, sxx0, syy0, szz0, sxy0, sxz0, syz0, 0.0, 0.0]
    return np.array(side + side, dtype=np.float64)


def fault_arrays(cfg):
    """Initial fault data over the faces with ghosts (RSIBC::initializeBdry, RSIBC.cpp:115-163;
    slip weakening: all zero, rupture time 1e10)."""
    nx = cfg["nx"]
    B = np.zeros((NB[cfg["fault"]], nx + 2 * GHOST), dtype=np.float64)
    if cfg["fault"] == "rs":
        B[9] = 1e10
        psi = cfg.get("faces", {}).get("psi")
        if psi is None:
            B[10] = cfg["rs"]["psi"]
        else:
            lo = cfg["faces"].get("lo", 0)
            ii = np.clip(np.arange(-GHOST, nx + GHOST) - lo, 0, len(psi) - 1)
            B[10] = np.asarray(psi)[ii]
    else:
        B[8] = 1e10
    return B


def level_arrays(cfg):
    """Initial state (LEINITF, LinElastPhysicsF.ChF:1215-1263) and fault data (fault_arrays)."""
    nx, ny = cfg["nx"], cfg["ny"]
    U = np.empty((NCOMP, ny + 2 * GHOST, nx + 2 * GHOST), dtype=np.float64)
    U[:] = background_vector(cfg)[:, None, None]
    return U, fault_arrays(cfg)


def perturb(U, seed, amp_v=1.0, amp_s=5.0, antiplane=False):
    """Add seeded U(-amp,amp) perturbations to the velocities and stresses (kernel micro-benchmarks)."""
    rng = np.random.default_rng(seed)
    comps_v = [0, 1] + ([2] if antiplane else [])
    comps_s = [3, 4, 5, 6] + ([7, 8] if antiplane else [])
    for s in (0, 11):
        for c in comps_v:
            U[s + c] += rng.uniform(-amp_v, amp_v, size=U.shape[1:])
        for c in comps_s:
            U[s + c] += rng.uniform(-amp_s, amp_s, size=U.shape[1:])
    return U


def split_boxes(nx, ny, maxbox):
    """Uniform decomposition of [0,nx) x [0,ny) into boxes of at most maxbox cells per side."""
    boxes = []
    for y0 in range(0, ny, maxbox):
        for x0 in range(0, nx, maxbox):
            boxes.append((x0, y0, min(x0 + maxbox, nx) - 1, min(y0 + maxbox, ny) - 1))
    return boxes


def box_view(U, box, ghost=GHOST):
    """Ghosted sub-array of a level array for a box (x0, y0, x1, y1): always a C-contiguous copy
    (also for a box that covers the whole level, whose slice is already contiguous)."""
    x0, y0, x1, y1 = box
    return np.array(U[:, y0 - ghost + GHOST:y1 + ghost + GHOST + 1, x0 - ghost + GHOST:x1 + ghost + GHOST + 1],
                    order="C", copy=True)


def box_bview(B, box, ghost=GHOST):
    """Fault-data row of a box with its x ghosts: always a C-contiguous copy."""
    x0, _, x1, _ = box
    return np.array(B[:, x0 - ghost + GHOST:x1 + ghost + GHOST + 1], order="C", copy=True)
