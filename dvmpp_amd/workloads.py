"""Synthetic workloads of the benchmark configurations: the plain cylinder run,
the uneven ensemble of bodies and the served request trace (all seeded)."""
import numpy as np

from . import bodies
from .sim import sheet_params


def plain_cylinder():
    """Circular cylinder r = 0.5, 256 equal panels, U = (1, 0), rho = 1,
    nu = 1e-3 (Re = 1000), cutoff_exp = 0.675, max_NumPanelVort = 3, Euler,
    dt = 0.005, REFLECT, seed 0."""
    px, pz = bodies.cylinder(256, 0.5)
    return px, pz, sheet_params(rho=1.0, nu=1e-3, dt=0.005, ux=1.0, uz=0.0, cutoff_exp=0.675,
                                max_npv=3)


def ensemble(count=128, seed=7, steps_range=(200, 2000), panels_range=(128, 512)):
    """Uneven bodies: ellipses (aspect ~ U[1,4], angle of attack ~ U[0,30] deg)
    and rounded squares, panels ~ U{128..512}, nu log-uniform in [2e-4, 2e-3],
    dt = 0.005, steps ~ U{200..2000}.  Returns a list of
    (px, pz, sheet_params, seed, steps)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nl = int(rng.integers(panels_range[0], panels_range[1] + 1))
        if rng.random() < 0.5:
            px, pz = bodies.ellipse(nl, aspect=float(rng.uniform(1.0, 4.0)),
                                    aoa_deg=float(rng.uniform(0.0, 30.0)))
        else:
            px, pz = bodies.rounded_square(nl, side=1.0, radius=float(rng.uniform(0.1, 0.3)))
        nu = float(np.exp(rng.uniform(np.log(2e-4), np.log(2e-3))))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        out.append((px, pz, sheet_params(nu=nu, dt=0.005), 1000 + k, steps))
    return out


def request_trace(count=16, rate=20.0, seed=11, steps_range=(500, 3000), panels_range=(128, 512)):
    """Poisson arrivals at `rate` requests/s drawn from the ensemble
    distribution, step counts ~ U{steps_range}: list of (arrival_s, job)."""
    rng = np.random.default_rng(seed)
    jobs = ensemble(count, seed=seed + 1, steps_range=steps_range, panels_range=panels_range)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    return list(zip(t.tolist(), jobs))


def projected_blobs(nl, steps, max_npv=3):
    """Blob count a simulation reaches without remesh: every panel releases at
    least one blob per step and about 2.3 on average for a bluff body
    (src/VortexSheet.cpp:285-287)."""
    return int(2.3 * nl * steps) + nl * (max_npv + 1)
