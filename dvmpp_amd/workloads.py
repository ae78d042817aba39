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


def rk3_cylinder():
    """RK3 cylinder: r = 0.5, 512 equal panels, U = (1, 0), nu = 5e-4,
    dt = 0.0025, cutoff_exp = 0.675 (nascent sigma = ds^0.675 = 0.0321),
    max_npv = 3, REFLECT; remesh with M4' every 20 steps at dx = dz = 0.01 on
    [-5, 25] x [-10, 10] with the nascent sigma."""
    px, pz = bodies.cylinder(512, 0.5)
    ds = float(np.hypot(px[1] - px[0], pz[1] - pz[0]))
    return px, pz, sheet_params(rho=1.0, nu=5e-4, dt=0.0025, ux=1.0, uz=0.0, cutoff_exp=0.675,
                                max_npv=3), ds ** 0.675


def wake_snapshot(n_target=1_500_000, dx=0.01, sigma=0.0321, seed=5):
    """A remeshed-wake state of about n_target blobs: blobs on the mesh nodes
    (spacing dx) of a rectangle behind and around the r = 0.5 body, outside
    it, with the circulation of two shed shear layers plus a vortex street
    (omega * dx^2).  Returns dict(x, z, circ, sigma, id, u, w, uvs, wvs)."""
    rng = np.random.default_rng(seed)
    width = 10.0
    length = n_target * dx * dx / width
    xs = np.arange(-0.7, -0.7 + length, dx)
    zs = np.arange(-0.5 * width, 0.5 * width, dx)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    x, z = X.ravel(), Z.ravel()
    keep = x * x + z * z > 0.55 ** 2
    x, z = x[keep], z[keep]
    layer = np.exp(-((np.abs(z) - 0.5 - 0.05 * np.maximum(x, 0.0)) / 0.15) ** 2)
    street = np.sin(2.0 * np.pi * x / 4.0) * np.exp(-z * z / 2.0)
    omega = -np.sign(z) * 8.0 * layer * np.exp(-0.1 * np.maximum(x, 0.0)) + 2.0 * street
    omega += rng.normal(0.0, 0.05, x.size)
    n = x.size
    zero = np.zeros(n)
    return {"x": x, "z": z, "circ": omega * dx * dx, "sigma": np.full(n, sigma),
            "id": np.arange(n, dtype=np.uint32), "u": zero, "w": zero.copy(),
            "uvs": zero.copy(), "wvs": zero.copy()}
