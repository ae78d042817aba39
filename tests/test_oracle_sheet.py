"""Pins the oracle's vortex-sheet restatement (src/VortexSheet.cpp).

Off-diagonal influence coefficients and sheet-on-blob velocities are checked
against Gauss-Legendre quadrature of a linear-strength vortex panel; the
least-squares solve against its normal equations; release, crossing and loads
against hand-worked cases and conservation laws.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import load_body

RP2 = 1.0 / (2.0 * math.pi)
GLX, GLW = np.polynomial.legendre.leggauss(400)


def panel_velocity(px, pz, x0, z0, th, ds, ga, gb):
    """Velocity at (px, pz) of a straight panel from (x0, z0) along angle th,
    length ds, with strength varying linearly from ga to gb (counter-clockwise
    point-vortex convention, u = -G dz / (2 pi r^2))."""
    s = 0.5 * (GLX + 1.0) * ds
    wq = 0.5 * ds * GLW
    yx = x0 + s * math.cos(th)
    yz = z0 + s * math.sin(th)
    g = ga * (1 - s / ds) + gb * (s / ds)
    dx = px - yx
    dz = pz - yz
    r2 = dx * dx + dz * dz
    return np.sum(wq * g * (-dz) / r2) * RP2, np.sum(wq * g * dx / r2) * RP2


@pytest.fixture
def params():
    return oracle.make_params()


def test_geometry_cylinder(params):
    nl = 64
    x, z = oracle.cylinder_points(nl)
    sh = oracle.Sheet(x, z, params)
    g = sh.geometry()
    assert sh.nl == nl
    # chord length and outward normal at the chord midpoint angle
    assert np.allclose(g["ds"], 2 * 0.5 * np.sin(np.pi / nl), rtol=1e-14)
    mid = 2 * np.pi * (np.arange(nl) + 0.5) / nl
    assert np.allclose(g["enx"], np.cos(mid), atol=1e-14)
    assert np.allclose(g["enz"], np.sin(mid), atol=1e-14)
    assert np.array_equal(g["etx"], -g["enz"]) and np.array_equal(g["etz"], g["enx"])
    assert g["aref"] == pytest.approx(1.0, abs=1e-15)


def test_influence_matches_quadrature(params):
    # A(i,j) = -(normal velocity at collocation i of unit gamma_j spread over
    # panel j (start end) and panel j-1 (finish end)); last row = ds.
    for name in ("simple_body.txt", "bodycoordinatesN40.txt"):
        x, z = load_body(name)
        sh = oracle.Sheet(x, z, params)
        g = sh.geometry()
        A = sh.influence()
        nl = sh.nl
        assert A.shape == (nl + 1, nl)
        assert np.array_equal(A[nl], g["ds"])
        worst = 0.0
        for i in range(nl):
            for j in range(nl):
                jm = (j - 1) % nl
                if i in (j, jm):
                    continue  # principal-value (self) panels: diagonal convention p=-1, q=1
                ua, wa = panel_velocity(g["xc"][i], g["zc"][i], x[j], z[j], g["theta"][j],
                                        g["ds"][j], 1.0, 0.0)
                ub, wb = panel_velocity(g["xc"][i], g["zc"][i], x[jm], z[jm], g["theta"][jm],
                                        g["ds"][jm], 0.0, 1.0)
                vn = (ua + ub) * g["enx"][i] + (wa + wb) * g["enz"][i]
                worst = max(worst, abs(A[i, j] + vn))
        assert worst < 1e-12, (name, worst)


def test_influence_diagonal_convention(params):
    # p(i,i) = -1, q(i,i) = 1 (src/VortexSheet.cpp:123-125) enter A(i,i) and A(i,i+1)
    x, z = oracle.cylinder_points(8)
    sh = oracle.Sheet(x, z, params)
    g = sh.geometry()
    A = sh.influence()
    i = 2
    j = 3  # A(i, i+1) = -rp (p(i,i+1) + q(i,i)) with q(i,i) = 1
    ua, wa = panel_velocity(g["xc"][i], g["zc"][i], x[j], z[j], g["theta"][j], g["ds"][j], 1, 0)
    vn = ua * g["enx"][i] + wa * g["enz"][i]
    assert abs(A[i, j] - (-vn - RP2 * 1.0)) < 1e-12


def test_sheet_bc_matches_quadrature(params):
    # vortexsheetbc = minus the velocity of midpoint-origin linear panels
    # (src/VortexSheet.cpp:230-231 places panel j at the collocation point).
    x, z = load_body("bodycoordinatesN40.txt")
    sh = oracle.Sheet(x, z, params)
    g = sh.geometry()
    nl = sh.nl
    gam = np.random.default_rng(11).normal(size=nl)
    sh.gamma = gam
    bx = np.array([1.3, -0.2, 0.9, 0.55, 3.0])
    bz = np.array([0.4, 1.1, -0.8, 0.1, -2.0])
    uvs, wvs = sh.bc(bx, bz)
    for k in range(bx.size):
        u = w = 0.0
        for j in range(nl):
            du, dw = panel_velocity(bx[k], bz[k], g["xc"][j], g["zc"][j], g["theta"][j],
                                    g["ds"][j], gam[j], gam[(j + 1) % nl])
            u += du
            w += dw
        assert abs(uvs[k] + u) < 1e-12 and abs(wvs[k] + w) < 1e-12


def test_solve_normal_equations(params):
    x, z = oracle.cylinder_points(256)
    sh = oracle.Sheet(x, z, params)
    A = sh.influence()
    rng = np.random.default_rng(1)
    bx = rng.uniform(1, 3, 200)
    bz = rng.uniform(-1, 1, 200)
    bc = rng.normal(0, 1e-3, 200)
    bsg = np.full(200, 0.05)
    b, gam = sh.solve(bx, bz, bc, bsg)
    r = A.T @ (A @ gam) - A.T @ b
    assert np.max(np.abs(r)) < 1e-13 * np.max(np.abs(A.T @ b))
    # RHS rows: (U + u) . n with u the unguarded collocation sum; last = -sum circ
    g = sh.geometry()
    u, w = oracle.bs_targets(bx, bz, bc, bsg, g["xc"], g["zc"])
    assert np.allclose(b[:-1], (1.0 + u) * g["enx"] + (0.0 + w) * g["enz"], rtol=0, atol=1e-15)
    assert b[-1] == pytest.approx(-bc.sum(), abs=1e-17)


def test_solve_no_blobs_potential_flow(params):
    # with no blobs, b = U . n; gamma tends to the potential-flow surface
    # vorticity 2 U sin(theta) in magnitude as the panel count grows
    x, z = oracle.cylinder_points(512)
    sh = oracle.Sheet(x, z, params)
    b, gam = sh.solve([], [], [], [])
    g = sh.geometry()
    assert np.array_equal(b[:-1], 1.0 * g["enx"] + 0.0 * g["enz"])
    assert b[-1] == 0.0
    mid = 2 * np.pi * (np.arange(512) + 0.5) / 512
    assert np.max(np.abs(np.abs(gam) - 2.0 * np.abs(np.sin(mid)))) < 2e-2
    assert abs(np.sum(gam * g["ds"])) < 1e-12  # zero net circulation


def test_inside_body_crossing_number(params):
    x, z = load_body("simple_body.txt")
    sh = oracle.Sheet(x, z, params)
    assert sh.inside(0.0, 0.0) and sh.inside(0.3, -0.2)
    assert not sh.inside(0.6, 0.0) and not sh.inside(0.0, 0.51)
    # a concave "C" body: a point in the notch has crossing number 2 -> outside,
    # and a point inside the body's material has crossing number 1 -> inside
    cx = np.array([0, 3, 3, 1, 1, 3, 3, 0, 0], dtype=float)
    cz = np.array([0, 0, 1, 1, 2, 2, 3, 3, 0], dtype=float)
    c = oracle.Sheet(cx, cz, params)
    assert c.inside(0.5, 1.5)
    assert not c.inside(2.0, 1.5)
    assert not c.inside(-0.5, 1.5)
    # half-open rule at a vertex height: z == z_i counts once
    assert c.inside(0.5, 1.0)


def test_release_counts_hand_worked():
    # 4-panel unit square, gamma chosen so gamma*ds = [1, -3, 0.5, 0]:
    # counts = floor(|g ds / max(g ds) * 3|) + 1 = [4, 10, 2, 1] (signed max)
    p = oracle.make_params(max_npv=3, nu=1e-3, dt=0.01, seed=5)
    x = np.array([0, 1, 1, 0, 0], dtype=float)
    z = np.array([0, 0, 1, 1, 0], dtype=float)
    sim = oracle.Sim(x, z, p)
    g = np.array([1.0, -3.0, 0.5, 0.0])
    sim.set_gamma(g)
    draws0 = sim.rng_draws()
    assert sim.release() == 0
    st = sim.get_state()
    assert st["x"].size == 17
    assert sim.rng_draws() - draws0 == 17
    counts = [4, 10, 2, 1]
    off = 0
    for i, n in enumerate(counts):
        seg = slice(off, off + n)
        assert np.all(st["circ"][seg] == g[i] / n)
        assert np.all(st["id"][seg] == 0)  # appended IDs = Nold + 0
        off += n
    assert abs(st["circ"].sum() - g.sum()) < 1e-15
    assert np.all(st["sigma"] == 1.0)  # ds^cutoff_exp with ds = 1
    # panel 0 (bottom edge, outward normal (0,-1)): every nascent blob of the
    # panel shares xm = 0 + 1 * 1 / (4 + 1); offsets are Rayleigh along -z
    assert np.all(st["x"][:4] == 0.2) and np.all(st["z"][:4] <= 0.0)
    # panel 1 (right edge, normal (1,0)): zm = 0 + 1/(10+1), x >= 1
    assert np.all(st["z"][4:14] == 1.0 / 11.0) and np.all(st["x"][4:14] >= 1.0)
    # the RNG stream: same seed, same uniforms -> offsets r = sqrt(4 nu dt ln(1/R))
    rng = oracle.Rng(5)
    R = rng.fill(17)
    r = np.abs(np.sqrt(4.0 * 1e-3 * 0.01 * np.log(1.0 / R)))
    assert np.allclose(-st["z"][:4], r[:4], rtol=1e-15, atol=0)
    assert np.allclose(st["x"][4:14] - 1.0, r[4:14], rtol=1e-14, atol=1e-16)


def test_release_nan_raises():
    # zero free stream -> gamma = 0 -> max(gamma ds) = 0 -> NaN counts (VortexException)
    p = oracle.make_params(ux=0.0, uz=0.0)
    x, z = oracle.cylinder_points(16)
    sim = oracle.Sim(x, z, p)
    assert sim.step() == 3


def test_reflect_and_delete(params):
    x, z = oracle.cylinder_points(64)
    for crossing in (0, 2):
        p = oracle.make_params(crossing=crossing)
        sim = oracle.Sim(x, z, p)
        n = 6
        st = {k: np.zeros(n) for k in oracle.STATE_FIELDS}
        st["x"] = np.array([0.1, 0.8, -0.3, 0.0, 0.45, 2.0])
        st["z"] = np.array([0.0, 0.0, 0.2, -0.49, 0.1, 1.0])
        st["circ"] = np.arange(1.0, 7.0)
        st["sigma"] = np.full(n, 0.05)
        st["id"] = np.arange(n, dtype=np.uint32) + 10
        sim.set_state(st)
        sim.crossing()
        out = sim.get_state()
        if crossing == 0:
            # inside: 0, 2, 3, 4 -> deleted, stable order, IDs renumbered
            assert np.array_equal(out["circ"], [2.0, 6.0])
            assert np.array_equal(out["id"], [0, 1])
        else:
            sh = oracle.Sheet(x, z, p)
            assert np.array_equal(out["circ"], st["circ"])
            assert np.array_equal(out["id"], st["id"])
            for k in range(n):
                assert not sh.inside(out["x"][k], out["z"][k])
            # blob 3 at (0, -0.49): nearest panel is the bottom one; moved along
            # that panel's outward normal by 2 d + 1e-6
            r = math.hypot(out["x"][3], out["z"][3])
            assert r > 0.49


def test_loads_hand_worked():
    # 4-panel square, rho = 2, dt = 0.5, gamma = [1, 2, 3, 4], ds = 1:
    # Dp = -rho/dt * gamma ds = [-4, -8, -12, -16]; P = [0, -4, -12, -24]
    # averaged: [-2, -8, -18, (-24 + -2)/2 = -13]; pmax = -2; pref = 0.5*2*1 = 1
    # P += 1 - (-2) = 3 -> [1, -5, -15, -10]
    # normals: bottom (0,-1), right (1,0), top (0,1), left (-1,0)
    # Fx = -(P1*1 + P3*(-1)) = -(-5 + 10) = -5 ; Fz = -(P0*(-1) + P2*1) = -(-1 - 15) = 16
    p = oracle.make_params(rho=2.0, dt=0.5, ux=1.0, uz=0.0)
    x = np.array([0, 1, 1, 0, 0], dtype=float)
    z = np.array([0, 0, 1, 1, 0], dtype=float)
    sh = oracle.Sheet(x, z, p)
    sh.gamma = np.array([1.0, 2.0, 3.0, 4.0])
    fx, fz, cd, cl = sh.loads(1.0)
    assert fx == -5.0 and fz == 16.0
    assert cd == -5.0 / (0.5 * 2.0 * 1.0) and cl == 16.0 / (0.5 * 2.0 * 1.0)


def test_probe_quirk_runs(params):
    x, z = oracle.cylinder_points(32)
    sh = oracle.Sheet(x, z, params)
    sh.gamma = np.linspace(-1, 1, 32)
    u, w, uvs, wvs = sh.probe_quirk([1.0, 2.0], [0.0, 0.5], [0.1, 0.2], [0.05, 0.05],
                                    [3.0], [0.0])
    # blob 0 is skipped (loop from j = 1): only blob 1 contributes to the blob term
    ub, wb = oracle.bs_targets([2.0], [0.5], [0.2], [0.05], [3.0], [0.0])
    assert abs((u[0] - uvs[0] - 1.0) - ub[0]) < 1e-15
    assert abs((w[0] - wvs[0] - 0.0) - wb[0]) < 1e-15


def test_probe_quirk_sheet_term(params):
    # probe_velocities' sheet term is that of panel Nl-1 alone, paired with
    # gamma_0 (the loop assigns, src/VortexSheet.cpp:628-636).  Its u part is
    # 2 pi times the vortexsheetbc coefficient pair, i.e. minus 2 pi times the
    # quadrature velocity of the midpoint-origin linear panel (no 1/(2 pi)).
    x, z = load_body("bodycoordinatesN40.txt")
    sh = oracle.Sheet(x, z, params)
    g = sh.geometry()
    nl = sh.nl
    gam = np.random.default_rng(3).normal(size=nl)
    sh.gamma = gam
    prx = np.array([1.3, -0.2, 0.9, 3.0])
    prz = np.array([0.4, 1.1, -0.8, -2.0])
    _, _, uvs, wvs = sh.probe_quirk([], [], [], [], prx, prz)
    for k in range(prx.size):
        du, _ = panel_velocity(prx[k], prz[k], g["xc"][nl - 1], g["zc"][nl - 1],
                               g["theta"][nl - 1], g["ds"][nl - 1], gam[nl - 1], gam[0])
        assert abs(uvs[k] + du / RP2) < 1e-11
    # the term does not depend on gamma_1 .. gamma_{Nl-2}
    gam2 = gam.copy()
    gam2[1:nl - 1] = 7.0
    sh.gamma = gam2
    _, _, uvs2, wvs2 = sh.probe_quirk([], [], [], [], prx, prz)
    assert np.array_equal(uvs, uvs2) and np.array_equal(wvs, wvs2)


def _quirk_w_from_geometry(prx, prz, xa, za, th, ds, ga, g0, variant=None):
    """The w part of probe_velocities' sheet term (src/VortexSheet.cpp:611-637),
    built from the panel's geometry and the vortexsheetbc coefficients.

    vortexsheetbc's (py, qy) pair of the panel (:248-250) is minus 2 pi times
    the quadrature w velocity of a unit linear strength at either end, which
    test_sheet_bc_matches_quadrature pins.  probe_velocities changes three
    things against it: c9 takes +dz sin(-2 theta) where vortexsheetbc takes
    -dz sin(-2 theta) (:619-620 against :240), qy's c7 factor is
    (c1 cos - c5 sin) where vortexsheetbc has (c1 sin - c5 cos) (:624-625
    against :248-249), and py's c6 term has the opposite sign (:626 against
    :250); the gamma_0 product is subtracted (:634).  With s and n the
    probe's tangential and normal coordinates about the panel origin A and B
    the panel's far end, c1 = -s, c5 = n, c6 = ln(|P-B|^2 / |P-A|^2) and c7
    is the signed angle from P-B to P-A (the angle AB subtends at P), so

        qy' = qy - dz sin(2 theta) c6 / ds + (n - s)(cos theta - sin theta) c7 / ds
        py' = py - c6 cos theta - (qy' - qy)
        w   = py' ga - qy' g0.

    `variant` plants one plausible transcription mistake in this formula so a
    test can show that the tolerance separates it.
    """
    c, s_ = math.cos(th), math.sin(th)
    xb, zb = xa + ds * c, za + ds * s_
    pa = np.array([prx - xa, prz - za])
    pb = np.array([prx - xb, prz - zb])
    c6 = math.log(float(pb @ pb) / float(pa @ pa))
    c7 = math.atan2(pb[0] * pa[1] - pb[1] * pa[0], float(pa @ pb))
    t = pa[0] * c + pa[1] * s_          # tangential coordinate of P
    nn = pa[0] * s_ - pa[1] * c         # normal coordinate (c5's sign)
    _, w10 = panel_velocity(prx, prz, xa, za, th, ds, 1.0, 0.0)
    _, w01 = panel_velocity(prx, prz, xa, za, th, ds, 0.0, 1.0)
    py = -w10 / RP2
    qy = -w01 / RP2
    dq = -pa[1] * math.sin(2 * th) * c6 / ds + (nn - t) * (c - s_) * c7 / ds
    dp = -c6 * c
    if variant == "c9_as_bc":
        dq = (nn - t) * (c - s_) * c7 / ds
    elif variant == "qy_c7_as_bc":
        dq = -pa[1] * math.sin(2 * th) * c6 / ds
    elif variant == "py_c6_as_bc":
        dp = 0.0
    elif variant == "gamma0_added":
        return (py + dp - dq) * ga + (qy + dq) * g0
    return (py + dp - dq) * ga - (qy + dq) * g0


def test_probe_quirk_w_term_geometry(params):
    # Pins the w component of probe_velocities' sheet term, which no closed
    # form of a panel velocity gives: it is vortexsheetbc's pair plus the
    # terms the source's sign changes add, each from the geometry.
    rng = np.random.default_rng(17)
    for body in ("bodycoordinatesN40.txt", "simple_body.txt"):
        x, z = load_body(body)
        sh = oracle.Sheet(x, z, params)
        g = sh.geometry()
        nl = sh.nl
        gam = rng.normal(size=nl)
        sh.gamma = gam
        ang = rng.uniform(0, 2 * math.pi, 12)
        rad = rng.uniform(0.7, 4.0, 12)
        prx = g["xc"].mean() + rad * np.cos(ang)
        prz = g["zc"].mean() + rad * np.sin(ang)
        _, w, _, wvs = sh.probe_quirk([], [], [], [], prx, prz)
        L = nl - 1
        args = (g["xc"][L], g["zc"][L], g["theta"][L], g["ds"][L], gam[L], gam[0])
        scale = max(1.0, np.abs(wvs).max())
        for k in range(prx.size):
            ref = _quirk_w_from_geometry(prx[k], prz[k], *args)
            assert abs(wvs[k] - ref) < 1e-10 * scale, (body, k, wvs[k], ref)
            # no blobs: w = sheet term + Uz (Uz = 0 here)
            assert w[k] == wvs[k] + params.uz
        # each plausible slip in the three changed lines moves w by far more
        # than the tolerance at these probes
        for variant in ("c9_as_bc", "qy_c7_as_bc", "py_c6_as_bc", "gamma0_added"):
            worst = max(abs(wvs[k] - _quirk_w_from_geometry(prx[k], prz[k], *args, variant=variant))
                        for k in range(prx.size))
            assert worst > 1e-4 * scale, variant
