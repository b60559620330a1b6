"""Closed forms for the oracle's transient terms (TEST INFRASTRUCTURE checks).

Each check sets up a state on which the reference's discrete scheme has an
exact solution derived from the equations it discretises, not from the
oracle's code:

* hydro_intern's element column (hydrodynamic.f90:1132-1416) is a
  theta-weighted implicit step of  dU/dt = -r U + f x U + d/dz(nu dU/dz).
  On a flat, level basin each of its pieces decays or turns every vertical
  mode of U by a known factor: bottom friction of each law in chezy.f90
  (backward Euler: U/(1 + dt r)), Coriolis with the afpar weight (a complex
  rotation), vertical viscosity with the atpar weight (the eigenvalues of the
  three-point layer Laplacian with no-stress ends) and the three together.
* The whole 2D step (hydro_zeta + solve + hydro_transports_final, :488-754,
  :1468-1578) on a closed, flat basin moves each eigenmode K v = lambda M v
  of the P1 stiffness / lumped mass pair by the two-level recurrence
  u' = u - dt g H (am z' + (1-am) z),  z' = z + dt lambda (az u' + (1-az) u),
  with K and M assembled here from the coordinates alone.
* conz3d's vertical diffusion (concentration_omp.f90:583-638, :726-734,
  :855-870) on a resting basin decays each cosine layer mode by the theta
  factor of adpar, the 1D diffusion of a Gaussian the reference's comment
  describes (concentration_omp.f90:108-120) in discrete form.
* scalar_stability (stability.f90:220-260, conzstab concentration.f90:
  1573-2097): the explicit vertical-diffusion number 2 D dt / h^2, the
  uniform-flow index on the structured basin derived by hand from the
  element signs (see the test), and point sources.

A dropped term, a wrong weight (az for am, at for af), a wrong sign of the
Coriolis term or a wrong friction exponent fails one of these.
"""
import dataclasses

import numpy as np
import pytest

import pyoracle
from shyhip import mesh

G = 9.81


def _closed(b):
    return dataclasses.replace(b, ibnd=np.zeros(b.nkn, np.int32), ibtyp=np.array([0], np.int32),
                               bnd_nodes=np.zeros(0, np.int32))


def _flat(nlv=1, depth=10.0, nx=11, ny=7):
    """Closed, flat rectangular basin; nlv layers of 1 m when nlv > 1."""
    return _closed(mesh.tidal_basin(nx=nx, ny=ny, depth=depth if nlv == 1 else float(nlv), noise=0.0, nlv=nlv))


def _column_model(b, prec="f64", **kw):
    p = dict(dt=300.0, ireib=0, itpre=1, vismol=0.0)
    p.update(kw.pop("params", {}))
    return pyoracle.OracleModel(b, prec, params=p, **kw)


def _set_transports(om, U, V):
    om.a["utlov"][:] = U
    om.a["vtlov"][:] = V


# ---------------------------------------------------------------- element columns

def _friction_r(ireib, cz, U, V, h, uvmin=0.2):
    """The friction coefficient r of each law as chezy.f90 documents it
    (U, V transports of the bottom layer, h its thickness)."""
    uv = np.hypot(U, V)
    if ireib == 1:
        return np.full_like(uv, cz)
    if ireib == 2:                                   # Strickler: cd = g / (ks^2 h^(1/3))
        return G / (cz ** 2 * h ** (1 / 3)) * uv / h ** 2
    if ireib == 3:                                   # Chezy: cd = g / C^2
        return G / cz ** 2 * uv / h ** 2
    if ireib == 4:                                   # linear: r = cz / h
        return np.full_like(uv, cz / h)
    if ireib == 5:                                   # quadratic, constant cd
        return cz * uv / h ** 2
    if ireib == 6:                                   # roughness length: cd = (kappa / ln((z0 + h/2) / z0))^2
        return (0.4 / np.log((cz + 0.5 * h) / cz)) ** 2 * uv / h ** 2
    if ireib == 10:                                  # quadratic above uvmin, linear below
        return np.where(uv > uvmin * h, cz * uv / h ** 2, cz * uvmin / h)
    raise ValueError(ireib)


@pytest.mark.parametrize("ireib,cz", [(1, 1e-3), (2, 30.0), (3, 50.0), (4, 2e-3), (5, 2.5e-3), (6, 0.01),
                                      (10, 2.5e-3)])
def test_friction_decay_single_layer(ireib, cz):
    """One layer, level surface, no Coriolis: U' = U / (1 + dt r) with r of
    the law; ddx = h / (1 + dt r) in the u row and 0 in the v row."""
    b = _flat(nlv=1, depth=10.0)
    om = _column_model(b, params=dict(ireib=ireib), czv=cz)
    rng = np.random.default_rng(ireib)
    U = rng.uniform(-3.0, 3.0, b.nel)
    V = rng.uniform(-3.0, 3.0, b.nel)
    _set_transports(om, U[:, None], V[:, None])
    om.transports()
    h = 10.0
    fac = 1.0 / (1.0 + 300.0 * _friction_r(ireib, cz, U, V, h))
    assert np.allclose(om.a["utlnv"][:, 0], U * fac, rtol=1e-13, atol=0)
    assert np.allclose(om.a["vtlnv"][:, 0], V * fac, rtol=1e-13, atol=0)
    assert np.allclose(om.a["ddxv"][:, 0], h * fac, rtol=1e-13) and np.all(om.a["ddxv"][:, 1] == 0)
    assert np.allclose(om.a["ddyv"][:, 1], h * fac, rtol=1e-13) and np.all(om.a["ddyv"][:, 0] == 0)


def test_quadratic_friction_acts_on_bottom_layer_only():
    """Three layers without viscosity: only the bottom layer feels friction,
    with r from its own transport and thickness (here 0.6 m)."""
    b = _closed(mesh.tidal_basin(nx=9, ny=5, depth=2.6, noise=0.0, nlv=3))
    assert np.all(b.ilhv == 3)
    om = _column_model(b, params=dict(ireib=5), czv=2.5e-3)
    rng = np.random.default_rng(3)
    U = rng.uniform(-1.0, 1.0, (b.nel, 3))
    V = rng.uniform(-1.0, 1.0, (b.nel, 3))
    _set_transports(om, U, V)
    om.transports()
    hb = om.a["hdeov"][:, 2].astype(np.float64)
    assert np.allclose(hb, 0.6, rtol=1e-6)
    fac = 1.0 / (1.0 + 300.0 * 2.5e-3 * np.hypot(U[:, 2], V[:, 2]) / hb ** 2)
    assert np.allclose(om.a["utlnv"][:, :2], U[:, :2], rtol=1e-14, atol=0)
    assert np.allclose(om.a["utlnv"][:, 2], U[:, 2] * fac, rtol=1e-13)
    assert np.allclose(om.a["vtlnv"][:, 2], V[:, 2] * fac, rtol=1e-13)


@pytest.mark.parametrize("af", [0.5, 0.7])
def test_inertial_rotation_single_layer(af):
    """Coriolis alone: W = U + iV turns by (1 - i(1-af) f dt) / (1 + i af f dt),
    clockwise for f > 0, without loss for af = 0.5; ddx, ddy are the columns
    of the inverse rotation matrix times h."""
    f = 1.0e-4
    b = _flat(nlv=1, depth=10.0)
    om = _column_model(b, params=dict(afpar=af), fcor=f)
    rng = np.random.default_rng(5)
    W = rng.uniform(-2, 2, b.nel) + 1j * rng.uniform(-2, 2, b.nel)
    _set_transports(om, W.real[:, None], W.imag[:, None])
    om.transports()
    tau = 300.0 * f
    Wn = W * (1 - 1j * (1 - af) * tau) / (1 + 1j * af * tau)
    got = om.a["utlnv"][:, 0] + 1j * om.a["vtlnv"][:, 0]
    assert np.allclose(got, Wn, rtol=1e-13, atol=1e-15)
    if af == 0.5:
        assert np.allclose(np.abs(got), np.abs(W), rtol=1e-13)
        assert np.all(np.angle(got / W) < 0)                         # clockwise
    gam = af * tau
    assert np.allclose(om.a["ddxv"][:, 0], 10.0 / (1 + gam ** 2), rtol=1e-13)
    assert np.allclose(om.a["ddxv"][:, 1], -gam * 10.0 / (1 + gam ** 2), rtol=1e-13)
    assert np.allclose(om.a["ddyv"][:, 0], gam * 10.0 / (1 + gam ** 2), rtol=1e-13)
    assert np.allclose(om.a["ddyv"][:, 1], 10.0 / (1 + gam ** 2), rtol=1e-13)


def _layer_modes(L):
    """Cosine modes of the three-point layer Laplacian with no-flux ends:
    columns k = 0..L-1 of (L, L), and 4 sin^2(pi k / 2L)."""
    l = np.arange(L) + 0.5
    k = np.arange(L)
    Q = np.cos(np.pi * np.outer(l, k) / L)
    return Q, 4.0 * np.sin(np.pi * k / (2 * L)) ** 2


@pytest.mark.parametrize("at", [1.0, 0.6])
def test_vertical_viscosity_modes(at):
    """Six 1 m layers, viscosity nu at every interface (plus vismol), level
    surface: mode k of U is multiplied by (1 - (1-at) dt mu_k) / (1 + at dt
    mu_k), mu_k = nu/h^2 * 4 sin^2(pi k / 2L); ddx = h in every layer."""
    L, nu, vismol, dt = 6, 4e-3, 1e-6, 300.0
    b = _flat(nlv=L)
    om = _column_model(b, params=dict(atpar=at, vismol=vismol), visv=nu)
    Q, lam = _layer_modes(L)
    rng = np.random.default_rng(7)
    cu = rng.uniform(-1, 1, (b.nel, L))
    cv = rng.uniform(-1, 1, (b.nel, L))
    _set_transports(om, cu @ Q.T, cv @ Q.T)
    om.transports()
    nue = float(np.float64(vismol)) + nu
    mu = nue * lam
    g = (1 - (1 - at) * dt * mu) / (1 + at * dt * mu)
    assert np.allclose(om.a["utlnv"], (cu * g) @ Q.T, rtol=0, atol=1e-12)
    assert np.allclose(om.a["vtlnv"], (cv * g) @ Q.T, rtol=0, atol=1e-12)
    assert np.allclose(om.a["ddxv"][:, 0::2], 1.0, rtol=1e-13) and np.all(om.a["ddxv"][:, 1::2] == 0)


@pytest.mark.parametrize("at,af", [(1.0, 0.5), (0.7, 0.6)])
def test_coriolis_with_viscosity_modes(at, af):
    """The coupled pentadiagonal column (Coriolis + viscosity, four layers):
    W_k' = W_k (1 - dt (mu_k + i f) / (1 + at dt mu_k + i af dt f))."""
    L, nu, f, dt = 4, 6e-3, 1.2e-4, 300.0
    b = _flat(nlv=L)
    om = _column_model(b, params=dict(atpar=at, afpar=af), visv=nu, fcor=f)
    Q, lam = _layer_modes(L)
    rng = np.random.default_rng(11)
    cw = rng.uniform(-1, 1, (b.nel, L)) + 1j * rng.uniform(-1, 1, (b.nel, L))
    W = cw @ Q.T
    _set_transports(om, W.real, W.imag)
    om.transports()
    mu = nu * lam
    g = 1 - dt * (mu + 1j * f) / (1 + at * dt * mu + 1j * af * dt * f)
    got = om.a["utlnv"] + 1j * om.a["vtlnv"]
    assert np.allclose(got, (cw * g) @ Q.T, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- whole 2D step

def _p1_stiffness_mass(b):
    x = b.xgv.astype(np.float64)
    y = b.ygv.astype(np.float64)
    n = b.nkn
    K = np.zeros((n, n))
    M = np.zeros(n)
    for tri in b.nen3v:
        xs, ys = x[tri], y[tri]
        A = 0.5 * ((xs[1] - xs[0]) * (ys[2] - ys[0]) - (xs[2] - xs[0]) * (ys[1] - ys[0]))
        bb = np.array([ys[1] - ys[2], ys[2] - ys[0], ys[0] - ys[1]]) / (2 * A)
        cc = np.array([xs[2] - xs[1], xs[0] - xs[2], xs[1] - xs[0]]) / (2 * A)
        K[np.ix_(tri, tri)] += A * (np.outer(bb, bb) + np.outer(cc, cc))
        M[tri] += A / 3
    return K, M


@pytest.mark.parametrize("az,am", [(0.5, 0.5), (0.6, 0.8)])
def test_standing_wave_recurrence(az, am):
    """Closed flat basin 10 m deep, linearised continuity (iclin=1), no
    friction: a start level in eigenmode v (K v = lambda M v) stays in that
    mode, and its amplitude z and transport amplitude u follow the two-level
    recurrence of the semi-implicit scheme exactly (to the solve)."""
    H, dt, eps = 10.0, 300.0, 0.01
    b = _flat(nlv=1, depth=H, nx=21, ny=11)
    K, M = _p1_stiffness_mass(b)
    Mi = 1 / np.sqrt(M)
    w, vec = np.linalg.eigh(Mi[:, None] * K * Mi[None, :])
    lam = w[1]                                                # first seiche
    v = Mi * vec[:, 1]
    v = v / np.abs(v).max()
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=dt, ireib=0, iclin=1, azpar=az, ampar=am, itpre=1,
                                                    rtol=1e-15, atol=1e-15, vismol=0.0))
    om.set_initial_level(eps * v)
    P = np.array([[1.0, -dt * lam * az], [dt * G * H * am, 1.0]])
    Q = np.array([[1.0, dt * lam * (1 - az)], [-dt * G * H * (1 - am), 1.0]])
    zu = np.array([eps, 0.0])
    for _ in range(40):
        om.step(None)
        zu = np.linalg.solve(P, Q @ zu)
        z = om.a["znv"].astype(np.float64)
        amp = (v * M) @ z / ((v * M) @ v)
        assert abs(amp - zu[0]) < 1e-9 * eps
        assert np.abs(z - amp * v).max() < 1e-9 * eps
    # the weights matter: the amplitude after 40 steps differs between schemes
    assert abs(zu[0]) < eps


# ---------------------------------------------------------------- conz3d vertical diffusion

@pytest.mark.parametrize("ad", [1.0, 0.6])
def test_vertical_diffusion_gaussian(ad):
    """A resting basin of eight 1 m layers, tracer uniform in the horizontal
    with a Gaussian profile: each cosine mode of the profile decays by
    (1 - (1-ad) dt mu_k) / (1 + ad dt mu_k), mu_k = D/h^2 4 sin^2(pi k/2L),
    D = difv + difmol, so the discrete profile is known at every step."""
    L, D, difmol, dt = 8, 2e-3, 1e-6, 300.0
    b = _flat(nlv=L, nx=7, ny=5)
    z = np.arange(L) + 0.5
    prof = 20.0 + 5.0 * np.exp(-((z - 3.0) / 1.2) ** 2)
    c0 = np.tile(prof, (b.nkn, 1))
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=dt, ireib=0, itpre=1, adpar=ad, aapar=0.0,
                                                    difmol=difmol, rtol=1e-15, atol=1e-15), difv=D,
                              tracers=dict(salt=c0))
    Q, lam = _layer_modes(L)
    coef = np.linalg.solve(Q, prof)
    mu = (D + float(np.float64(difmol))) * lam
    g = (1 - (1 - ad) * dt * mu) / (1 + ad * dt * mu)
    for n in range(1, 6):
        om.step(None, istot=1)
        want = Q @ (coef * g ** n)
        assert np.allclose(om.tracers["salt"], want[None, :], rtol=0, atol=1e-11)
    assert abs(om.tracers["salt"].sum() - c0.sum()) < 1e-9 * c0.sum()


# ---------------------------------------------------------------- scalar_stability

def test_stability_explicit_vertical_diffusion():
    """adpar = 0 (explicit vertical diffusion): the index is 2 D / h^2 in the
    inner layers (one neighbour each at top and bottom), difmol not counted;
    istot = 1 + int(dt * index).  adpar = 1 leaves nothing explicit."""
    L, D = 5, 4e-3
    b = _flat(nlv=L, nx=7, ny=5)
    for ad, want in [(0.0, 2 * D), (1.0, 0.0)]:
        om = pyoracle.OracleModel(b, "f64", params=dict(dt=300.0, adpar=ad, aapar=0.0, itpre=1, difmol=1e-3),
                                  difv=D)
        for dt in (100.0, 300.0, 1000.0):
            istot, ri = om.scalar_stability(dt=dt)
            # the index passes through a default real (shympi_max of raux)
            assert abs(ri - dt * want) <= 2e-7 * max(1.0, dt * want)
            assert istot == 1 + int(dt * want)


def test_stability_uniform_flow_structured_basin():
    """One layer of depth h, the same transport U in x in every element, at
    both time levels, on the structured basin (cells split along (i,j)-
    (i+1,j+1)).  conzstab's advective term per element carries the last in-
    and out-flux from vertex to vertex; working it out for the two triangle
    kinds gives every vertex but the lower-left one of the upper triangle a
    term dt dx U / 2.  Summed over a node's triangles and divided by its
    volume: 2.5 U/(h dx) inside, 2 on the bottom and left walls, 3 on the top
    and right walls and the corners -- so the index is 3 U / (h dx)."""
    h, dx, U = 10.0, 250.0, 0.8
    b = _flat(nlv=1, depth=h, nx=11, ny=7)
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=300.0, itpre=1))
    om.a["utlnv"][:] = U
    om.a["utlov"][:] = U
    for dt in (60.0, 300.0, 3000.0):
        istot, ri = om.scalar_stability(dt=dt)
        want = dt * 3 * U / (h * dx)
        assert abs(ri - want) < 2e-7 * want
        assert istot == 1 + int(want)


def test_stability_point_source():
    """A discharge q into one node at rest: index q / (node volume)."""
    h = 10.0
    b = _flat(nlv=1, depth=h, nx=11, ny=7)
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=300.0, itpre=1))
    k = 5 + 11 * 3
    q = 7.5
    om.a["mfluxv"][k, 0] = q
    istot, ri = om.scalar_stability(dt=300.0)
    area = om.areakv[k, 0]
    assert abs(ri - 300.0 * q / (area * h)) < 2e-7 * ri
    om.a["mfluxv"][k, 0] = -q                                  # a sink does not count
    assert om.scalar_stability(dt=300.0) == (1, 0.0)


def test_scal3sh_takes_istot_from_stability():
    """scal3sh with istot from scalar_stability (istot <= 0) gives the same
    bits as the same number passed explicitly, and differs from one substep."""
    L, D = 5, 4e-3
    b = _flat(nlv=L, nx=7, ny=5)
    prof = 20.0 + np.arange(L, dtype=np.float64)
    c0 = np.tile(prof, (b.nkn, 1))
    kw = dict(params=dict(dt=300.0, adpar=0.0, aapar=0.0, itpre=1, rtol=1e-15, atol=1e-15), difv=D)
    runs = {}
    for istot in (0, 3, 1):
        om = pyoracle.OracleModel(b, "f64", tracers=dict(salt=c0), **kw)
        if istot == 0:
            assert om.scalar_stability()[0] == 3                  # 1 + int(300 * 2 * 4e-3)
        om.step(None, istot=istot)
        runs[istot] = om.tracers["salt"].copy()
    assert np.array_equal(runs[0], runs[3])
    assert not np.allclose(runs[0], runs[1], rtol=0, atol=1e-6)
