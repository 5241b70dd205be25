"""Closed-form pins of the oracle's right-hand side (src/cu/pde.cu:139-787,
src/cu/turbulence.cu:255-454, src/cu/pde.cu:14-99).

The expected values are not stencils applied to grid arrays: they are the
symbols of the difference operators on trigonometric fields (e.g. the central
difference of sin(t) is cos(t) sin(h)/dx), exact values of polynomial fields
on a stretched z-grid, and closed forms of the source terms.  A wrong
neighbour index, sign, weight or dropped term changes them."""
import numpy as np
import pytest

from conftest import make_params

from oracle import oracle as orc

SIGMA = 5.670374419e-8


def _state(P, u=None, v=None, w=None, T=None, Y=None):
    shape = (P.Nx, P.Ny, P.Nz)
    f = [np.zeros(shape) if a is None else np.broadcast_to(a, shape).astype(float) for a in (u, v, w)]
    f.append(np.full(shape, P.T_inf) if T is None else np.broadcast_to(T, shape).astype(float))
    Yf = np.zeros((P.Nx, P.Ny, P.Nz_Y_max)) if Y is None else np.broadcast_to(Y, (P.Nx, P.Ny, P.Nz_Y_max)).astype(float)
    return np.concatenate([a.ravel() for a in f] + [Yf.ravel()])


def _rhs(P, y):
    Rn = np.zeros_like(y)
    Rt = np.zeros(18 * P.N + P.Nx * P.Ny)
    orc.Oracle(P.ptr).rhs(y, Rn, Rt)
    return Rn, Rt


def _plane(P, Rt, name):
    ti = P.p.turbulence_indexes
    off = getattr(ti, name)
    n = P.Nx * P.Ny if name == "u_tau_wall" else P.N
    a = Rt[off:off + n]
    return a.reshape(P.Nx, P.Ny) if name == "u_tau_wall" else a.reshape(P.Nx, P.Ny, P.Nz)


def test_rest_state_is_steady(tmp_path):
    """Uniform T = T_inf at rest: every tendency vanishes up to the rounding of
    the non-uniform-z stencil weights (their sum is zero only in exact
    arithmetic: about 1e-12 K/m^2 for a 300 K field)."""
    P = make_params(tmp_path, "small_flat")
    y = _state(P)
    R = orc.Oracle(P.ptr).phi(y)
    u, v, w, T, Y = P.split(R)
    assert not u.any() and not v.any() and not Y.any()
    assert np.abs(w).max() <= 1e-20 and np.abs(T).max() <= 1e-15


def test_buoyancy_of_uniform_hot_gas(tmp_path):
    P = make_params(tmp_path, "small_flat")
    T1 = 450.0
    y = _state(P, T=T1)
    R = orc.Oracle(P.ptr).phi(y)
    u, v, w, T, _ = P.split(R)
    rho = P.T_inf * P.rho_inf / T1
    ref = P.g * (rho - P.rho_inf) / rho
    assert np.allclose(w[:, :, 1:-1], ref, rtol=1e-14, atol=0)
    assert np.allclose(w[:, :, -1], ref, rtol=1e-13, atol=0)       # top: zero-gradient extrapolation
    assert not w[:, :, 0].any() and not u.any() and not v.any()    # wall node
    assert np.abs(T).max() <= 1e-9


def test_upwind_advection_diffusion_radiation_in_x(tmp_path):
    """T = T_inf + a sin(t), t = 2 pi m x/Lx, transported by u = U > 0."""
    P = make_params(tmp_path, "small_flat")
    Lx, dx = P.x_max - P.x_min, P.dx
    m, a, U = 2, 6.0, 1.7
    t = 2 * np.pi * m * P.x / Lx
    h = 2 * np.pi * m * dx / Lx
    T = (P.T_inf + a * np.sin(t))[:, None, None]
    y = _state(P, u=U, T=T)
    Rn, Rt = _rhs(P, y)
    s, c = np.sin(t), np.cos(t)
    Tx = a * c * np.sin(h) / dx
    Txx = -4 * a * s * np.sin(h / 2) ** 2 / dx ** 2
    # backward one-sided 2nd-order difference (3T_i - 4T_{i-1} + T_{i-2})/(2dx)
    T_im = a * (s * (3 - 4 * np.cos(h) + np.cos(2 * h)) + c * (4 * np.sin(h) - np.sin(2 * h))) / (2 * dx)
    Tn = P.T_inf + a * s
    rho = P.T_inf * P.rho_inf / Tn
    kappa = P.k + 4 * SIGMA * P.delta * Tn ** 3
    ref = (12 * SIGMA * P.delta * Tn ** 2 * Tx ** 2 + kappa * Txx) / (rho * P.c_p) - U * T_im
    got = P.split(Rn)[3]
    scale = np.abs(ref).max()
    for k in range(1, P.Nz - 1):
        assert np.allclose(got[:, :, k], ref[:, None], rtol=0, atol=1e-12 * scale)
    assert np.allclose(_plane(P, Rt, "Tx")[:, 3, 5], Tx, rtol=0, atol=1e-12 * np.abs(Tx).max())
    assert np.allclose(_plane(P, Rt, "Txx")[:, 3, 5], Txx, rtol=0, atol=1e-11 * np.abs(Txx).max())


def test_downwind_side_and_y_direction(tmp_path):
    """u = -U < 0 selects the forward one-sided difference; the same field
    along y with v > 0 exercises the y stencils."""
    P = make_params(tmp_path, "small_flat")
    Ly, dy = P.y_max - P.y_min, P.dy
    m, a, U = 1, 4.0, 0.9
    t = 2 * np.pi * m * P.y / Ly
    h = 2 * np.pi * m * dy / Ly
    y = _state(P, u=-U, v=U, T=(P.T_inf + a * np.sin(t))[None, :, None])
    Rn, Rt = _rhs(P, y)
    s, c = np.sin(t), np.cos(t)
    T_jm = a * (s * (3 - 4 * np.cos(h) + np.cos(2 * h)) + c * (4 * np.sin(h) - np.sin(2 * h))) / (2 * dy)
    Ty = a * c * np.sin(h) / dy
    Tyy = -4 * a * s * np.sin(h / 2) ** 2 / dy ** 2
    Tn = P.T_inf + a * s
    rho = P.T_inf * P.rho_inf / Tn
    ref = (12 * SIGMA * P.delta * Tn ** 2 * Ty ** 2 + (P.k + 4 * SIGMA * P.delta * Tn ** 3) * Tyy) / (rho * P.c_p)
    ref = ref - U * T_jm       # u < 0 has no x-gradient to advect
    got = P.split(Rn)[3][4, :, 3]
    assert np.allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("z_domain", [0, 1])
def test_z_stencils_exact_for_polynomials(tmp_path, z_domain):
    """T = T_inf + c1 z + c2 z^2 + c3 z^3 on a stretched grid: the 4-point
    one-sided second derivatives at k = 0 and k = Nz-1 are exact for cubics,
    the interior 3-point second derivative and the one-sided first
    derivatives are exact for quadratics (src/cu/pde.cu:485-575)."""
    over = dict(z_domain=1, k_nu_grid=1.8) if z_domain else {}
    P = make_params(tmp_path, "small_flat", **over)
    z = P.z
    c1, c2, c3 = 3.0, -0.8, 0.05
    y = _state(P, T=(P.T_inf + c1 * z + c2 * z ** 2 + c3 * z ** 3)[None, None, :])
    _, Rt = _rhs(P, y)
    Tzz = _plane(P, Rt, "Tzz")[2, 3]
    exact = 2 * c2 + 6 * c3 * z
    assert abs(Tzz[0] - exact[0]) <= 1e-9 and abs(Tzz[-1] - exact[-1]) <= 1e-9
    y = _state(P, T=(P.T_inf + c1 * z + c2 * z ** 2)[None, None, :])
    _, Rt = _rhs(P, y)
    Tz = _plane(P, Rt, "Tz")[2, 3]
    Tzz = _plane(P, Rt, "Tzz")[2, 3]
    assert np.allclose(Tzz, 2 * c2, rtol=0, atol=1e-9)
    assert abs(Tz[0] - c1) <= 1e-10 and abs(Tz[-1] - (c1 + 2 * c2 * z[-1])) <= 1e-10
    if z_domain == 0:
        assert np.allclose(Tz[1:-1], c1 + 2 * c2 * z[1:-1], rtol=0, atol=1e-10)


def test_viscous_term_and_wall_shear(tmp_path):
    """u = c1 z + c2 z^2 with C_s = 0 (no SGS viscosity): u_RHS = nu u_zz at
    interior nodes and u_tau = sqrt(mu |u_z(0)| / rho) on the wall plane."""
    P = make_params(tmp_path, "small_flat", C_s=0.0, z_domain=1, k_nu_grid=1.5)
    z = P.z
    c1, c2 = 0.7, 0.2
    y = _state(P, u=(c1 * z + c2 * z ** 2)[None, None, :])
    Rn, Rt = _rhs(P, y)
    nu = P.mu / (P.T_inf * P.rho_inf / P.T_inf)
    got = P.split(Rn)[0][5, 6]
    assert np.allclose(got[1:-1], nu * 2 * c2, rtol=1e-9, atol=0)
    ut = _plane(P, Rt, "u_tau_wall")
    assert np.allclose(ut, np.sqrt(P.mu * c1 / P.rho_inf), rtol=1e-12)


def test_strain_and_smagorinsky_viscosity(tmp_path):
    """u = a sin(2 pi y / Ly): S_12 = u_y / 2, |S| = |u_y|, every other strain
    entry zero, mu_sgs = rho (C_s (dx dy dz)^(1/3))^2 |S| (src/cu/pde.cu:632-745)."""
    P = make_params(tmp_path, "small_flat")
    Ly, dy = P.y_max - P.y_min, P.dy
    a = 1.5
    t = 2 * np.pi * P.y / Ly
    y = _state(P, u=(a * np.sin(t))[None, :, None])
    _, Rt = _rhs(P, y)
    uy = a * np.cos(t) * np.sin(2 * np.pi * dy / Ly) / dy
    S12 = _plane(P, Rt, "S_12")[3, :, 4]
    assert np.allclose(S12, 0.5 * uy, rtol=0, atol=1e-14)
    for name in ("S_11", "S_22", "S_33", "S_13", "S_23"):
        assert np.abs(_plane(P, Rt, name)).max() <= 1e-14, name
    dz = P.z[5] - P.z[4]
    Delta2 = (P.C_s * (P.dx * P.dy * dz) ** (1.0 / 3.0)) ** 2
    mu = _plane(P, Rt, "mu_sgs")[3, :, 4]
    assert np.allclose(mu, P.rho_inf * Delta2 * np.abs(uy), rtol=1e-12, atol=1e-15)


def test_arrhenius_fuel_consumption_and_heat_release(tmp_path):
    """Uniform T = 700 K > T_pc with Y = 0.5 in the fuel layer:
    dY/dt = -Y_f A exp(-T_act/T) Y and T_RHS = H_C Y K(T) / c_p on fuel nodes
    (no gradients); zero fuel RHS where T <= T_pc."""
    P = make_params(tmp_path, "small_flat")
    T1, Y1 = 700.0, 0.5
    y = _state(P, T=T1, Y=Y1)
    Rn, _ = _rhs(P, y)
    K = P.A * np.exp(-P.T_act / T1)
    Yr = P.split(Rn)[4]
    nzy = P.Nz_Y.reshape(P.Nx, P.Ny)
    mask = np.arange(P.Nz_Y_max)[None, None, :] < nzy[:, :, None]
    assert np.allclose(Yr[mask], -P.Y_f * K * Y1, rtol=1e-14)
    Tr = P.split(Rn)[3]
    assert np.allclose(Tr[:, :, 1][nzy > 1], P.H_C * Y1 * K / P.c_p, rtol=1e-12)
    y = _state(P, T=500.0, Y=Y1)
    Rn, _ = _rhs(P, y)
    assert not np.any(P.split(Rn)[4][mask])


def test_cut_node_and_top_conditions(tmp_path):
    """Phi on the hill case: u = v = w = 0 at each column's cut node, dead-node
    values below it, and the zero-gradient extrapolations of T above the cut
    node and of u,v,w,T below the top (src/cu/pde.cu:45-97)."""
    P = make_params(tmp_path, "small_hill")
    rng = np.random.default_rng(2)
    y = _state(P, u=1.0 + 0.1 * rng.standard_normal((P.Nx, P.Ny, P.Nz)),
               T=P.T_inf + 20 * rng.random((P.Nx, P.Ny, P.Nz)))
    R = orc.Oracle(P.ptr).phi(y)
    u, v, w, T, Y = P.split(R)
    z = P.z
    cut = P.cut_nodes.reshape(P.Nx, P.Ny)
    for i, j in ((0, 0), (12, 10), (11, 9), (5, 14)):
        k = cut[i, j]
        assert u[i, j, k] == 0 and v[i, j, k] == 0 and w[i, j, k] == 0
        assert np.all(T[i, j, :k] == P.T_dead_nodes)
        h0, h1 = z[k + 1] - z[k], z[k + 2] - z[k + 1]
        # the extrapolation makes the one-sided derivative at the cut node vanish
        d = -(2 * h0 + h1) * T[i, j, k] / (h0 * (h0 + h1)) + (h0 + h1) * T[i, j, k + 1] / (h0 * h1) \
            - h0 * T[i, j, k + 2] / (h1 * (h0 + h1))
        assert abs(d) <= 1e-9 * np.abs(T[i, j, k:k + 3]).max()
        hm1, hm2 = z[-1] - z[-2], z[-2] - z[-3]
        for f in (u, v, w, T):
            d = hm1 * f[i, j, -3] / (hm2 * (hm2 + hm1)) - (hm2 + hm1) * f[i, j, -2] / (hm2 * hm1) \
                + (hm2 + 2 * hm1) * f[i, j, -1] / (hm1 * (hm2 + hm1))
            assert abs(d) <= 1e-9 * max(1.0, np.abs(f[i, j, -3:]).max())
    assert cut.max() >= 2
