"""Pins of the oracle's temperature equation (runs without a GPU).

The reference declares thermal properties and a temperature BC table (config.h:19-23, 45-63)
but never solves for temperature, so there is no reference output to match.  The build's
equation is the momentum discretisation with mu -> k / c_p and the TempBCs rows; it is pinned by
(1) bitwise identity with the independently pinned momentum matrix when k / c_p = mu and the
temperature BCs mirror the velocity BCs, (2) the exact linear conduction profile between two
Dirichlet walls, in 2D and 3D, (3) the closed-form ratio of successive differences of 1D upwind
advection-diffusion, which fixes the sign and size of the convective coefficient, and (4) the
energy switch leaving the flow untouched.
"""
import numpy as np
import pytest

from helpers import cavity, cavity3d, channel, duct3d
from incompflow import _abi as A
from oracle.oracle import Oracle


def with_energy(cfg, gam=None, types=None, vals=None, n=10):
    c = cfg.copy()
    c.energy = 1
    c.n_iter_t = n
    c.relax_t = 1.0
    c.tol_t = 0.0
    c.specific_heat = 2.0
    c.conductivity = 2.0 * (cfg.viscosity if gam is None else gam)
    for f in range(6):
        c.temp_bc_type[f] = A.TEMP_NEUMANN if types is None else types[f]
        c.temp_bc_val[f] = 0.0 if vals is None else vals[f]
    return c


def mirrored(cfg):
    """Temperature BCs that mirror the u-velocity BCs: Dirichlet with u's value on wall and inlet
    faces, Neumann on outlets."""
    types = [A.TEMP_NEUMANN if cfg.bc_type[f] == A.OUTLET else A.TEMP_DIRICHLET for f in range(6)]
    vals = [cfg.bc_val[f][0] for f in range(6)]
    return types, vals


@pytest.mark.parametrize("mk", [lambda: channel(12, 6), lambda: cavity(9, 7), lambda: duct3d()])
def test_matrix_equals_momentum_matrix_when_mirrored(mk):
    cfg = mk()
    cfg.bc_val[A.NORTH][0] = 0.3 if cfg.bc_type[A.NORTH] == A.WALL else cfg.bc_val[A.NORTH][0]
    types, vals = mirrored(cfg)
    c = with_energy(cfg, types=types, vals=vals)
    # k / c_p must equal mu in float arithmetic: specific_heat 2 and conductivity 2 mu are exact
    o = Oracle(c)
    rng = np.random.default_rng(11)
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_W, A.FIELD_UF, A.FIELD_VF, A.FIELD_WF):
        o.set(f, rng.uniform(-1, 1, c.field_size(f)))
    o.stage(A.STAGE_MOM_ASSEMBLE)           # p = 0: the momentum source is the BC rows alone
    assert np.array_equal(o.get(A.FIELD_TCOEF).view(np.uint32), o.get(A.FIELD_MCOEF).view(np.uint32))
    assert np.array_equal(o.get(A.FIELD_BT).view(np.uint32), o.get(A.FIELD_BU).view(np.uint32))


@pytest.mark.parametrize("dim", [2, 3])
def test_linear_conduction_between_dirichlet_walls(dim):
    """Fluid at rest, south wall 373, north wall 293, other faces adiabatic: the discrete solution
    is exactly linear in y (the wall row is exact for quadratics)."""
    cfg = cavity(6, 10) if dim == 2 else cavity3d(5, 8, 4)
    cfg.bc_val[A.NORTH][0] = 0.0
    cfg.bc_val[A.TOP][0] = 0.0
    types = [A.TEMP_NEUMANN] * 6
    types[A.NORTH] = types[A.SOUTH] = A.TEMP_DIRICHLET
    vals = [0.0] * 6
    vals[A.NORTH], vals[A.SOUTH] = 293.0, 373.0
    c = with_energy(cfg, gam=0.37, types=types, vals=vals, n=4000)
    o = Oracle(c, fp64=True)
    o.stage(A.STAGE_MOM_ASSEMBLE)
    o.stage(A.STAGE_SMOOTH_T)
    T = o.get(A.FIELD_TEMP)
    H = cfg.ymax - cfg.ymin
    y = (np.arange(cfg.ny) + 0.5) * H / cfg.ny
    exact = 373.0 + (293.0 - 373.0) * y / H
    np.testing.assert_allclose(T, np.broadcast_to(exact[None, :, None], T.shape), rtol=1e-10)
    res = o.stage(A.STAGE_MOM_RESIDUAL)
    assert res[3] < 1e-8


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_upwind_advection_diffusion_ratio(sign):
    """Uniform flow U along x between a Dirichlet west face and a Dirichlet east face: interior
    cells satisfy (F + 2D) T_i = (F + D) T_upstream + D T_downstream, whose solutions have
    successive-difference ratio 1 + F/D (F = rho |U| A, D = gamma A / dx) downstream."""
    cfg = channel(24, 4)
    cfg.density = 1.0
    U = 0.05 * sign
    gam = 0.01
    for f in range(6):
        cfg.bc_type[f] = A.WALL
        for q in range(3):
            cfg.bc_val[f][q] = 0.0
    cfg.bc_type[A.WEST] = A.INLET
    cfg.bc_type[A.EAST] = A.INLET
    cfg.bc_val[A.WEST][0] = U
    cfg.bc_val[A.EAST][0] = U
    types = [A.TEMP_NEUMANN] * 6
    types[A.WEST] = types[A.EAST] = A.TEMP_DIRICHLET
    vals = [0.0] * 6
    vals[A.WEST], vals[A.EAST] = 1.0, 0.0
    c = with_energy(cfg, gam=gam, types=types, vals=vals, n=6000)
    o = Oracle(c, fp64=True)
    o.set(A.FIELD_U, np.full(c.field_size(A.FIELD_U), U))
    o.set(A.FIELD_UF, np.full(c.field_size(A.FIELD_UF), U))
    o.stage(A.STAGE_MOM_ASSEMBLE)
    o.stage(A.STAGE_SMOOTH_T)
    T = o.get(A.FIELD_TEMP)[0, 1]                 # one row (uniform in y)
    dx = (cfg.xmax - cfg.xmin) / cfg.nx
    Ay = (cfg.ymax - cfg.ymin) / cfg.ny
    F, D = cfg.density * abs(U) * Ay, gam * Ay / dx
    d = np.diff(T)
    ratio = d[2:-2][1:] / d[2:-2][:-1] if sign > 0 else d[2:-2][:-1] / d[2:-2][1:]
    np.testing.assert_allclose(ratio, 1.0 + F / D, rtol=1e-8)
    assert np.all(np.diff(T) < 0)                 # monotone from 1 (west) to 0 (east)


def test_energy_off_leaves_flow_and_temperature_alone():
    cfg = channel(16, 8, 5)
    o0 = Oracle(cfg)
    o0.solve(5, history=False)
    c = with_energy(cfg, types=[A.TEMP_DIRICHLET] * 6, vals=[300.0] * 6)
    o1 = Oracle(c)
    r = o1.solve(5, history=True)
    for f in (A.FIELD_U, A.FIELD_V, A.FIELD_P, A.FIELD_UF, A.FIELD_VF):
        assert np.array_equal(o0.get(f).view(np.uint32), o1.get(f).view(np.uint32))
    assert np.all(o0.get(A.FIELD_TEMP) == np.float32(cfg.init[4]))
    assert not np.all(o1.get(A.FIELD_TEMP) == np.float32(cfg.init[4]))
    assert all(x.sweeps_t == 10 for x in r["history"]) and all(x.res_t > 0 for x in r["history"])
    with pytest.raises(Exception):
        o0.get(A.FIELD_TCOEF)


def test_scaled_residual_includes_temperature():
    cfg = cavity(12, 12)
    cfg.stop_rule = A.STOP_SCALED
    types = [A.TEMP_DIRICHLET] * 6
    vals = [300.0, 300.0, 310.0, 290.0, 0.0, 0.0]
    c = with_energy(cfg, types=types, vals=vals, n=1)
    c.relax_t = 0.05
    o = Oracle(c, fp64=True)
    r = o.solve(40, history=True)
    h = r["history"]
    norm = [max(getattr(x, k) for x in h[:5]) for k in ("res_cont", "res_u", "res_v", "res_t")]
    for x in h:
        want = max(x.res_cont / norm[0], x.res_u / norm[1], x.res_v / norm[2], x.res_t / norm[3])
        assert x.scaled == pytest.approx(want, rel=1e-12)
    # with one heavily under-relaxed sweep per iteration the temperature lags the flow and sets the
    # scaled residual
    assert any(x.res_t / norm[3] > max(x.res_cont / norm[0], x.res_u / norm[1], x.res_v / norm[2]) for x in h[5:])
