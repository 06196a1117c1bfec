"""Pins of the oracle's discrete adjoint (oracle/adjoint_oracle.hpp), on the CPU.

* transpose: <G dx + G_t dth, l> = <dx, G^T l> + <dth, G_t^T l> for random vectors, on 2D and 3D
  cases whose face velocities take both signs (every upwind branch) and with wall, inlet and
  outlet faces;
* tangent: central differences of the oracle's own nonlinear outer iteration (the reference's
  algorithm) at a converged FP64 state;
* gradient: central differences of converged FP64 solves, and the Stokes-limit identities
  mu dJ/dmu = J and sum_j theta_j dJ/dtheta_j = J (the pressure drop of creeping flow is linear
  in mu and in the inflow), which a dropped term or a wrong sign breaks.
"""
import numpy as np
import pytest

from incompflow import _abi
from oracle.oracle import Oracle

from helpers import cavity, cavity3d, duct3d


def layout(cfg):
    """Slices of the flat adjoint vectors and masks of their state / parameter entries."""
    nx, ny, nz, dim = cfg.nx, cfg.ny, cfg.nz, cfg.dim
    N = nx * ny * nz
    nf = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
    mask = np.zeros(4 * N + sum(nf), bool)
    for d in range(dim):
        mask[d * N:(d + 1) * N] = True
    mask[3 * N:4 * N] = True
    off = 4 * N
    for d in range(3):
        if d < dim:
            f = np.zeros((nz + (d == 2), ny + (d == 1), nx + (d == 0)), bool)
            if d == 0:
                f[:, :, 1:nx] = True
            elif d == 1:
                f[:, 1:ny, :] = True
            else:
                f[1:nz, :, :] = True
            mask[off:off + nf[d]] = f.ravel()
        off += nf[d]
    faces = [ny * nz, ny * nz, nx * nz, nx * nz, nx * ny, nx * ny]
    pmask = np.zeros(sum(faces) + 1, bool)
    poff = np.cumsum([0] + faces)
    for f in range(2 * dim):
        if cfg.bc_type[f] == _abi.INLET:
            pmask[poff[f]:poff[f + 1]] = True
    pmask[-1] = True
    return mask, pmask, poff


def channel_case(nx=16, ny=8, mu=1e-3, noise=1e-3, seed=3):
    c = _abi.adjoint_channel_config(nx, ny)
    c.viscosity = mu
    H = c.ymax - c.ymin
    y = (np.arange(ny) + 0.5) * (H / ny)
    rng = np.random.default_rng(seed)
    prof = 4 * 0.02 * y * (H - y) / H ** 2 + noise * 0.02 * rng.standard_normal(ny)
    return c, prof


def converged(c, prof=None, tol=1e-13, cap=200000):
    c = c.copy()
    c.stop_rule = _abi.STOP_SCALED
    c.tol_residual = tol
    c.num_outer_iter = cap
    o = Oracle(c, fp64=True)
    if prof is not None:
        o.set_face_profile(_abi.WEST, prof)
    r = o.solve(cap, history=False)
    assert r["converged"], "FP64 solve did not converge"
    return o


def pressure_drop(o, c):
    p = o.get(_abi.FIELD_P)
    return float(p[:, :, 0].mean() - p[:, :, c.nx - 1].mean())


# ---------------- transpose ----------------

def _unconverged(c, iters, prof=None):
    o = Oracle(c, fp64=True)
    if prof is not None:
        o.set_face_profile(_abi.WEST, prof)
    o.solve(iters, history=False)
    return o


def _dot_case(o, c, seed):
    o.adj_setup(1)                           # linear objective, zero weights
    mask, pmask, _ = layout(c)
    rng = np.random.default_rng(seed)
    dx = rng.standard_normal(o.adj_nvec) * mask
    dth = rng.standard_normal(o.adj_npar) * pmask
    lam = rng.standard_normal(o.adj_nvec) * mask
    t = o.adj_tangent(dx, dth)
    xb, tb = o.adj_transpose(lam)
    lhs = float(np.dot(t * mask, lam))
    rhs = float(np.dot(dx, xb * mask) + np.dot(dth, tb * pmask))
    scale = float(np.abs(t * mask * lam).sum() + np.abs(dx * xb * mask).sum())
    return lhs, rhs, scale, o


@pytest.mark.parametrize("name", ["channel", "cavity", "cavity3d", "duct3d", "outlet_north"])
def test_transpose_is_the_adjoint_of_the_tangent(name):
    if name == "channel":
        c, prof = channel_case(12, 8)
        o = _unconverged(c, 25, prof)
    elif name == "cavity":
        c = cavity(10, 8, re=400.0)
        o = _unconverged(c, 15)
    elif name == "cavity3d":
        c = cavity3d(6, 5, 4)
        o = _unconverged(c, 6)
    elif name == "duct3d":
        c = duct3d(8, 5, 4)
        o = _unconverged(c, 6)
    else:                                   # inlet west, outlets east and north, moving south wall
        c = cavity(10, 8, re=200.0)
        c.bc_type[_abi.WEST] = _abi.INLET
        c.bc_val[_abi.WEST][0] = 0.5
        c.bc_type[_abi.EAST] = _abi.OUTLET
        c.bc_type[_abi.NORTH] = _abi.OUTLET
        c.bc_val[_abi.NORTH][0] = 0.0
        c.bc_val[_abi.SOUTH][0] = 0.3
        o = _unconverged(c, 12)
    for seed in (1, 2):
        lhs, rhs, scale, _ = _dot_case(o, c, seed)
        assert abs(lhs - rhs) <= 1e-12 * scale, (name, seed, lhs, rhs)


def test_transpose_catches_a_planted_sign_error():
    """A perturbed transpose (scaled adjoint of one part) fails the dot test: the identity is
    not satisfied by accident."""
    c, prof = channel_case(12, 8)
    o = _unconverged(c, 25, prof)
    o.adj_setup(1)
    mask, pmask, _ = layout(c)
    rng = np.random.default_rng(5)
    dx = rng.standard_normal(o.adj_nvec) * mask
    dth = rng.standard_normal(o.adj_npar) * pmask
    lam = rng.standard_normal(o.adj_nvec) * mask
    t = o.adj_tangent(dx, dth)
    xb, tb = o.adj_transpose(lam)
    N = c.nx * c.ny
    xb_bad = xb.copy()
    xb_bad[3 * N:4 * N] *= -1.0                  # wrong sign on the pressure adjoint
    lhs = float(np.dot(t * mask, lam))
    bad = float(np.dot(dx, xb_bad * mask) + np.dot(dth, tb * pmask))
    assert abs(lhs - bad) > 1e-6 * abs(lhs)


# ---------------- tangent vs the nonlinear outer iteration ----------------

def _outer_from(c, mu, prof, fields):
    c2 = c.copy()
    c2.viscosity = mu
    c2.stop_rule = _abi.STOP_NONE
    o = Oracle(c2, fp64=True)
    o.set_face_profile(_abi.WEST, prof)
    for f, v in fields.items():
        o.set(f, v)
    o.set_iter(100)
    o.stage(_abi.STAGE_OUTER)
    return o


STATE = (_abi.FIELD_U, _abi.FIELD_V, _abi.FIELD_W, _abi.FIELD_P, _abi.FIELD_UF, _abi.FIELD_VF, _abi.FIELD_WF)


def _flat(o):
    return np.concatenate([o.get(f, shaped=False) for f in STATE])


def test_tangent_matches_central_differences_of_one_outer_iteration():
    c, prof = channel_case(16, 8)
    base = converged(c, prof)
    fields = {f: base.get(f, shaped=False) for f in STATE}
    x0 = _flat(base)
    base.adj_setup(1)
    mask, pmask, poff = layout(c)
    rng = np.random.default_rng(11)
    # relative perturbations: no face velocity changes sign (the upwind switch min(uf, 0) has a
    # kink at 0, where central differences do not converge to a derivative)
    dx = rng.standard_normal(base.adj_nvec) * mask * np.abs(x0)
    dth = np.zeros(base.adj_npar)
    dth[poff[_abi.WEST]:poff[_abi.WEST + 1]] = 0.02 * rng.standard_normal(c.ny)
    dth[-1] = 1e-3 * 0.3
    t = base.adj_tangent(dx, dth)
    h = 1e-6
    outs = []
    for sgn in (+1, -1):
        fp = {}
        off = 0
        for f in STATE:
            n = fields[f].size
            fp[f] = fields[f] + sgn * h * dx[off:off + n]
            off += n
        o = _outer_from(c, c.viscosity + sgn * h * dth[-1], prof + sgn * h * dth[poff[_abi.WEST]:poff[_abi.WEST + 1]], fp)
        outs.append(_flat(o))
    fd = (outs[0] - outs[1]) / (2 * h)
    err = np.abs((fd - t) * mask).max() / np.abs(t * mask).max()
    assert err < 1e-6, err


# ---------------- gradients ----------------

def _adjoint_gradient(o, tol=1e-12, chunk=250, cap=200):
    o.adj_setup(0)
    for _ in range(cap):
        if o.adj_iterate(chunk) < tol:
            break
    else:
        raise AssertionError("adjoint did not converge")
    return o.adj_gradient()


def test_gradient_matches_central_differences_of_converged_solves():
    c, prof = channel_case(16, 8, noise=1e-2)
    base = converged(c, prof)
    J0 = pressure_drop(base, c)
    g = _adjoint_gradient(base)
    _, _, poff = layout(c)
    gW = g[poff[_abi.WEST]:poff[_abi.WEST + 1]]
    # mu
    h = 1e-4 * c.viscosity
    Jp = pressure_drop(converged(_with_mu(c, c.viscosity + h), prof), c)
    Jm = pressure_drop(converged(_with_mu(c, c.viscosity - h), prof), c)
    fd_mu = (Jp - Jm) / (2 * h)
    assert abs(g[-1] - fd_mu) <= 1e-5 * abs(fd_mu), (g[-1], fd_mu)
    # three profile entries (bottom, middle, top row)
    for j in (0, c.ny // 2, c.ny - 1):
        hp = 1e-4 * 0.02
        pp, pm = prof.copy(), prof.copy()
        pp[j] += hp
        pm[j] -= hp
        fd = (pressure_drop(converged(c, pp), c) - pressure_drop(converged(c, pm), c)) / (2 * hp)
        assert abs(gW[j] - fd) <= 1e-5 * np.abs(gW).max(), (j, gW[j], fd)
    assert J0 > 0


def _with_mu(c, mu):
    c2 = c.copy()
    c2.viscosity = mu
    return c2


def test_stokes_limit_identities():
    """Creeping flow (Re ~ 1e-3): J is linear in mu and in the inflow, so mu dJ/dmu = J and
    sum_j theta_j dJ/dtheta_j = J to O(Re)."""
    c, prof = channel_case(16, 8, mu=1.0, noise=1e-2)
    o = converged(c, prof)
    J = pressure_drop(o, c)
    g = _adjoint_gradient(o)
    _, _, poff = layout(c)
    gW = g[poff[_abi.WEST]:poff[_abi.WEST + 1]]
    assert abs(c.viscosity * g[-1] - J) <= 1e-3 * abs(J), (c.viscosity * g[-1], J)
    assert abs(float(np.dot(prof, gW)) - J) <= 1e-3 * abs(J), (float(np.dot(prof, gW)), J)


def test_linear_objective_gradient_on_a_cavity():
    """J = sum of u over the middle column (weights on u): d/dmu against central differences."""
    c = cavity(12, 12, re=50.0)
    c.n_iter[3] = 40
    base = converged(c, None)
    mask, _, _ = layout(c)
    N = c.nx * c.ny
    w = np.zeros(mask.size)
    wu = np.zeros((c.ny, c.nx))
    wu[:, c.nx // 2] = 1.0
    w[:N] = wu.ravel()

    def J(o):
        return float(o.get(_abi.FIELD_U)[0][:, c.nx // 2].sum())

    base.adj_setup(1)
    base.adj_set_weights(w)
    for _ in range(400):
        if base.adj_iterate(250) < 1e-12:
            break
    g = base.adj_gradient()
    h = 1e-4 * c.viscosity
    fd = (J(converged(_with_mu(c, c.viscosity + h))) - J(converged(_with_mu(c, c.viscosity - h)))) / (2 * h)
    assert abs(g[-1] - fd) <= 1e-5 * abs(fd), (g[-1], fd)
