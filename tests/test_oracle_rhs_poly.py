"""The oracle's ACM right-hand side on polynomial fields, every scheme, both
forms of the nonlinear term, 2D and 3D: closed forms that pin what the
reference's regression outputs leave open (they run the skew-symmetric form
only, rhs_ACM.f90:375-450 FD2, :450-530 Tam-Webb, :534-614 FD4, :690-870 FD6).

A central first-derivative stencil of order p is exact on polynomials of
degree <= p, the matching second-derivative stencil on degree <= p + 1
(FD_4th_central_optimized uses FD4's second derivative, rhs_ACM.f90:459-470).
So on polynomial fields of a low enough degree the discrete RHS equals the
analytic one up to rounding:

    rhs_u_c = -sum_d u_d du_c/dx_d - dp/dx_c + nu lap u_c       (non-skew)
    rhs_u_c = -1/2 sum_d (d(u_c u_d)/dx_d + u_d du_c/dx_d) - ...   (skew)
    rhs_p   = -c0^2 div u - gamma p

The Tam-Webb coefficients (8 decimal digits, T&W 1993) make the first
derivative exact on quadratics only; on x^3 it returns 3x^2 + h^2 M3 with
M3 = sum_j a_j j^3 of the published table -- a closed form that pins every
coefficient.  The fields differ per component and direction (and so do the
spacings), so a transposed operand or a swapped direction fails.  Each
planted mistake in the oracle (or_acm_params.plant) fails its check.
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po
from wabbit_amd.params import ACMParams

BS, G = 8, 3
DX = (0.1, 0.07, 0.05)
X0 = (0.3, -0.2, 0.15)
# polynomial degree on which the scheme's first derivative is exact
EXACT_DEG = {2: 2, 4: 4, 40: 2, 6: 6}
SCHEMES = [2, 4, 40, 6]


def _monomials(dim, deg):
    return [e for e in itertools.product(range(deg + 1), repeat=dim) if sum(e) <= deg]


class Poly:
    """sum_k c_k prod_d x_d^e_kd"""

    def __init__(self, dim, deg, rng):
        self.terms = [(e, float(rng.uniform(-1, 1))) for e in _monomials(dim, deg)]
        self.dim = dim

    def __call__(self, X):
        out = np.zeros_like(X[0])
        for e, c in self.terms:
            t = np.full_like(X[0], c)
            for d in range(self.dim):
                t = t * X[d] ** e[d]
            out += t
        return out

    def deriv(self, X, d, k=1):
        out = np.zeros_like(X[0])
        for e, c in self.terms:
            if e[d] < k:
                continue
            f = c * np.prod([e[d] - i for i in range(k)])
            t = np.full_like(X[0], f)
            for dd in range(self.dim):
                t = t * X[dd] ** (e[dd] - (k if dd == d else 0))
            out += t
        return out


def _grid(dim):
    n = BS + 2 * G
    ax = [X0[d] + (np.arange(n) - G) * DX[d] for d in range(dim)]
    if dim == 2:
        Y, Xg = np.meshgrid(ax[1], ax[0], indexing="ij")     # arrays [y][x]
        return [Xg[None], Y[None]]
    Z, Y, Xg = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return [Xg, Y, Z]


def _case(dim, scheme, skew, seed=7, nu=0.013, c0=3.0, gamma=0.7, plant=0):
    rng = np.random.default_rng(seed + 97 * scheme + 13 * dim + int(skew))
    deg = EXACT_DEG[scheme] // (2 if skew else 1)
    X = _grid(dim)
    polys = [Poly(dim, deg, rng) for _ in range(dim + 1)]
    u = np.stack([p(X) for p in polys])[None]              # (1, neq, nz, ny, nx)
    acm = ACMParams(dim=dim, scheme=scheme, skew_symmetry=skew, penalization=False, c_0=c0, nu=nu,
                    gamma_p=gamma, domain_size=[1.0, 1.0, 1.0])
    p = po.make_params(acm)
    p.plant = plant
    got = po.rhs_block(p, [BS] * dim, G, DX[:dim], u[0])
    # analytic right-hand side
    U = [polys[c](X) for c in range(dim)]
    P = polys[dim]
    want = np.zeros_like(u[0])
    for c in range(dim):
        conv = sum(U[d] * polys[c].deriv(X, d) for d in range(dim))
        if skew:
            prod = sum(U[d] * polys[c].deriv(X, d) + U[c] * polys[d].deriv(X, d) for d in range(dim))
            conv = 0.5 * (prod + conv)
        lap = sum(polys[c].deriv(X, d, 2) for d in range(dim))
        want[c] = -conv - P.deriv(X, c) + nu * lap
    div = sum(polys[d].deriv(X, d) for d in range(dim))
    want[dim] = -c0 ** 2 * div - gamma * P(X)
    sl = (slice(None), slice(G, G + BS) if dim == 3 else slice(0, 1), slice(G, G + BS), slice(G, G + BS))
    return got[sl], want[sl]


def _err(got, want):
    return np.max(np.abs(got - want)) / np.max(np.abs(want))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("skew", [False, True])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_rhs_exact_on_polynomials(dim, scheme, skew):
    got, want = _case(dim, scheme, skew)
    assert _err(got, want) <= 1e-11


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scheme", [2, 4, 6])
def test_rhs_order_is_sharp(dim, scheme):
    """One degree more than the stencil is exact on breaks exactness: the
    check above is not passed by accident of the fields."""
    got, want = _case(dim, scheme, False, seed=11)
    deg = EXACT_DEG[scheme]
    EXACT_DEG[scheme] = deg + 1
    try:
        got1, want1 = _case(dim, scheme, False, seed=11)
    finally:
        EXACT_DEG[scheme] = deg
    assert _err(got, want) <= 1e-11
    assert _err(got1, want1) > 1e-9


# the published Tam-Webb table (T&W 1993), as the reference types it
A_TW = {-3: -0.02651995, -2: 0.18941314, -1: -0.79926643, 1: 0.79926643, 2: -0.18941314, 3: 0.02651995}


@pytest.mark.parametrize("dim", [2, 3])
def test_tam_webb_cubic_moment(dim):
    """p = x_d^3 (u = 0, gamma = 0): rhs_u_d = -dp/dx_d evaluates the
    stencil on a cubic, which returns 3x^2 + h^2 M3 (even moments and
    sum_j a_j vanish by antisymmetry, sum_j a_j j = 1)."""
    m1 = sum(a * j for j, a in A_TW.items())
    m3 = sum(a * j ** 3 for j, a in A_TW.items())
    assert abs(m1 - 1.0) < 1e-15 and abs(m3) > 1e-8
    X = _grid(dim)
    for d in range(dim):
        u = np.zeros((dim + 1,) + X[0].shape)
        u[dim] = X[d] ** 3
        acm = ACMParams(dim=dim, scheme=40, skew_symmetry=False, penalization=False, c_0=1.0, nu=0.0,
                        gamma_p=0.0, domain_size=[1.0, 1.0, 1.0])
        got = po.rhs_block(po.make_params(acm), [BS] * dim, G, DX[:dim], u)
        h = DX[d]
        want = -(3 * X[d] ** 2 + h * h * m3)
        sl = (slice(G, G + BS) if dim == 3 else slice(0, 1), slice(G, G + BS), slice(G, G + BS))
        assert np.max(np.abs(got[d][sl] - want[sl])) <= 1e-12
        # the moment term is resolved: without it the check fails
        assert np.max(np.abs(got[d][sl] + 3 * X[d][sl] ** 2)) > 100 * 1e-12


@pytest.mark.parametrize("plant,scheme,skew", [(1, 4, False), (1, 2, False), (2, 6, False), (2, 6, True),
                                               (4, 40, False), (8, 2, False)])
@pytest.mark.parametrize("dim", [2, 3])
def test_planted_mistakes_fail(plant, scheme, skew, dim):
    got, want = _case(dim, scheme, skew, plant=plant)
    assert _err(got, want) > 1e-9


# degree on which the scheme's second derivative is exact (Tam-Webb: FD4's)
LAP_DEG = {2: 3, 4: 5, 40: 5, 6: 7}


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_viscous_term_stencil(dim, scheme):
    """Only u_c = f(x_d), d != c, non-zero, p = 0: every convective product
    (u_c du_c/dx_c) and the divergence vanish identically, so rhs_u_c =
    nu f'' evaluates the second-derivative stencil alone, exact for
    deg f <= LAP_DEG and not one degree above."""
    X = _grid(dim)
    nu = 0.37
    for c in range(dim):
        d = (c + 1) % dim
        for deg, exact in ((LAP_DEG[scheme], True), (LAP_DEG[scheme] + 1, False)):
            co = np.random.default_rng(5 + deg + 10 * c).uniform(0.5, 1.0, deg + 1)
            u = np.zeros((dim + 1,) + X[0].shape)
            want = np.zeros_like(u)
            u[c] = sum(co[k] * X[d] ** k for k in range(deg + 1))
            want[c] = nu * sum(co[k] * k * (k - 1) * X[d] ** (k - 2) for k in range(2, deg + 1))
            acm = ACMParams(dim=dim, scheme=scheme, skew_symmetry=False, penalization=False, c_0=2.0, nu=nu,
                            gamma_p=0.0, domain_size=[1.0, 1.0, 1.0])
            got = po.rhs_block(po.make_params(acm), [BS] * dim, G, DX[:dim], u)
            sl = (slice(None), slice(G, G + BS) if dim == 3 else slice(0, 1), slice(G, G + BS), slice(G, G + BS))
            err = _err(got[sl], want[sl])
            assert (err <= 1e-11) if exact else (err > 1e-9), (c, deg, err)
