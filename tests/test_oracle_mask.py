"""Oracle: the rectangular sponge and the cylinder of or_mask_block on a
non-square domain (the channel workload's 4 x 2), checked by properties the
definitions fix (sponge.f90:15-29, module_geometry.f90:315-381): the
reference's saved masks pin only the p-norm sponge on a square domain.

- the rect sponge is mirror-symmetric in x about Lx/2 and in y about Ly/2
  (a swapped domain length or axis breaks one of these on a 4 x 2 domain);
- it is 1 on the domain boundary lines x = 0, y = 0 and 0 at points farther
  than L_sponge from every boundary;
- it is monotone non-increasing moving away from the nearest boundary;
- chi of the cylinder is 1 inside R - h, 0 outside R + h, rotationally
  symmetric about the centre when the centre is on a grid point.
"""
import numpy as np

from oracle import pyoracle as po

BS, G, J = 16, 6, 3
LX, LY = 4.0, 2.0


def global_mask(mp):
    n1 = 2 ** J
    dx = (LX / (n1 * BS), LY / (n1 * BS))
    N = n1 * BS
    out = np.zeros((6, N, N))
    for by in range(n1):
        for bx in range(n1):
            x0 = (bx * BS * dx[0], by * BS * dx[1], 0.0)
            m = po.mask_block(mp, 2, [BS, BS], G, x0, (dx[0], dx[1], 1.0))
            out[:, by * BS:(by + 1) * BS, bx * BS:(bx + 1) * BS] = m[:, 0, G:G + BS, G:G + BS]
    return out, dx


def test_rect_sponge_properties():
    L = 0.25
    mp = po.mask_params(use_sponge=True, sponge_type=0, L_sponge=L, domain=(LX, LY, 1.0))
    m, dx = global_mask(mp)
    s = m[5]
    N = s.shape[0]
    i = np.arange(1, N)
    assert np.array_equal(s[:, i], s[:, N - i])          # x -> Lx - x
    assert np.array_equal(s[i, :], s[N - i, :])          # y -> Ly - y
    assert np.all(s[0, :] == 1.0) and np.all(s[:, 0] == 1.0)
    x = np.arange(N) * dx[0]
    y = np.arange(N) * dx[1]
    X, Y = np.meshgrid(x, y)
    dist = np.minimum.reduce([X, Y, LX - X, LY - Y])
    assert np.all(s[dist >= L] == 0.0)
    assert np.all((s > 0.0) & (s < 1.0) == ((dist > 0.0) & (dist < L)))
    # monotone along a row (y = Ly/2) away from x = 0
    row = s[N // 2, : N // 2]
    assert np.all(np.diff(row) <= 0.0)
    # the colour is 1 everywhere, no geometry: chi = 0
    assert np.all(m[4] == 1.0) and np.all(m[0] == 0.0)


def test_cylinder_on_channel():
    n1 = 2 ** J
    dx = (LX / (n1 * BS), LY / (n1 * BS))
    xc = (1.0, 1.0, 0.0)                                 # on a grid point
    R, h = 0.15, 1.5 * min(dx)
    mp = po.mask_params(geometry=1, x_cntr=xc, R_cyl=R, smoothing_width=h, smoothing_safety=h,
                        domain=(LX, LY, 1.0))
    m, _ = global_mask(mp)
    chi = m[0]
    N = chi.shape[0]
    X, Y = np.meshgrid(np.arange(N) * dx[0], np.arange(N) * dx[1])
    r = np.sqrt((X - xc[0]) ** 2 + (Y - xc[1]) ** 2)
    assert np.all(chi[r <= R - h] == 1.0)
    assert np.all(chi[r >= R + h] == 0.0)
    assert np.any((chi > 0) & (chi < 1))
    # symmetric under x -> 2 xc - x and y -> 2 yc - y
    ic, jc = int(round(xc[0] / dx[0])), int(round(xc[1] / dx[1]))
    k = np.arange(1, 40)
    assert np.allclose(chi[jc, ic + k], chi[jc, ic - k], rtol=0, atol=1e-15)
    assert np.allclose(chi[jc + k, ic], chi[jc - k, ic], rtol=0, atol=1e-15)
