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


# ---------------------------------------------------------------------------
# 3D: create_mask_3D_acm "sphere-fixed" (draw_sphere, module_geometry.f90:
# 49-117) and sponge_3D (sponge.f90:61-233).  The plane z = z_c through the
# sphere's centre must equal the 2D circle of the same centre and radius bit
# for bit (the z term adds an exact 0), and for the p-norm sponge the plane
# z = L/2 must equal the 2D p-norm sponge (pinned by the reference's saved
# masks); the rect sponge equals the 2D one on planes farther than L_sponge
# from the z walls and is symmetric under z -> L - z.
# ---------------------------------------------------------------------------
BS3, G3 = 12, 4


def _block3(mp, x0, dx):
    return po.mask_block(mp, 3, [BS3] * 3, G3, x0, dx)


def _block2(mp, x0, dx):
    return po.mask_block(mp, 2, [BS3] * 2, G3, x0[:2], dx[:2])


def test_sphere_centre_plane_is_the_circle():
    dx = (1.0 / 48, 1.0 / 48, 1.0 / 48)
    c = (0.5, 0.5, 0.5)
    R, h = 0.15, 1.5 * dx[0]
    x0 = (0.375, 0.375, 0.375)            # the centre plane z = 0.5 is interior index G3 + 6
    kz = G3 + int(round((c[2] - x0[2]) / dx[2]))
    mp3 = po.mask_params(geometry=2, x_cntr=c, R_cyl=R, smoothing_width=h, smoothing_safety=h,
                         domain=(1.0, 1.0, 1.0))
    mp2 = po.mask_params(geometry=1, x_cntr=c, R_cyl=R, smoothing_width=h, smoothing_safety=h,
                         domain=(1.0, 1.0, 1.0))
    m3 = _block3(mp3, x0, dx)
    m2 = _block2(mp2, x0, dx)
    assert np.array_equal(m3[0, kz], m2[0, 0]) and np.array_equal(m3[4, kz], m2[4, 0])
    assert np.any((m3[0] > 0) & (m3[0] < 1)) and np.any(m3[0] == 1.0)
    # radial: chi depends on |x - c| only -- the octant mirror images agree
    n = BS3 + 2 * G3
    ar = (np.arange(n) - G3) * dx[0] + x0[0]
    Z, Y, X = np.meshgrid(ar, ar, ar, indexing="ij")
    r = np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    inner = (slice(G3, G3 + BS3 + 1),) * 3
    chi, rr = m3[0][inner], r[inner]
    assert np.all(chi[rr <= R - h] == 1.0) and np.all(chi[rr >= R + h] == 0.0)
    ic = G3 + 6
    for k in range(1, 6):
        a = m3[0, ic + k, ic, ic]
        assert a == m3[0, ic - k, ic, ic] == m3[0, ic, ic + k, ic] == m3[0, ic, ic, ic - k]
    # no sponge: component 5 is 0; u_s is 0; colour 1
    assert np.all(m3[5] == 0.0) and np.all(m3[1:4] == 0.0) and np.all(m3[4] == 1.0)


def test_sponge_3d_planes_are_the_2d_sponges():
    dx = (1.0 / 64, 1.0 / 64, 1.0 / 64)
    L = 1.0
    # p-norm on the plane z = L/2
    mp = po.mask_params(use_sponge=True, sponge_type=1, L_sponge=0.2, p_sponge=8.0, domain=(L, L, L))
    for x0 in ((0.0, 0.0, 0.40625), (0.8125, 0.625, 0.40625)):
        kz = G3 + int(round((0.5 - x0[2]) / dx[2]))
        m3 = _block3(mp, x0, dx)
        m2 = _block2(mp, x0, dx)
        assert np.array_equal(m3[5, kz], m2[5, 0])
        assert np.any(m3[5] > 0.0)
    # rect: planes farther than L_sponge from both z walls equal the 2D
    # sponge of a domain with the same x, y lengths
    mp = po.mask_params(use_sponge=True, sponge_type=0, L_sponge=0.125, domain=(L, L, L))
    x0 = (0.0, 0.8125, 0.375)
    m3 = _block3(mp, x0, dx)
    m2 = _block2(mp, x0, dx)
    for k in range(G3, G3 + BS3):
        assert np.array_equal(m3[5, k], m2[5, 0])
    # symmetric under z -> L - z: the blocks at z0 and L - z0 - Bs dx mirror
    a = _block3(mp, (0.0, 0.0, 0.0), dx)[5, G3:G3 + BS3]
    b = _block3(mp, (0.0, 0.0, L - BS3 * dx[2]), dx)[5, G3:G3 + BS3]
    assert np.array_equal(a[1:], b[::-1][:-1])
    assert np.all(a[0][G3:G3 + BS3, G3:G3 + BS3] == 1.0)          # on the wall z = 0
