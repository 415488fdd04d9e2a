"""wh_wavelet_setup (host-only part of the C ABI) against the oracle's
restatement of setup_wavelet: every table entry and width identical."""
import numpy as np
import pytest

from oracle import pyoracle as po
from wabbit_amd import wavelets as W

NAMES = ["CDF20", "CDF22", "CDF40", "CDF42", "CDF44", "CDF60", "CDF62", "CDF64", "CDF66"]


@pytest.mark.parametrize("fd", [0, 1, 2, 3])
@pytest.mark.parametrize("name", NAMES)
def test_setup_matches_oracle(name, fd):
    a = W.setup(name, fd)
    b = po.wavelet(name, fd)
    for f in ("X", "Y", "lifted", "specialised", "g", "g_rhs", "order", "Nscl", "Nscr", "Nwcl",
              "Nwcr", "Nreconl", "Nreconr", "Ndep2l", "Ndep2r"):
        assert getattr(a, f) == getattr(b, f), f
    for filt in ("HD", "GD", "HR", "GR"):
        la, ca = a.filt(filt)
        lb, cb = b.filt(filt)
        assert la == lb and np.array_equal(ca, cb), filt


@pytest.mark.parametrize("bad", ["CDF24", "CDF80", "XYZ44", "CDF4", "CDF444"])
def test_setup_rejects(bad):
    with pytest.raises(Exception):
        W.setup(bad)
