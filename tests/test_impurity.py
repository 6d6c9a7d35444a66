"""The impurity-He pair potentials of Select_Pot (V_impur.f90:84-929) behind
hetd_vimp / hetd_build_potion, against the reference's own Select_Pot
(compiled from V_impur.f90 by oracle/Makefile.ref, run by
tests/golden/make_golden.py into impurity_potentials.npz).

Every name the reference accepts is covered on 509 distances that include
the branch points of the fits.  The bar is bitwise: the library evaluates in
the reference's kinds and operation order (binary128 sums, binary32-rounded
default-real constants, the compiler's power expansions)."""
import ctypes
import os

import numpy as np
import pytest

import golden_util as G
from hetd import build_potion, capi


@pytest.fixture(scope="module")
def table():
    z = np.load(os.path.join(G.GOLDEN, "impurity_potentials.npz"))
    return z["r"], [str(n) for n in z["names"]], z["v"]


def vimp(L, name, r, r_cutoff=0.0, umax=0.0):
    v = ctypes.c_double()
    rc = L.hetd_vimp(name.encode(), float(r), float(r_cutoff), float(umax), ctypes.byref(v))
    assert rc == 0, name
    return v.value


def test_every_reference_potential_is_bitwise(table):
    r, names, ref = table
    L = capi.load()
    assert len(names) == 66
    bad = []
    for i, n in enumerate(names):
        ours = np.array([vimp(L, n, x) for x in r])
        if not np.array_equal(ours, ref[i]):
            k = int(np.argmax(ours != ref[i]))
            bad.append((n, float(r[k]), ours[k], ref[i][k]))
    assert not bad, bad[:5]


def test_select_pot_rules(table):
    _, names, _ = table
    L = capi.load()
    v = ctypes.c_double()
    for n in names:
        if n == "null":
            continue
        # r < r_cutoff gives umax, r = r_cutoff the potential (Select_Pot's .lt.)
        assert vimp(L, n, 2.0, 2.5, 123.25) == 123.25
        assert vimp(L, n, 2.5, 2.5, 123.25) == vimp(L, n, 2.5)
    assert vimp(L, "null", 0.1, 1.0, 5.0) == 0.0
    # Fortran Trim: trailing blanks do not matter
    assert vimp(L, "RB   ", 6.0) == vimp(L, "RB", 6.0)
    assert L.hetd_vimp(b"rb", 6.0, 0.0, 0.0, ctypes.byref(v)) == capi.HETD_ERR_ARG
    assert L.hetd_vimp(b"Pascale_Rb_6p0", 6.0, 0.0, 0.0, ctypes.byref(v)) == capi.HETD_ERR_ARG


def test_potion_table_is_select_pot_on_its_grid():
    # potenimpini (potenimpini.f90:11-25): npot = 100 nmax + 1 points on [0, 4 xmaxbox]
    L = capi.load()
    tab, d = build_potion("Ba_plus_gs", 1.05, 6.836177e3, 64, 25.6)
    assert len(tab) == 6401 and d == 4 * 25.6 / 6400
    for i in (0, 1, 2, 3, 100, 999, 6400):
        assert tab[i] == vimp(L, "Ba_plus_gs", i * d, 1.05, 6.836177e3)


def test_surface_potentials_are_bitwise():
    """Morse_HeTiO2_1D and Morse_HeTiO2_3D (morse.f90) behind hetd_vsurf,
    against the reference's own functions on the same points."""
    from hetd import vsurf
    z = np.load(os.path.join(G.GOLDEN, "impurity_potentials.npz"))
    r = z["r"]
    one = np.array([vsurf(0, 0.0, 0.0, x) for x in r])
    three = np.array([vsurf(1, 0.3 * x, 1.7 - 0.2 * x, x) for x in r])
    assert np.array_equal(one, z["morse1d"])
    assert np.array_equal(three, z["morse3d"])
    # the 1D form ignores x and y
    assert vsurf(0, 3.0, -2.0, 5.0) == vsurf(0, 0.0, 0.0, 5.0)
