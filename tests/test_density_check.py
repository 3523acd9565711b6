"""The density checks every engine runs before a build (host logic).

The pair lists drop blocks whose bound Qmax * Qglob * d_bound is below the
threshold, which matches the reference's per-quartet screening only while
max|D| <= d_bound; a larger density must be refused, not silently screened.
A non-symmetric density must be refused too (the engines symmetrise D and some
exchange classes read D_jl as D[l][j])."""

import numpy as np
import pytest

import _util as U
from vlxmi import configs
from vlxmi import pairs as P


@pytest.fixture(scope="module")
def small8():
    mol, D = configs.make_small8()
    fl = P.build_lists(mol, U.oracle_schwarz)
    return mol, D, fl


def _stats(mol, D):
    Dc = U.cart_density(mol, D)
    return float(np.abs(Dc).max()), float(np.abs(D - D.T).max()), float(np.abs(D).max())


def test_density_within_bound_accepted(small8):
    mol, D, fl = small8
    P.check_density(fl, *_stats(mol, D))


def test_scaled_density_refused(small8):
    mol, D, fl = small8
    scale = 1.5 * fl.d_bound / np.abs(U.cart_density(mol, D)).max()
    with pytest.raises(P.DensityError, match="exceeds the bound"):
        P.check_density(fl, *_stats(mol, D * scale))


def test_bound_is_what_the_lists_assume(small8):
    """A block dropped by the list builder cannot pass the reference screen for
    any density with max|D| <= d_bound: Qmax_IJ * Qglob * d_bound <= thr."""
    mol, D, fl = small8
    assert fl.pair_cut * fl.qglob * fl.d_bound == pytest.approx(fl.thr, rel=1e-12)
    for pc, ps in fl.pairsets.items():
        assert np.all(ps.qmax > fl.pair_cut), pc


def test_nonsymmetric_density_refused(small8):
    mol, D, fl = small8
    rng = np.random.default_rng(3)
    Dn = D + 1e-6 * rng.standard_normal(D.shape)
    with pytest.raises(P.DensityError, match="not symmetric"):
        P.check_density(fl, *_stats(mol, Dn))
    # rounding-level asymmetry is accepted
    Dr = D.copy()
    Dr[0, 1] = np.nextafter(Dr[0, 1], np.inf)
    P.check_density(fl, *_stats(mol, Dr))


def test_nan_density_refused(small8):
    mol, D, fl = small8
    Dn = D.copy()
    Dn[2, 2] = np.nan
    with pytest.raises(P.DensityError):
        P.check_density(fl, *_stats(mol, Dn))
