"""Host-side SH basis and direction sets against the reference's fixtures and closed forms."""

from pathlib import Path

import numpy as np
import pytest

from larmor_amd.sd.sh import BASES, build_basis_matrix, convert_sh_basis, fit_sh, sh_layout
from larmor_amd.sd.sphere import DIRS_300, get_dirs, get_theta_phi

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_tournier_matches_mrtrix_sh2amp():
    """Every basis function up to lmax 8 at 60 directions against MRtrix's sh2amp
    (the reference's tests/data/sh_tournier_ref.npz, float32 output, hence atol 1e-6)."""
    ref = np.load(GOLDEN / "sh_tournier_ref.npz")
    basis = build_basis_matrix(*get_theta_phi(ref["dirs"]), int(ref["lmax"]), "tournier")
    np.testing.assert_allclose(basis, ref["basis"], atol=1e-6)


def test_basis_orthonormal_on_dense_sphere():
    """The real SH are orthonormal: a quadrature on a dense spiral gives ~identity."""
    n = 20000
    k = np.arange(n) + 0.5
    z = 1 - 2 * k / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (1 + np.sqrt(5)) * k
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)
    for basis in BASES:
        b = build_basis_matrix(*get_theta_phi(dirs), 6, basis, full=True)
        gram = b.T @ b * (4 * np.pi / n)
        np.testing.assert_allclose(gram, np.eye(b.shape[1]), atol=2e-3)


@pytest.mark.parametrize("src", BASES)
@pytest.mark.parametrize("dst", BASES)
def test_conversion_preserves_the_function(src, dst):
    rng = np.random.default_rng(0)
    coeffs = rng.normal(size=(4, 45))
    converted = convert_sh_basis(coeffs, src, dst)
    dirs = DIRS_300
    a = coeffs @ build_basis_matrix(*get_theta_phi(dirs), 8, src).T
    b = converted @ build_basis_matrix(*get_theta_phi(dirs), 8, dst).T
    np.testing.assert_allclose(a, b, atol=1e-10)


def test_fit_sh_recovers_coefficients():
    rng = np.random.default_rng(1)
    coeffs = rng.normal(size=(3, 45))
    amps = coeffs @ build_basis_matrix(*get_theta_phi(DIRS_300), 8).T
    np.testing.assert_allclose(fit_sh(amps, DIRS_300, 8), coeffs, atol=1e-10)


def test_sh_layout():
    assert sh_layout(45) == (8, False)
    assert sh_layout(1) == (0, False)
    assert sh_layout(81) == (8, True)
    with pytest.raises(ValueError):
        sh_layout(7)


def test_get_dirs_spiral():
    d = get_dirs(45)
    assert d.shape == (45, 3)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
    assert (d[:, 2] >= 0).all()
    np.testing.assert_allclose(d[0], [0, 0, 1], atol=1e-12)
    np.testing.assert_allclose(d[-1], [0, 0, 1], atol=1e-12)


def test_dirs_300_table():
    assert DIRS_300.shape == (300, 3)
    np.testing.assert_allclose(np.linalg.norm(DIRS_300, axis=1), 1.0, atol=1e-6)
