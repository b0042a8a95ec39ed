"""Real spherical-harmonic bases and least-squares SH fits.

Mirrors larmor's `larmor.sd.sh` (src/larmor/sd/sh.py): `BASES`, `get_sh_indices`,
`get_sh_coeffs`, `build_basis_matrix`, `fit_sh`, `convert_sh_basis`, `sh_layout`.  The
conventions are fixed by the file formats they serve: "tournier" is MRtrix's basis (associated
Legendre functions with the Condon-Shortley phase, sqrt(2) sin(|m| phi) for m < 0 and
sqrt(2) cos(m phi) for m > 0), "descoteaux" swaps sine and cosine with a (-1)^m sign on m < 0,
and "leemans" is tournier evaluated at azimuth phi - pi / 2 (ExploreDTI / MRIToolkit).
"""

import math

import numpy as np
from scipy.special import lpmv

from larmor_amd.sd.sphere import get_theta_phi

BASES = ("tournier", "descoteaux", "leemans")
_SQRT2 = np.sqrt(2.0)


def get_sh_indices(lmax: int, full: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """Degree and order of each coefficient: l ascending (even only unless full), m = -l..l."""
    degrees = range(0, lmax + 1, 1 if full else 2)
    ls = [ell for ell in degrees for _ in range(2 * ell + 1)]
    ms = [m for ell in degrees for m in range(-ell, ell + 1)]
    return np.array(ls, dtype=np.int64), np.array(ms, dtype=np.int64)


def get_sh_coeffs(l: int, m: int, theta: np.ndarray, phi: np.ndarray,  # noqa: E741
                  basis: str = "tournier") -> np.ndarray:
    """Y_lm of `basis` at polar angles theta and azimuths phi."""
    if basis == "leemans":
        return get_sh_coeffs(l, m, theta, phi - np.pi / 2, "tournier")
    a = abs(m)
    norm = np.sqrt((2 * l + 1) / (4.0 * np.pi) * (math.factorial(l - a) / math.factorial(l + a)))
    value = norm * lpmv(a, l, np.cos(theta))
    if m == 0:
        if basis in ("tournier", "descoteaux"):
            return value
        raise ValueError(f"Unsupported basis {basis}")
    if basis == "tournier":
        return _SQRT2 * value * (np.sin(a * phi) if m < 0 else np.cos(m * phi))
    if basis == "descoteaux":
        if m < 0:
            return _SQRT2 * ((-1) ** a) * value * np.cos(a * phi)
        return _SQRT2 * value * np.sin(m * phi)
    raise ValueError(f"Unsupported basis {basis}")


def build_basis_matrix(theta: np.ndarray, phi: np.ndarray, lmax: int, basis: str = "tournier",
                       full: bool = False) -> np.ndarray:
    """The basis at each direction, (n_pts, n_coeffs), columns in get_sh_indices order."""
    ls, ms = get_sh_indices(lmax, full)
    out = np.zeros((len(theta), len(ls)), dtype=np.float64)
    for j, (ell, m) in enumerate(zip(ls, ms, strict=True)):
        out[:, j] = get_sh_coeffs(int(ell), int(m), theta, phi, basis)
    return out


def fit_sh(amplitudes, directions, lmax: int, *, basis: str = "tournier",
           full: bool = False) -> np.ndarray:
    """Least-squares SH coefficients (..., C) of amplitudes (..., N) on directions (N, 3)."""
    values = np.asarray(amplitudes, dtype=np.float64)
    dirs = np.asarray(directions, dtype=np.float64)
    if values.ndim == 0 or dirs.ndim != 2 or dirs.shape[1] != 3 or values.shape[-1] != len(dirs):
        raise ValueError(
            f"Expected amplitudes (..., N) for N directions (N, 3), got amplitudes "
            f"{values.shape} and directions {dirs.shape}."
        )
    matrix = build_basis_matrix(*get_theta_phi(dirs), lmax, basis, full=full)
    flat = values.reshape(-1, values.shape[-1])
    coeffs = np.linalg.lstsq(matrix, flat.T, rcond=None)[0].T
    return coeffs.reshape(*values.shape[:-1], matrix.shape[1])


def projection_matrix(directions, lmax: int, basis: str = "tournier") -> np.ndarray:
    """The matrix P (C, N) with fit_sh(a) = P a: the pseudo-inverse of the basis on directions."""
    matrix = build_basis_matrix(*get_theta_phi(np.asarray(directions, dtype=np.float64)), lmax,
                                basis)
    return np.linalg.pinv(matrix)


def _alternating(k: np.ndarray) -> np.ndarray:
    return 1 - 2 * (k % 2)


def _tournier_perm_sign(lmax: int, full: bool, basis: str) -> tuple[np.ndarray, np.ndarray]:
    """x[i] = t[perm[i]] sign[i] maps tournier coefficients t to `basis` coefficients x."""
    _, ms = get_sh_indices(lmax, full)
    idx = np.arange(len(ms))
    a = np.abs(ms)
    if basis == "tournier":
        return idx, np.ones(len(ms), dtype=np.int8)
    if basis == "descoteaux":
        return idx - 2 * ms, np.where(ms < 0, _alternating(a), 1).astype(np.int8)
    if basis == "leemans":
        odd = a % 2 == 1
        perm = np.where(odd, idx - 2 * ms, idx)
        sign = np.where(odd, np.where(ms < 0, -1, 1) * _alternating((a - 1) // 2),
                        _alternating(a // 2))
        return perm, sign.astype(np.int8)
    raise ValueError(f"Invalid basis {basis}, expected one of {', '.join(BASES)}")


def sh_layout(num_coeffs: int) -> tuple[int, bool]:
    """lmax of num_coeffs coefficients and whether all degrees are stored (asymmetric)."""
    err = ValueError(
        f"{num_coeffs} SH coefficients fit no layout: a symmetric FOD has "
        f"(lmax + 1)(lmax + 2) / 2 (1, 6, 15, 28, 45, ...), an asymmetric one "
        f"(lmax + 1)^2 (4, 9, 16, 25, ...)."
    )
    if num_coeffs < 1:
        raise err
    lmax = (-3 + math.isqrt(8 * num_coeffs + 1)) // 2
    if (lmax + 1) * (lmax + 2) // 2 == num_coeffs and lmax % 2 == 0:
        return lmax, False
    lmax = math.isqrt(num_coeffs) - 1
    if (lmax + 1) ** 2 == num_coeffs:
        return lmax, True
    raise err


def convert_sh_basis(coeffs: np.ndarray, from_basis: str, to_basis: str) -> np.ndarray:
    """Converts coefficients (..., C) between bases."""
    for b in (from_basis, to_basis):
        if b not in BASES:
            raise ValueError(f"Invalid basis {b}, expected one of {', '.join(BASES)}")
    lmax, full = sh_layout(coeffs.shape[-1])
    if from_basis == to_basis:
        return np.copy(coeffs)
    p_from, s_from = _tournier_perm_sign(lmax, full, from_basis)
    p_to, s_to = _tournier_perm_sign(lmax, full, to_basis)
    tournier = coeffs[..., p_from] * s_from[p_from]
    return tournier[..., p_to] * s_to
