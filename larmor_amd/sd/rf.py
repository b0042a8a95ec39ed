"""Response functions: the signal of one tissue compartment along each orientation.

Mirrors the response-function classes of larmor's `larmor.sd.rf` (src/larmor/sd/rf.py):
`ResponseFunction`, `IsotropicRF`, `AnisotropicRF`, `NoddiRF` and `ZonalSHRF`, with the same
constructor arguments, attributes and `signal(bvals, bvecs, dirs)` -> (n_meas, n_cols) contract,
and the white matter response estimated from data, `estimate_wm_rf`, with the response list of an
`SDModel` fit, `build_response_functions` (src/larmor/sd/rf.py:166-331).
"""

from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from typing import ClassVar

from collections.abc import Sequence

import numpy as np
from scipy.special import erf, erfi, eval_legendre

from larmor_amd._log import LarmorLogger
from larmor_amd.dwi.fitting import fit_log_signal
from larmor_amd.dwi.inputs import as_mask, signal_chunks
from larmor_amd.dwi.shells import SHELL_TOLERANCE
from larmor_amd.sd.sphere import get_theta_phi


@dataclass
class ResponseFunction(ABC):
    """Base class of the response functions (src/larmor/sd/rf.py:ResponseFunction)."""

    # Spiral points of an anisotropic response's low-resolution directions (rf.py:30).
    low_resolution_dirs: ClassVar[int] = 45

    @abstractmethod
    def signal(self, bvals: np.ndarray, bvecs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        """Returns the signal, one row per measurement and one column per orientation."""


@dataclass
class IsotropicRF(ResponseFunction):
    """exp(-b D), one column whatever the directions (src/larmor/sd/rf.py:IsotropicRF)."""

    diffusivity: float

    def signal(self, bvals: np.ndarray, bvecs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        return np.exp(-np.asarray(bvals, dtype=np.float64) * self.diffusivity)[:, None]


def _fibre_frames(dirs: np.ndarray) -> np.ndarray:
    """Rotations whose third column is each direction, (n, 3, 3).

    The frame is the one larmor builds from the polar angles (rf.py:AnisotropicRF.signal): the
    columns are e_theta-like, e_phi and the direction itself, so a tensor diag(l2, l3, l1) in it
    has its first eigenvalue along the direction.
    """
    theta, phi = get_theta_phi(dirs)
    ct, st, cp, sp = np.cos(theta), np.sin(theta), np.cos(phi), np.sin(phi)
    frames = np.empty((len(theta), 3, 3))
    frames[:, 0] = np.stack([cp * ct, -sp, cp * st], axis=-1)
    frames[:, 1] = np.stack([sp * ct, cp, sp * st], axis=-1)
    frames[:, 2] = np.stack([-st, np.zeros_like(theta), ct], axis=-1)
    return frames


@dataclass
class AnisotropicRF(ResponseFunction):
    """Single-fibre tensor response with an isotropic kurtosis term.

    S(b, g) = exp(-b g^T D g + b^2 MD^2 K / 6) + signal_offset, D = R diag(l2, l3, l1) R^T with
    R's third column along the fibre (src/larmor/sd/rf.py:AnisotropicRF).
    """

    eigen_vals: np.ndarray | list[float]
    mean_kurtosis: float
    signal_offset: float = field(default=0.0, kw_only=True)

    def signal(self, bvals: np.ndarray, bvecs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        bvals = np.asarray(bvals, dtype=np.float64)
        bvecs = np.asarray(bvecs, dtype=np.float64)
        l1, l2, l3 = (float(v) for v in self.eigen_vals)
        frames = _fibre_frames(np.asarray(dirs, dtype=np.float64))
        tensors = frames @ np.diag([l2, l3, l1]) @ np.swapaxes(frames, -1, -2)
        quadratic = np.einsum("mi,nij,mj->mn", bvecs, tensors, bvecs)
        md_sq = float(np.mean([l1, l2, l3])) ** 2
        b = bvals[:, None]
        return np.exp(-b * quadratic + (1.0 / 6.0) * (b**2) * md_sq * self.mean_kurtosis) + (
            self.signal_offset
        )


# NODDI helpers (src/larmor/sd/rf.py:_watson_sh_coefficients, _legendre_gaussian_integrals,
# _watson_hindered_diffusivities): the Watson SH series, the Legendre-Gaussian integrals and the
# dispersed tensor.  The polynomial table for large kappa is the NODDI toolbox's fit.
# fmt: off
_WATSON_LARGE_KAPPA = np.array([
    [7.52308, 0.411538, -0.214588, 0.0784091, -0.023981, 0.00731537, -0.0026467],
    [8.93718, 1.62147, -0.733421, 0.191568, -0.0202906, -0.00779095, 0.00574847],
    [8.87905, 3.35689, -1.15935, 0.0673053, 0.121857, -0.066642, 0.0180215],
    [7.84352, 5.03178, -1.0193, -0.426362, 0.328816, -0.0688176, -0.0229398],
    [6.30113, 6.09914, -0.16088, -1.05578, 0.338069, 0.0937157, -0.106935],
    [4.65678, 6.30069, 1.13754, -1.38393, -0.0134758, 0.331686, -0.105954],
])
# fmt: on


def _watson_sh(kappa: float) -> np.ndarray:
    """Zonal SH coefficients of degrees 0, 2, ..., 12 of a Watson density of concentration kappa."""
    c = np.zeros(7)
    rpi = np.sqrt(np.pi)
    c[0] = 2.0 * rpi
    if kappa <= 0.1:
        k = kappa
        c[1] = (4 / 3 * k + 8 / 63 * k**2) * np.sqrt(np.pi / 5)
        c[2] = (8 / 21 * k**2 + 32 / 693 * k**3) * 0.2 * rpi
        c[3] = (16 / 693 * k**3 + 32 / 10395 * k**4) * np.sqrt(np.pi / 13)
        c[4] = 32 / 19305 * k**4 * np.sqrt(np.pi / 17)
        c[5] = 64 * np.sqrt(np.pi / 21) * k**5 / 692835
        c[6] = 128 * rpi * k**6 / 152108775
        return c
    if kappa > 30:
        powers = (np.log(kappa) - np.log(30)) ** np.arange(7)
        c[1:] = _WATSON_LARGE_KAPPA @ powers
        return c
    k = [kappa**p for p in range(7)]
    rk = np.sqrt(kappa)
    rkp = [rk * kappa**p for p in range(7)]
    ei = erfi(rk)
    ek = np.exp(kappa)
    daw = 0.5 * rpi * ei / ek
    c[1] = np.sqrt(5) * (3 * rkp[0] - (3 + 2 * k[1]) * daw) * ek / ei / k[1]
    c[2] = 0.375 * ((105 + 60 * k[1] + 12 * k[2]) * daw - 105 * rkp[0] + 10 * rkp[1]) * ek / k[2] / ei
    c[3] = (
        ((-3465 - 1890 * k[1] - 420 * k[2] - 40 * k[3]) * daw + 3465 * rkp[0] - 420 * rkp[1]
         + 84 * rkp[2]) * np.sqrt(13 * np.pi) / 64 / k[3] / daw
    )
    c[4] = (
        np.sqrt(17)
        * ((675675 + 360360 * k[1] + 83160 * k[2] + 10080 * k[3] + 560 * k[4]) * daw
           - 675675 * rkp[0] + 90090 * rkp[1] - 23100 * rkp[2] + 744 * rkp[3])
        * ek / 512 / k[4] / ei
    )
    c[5] = (
        np.sqrt(21 * np.pi)
        * ((-43648605 - 22972950 * k[1] - 5405400 * k[2] - 720720 * k[3] - 55440 * k[4]
            - 2016 * k[5]) * daw
           + 43648605 * rkp[0] - 6126120 * rkp[1] + 1729728 * rkp[2] - 82368 * rkp[3]
           + 5104 * rkp[4])
        / 4096 / k[5] / daw
    )
    c[6] = (
        5
        * ((7027425405 + 3666482820 * k[1] + 872972100 * k[2] + 122522400 * k[3]
            + 10810800 * k[4] + 576576 * k[5] + 14784 * k[6]) * daw
           - 7027425405 * rkp[0] + 1018467450 * rkp[1] - 302630328 * rkp[2]
           + 17153136 * rkp[3] - 1553552 * rkp[4] + 25376 * rkp[5])
        * ek / 16384 / k[6] / ei
    )
    return c


def _even_legendre_monomials() -> np.ndarray:
    """Coefficients of t^0, t^2, ..., t^12 in P_0, P_2, ..., P_12, one row per degree."""
    table = np.zeros((7, 7))
    for d in range(7):
        unit = np.zeros(2 * d + 1)
        unit[-1] = 1.0
        table[d, : d + 1] = np.polynomial.legendre.leg2poly(unit)[::2]
    return table


_LEGENDRE_MONOMIALS = _even_legendre_monomials()


def _legendre_gaussian(x: np.ndarray) -> np.ndarray:
    """Integrals of exp(-x t^2) P_2l(t) over [-1, 1] for l = 0..6, shape (n, 7)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros((len(x), 7))
    big = x > 0.05
    xb = x[big]
    moments = np.zeros((len(xb), 7))
    sq = np.sqrt(xb)
    moments[:, 0] = np.sqrt(np.pi) * erf(sq) / sq
    for j in range(1, 7):
        moments[:, j] = (-np.exp(-xb) + (j - 0.5) * moments[:, j - 1]) / xb
    out[big] = moments @ _LEGENDRE_MONOMIALS.T
    xs = x[~big]
    p = [xs**n for n in range(7)]
    out[~big, 0] = 2 - 2 * p[1] / 3 + p[2] / 5 - p[3] / 21 + p[4] / 108
    out[~big, 1] = -4 * p[1] / 15 + 4 * p[2] / 35 - 2 * p[3] / 63 + 2 * p[4] / 297
    out[~big, 2] = 8 * p[2] / 315 - 8 * p[3] / 693 + 4 * p[4] / 1287
    out[~big, 3] = -16 * p[3] / 9009 + 16 * p[4] / 19305
    out[~big, 4] = 32 * p[4] / 328185
    out[~big, 5] = -64 * p[5] / 14549535
    out[~big, 6] = 128 * p[6] / 760543875
    return out


def _watson_dispersed_tensor(parallel: float, perpendicular: float, kappa: float) -> tuple[float, float]:
    """Parallel and perpendicular diffusivities of a tensor averaged over a Watson density."""
    diff = parallel - perpendicular
    if kappa < 1e-5:
        mean = (parallel + 2 * perpendicular) / 3
        return (
            mean + 4 * diff * kappa / 45 + 8 * diff * kappa**2 / 945,
            mean - 2 * diff * kappa / 45 - 4 * diff * kappa**2 / 945,
        )
    rk = np.sqrt(kappa)
    daw = 0.5 * np.exp(-kappa) * np.sqrt(np.pi) * erfi(rk)
    f = rk / daw
    par = (-diff + 2 * perpendicular * kappa + diff * f) / (2 * kappa)
    perp = (diff + 2 * (parallel + perpendicular) * kappa - diff * f) / (4 * kappa)
    return float(par), float(perp)


@dataclass(kw_only=True)
class NoddiRF(ResponseFunction):
    """NODDI response: Watson-dispersed sticks, a tortuous extra-neurite tensor and free water.

    Same parameters and defaults as src/larmor/sd/rf.py:NoddiRF (MRIToolkit's mFOD grey matter).
    """

    low_resolution_dirs: ClassVar[int] = 90

    intra_fraction: float = 0.4
    parallel_diffusivity: float = 1.7e-3
    concentration: float = 1.0
    iso_fraction: float = 0.0
    iso_diffusivity: float = 3e-3
    s0: float = 1.0

    def signal(self, bvals: np.ndarray, bvecs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        bvals = np.asarray(bvals, dtype=np.float64)
        cosines = np.clip(np.asarray(bvecs, dtype=np.float64) @ np.asarray(dirs).T, -1.0, 1.0)
        b = bvals[:, None]
        weights = _legendre_gaussian(bvals * self.parallel_diffusivity) * _watson_sh(
            self.concentration
        )
        intra = np.zeros(cosines.shape)
        for d in range(7):
            intra += weights[:, d, None] * (
                np.sqrt((4 * d + 1) / (4 * np.pi)) * eval_legendre(2 * d, cosines)
            )
        intra *= 0.5
        # The truncated series can dip to or below zero where the true signal is tiny.
        for column in intra.T:
            bad = column <= 0
            if bad.any() and (~bad).any():
                column[bad] = 0.1 * column[~bad].min()
        tortuous = self.parallel_diffusivity * (1 - self.intra_fraction)
        par, perp = _watson_dispersed_tensor(self.parallel_diffusivity, tortuous, self.concentration)
        extra = np.exp(-b * ((par - perp) * cosines**2 + perp))
        neurite = (1 - self.intra_fraction) * extra + self.intra_fraction * intra
        free = np.exp(-b * self.iso_diffusivity)
        return self.s0 * ((1 - self.iso_fraction) * neurite + self.iso_fraction * free)


@dataclass(kw_only=True)
class ZonalSHRF(ResponseFunction):
    """Response given by zonal SH coefficients per shell (MRtrix's response format).

    On a shell of b-value b the signal at angle t to the axis is
    sum_l c_l(b) sqrt((2l + 1) / (4 pi)) P_l(cos t) over even l; a measurement takes the row of
    the nearest shell, which must lie within 100 s/mm^2 (src/larmor/sd/rf.py:ZonalSHRF).
    """

    bvals: np.ndarray
    coefficients: np.ndarray

    def __post_init__(self) -> None:
        self.bvals = np.atleast_1d(np.asarray(self.bvals, dtype=np.float64))
        self.coefficients = np.atleast_2d(np.asarray(self.coefficients, dtype=np.float64))
        if self.bvals.ndim != 1 or self.coefficients.shape[0] != len(self.bvals):
            raise ValueError(
                f"Expected one row of coefficients per b-value, got {len(self.bvals)} "
                f"b-values and coefficients of shape {self.coefficients.shape}."
            )
        if not np.isfinite(self.coefficients).all():
            raise ValueError("The response coefficients must be finite numbers.")

    @property
    def lmax(self) -> int:
        return 2 * (int(self.coefficients.shape[1]) - 1)

    def shell_coefficients(self, bval: float) -> np.ndarray:
        distances = np.abs(self.bvals - bval)
        nearest = int(np.argmin(distances))
        if distances[nearest] >= SHELL_TOLERANCE:
            shells = ", ".join(f"{b:g}" for b in self.bvals)
            raise ValueError(
                f"The response has no shell near b = {bval:g} s/mm^2; its shells are {shells}."
            )
        return self.coefficients[nearest]

    def signal(self, bvals: np.ndarray, bvecs: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        cosines = np.clip(np.asarray(bvecs, dtype=np.float64) @ np.asarray(dirs).T, -1.0, 1.0)
        degrees = np.arange(0, self.lmax + 1, 2)
        rows = np.array([self.shell_coefficients(b) for b in np.asarray(bvals)])
        weights = rows * np.sqrt((2 * degrees + 1) / (4 * np.pi))
        out = np.zeros(cosines.shape)
        for j, degree in enumerate(degrees):
            out += weights[:, j, None] * eval_legendre(degree, cosines)
        return out


# The white matter response from the data (src/larmor/sd/rf.py:157-242), after MRIToolkit's
# Eigenval_IsoK_WM_FromData: a diffusion tensor, with an isotropic kurtosis term when two shells
# lie above KURTOSIS_MIN_B, is fitted to every usable mask voxel by weighted linear least squares
# of the log signal, ln S = ln S0 - b g^T D g + (b / 1000)^2 p_K / 6; the voxels of fractional
# anisotropy above WM_MIN_FA are single-fibre white matter.  The response takes the medians of
# their eigenvalues, the two smaller averaged into one (an axially symmetric tensor), and the
# mean of their kurtosis values K = 1e-6 p_K / MD^2 that lie inside KURTOSIS_RANGE.
WM_MIN_FA = 0.7
KURTOSIS_MIN_B = 500.0
KURTOSIS_RANGE = (0.0, 4.0)


def _wm_design(gtab, with_kurtosis: bool) -> np.ndarray:
    """Columns ln S0, the six tensor elements (the b-matrix, negated) and, with kurtosis,
    (b / 1000)^2 / 6."""
    columns = [np.ones((len(gtab.bvals), 1)), -gtab.b_matrix]
    if with_kurtosis:
        columns.append(((gtab.bvals / 1000.0) ** 2 / 6.0)[:, None])
    return np.hstack(columns)


def _eigenvalues(elements: np.ndarray) -> np.ndarray:
    """Eigenvalues, largest first, (n, 3), of the symmetric tensors given by their elements
    (xx, xy, xz, yy, yz, zz) as columns of elements (6, n)."""
    xx, xy, xz, yy, yz, zz = elements
    tensors = np.stack([np.stack([xx, xy, xz], -1), np.stack([xy, yy, yz], -1),
                        np.stack([xz, yz, zz], -1)], -2)
    return np.linalg.eigvalsh(tensors)[:, ::-1]


def _fractional_anisotropy(ev: np.ndarray) -> np.ndarray:
    """sqrt(3/2) |ev - mean| / |ev| per row; 0 for a zero tensor."""
    spread = ev - ev.mean(axis=1, keepdims=True)
    num = np.sqrt(1.5 * np.einsum("ij,ij->i", spread, spread))
    den = np.sqrt(np.einsum("ij,ij->i", ev, ev))
    fa = np.zeros_like(num)
    np.divide(num, den, out=fa, where=den > 0)
    return fa


def estimate_wm_rf(volume, gtab, mask) -> AnisotropicRF:
    """The white matter response of the mask's single-fibre voxels (FA > 0.7): the median
    eigenvalues [l1, l23, l23] and the mean voxel kurtosis in (0, 4) (0 with fewer than two
    shells above b = 500).  Raises ValueError when the mask does not match the volume or no
    usable voxel has FA > 0.7."""
    inside = as_mask(volume, mask)
    with_kurtosis = gtab.count_shells(above=KURTOSIS_MIN_B) >= 2
    design = _wm_design(gtab, with_kurtosis)
    fits = [np.empty((design.shape[1], 0))]
    for _, s in signal_chunks(volume, np.nonzero(inside)):
        usable = np.all(np.isfinite(s) & (s > 0), axis=1)
        fits.append(fit_log_signal(design, s[usable], weighted=True)[0])
    params = np.concatenate(fits, axis=1)
    ev = _eigenvalues(params[1:7])
    single_fibre = _fractional_anisotropy(ev) > WM_MIN_FA
    n_selected = int(np.count_nonzero(single_fibre))
    if n_selected == 0:
        raise ValueError(
            f"No usable voxel in the mask has FA > {WM_MIN_FA} to estimate the white "
            "matter response from."
        )
    median = np.median(ev[single_fibre], axis=0)
    axial, radial = float(median[0]), float(np.mean(median[1:]))
    kurtosis = 0.0
    if with_kurtosis:
        voxel_k = 1e-6 * params[7] / ev.mean(axis=1) ** 2
        low, high = KURTOSIS_RANGE
        kept = voxel_k[single_fibre & (voxel_k > low) & (voxel_k < high)]
        if kept.size:
            kurtosis = float(kept.mean())
        else:
            LarmorLogger.warning(
                "No voxel with FA > %s has a kurtosis in (%s, %s); the white matter "
                "response gets kurtosis 0.", WM_MIN_FA, low, high)
    LarmorLogger.info(
        f"Estimated WM eigenvalues: [{axial:.2e}, {radial:.2e}, {radial:.2e}], "
        f"kurtosis {kurtosis:.3f}, from {n_selected} voxels."
    )
    return AnisotropicRF(eigen_vals=[axial, radial, radial], mean_kurtosis=kurtosis)


def build_response_functions(volume, gtab, mask, *, wm_eigen_vals: Sequence[float] | None = None,
                             mean_kurtosis: float | None = None,
                             gm_diffusivity: float | None = None,
                             csf_diffusivity: float | None = None,
                             gm_noddi: bool = False) -> list[ResponseFunction]:
    """The responses of an SDModel fit, white matter first, then grey matter (NODDI with
    gm_noddi, isotropic with gm_diffusivity) and CSF (csf_diffusivity), each left out when not
    asked for (src/larmor/sd/rf.py:271-331).  White matter comes from wm_eigen_vals, or from
    the data by estimate_wm_rf; mean_kurtosis, when given, replaces its kurtosis (0 for given
    eigenvalues otherwise)."""
    if gm_noddi and gm_diffusivity is not None:
        raise ValueError("Give either a grey matter diffusivity or gm_noddi, not both.")
    if wm_eigen_vals is None:
        wm = estimate_wm_rf(volume, gtab, mask)
        if mean_kurtosis is not None:
            wm.mean_kurtosis = mean_kurtosis
    else:
        wm = AnisotropicRF(eigen_vals=list(wm_eigen_vals),
                           mean_kurtosis=0.0 if mean_kurtosis is None else mean_kurtosis)
    responses: list[ResponseFunction] = [wm]
    if gm_noddi:
        responses.append(NoddiRF())
    responses.extend(IsotropicRF(diffusivity=d) for d in (gm_diffusivity, csf_diffusivity)
                     if d is not None)
    return responses
