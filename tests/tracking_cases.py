"""Shared inputs of the tracking tests: the tckgen fixture (tests/golden/tracking_fixture.npz) and
synthetic FOD fields, with the splitting of a bidirectional streamline at its seed."""

import math
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "tracking_fixture.npz"
STEP = 0.2
CUTOFF = 0.3


def fixture():
    return np.load(GOLDEN)


def max_steps_for(max_length: float, step: float) -> int:
    """The tracker's step budget: ceil(round(max_length / step, 9)) (tracker.py _prepare)."""
    return math.ceil(round(max_length / step, 9))


def tracks(points, offsets):
    return [points[offsets[i]:offsets[i + 1]].astype(np.float64) for i in range(len(offsets) - 1)]


def halves(points, k: int):
    """(away, back): the streamline split at its seed point k, both halves starting at the seed."""
    return points[: k + 1][::-1], points[k:]


def det_reference(z, name: str):
    """(refs, seed point indices, seeds) of a tckgen SD_Stream fixture track set."""
    refs = tracks(z[f"{name}_points"], z[f"{name}_offsets"])
    k = z[f"{name}_seeds"][:, 1].astype(np.int64)
    seeds = np.array([r[j] for r, j in zip(refs, k)])
    return refs, k, seeds


def split_ours(positions, counts, seeds):
    """Our streamlines split at their (float32) seed points."""
    out = []
    for i in range(len(counts)):
        p = positions[i, : counts[i]].astype(np.float64)
        at = np.flatnonzero(np.all(p == seeds[i].astype(np.float32), axis=1))
        assert len(at) >= 1, i
        out.append(halves(p, int(at[0])))
    return out


def fibre_field(shape=(12, 12, 8), lmax=8, seed=0, full=False):
    """A smooth synthetic FOD field: per voxel one or two lobes exp(-8 (1 - (u.a)^2)) projected on
    the tournier basis by quadrature, with directions that turn slowly across the volume, plus an
    isotropic floor; float32 (X, Y, Z, C) and a 2 mm affine with an offset."""
    from larmor_amd.sd.sh import build_basis_matrix

    rng = np.random.default_rng(seed)
    g = rng.normal(size=(4000, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    theta = np.arccos(np.clip(g[:, 2], -1, 1))
    phi = np.arctan2(g[:, 1], g[:, 0])
    basis = build_basis_matrix(theta, phi, lmax, "tournier", full=full)
    proj = np.linalg.pinv(basis)
    xs, ys, zs = np.meshgrid(*[np.linspace(0, 1, n) for n in shape], indexing="ij")
    a1 = np.stack([np.cos(xs * 1.2), np.sin(xs * 1.2), 0.3 * zs], -1)
    a2 = np.stack([0.2 * ys, np.cos(ys * 0.8), np.sin(ys * 0.8)], -1)
    a1 /= np.linalg.norm(a1, axis=-1, keepdims=True)
    a2 /= np.linalg.norm(a2, axis=-1, keepdims=True)
    w2 = (ys > 0.4).astype(np.float64) * 0.7
    lobes = (np.exp(-8 * (1 - (a1.reshape(-1, 3) @ g.T) ** 2))
             + w2.reshape(-1, 1) * np.exp(-8 * (1 - (a2.reshape(-1, 3) @ g.T) ** 2)) + 0.05)
    if full:
        lobes = lobes * (1.0 + 0.3 * (a1.reshape(-1, 3) @ g.T))
    coeffs = lobes @ proj.T
    affine = np.diag([2.0, 2.0, 2.0, 1.0])
    affine[:3, 3] = [-11.0, -11.0, -7.0]
    return np.ascontiguousarray(coeffs.reshape(*shape, -1), dtype=np.float32), affine


def random_seeds(affine, shape, n, seed=1, margin=-0.6):
    """World seeds uniform over the voxel box widened by -margin voxels (some off the image)."""
    rng = np.random.default_rng(seed)
    vox = rng.uniform(margin, np.asarray(shape) - 1 - margin, size=(n, 3))
    return vox @ affine[:3, :3].T + affine[:3, 3]
