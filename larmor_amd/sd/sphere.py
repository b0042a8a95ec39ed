"""Direction sets on the sphere: the 300 FOD directions and the spiral of the low-resolution kernel.

Mirrors larmor's `larmor.sd.sphere` (src/larmor/sd/sphere.py): `DIRS_300`, `get_theta_phi`,
`get_dirs`.
"""

from importlib.resources import files

import numpy as np

# The 300 unit directions FOD amplitudes are solved on, shape (300, 3); the same table as
# larmor's src/larmor/data/dir300.txt.
DIRS_300 = np.loadtxt(str(files("larmor_amd") / "data" / "dir300.txt"))


def get_theta_phi(dirs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Returns the polar angle (from +z) and azimuth (from +x, about z) of unit vectors (n, 3).

    Same convention as src/larmor/sd/sphere.py:get_theta_phi.
    """
    dirs = np.asarray(dirs, dtype=np.float64)
    return np.arccos(dirs[:, 2]), np.arctan2(dirs[:, 1], dirs[:, 0])


def get_dirs(n_dirs: int = 45) -> np.ndarray:
    """Returns n_dirs points of a Saff-Kuijlaars spiral, flipped onto the z >= 0 hemisphere.

    Point k (1-based) sits at height h_k = -1 + 2 (k - 1) / (n - 1); its azimuth advances by
    3.6 / sqrt(n (1 - h_k^2)) from the previous point, modulo 2 pi, and both poles take azimuth
    0.  Each point is multiplied by the sign of its z, so the first and last both become +z.
    Behaviour of src/larmor/sd/sphere.py:get_dirs.
    """
    heights = -1.0 + 2.0 * np.arange(n_dirs) / (n_dirs - 1)
    azimuths = np.zeros(n_dirs)
    two_pi = 2.0 * np.pi
    # The azimuth recursion is sequential (each step reduces modulo 2 pi), so it is a loop.
    for k in range(1, n_dirs - 1):
        step = 3.6 / np.sqrt(n_dirs * (1.0 - heights[k] ** 2))
        azimuths[k] = (azimuths[k - 1] + step) % two_pi
    polar = np.arccos(heights)
    sin_polar = np.sin(polar)
    points = np.stack(
        [np.cos(azimuths) * sin_polar, np.sin(azimuths) * sin_polar, np.cos(polar)], axis=1
    )
    return points * np.sign(points[:, 2:3])
