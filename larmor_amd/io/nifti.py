"""NIfTI-1 and NIfTI-2 images (one file, `.nii` or `.nii.gz`): read into a `Volume`, written from
one.

The layout is the published NIfTI-1.1 / NIfTI-2 header (nifti.nimh.nih.gov); the reference reads
and writes the same format in src/larmor/io/formats/nifti.py.  The voxel-to-world matrix is the
sform when its code is set, else the qform (a quaternion, the voxel sizes and qfac), else the
voxel sizes alone.  The data are stored x fastest; the returned array is indexed [x, y, z, ...].
"""

import gzip
import math
import struct
from pathlib import Path

import numpy as np

from larmor_amd.types.volume import Volume

# datatype code -> numpy kind (byte order added from the header)
_DTYPES = {2: "u1", 4: "i2", 8: "i4", 16: "f4", 64: "f8", 256: "i1", 512: "u2", 768: "u4",
           1024: "i8", 1280: "u8"}
_CODES = {np.dtype(v).str[1:]: k for k, v in _DTYPES.items()}


def _open(path: Path, mode: str):
    return gzip.open(path, mode) if path.suffix == ".gz" else open(path, mode)


def _quaternion_affine(b, c, d, qfac, pixdim, offset) -> np.ndarray:
    """The qform: rotation from the unit quaternion (a implied), columns scaled by the voxel
    sizes, the third by qfac."""
    a = math.sqrt(max(0.0, 1.0 - (b * b + c * c + d * d)))
    rot = np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b],
    ])
    scale = np.array([pixdim[0], pixdim[1], pixdim[2] * (-1.0 if qfac < 0 else 1.0)])
    affine = np.eye(4)
    affine[:3, :3] = rot * scale
    affine[:3, 3] = offset
    return affine


def read_nifti(path) -> Volume:
    """Reads a NIfTI-1 or NIfTI-2 file.

    Raises:
        ValueError: if the file is not NIfTI or its datatype is not one of the integer or
            float32/float64 types.
    """
    path = Path(path)
    with _open(path, "rb") as f:
        raw = f.read()
    if len(raw) < 348:
        raise ValueError(f"{path} is too short to be NIfTI")
    for order in "<>":
        size = struct.unpack(order + "i", raw[:4])[0]
        if size in (348, 540):
            break
    else:
        raise ValueError(f"{path} is not a NIfTI file")
    if size == 348:
        dim = struct.unpack(order + "8h", raw[40:56])
        datatype = struct.unpack(order + "h", raw[70:72])[0]
        pixdim = struct.unpack(order + "8f", raw[76:108])
        vox_offset = int(struct.unpack(order + "f", raw[108:112])[0])
        slope, inter = struct.unpack(order + "2f", raw[112:120])
        qform_code, sform_code = struct.unpack(order + "2h", raw[252:256])
        quat = struct.unpack(order + "6f", raw[256:280])
        srows = struct.unpack(order + "12f", raw[280:328])
    else:
        datatype = struct.unpack(order + "h", raw[12:14])[0]
        dim = struct.unpack(order + "8q", raw[16:80])
        pixdim = struct.unpack(order + "8d", raw[104:168])
        vox_offset = struct.unpack(order + "q", raw[168:176])[0]
        slope, inter = struct.unpack(order + "2d", raw[176:192])
        qform_code, sform_code = struct.unpack(order + "2i", raw[344:352])
        quat = struct.unpack(order + "6d", raw[352:400])
        srows = struct.unpack(order + "12d", raw[400:496])
    if datatype not in _DTYPES:
        raise ValueError(f"{path}: unsupported NIfTI datatype {datatype}")
    ndim = int(dim[0])
    if not 1 <= ndim <= 7:
        raise ValueError(f"{path}: bad number of dimensions {ndim}")
    shape = tuple(int(n) for n in dim[1:1 + ndim])
    dtype = np.dtype(order + _DTYPES[datatype])
    count = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dtype, count=count, offset=vox_offset)
    data = data.reshape(shape, order="F").astype(dtype.newbyteorder("="))
    if sform_code > 0:
        affine = np.eye(4)
        affine[:3, :] = np.asarray(srows, dtype=np.float64).reshape(3, 4)
    elif qform_code > 0:
        affine = _quaternion_affine(*quat[:3], pixdim[0], pixdim[1:4], quat[3:6])
    else:
        affine = np.diag([pixdim[1], pixdim[2], pixdim[3], 1.0]).astype(np.float64)
    if not math.isfinite(slope) or slope == 0.0:
        slope, inter = 1.0, 0.0
    return Volume(data=data, affine=affine, slope=float(slope), intercept=float(inter))


def _affine_quaternion(affine: np.ndarray):
    """(b, c, d, qfac, voxel sizes) of the rotation part of an affine (the NIfTI-1 recipe:
    columns normalised, a left-handed frame flips the third column and sets qfac -1)."""
    m = np.asarray(affine, dtype=np.float64)[:3, :3].copy()
    sizes = np.linalg.norm(m, axis=0)
    sizes[sizes == 0] = 1.0
    r = m / sizes
    qfac = 1.0
    if np.linalg.det(r) < 0:
        qfac = -1.0
        r[:, 2] = -r[:, 2]
    # The nearest rotation (polar decomposition) before the quaternion.
    u, _, vt = np.linalg.svd(r)
    r = u @ vt
    trace = r[0, 0] + r[1, 1] + r[2, 2]
    if trace > 0:
        s = 0.5 / math.sqrt(trace + 1.0)
        a, b, c, d = 0.25 / s, (r[2, 1] - r[1, 2]) * s, (r[0, 2] - r[2, 0]) * s, (r[1, 0] - r[0, 1]) * s
    elif r[0, 0] > r[1, 1] and r[0, 0] > r[2, 2]:
        s = 2.0 * math.sqrt(1.0 + r[0, 0] - r[1, 1] - r[2, 2])
        a, b, c, d = (r[2, 1] - r[1, 2]) / s, 0.25 * s, (r[0, 1] + r[1, 0]) / s, (r[0, 2] + r[2, 0]) / s
    elif r[1, 1] > r[2, 2]:
        s = 2.0 * math.sqrt(1.0 + r[1, 1] - r[0, 0] - r[2, 2])
        a, b, c, d = (r[0, 2] - r[2, 0]) / s, (r[0, 1] + r[1, 0]) / s, 0.25 * s, (r[1, 2] + r[2, 1]) / s
    else:
        s = 2.0 * math.sqrt(1.0 + r[2, 2] - r[0, 0] - r[1, 1])
        a, b, c, d = (r[1, 0] - r[0, 1]) / s, (r[0, 2] + r[2, 0]) / s, (r[1, 2] + r[2, 1]) / s, 0.25 * s
    if a < 0:
        b, c, d = -b, -c, -d
    return b, c, d, qfac, sizes


def write_nifti(volume: Volume, path) -> None:
    """Writes a NIfTI-1 single file (gzip-compressed for `.gz`), little endian, with the affine
    as both sform and qform (codes 1, scanner) and the volume's slope and intercept."""
    path = Path(path)
    data = np.asarray(volume.data)
    if data.dtype == np.bool_:
        data = data.astype(np.uint8)
    kind = data.dtype.newbyteorder("<").str[1:]
    if kind not in _CODES:
        raise ValueError(f"cannot write {data.dtype} to NIfTI")
    if not 1 <= data.ndim <= 7:
        raise ValueError(f"cannot write a {data.ndim}-D array to NIfTI")
    affine = np.asarray(volume.affine, dtype=np.float64)
    b, c, d, qfac, sizes = _affine_quaternion(affine)
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    dims = [data.ndim, *data.shape] + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<2h", hdr, 70, _CODES[kind], data.dtype.itemsize * 8)
    pix = [qfac, *sizes] + [1.0] * 4
    struct.pack_into("<8f", hdr, 76, *pix)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<2f", hdr, 112, float(volume.slope), float(volume.intercept))
    struct.pack_into("<B", hdr, 123, 10)  # xyzt_units: mm and seconds
    struct.pack_into("<2h", hdr, 252, 1, 1)
    struct.pack_into("<6f", hdr, 256, b, c, d, *affine[:3, 3])
    struct.pack_into("<12f", hdr, 280, *affine[:3, :].ravel())
    hdr[344:348] = b"n+1\x00"
    payload = np.asarray(data, dtype=np.dtype("<" + kind)).tobytes(order="F")
    with _open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(payload)
