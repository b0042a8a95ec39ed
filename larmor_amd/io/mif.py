"""MRtrix MIF images (`.mif`, `.mif.gz`, and `.mih` headers with their data file): read into a
`Volume`, written from one, with the diffusion gradient table of the `dw_scheme` lines.

The format is MRtrix3's published image format (mrtrix.readthedocs.io, "MRtrix image formats");
the reference reads and writes it in src/larmor/io/formats/mif.py.  A header is "mrtrix image",
then "key: value" lines, then "END"; the keys read here:
  dim        the size of each axis;
  vox        the voxel size of each axis;
  layout     per axis its rank in the storage order (0 fastest) and a sign, "-" when the axis is
             stored from its last index down;
  datatype   Bit, Int8, UInt8, (U)Int16/32/64 and Float32/64 with LE or BE;
  transform  three rows of a 3 x 4 matrix: the unit directions of the axes and the position of
             voxel 0 in world mm (the voxel-to-world affine is transform * diag(vox));
  scaling    "offset,scale" (value = offset + scale * stored);
  dw_scheme  one row x,y,z,b per volume, in world coordinates;
  file       the data file ("." for the same file) and the byte offset of the data.
The returned array is indexed in the header's axis order, whatever the storage order.
"""

import gzip
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from larmor_amd.types.gradient_table import GradientTable
from larmor_amd.types.volume import Volume

MAGIC = "mrtrix image"

_TYPES = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4",
          "UInt32": "u4", "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}
_NAMES = {v: k for k, v in _TYPES.items()}


@dataclass
class MifHeader:
    fields: dict
    shape: tuple
    vox: tuple
    layout: list
    datatype: str
    affine: np.ndarray
    data_path: Path
    offset: int


def _is_gz(path: Path) -> bool:
    return path.name.endswith(".gz")


def _raw(path: Path) -> bytes:
    data = path.read_bytes()
    return gzip.decompress(data) if _is_gz(path) else data


def _split(raw: bytes, path: Path, need_end: bool):
    """(header lines, bytes after the END line)."""
    end = raw.find(b"\nEND\n")
    if end < 0:
        if need_end:
            raise ValueError(f"{path}: the header has no END line.")
        text = raw.decode("latin-1")
        return text.splitlines(), b""
    return raw[:end].decode("latin-1").splitlines(), raw[end + 5:]


def _parse(lines: list[str], path: Path) -> dict:
    if not lines or lines[0].strip() != MAGIC:
        raise ValueError(f"{path} is not an MRtrix image: no {MAGIC!r} line.")
    fields: dict[str, list[str]] = {}
    for line in lines[1:]:
        line = line.strip()
        if not line or line.startswith("#") or line == "END":
            continue
        key, sep, value = line.partition(":")
        if not sep:
            raise ValueError(f"{path}: bad header line {line!r}.")
        fields.setdefault(key.strip(), []).append(value.strip())
    return fields


def _floats(text: str) -> list[float]:
    return [float(x) for x in text.split(",")]


def read_mif_header(path) -> MifHeader:
    path = Path(path)
    raw = _raw(path)
    lines, _ = _split(raw, path, need_end=not path.name.endswith(".mih"))
    fields = _parse(lines, path)
    try:
        dims = [int(x) for x in fields["dim"][0].split(",")]
        vox = _floats(fields["vox"][0]) if "vox" in fields else [1.0] * len(dims)
        datatype = fields["datatype"][0]
        file_entry = fields["file"][0]
    except (KeyError, ValueError) as exc:
        raise ValueError(f"{path}: bad or missing header entry ({exc}).") from exc
    if len(dims) > 4:
        raise ValueError(f"{path}: images with more than 4 dimensions are not read.")
    layout_text = fields.get("layout", [",".join(f"+{a}" for a in range(len(dims)))])[0]
    layout = [x.strip() for x in layout_text.split(",")]
    while len(dims) < 3:  # fewer than three axes: size-one axes after them
        dims.append(1)
        vox.append(1.0)
        layout.append(f"+{len(layout)}")
    affine = np.eye(4)
    if "transform" in fields:
        rows = np.array([_floats(r) for r in fields["transform"][:3]])
        affine[:3, :3] = rows[:, :3] * np.asarray(vox[:3])
        affine[:3, 3] = rows[:, 3]
    else:
        affine[:3, :3] = np.diag(vox[:3])
    name, _, offset_text = file_entry.rpartition(" ")
    if not name:
        name, offset_text = offset_text, "0"
    try:
        offset = int(offset_text)
    except ValueError:
        name, offset = file_entry, 0
    name = name.strip()
    data_path = path if name == "." else path.parent / name
    return MifHeader(fields, tuple(dims), tuple(vox), layout, datatype, affine, data_path, offset)


def _decode(buf: bytes, h: MifHeader, path: Path) -> np.ndarray:
    n = int(np.prod(h.shape))
    dt = h.datatype
    if dt == "Bit":
        need = -(-n // 8)
        if len(buf) < need:
            raise ValueError(f"{path}: the data is shorter than the header says.")
        flat = np.unpackbits(np.frombuffer(buf, np.uint8, count=need), bitorder="big")[:n]
        flat = flat.astype(bool)
    else:
        base = dt[:-2] if dt.endswith(("LE", "BE")) else dt
        if base not in _TYPES:
            raise ValueError(f"{path}: unsupported datatype {dt}.")
        order = ">" if dt.endswith("BE") else "<"
        dtype = np.dtype(order + _TYPES[base])
        if len(buf) < n * dtype.itemsize:
            raise ValueError(f"{path}: the data is shorter than the header says.")
        flat = np.frombuffer(buf, dtype=dtype, count=n).astype(dtype.newbyteorder("="))
    # Storage order: rank 0 fastest.  Reshape slowest-first, then back to the axis order.
    ranks = [abs(int(x)) for x in h.layout]
    slow_to_fast = sorted(range(len(h.shape)), key=lambda a: -ranks[a])
    arr = flat.reshape([h.shape[a] for a in slow_to_fast])
    arr = np.transpose(arr, np.argsort(slow_to_fast))
    flips = tuple(a for a, x in enumerate(h.layout) if x.startswith("-"))
    if flips:
        arr = np.flip(arr, axis=flips)
    return np.ascontiguousarray(arr)


def read_mif(path) -> Volume:
    """Reads a MIF/MIH image; scaling goes to the volume's slope and intercept (not applied)."""
    path = Path(path)
    h = read_mif_header(path)
    if h.data_path == path:
        raw = _raw(path)
        buf = raw[h.offset:]
    else:
        if not h.data_path.is_file():
            raise ValueError(f"{path}: its data file {h.data_path} does not exist.")
        buf = _raw(h.data_path)[h.offset:]
    data = _decode(buf, h, path)
    slope, intercept = 1.0, 0.0
    if "scaling" in h.fields:
        intercept, slope = _floats(h.fields["scaling"][0])[:2]
    return Volume(data=data, affine=h.affine, slope=float(slope), intercept=float(intercept))


def read_mif_gradients(path) -> GradientTable | None:
    """The dw_scheme of a MIF header as a gradient table (world directions), or None."""
    path = Path(path)
    h = read_mif_header(path)
    rows = h.fields.get("dw_scheme")
    if not rows:
        return None
    if len(h.shape) < 4:
        raise ValueError(f"{path}: a dw_scheme needs a 4-D image.")
    try:
        table = np.array([_floats(r) for r in rows])
    except ValueError as exc:
        raise ValueError(f"{path}: bad dw_scheme ({exc}).") from exc
    if table.ndim != 2 or table.shape[1] != 4 or len(table) != h.shape[3]:
        raise ValueError(f"{path}: the dw_scheme needs one row x,y,z,b per volume.")
    return GradientTable(table[:, 3], table[:, :3])


def has_mif_gradients(path) -> bool:
    return bool(read_mif_header(path).fields.get("dw_scheme"))


def _number(x: float) -> str:
    return f"{x:.17g}"


def write_mif(volume: Volume, path, gradients: GradientTable | None = None) -> None:
    """Writes a MIF (`.mif`, `.mif.gz`) or MIH + `.dat` image: axes in the array's order stored
    x fastest (layout +0,+1,...), little endian (Bit packed most significant bit first), the
    affine as transform and vox, scaling when the volume has one, and the gradient table as
    dw_scheme lines."""
    path = Path(path)
    data = np.asarray(volume.data)
    if data.ndim < 1 or data.ndim > 4:
        raise ValueError(f"cannot write a {data.ndim}-D array as MIF")
    if data.dtype == np.bool_:
        datatype = "Bit"
    else:
        kind = data.dtype.newbyteorder("<").str[1:]
        if kind not in _NAMES:
            raise ValueError(f"cannot write {data.dtype} as MIF")
        datatype = _NAMES[kind] + ("" if data.dtype.itemsize == 1 else "LE")
    if gradients is not None:
        if data.ndim != 4:
            raise ValueError("a gradient table needs a 4-D image")
        if len(gradients.bvals) != data.shape[3]:
            raise ValueError(f"the gradient table has {len(gradients.bvals)} entries, but the "
                             f"image has {data.shape[3]} volumes")
    affine = np.asarray(volume.affine, dtype=np.float64)
    vox = np.linalg.norm(affine[:3, :3], axis=0)
    vox[vox == 0] = 1.0
    rot = affine[:3, :3] / vox
    vox_all = list(vox) + [1.0] * (data.ndim - 3)
    lines = [MAGIC,
             "dim: " + ",".join(str(n) for n in data.shape),
             "vox: " + ",".join(_number(v) for v in vox_all[: data.ndim]),
             "layout: " + ",".join(f"+{a}" for a in range(data.ndim)),
             f"datatype: {datatype}"]
    for r in range(3):
        lines.append("transform: " + ",".join(_number(v) for v in [*rot[r], affine[r, 3]]))
    if volume.slope != 1.0 or volume.intercept != 0.0:
        lines.append(f"scaling: {_number(volume.intercept)},{_number(volume.slope)}")
    if gradients is not None:
        for d, b in zip(np.asarray(gradients.bvecs), np.asarray(gradients.bvals)):
            lines.append("dw_scheme: " + ",".join(_number(v) for v in [*d, b]))
    if datatype == "Bit":
        payload = np.packbits(data.ravel(order="F").astype(np.uint8), bitorder="big").tobytes()
    else:
        payload = np.asarray(data, dtype=np.dtype("<" + data.dtype.newbyteorder("<").str[1:])) \
            .tobytes(order="F")
    if path.name.endswith(".mih"):
        dat = path.with_suffix(".dat")
        header = "\n".join(lines + [f"file: {dat.name} 0", "END"]) + "\n"
        path.write_text(header, encoding="latin-1")
        dat.write_bytes(payload)
        return
    head = "\n".join(lines) + "\nfile: . "
    offset = len(head) + len("\nEND\n") + 1
    offset += (-offset) % 4
    while len(f"{head}{offset}\nEND\n") > offset:
        offset += 4
    header = f"{head}{offset}\nEND\n".encode("latin-1")
    raw = header + b"\0" * (offset - len(header)) + payload
    if _is_gz(path):
        path.write_bytes(gzip.compress(raw))
    else:
        path.write_bytes(raw)
