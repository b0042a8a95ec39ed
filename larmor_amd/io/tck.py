"""MRtrix TCK tractograms (the published MRtrix3 track format; the reference reads and writes it
in src/larmor/io/formats/tck.py).

The file is a text header -- "mrtrix tracks", then "key: value" lines, "END" -- followed at the
byte offset of its "file: . <offset>" line by float32 triplets: each streamline's points, a NaN
triplet after each streamline and an Inf triplet at the end.
"""

from pathlib import Path

import numpy as np

from larmor_amd.types.tractogram import Tractogram

MAGIC = "mrtrix tracks"
# Header keys the writers set themselves.
_RESERVED = {"count", "datatype", "file"}


def read_tck_header(path) -> tuple[dict, int]:
    """(fields, payload offset) of a TCK file; repeated keys are joined with newlines."""
    fields: dict[str, str] = {}
    with open(path, "rb") as f:
        first = f.readline().decode("latin-1").strip()
        if first != MAGIC:
            raise ValueError(f"{path} is not a TCK file: no {MAGIC!r} line.")
        for raw in f:
            line = raw.decode("latin-1").strip()
            if line == "END":
                break
            key, sep, value = line.partition(":")
            if not sep:
                continue
            key, value = key.strip(), value.strip()
            fields[key] = value if key not in fields else fields[key] + "\n" + value
        else:
            raise ValueError(f"{path}: the header has no END line.")
    if "file" not in fields:
        raise ValueError(f"{path}: the header has no file line.")
    parts = fields["file"].split()
    if len(parts) != 2 or parts[0] != ".":
        raise ValueError(f"{path}: only single-file TCK is read, got file: {fields['file']}")
    return fields, int(parts[1])


def read_tck(path) -> Tractogram:
    fields, offset = read_tck_header(path)
    datatype = fields.get("datatype", "Float32LE")
    dtypes = {"Float32LE": "<f4", "Float32BE": ">f4", "Float64LE": "<f8", "Float64BE": ">f8"}
    if datatype not in dtypes:
        raise ValueError(f"{path}: unsupported TCK datatype {datatype}")
    raw = np.fromfile(path, dtype=dtypes[datatype], offset=offset)
    rows = raw[: len(raw) // 3 * 3].reshape(-1, 3).astype(np.float64)
    end = np.flatnonzero(np.isinf(rows[:, 0]))
    if len(end):
        rows = rows[: end[0]]
    breaks = np.flatnonzero(np.isnan(rows[:, 0]))
    keep = ~np.isnan(rows[:, 0])
    points = rows[keep]
    # Streamline s ends where its NaN row is; points before it, minus the s earlier NaN rows.
    ends = breaks - np.arange(len(breaks))
    offsets = np.concatenate([[0], ends]).astype(np.int64)
    if offsets[-1] != len(points):  # points after the last separator form no streamline
        points = points[: offsets[-1]]
    return Tractogram(points.astype(np.float32), offsets)


def write_tck(tractogram: Tractogram, path, extra: dict | None = None) -> None:
    """Writes Float32LE TCK with a count line and any extra header fields."""
    lines = [MAGIC]
    for key, value in (extra or {}).items():
        if key in _RESERVED:
            continue
        for part in str(value).split("\n"):
            lines.append(f"{key}: {part}")
    lines.append(f"count: {len(tractogram)}")
    lines.append("datatype: Float32LE")
    head = "\n".join(lines) + "\nfile: . "
    # The offset counts its own digits: grow it until the header fits.
    offset = len(head) + len("\nEND\n") + 1
    while len(f"{head}{offset}\nEND\n") > offset:
        offset += 1
    header = f"{head}{offset}\nEND\n".encode("latin-1")
    header += b"\0" * (offset - len(header))
    pts = np.asarray(tractogram.points, dtype="<f4")
    n = len(tractogram)
    rows = np.empty((len(pts) + n + 1, 3), dtype="<f4")
    idx = tractogram.offsets[1:] + np.arange(n)  # row of each NaN separator
    mask = np.ones(len(rows), dtype=bool)
    mask[idx] = False
    mask[-1] = False
    rows[mask] = pts
    rows[idx] = np.nan
    rows[-1] = np.inf
    with open(path, "wb") as f:
        f.write(header)
        f.write(rows.tobytes())


class TckWriter:
    """Writes a TCK file batch by batch: the header with a fixed-width count, each batch's rows
    appended before the end marker, and the count filled in at close."""

    COUNT_WIDTH = 20

    def __init__(self, path, extra: dict | None = None):
        self.path = Path(path)
        self.count = 0
        lines = [MAGIC]
        for key, value in (extra or {}).items():
            if key in _RESERVED:
                continue
            for part in str(value).split("\n"):
                lines.append(f"{key}: {part}")
        self._count_prefix = "\n".join(lines) + "\ncount: "
        head = self._count_prefix + "0".zfill(self.COUNT_WIDTH) + "\ndatatype: Float32LE\nfile: . "
        offset = len(head) + len("\nEND\n") + 1
        while len(f"{head}{offset}\nEND\n") > offset:
            offset += 1
        header = f"{head}{offset}\nEND\n".encode("latin-1")
        self._file = open(self.path, "wb")
        self._file.write(header + b"\0" * (offset - len(header)))
        self._file.write(np.full(3, np.inf, dtype="<f4").tobytes())

    def write(self, tractogram: Tractogram) -> None:
        n = len(tractogram)
        if n == 0:
            return
        pts = np.asarray(tractogram.points, dtype="<f4")
        rows = np.empty((len(pts) + n, 3), dtype="<f4")
        idx = tractogram.offsets[1:] - 1 + np.arange(1, n + 1)
        mask = np.ones(len(rows), dtype=bool)
        mask[idx] = False
        rows[mask] = pts
        rows[idx] = np.nan
        self._file.seek(-12, 2)
        self._file.write(rows.tobytes())
        self._file.write(np.full(3, np.inf, dtype="<f4").tobytes())
        self.count += n

    def close(self) -> None:
        if self._file.closed:
            return
        self._file.seek(len(self._count_prefix.encode("latin-1")))
        self._file.write(str(self.count).zfill(self.COUNT_WIDTH).encode("latin-1"))
        self._file.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
