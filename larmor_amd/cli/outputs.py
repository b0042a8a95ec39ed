"""Output paths of the commands: the prefix rule and the checks made before any work."""

from pathlib import Path


def written_files(path) -> list[Path]:
    """The files an output path writes (a `.mih` header also writes its `.dat`)."""
    p = Path(path)
    return [p, p.with_suffix(".dat")] if p.name.endswith(".mih") else [p]


def check_outputs(paths, force: bool) -> None:
    """Refuses two outputs that are one file, a missing folder, and an existing file without
    force."""
    files = [f for path in paths for f in written_files(path)]
    seen: dict[Path, Path] = {}
    for p in files:
        r = p.resolve()
        if r in seen:
            raise ValueError(f"Outputs {seen[r]} and {p} are the same file; give each output its "
                             "own path.")
        seen[r] = p
    for p in files:
        if not p.parent.is_dir():
            raise FileNotFoundError(f"Output folder {p.parent} does not exist.")
        if p.exists() and not force:
            raise FileExistsError(f"File {p} exists. Use --force to overwrite.")


def output_prefix(output: str) -> Path:
    """The output path without `.nii` / `.nii.gz`: `out`, `out.nii` and `out.nii.gz` all give
    `out`, to which the commands add their suffixes."""
    path = Path(output)
    for ext in (".nii.gz", ".nii"):
        if path.name.endswith(ext):
            return path.with_name(path.name.removesuffix(ext))
    return path


def with_suffix(prefix: Path, suffix: str) -> Path:
    return prefix.with_name(prefix.name + suffix)
