"""`lar fod` and `lar track` refusals, made before any file is read or any GPU work (the option
checks of src/larmor/cli/fod.py and track.py)."""

import pytest
from typer.testing import CliRunner

from larmor_amd.cli import app
from larmor_amd.cli.outputs import check_outputs, output_prefix, with_suffix

runner = CliRunner()


def _run(*args):
    return runner.invoke(app, [str(a) for a in args])


def _flat(output: str) -> str:
    """The command's message with the panel's borders and line breaks taken out."""
    return " ".join(output.replace("│", " ").split())


@pytest.fixture
def files(tmp_path):
    for name in ["dwi.nii.gz", "mask.nii.gz", "wm.txt", "fod.nii.gz", "s.txt"]:
        (tmp_path / name).write_bytes(b"")
    return tmp_path


@pytest.mark.parametrize("args,message", [
    (["--method", "GRL", "--response", "wm.txt"], "applies to --method CSD only"),
    (["--method", "CSD", "--wm-response", "wm.txt"], "applies to --method MSMT-CSD only"),
    (["--method", "CSD", "--solver", "dRL"], "CSD has its own solver"),
    (["--method", "CSD", "--mk", "0.5"], "does not apply to --method CSD"),
    (["--method", "MSMT-CSD", "--reg", "0.1"], "MSMT-CSD has its own solver"),
    (["--method", "MSMT-CSD", "--gm-response", "wm.txt"], "is needed with --gm-response"),
    (["--method", "GRL", "--fod-merge", "majority"], "applies only to a fit with two FODs"),
    (["--method", "GRL", "--solver", "dRL", "--reg", "0.1"],
     "applies to the L2NNLS and LASSO solvers only"),
])
def test_fod_refuses_options_of_other_methods(files, args, message):
    args = [files / a if a.endswith(".txt") else a for a in args]
    r = _run("fod", files / "dwi.nii.gz", files / "mask.nii.gz", files / "out", *args)
    assert r.exit_code != 0
    assert message in _flat(r.output)


def test_fod_refuses_an_existing_output_without_force(files):
    (files / "out.nii.gz").write_bytes(b"x")
    r = _run("fod", files / "dwi.nii.gz", files / "mask.nii.gz", files / "out", "--method",
             "CSD")
    assert r.exit_code != 0 and isinstance(r.exception, FileExistsError)


@pytest.mark.parametrize("args,hint", [
    (["out.trk", "--seed-file", "s.txt"], "must be a .tck file"),
    (["out.tck"], "give exactly one of --seed-image"),
    (["out.tck", "--seed-image", "mask.nii.gz"], "needs --select"),
    (["out.tck", "--seed-file", "s.txt", "--select", "0"], "must be at least 1"),
])
def test_track_refuses_bad_seed_sources(files, args, hint):
    out, *rest = args
    r = _run("track", files / "fod.nii.gz", files / out, *[
        files / a if a.endswith((".txt", ".gz")) else a for a in rest])
    assert r.exit_code != 0
    assert hint in _flat(r.output)


def test_output_paths(tmp_path):
    assert output_prefix("a/out.nii.gz").name == "out"
    assert output_prefix("a/out.nii").name == "out"
    assert with_suffix(output_prefix("out"), "_tissue_fraction.nii.gz").name == \
        "out_tissue_fraction.nii.gz"
    with pytest.raises(ValueError, match="are the same file"):
        check_outputs([tmp_path / "a.nii", tmp_path / "a.nii"], False)
    with pytest.raises(FileNotFoundError, match="does not exist"):
        check_outputs([tmp_path / "missing" / "a.nii"], False)
