"""`lar fod` and `lar track` end to end on the GPU: a synthetic DWI written as NIfTI with FSL
gradient files and as MIF with its dw_scheme, fitted by the command and by SDModel / CSDModel
directly (the same arrays), then tracked from the written FOD and re-tracked from the seeds file
the run wrote (the same streamlines)."""

import json

import numpy as np
import pytest
from typer.testing import CliRunner

from larmor_amd import synthetic
from larmor_amd.cli import app
from larmor_amd.io import load_volume, read_tck, save_volume, write_bvals_bvecs
from larmor_amd.sd import CSDModel, SDModel
from larmor_amd.sd.rf import AnisotropicRF, IsotropicRF
from larmor_amd.types import GradientTable, Volume

pytestmark = pytest.mark.gpu
runner = CliRunner()

AFFINE = np.array([[-2.0, 0.0, 0.0, 30.0], [0.0, 2.0, 0.0, -20.0], [0.0, 0.0, 2.0, -10.0],
                   [0.0, 0.0, 0.0, 1.0]])


def _run(*args):
    r = runner.invoke(app, [str(a) for a in args])
    assert r.exit_code == 0, (r.output, r.exception)
    return r


@pytest.fixture(scope="module")
def dwi(tmp_path_factory):
    d = tmp_path_factory.mktemp("dwi")
    ph = synthetic.phantom((10, 9, 6), 6, [(1000.0, 40), (2500.0, 40)], seed=3, snr=30.0,
                           mask_share=0.5)
    gtab = GradientTable(ph.bvals, ph.bvecs)
    vol = Volume(ph.data, AFFINE)
    save_volume(vol, d / "dwi.nii.gz")
    write_bvals_bvecs(gtab, d / "dwi.bval", d / "dwi.bvec", affine=AFFINE)
    save_volume(vol, d / "dwi.mif", gradients=gtab)
    save_volume(Volume(ph.mask.astype(np.uint8), AFFINE), d / "mask.nii.gz")
    return d, vol, gtab, ph.mask


def test_fod_mfod_writes_what_sdmodel_fits(dwi):
    d, vol, gtab, mask = dwi
    _run("fod", d / "dwi.nii.gz", d / "mask.nii.gz", d / "out", "--bvals", d / "dwi.bval",
         "--bvecs", d / "dwi.bvec", "--method", "mFOD", "--solver", "L2NNLS", "--wm-rf", "1.7e-3",
         "3e-4", "3e-4", "--gm-rf", "0.8e-3", "--csf-rf", "3e-3")
    fod = load_volume(d / "out.nii.gz")
    frac = load_volume(d / "out_tissue_fraction.nii.gz")
    np.testing.assert_allclose(fod.affine, AFFINE, atol=1e-5)
    # The same fit through the API (gradients read back from the FSL files: world -> FSL ->
    # world rounds their last bits, so the comparison is to 1e-5 of the FOD scale).
    from larmor_amd.io import read_bvals_bvecs

    g2 = read_bvals_bvecs(d / "dwi.bval", d / "dwi.bvec", affine=AFFINE)
    rfs = [AnisotropicRF([1.7e-3, 3e-4, 3e-4], 0.0), IsotropicRF(0.8e-3), IsotropicRF(3e-3)]
    from larmor_amd.sd import L2NNLS

    fit = SDModel(rfs, method="mFOD", solver=L2NNLS()).fit(vol, g2, mask=mask)
    assert np.array_equal(fod.data, fit.fods[0])
    assert np.array_equal(frac.data, fit.tissue_fractions)
    # The MIF input with its dw_scheme gives the same files.
    _run("fod", d / "dwi.mif", d / "mask.nii.gz", d / "out_mif", "--method", "mFOD", "--solver",
         "L2NNLS", "--wm-rf", "1.7e-3", "3e-4", "3e-4", "--gm-rf", "0.8e-3", "--csf-rf", "3e-3")
    mif_fod = load_volume(d / "out_mif.nii.gz").data
    scale = np.abs(fod.data).max()
    np.testing.assert_allclose(mif_fod, fod.data, rtol=0, atol=1e-5 * scale)


def test_fod_csd_with_a_response_file(dwi):
    d, vol, gtab, mask = dwi
    (d / "wm.txt").write_text("1.0 -0.45 0.12 -0.02 0.002\n")
    # One shell: the b = 1000 volumes and the b0s.
    keep = gtab.bvals < 1500
    v1 = Volume(vol.data[..., keep], AFFINE)
    g1 = GradientTable(gtab.bvals[keep], gtab.bvecs[keep])
    save_volume(v1, d / "dwi1.mif", gradients=g1)
    _run("fod", d / "dwi1.mif", d / "mask.nii.gz", d / "csd", "--method", "CSD", "--response",
         d / "wm.txt", "--lmax", "6")
    from larmor_amd.io import read_mrtrix_response
    from larmor_amd.sd import build_csd_response

    rf = build_csd_response(read_mrtrix_response(d / "wm.txt"), g1)
    fit = CSDModel(rf, lmax=6).fit(v1, g1, mask=mask)
    assert np.array_equal(load_volume(d / "csd.nii.gz").data, fit.fods[0])
    assert not (d / "csd_tissue_fraction.nii.gz").exists()


def test_track_from_the_written_fod_and_again_from_its_seeds(tmp_path):
    fod, affine, mask = synthetic.tracking_field(shape=(40, 40, 24))
    save_volume(Volume(fod, affine), tmp_path / "fod.nii.gz")
    save_volume(Volume(mask.astype(np.uint8), affine), tmp_path / "mask.nii.gz")
    _run("track", tmp_path / "fod.nii.gz", tmp_path / "a.tck", "--seed-image",
         tmp_path / "mask.nii.gz", "--select", "300", "--seed", "7", "--mask",
         tmp_path / "mask.nii.gz", "--output-seeds", tmp_path / "s.txt", "--json",
         tmp_path / "run.json", "--algorithm", "prob")
    first = read_tck(tmp_path / "a.tck")
    assert len(first) == 300
    record = json.loads((tmp_path / "run.json").read_text())
    assert record["num_streamlines"] == 300 and record["parameters"]["algorithm"] == "prob"
    _run("track", tmp_path / "fod.nii.gz", tmp_path / "b.tck", "--seed-file", tmp_path / "s.txt",
         "--seed", "7", "--mask", tmp_path / "mask.nii.gz", "--algorithm", "prob")
    again = read_tck(tmp_path / "b.tck")
    assert np.array_equal(again.offsets, first.offsets)
    assert np.array_equal(again.points, first.points)
