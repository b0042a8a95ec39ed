"""`lar fod`: fits FODs by spherical deconvolution on the GPU and writes them as NIfTI.

The command of the reference (src/larmor/cli/fod.py:367-553) with its arguments, options,
defaults, refusals and output files: DWI, MASK and an output prefix; --method GRL (default), mFOD,
CSD or MSMT-CSD; the GRL/mFOD solver, penalty, inner-shell weight, response options and FOD
merging; CSD's --response and MSMT-CSD's --wm/gm/csf-response files.  `out` writes the FOD to
out.nii.gz (out_fod1/2.nii.gz for an unmerged NODDI fit) and, for a multi-tissue fit, the tissue
fractions to out_tissue_fraction.nii.gz.
"""

import dataclasses
from enum import Enum
from pathlib import Path
from typing import Optional

import typer

from larmor_amd.cli.outputs import check_outputs, output_prefix, with_suffix


class Method(str, Enum):
    GRL = "GRL"
    MFOD = "mFOD"
    CSD = "CSD"
    MSMT_CSD = "MSMT-CSD"


class Solver(str, Enum):
    DRL = "dRL"
    L2NNLS = "L2NNLS"
    LASSO = "LASSO"


class FodMerge(str, Enum):
    WS = "ws"
    MAJORITY = "majority"
    NONE = "none"


class SHBasis(str, Enum):
    tournier = "tournier"
    descoteaux = "descoteaux"


class BvecFrame(str, Enum):
    fsl = "fsl"
    world = "world"


GM_NODDI = "noddi"
GM_AUTO = "auto"
GM_ISOTROPIC_DEFAULT = "0.7e-3"

# Options of SDModel's methods (refused with CSD and MSMT-CSD) and options of one method only.
SD_MODEL_OPTIONS = {"wm_rf": "--wm-rf", "mk": "--mk", "inner_shell_weight": "--inner-shell-weight",
                    "gm_rf": "--gm-rf", "csf_rf": "--csf-rf", "fod_merge": "--fod-merge"}
METHOD_OPTIONS = {"response": ("--response", "CSD"), "wm_response": ("--wm-response", "MSMT-CSD"),
                  "gm_response": ("--gm-response", "MSMT-CSD"),
                  "csf_response": ("--csf-response", "MSMT-CSD")}


def _given(ctx: typer.Context, name: str) -> bool:
    source = ctx.get_parameter_source(name)
    return source is not None and source.name != "DEFAULT"


def _diffusivity_or_none(value):
    if isinstance(value, str) and value.lower() == "none":
        return None
    return float(value)


def _gm_response(value: str):
    if value.lower() in (GM_NODDI, GM_AUTO):
        return value.lower()
    d = _diffusivity_or_none(value)
    return None if d is None else str(d)


def _solver(method: Method, name: Optional[Solver], reg: Optional[float]):
    from larmor_amd.sd import DRL, L2NNLS, Lasso, SDMethod

    solvers = {Solver.DRL: DRL, Solver.L2NNLS: L2NNLS, Solver.LASSO: Lasso}
    if method == Method.CSD:
        if name is not None or reg is not None:
            raise typer.BadParameter("CSD has its own solver; --solver and --reg do not apply",
                                     param_hint="--solver" if name is not None else "--reg")
        return None
    if reg is None:
        return None if name is None else solvers[name]()
    solver = SDMethod(method.value).default_solver() if name is None else solvers[name]()
    try:
        if isinstance(solver, L2NNLS):
            return L2NNLS(reg=reg)
        if isinstance(solver, Lasso):
            return Lasso(reg=reg)
    except ValueError as error:
        raise typer.BadParameter(str(error), param_hint="--reg") from error
    raise typer.BadParameter("applies to the L2NNLS and LASSO solvers only", param_hint="--reg")


def _fod_paths(prefix: Path, n_aniso: int) -> list[Path]:
    if n_aniso == 1:
        return [with_suffix(prefix, ".nii.gz")]
    return [with_suffix(prefix, f"_fod{i}.nii.gz") for i in range(1, n_aniso + 1)]


def _save_fit(image, results, sh_outs, frac_out) -> None:
    from larmor_amd.io import save_volume

    for fod, path in zip(results.fods, sh_outs):
        save_volume(image.with_data(fod), path)
    if results.tissue_fractions is not None:
        save_volume(image.with_data(results.tissue_fractions), frac_out)


def fod(
    ctx: typer.Context,
    volume: str = typer.Argument(..., help="Diffusion-weighted volume (NIfTI or MIF)."),
    mask: str = typer.Argument(..., help="Brain mask on the grid of the diffusion volume."),
    output: str = typer.Argument(..., help="Prefix of the output files: 'out' writes "
                                 "out.nii.gz (and out_tissue_fraction.nii.gz)."),
    bvals: Optional[str] = typer.Option(None, "--bvals", help="b-values file (FSL)."),
    bvecs: Optional[str] = typer.Option(None, "--bvecs", help="b-vectors file."),
    method: Method = typer.Option(Method.GRL, "--method", case_sensitive=False),
    solver: Optional[Solver] = typer.Option(None, "--solver", case_sensitive=False),
    reg: Optional[float] = typer.Option(None, "--reg"),
    inner_shell_weight: Optional[float] = typer.Option(None, "--inner-shell-weight"),
    bvec_frame: BvecFrame = typer.Option(BvecFrame.fsl, "--bvec-frame"),
    lmax: int = typer.Option(8, "--lmax"),
    wm_rf: Optional[tuple[float, float, float]] = typer.Option(None, "--wm-rf"),
    gm_rf: Optional[str] = typer.Option(GM_AUTO, "--gm-rf", parser=_gm_response,
                                        metavar="<float|none|noddi|auto>"),
    csf_rf: Optional[float] = typer.Option(3.0e-3, "--csf-rf", parser=_diffusivity_or_none,
                                           metavar="<float|none>"),
    mk: Optional[float] = typer.Option(None, "--mk"),
    response: Optional[Path] = typer.Option(None, "--response", exists=True, dir_okay=False),
    wm_response: Optional[Path] = typer.Option(None, "--wm-response", exists=True,
                                               dir_okay=False),
    gm_response: Optional[Path] = typer.Option(None, "--gm-response", exists=True,
                                               dir_okay=False),
    csf_response: Optional[Path] = typer.Option(None, "--csf-response", exists=True,
                                                dir_okay=False),
    basis: SHBasis = typer.Option(SHBasis.tournier, "--basis"),
    fod_merge: FodMerge = typer.Option(FodMerge.WS, "--fod-merge"),
    force: bool = typer.Option(False, "--force", "-f"),
    jobs: int = typer.Option(-1, "--nt"),
) -> None:
    """Estimate fibre orientation distributions (FODs) by spherical deconvolution."""
    from larmor_amd.io import load_dwi, load_volume, read_mrtrix_response
    from larmor_amd.sd import (CSDModel, MSMTCSDModel, SDModel, build_csd_response,
                               build_response_functions, build_shell_response, merge_fods)

    for name, (flag, owner) in METHOD_OPTIONS.items():
        if owner != method.value and _given(ctx, name):
            raise typer.BadParameter(f"applies to --method {owner} only", param_hint=flag)

    def refuse_sd_model_options():
        for name, flag in SD_MODEL_OPTIONS.items():
            if _given(ctx, name):
                raise typer.BadParameter(f"does not apply to --method {method.value}",
                                         param_hint=flag)

    prefix = output_prefix(output)
    frac_out = with_suffix(prefix, "_tissue_fraction.nii.gz")
    if method == Method.MSMT_CSD:
        refuse_sd_model_options()
        if solver is not None or reg is not None:
            raise typer.BadParameter("MSMT-CSD has its own solver; --solver and --reg do not apply",
                                     param_hint="--solver" if solver is not None else "--reg")
        responses = [r for r in (wm_response, gm_response, csf_response) if r is not None]
        if responses and wm_response is None:
            raise typer.BadParameter("is needed with --gm-response and --csf-response",
                                     param_hint="--wm-response")
        sh_outs = _fod_paths(prefix, 1)
        multi = not responses or len(responses) > 1
        check_outputs([*sh_outs, frac_out] if multi else sh_outs, force)
        image, gtab = load_dwi(volume, bvals, bvecs, frame=bvec_frame.value)
        mask_volume = load_volume(mask)
        rfs = [build_shell_response(read_mrtrix_response(p), gtab) for p in responses] or None
        results = MSMTCSDModel(rfs, lmax=lmax, basis=basis.value).fit(
            image, gtab, mask=mask_volume, n_jobs=jobs)
        _save_fit(image, results, sh_outs, frac_out)
        return

    fod_solver = _solver(method, solver, reg)
    if method == Method.CSD:
        refuse_sd_model_options()
        gm_rf = csf_rf = None
    if gm_rf == GM_AUTO:
        gm_rf = GM_NODDI if method == Method.MFOD else GM_ISOTROPIC_DEFAULT
    gm_noddi = gm_rf == GM_NODDI
    if not gm_noddi and _given(ctx, "fod_merge"):
        raise typer.BadParameter("applies only to a fit with two FODs (--gm-rf noddi)",
                                 param_hint="--fod-merge")
    merged = gm_noddi and fod_merge != FodMerge.NONE
    gm_diffusivity = None if gm_rf is None or gm_noddi else float(gm_rf)
    sh_outs = _fod_paths(prefix, 2 if gm_noddi and not merged else 1)
    outs = list(sh_outs)
    if gm_rf is not None or csf_rf is not None:
        outs.append(frac_out)
    check_outputs(outs, force)

    image, gtab = load_dwi(volume, bvals, bvecs, frame=bvec_frame.value)
    mask_volume = load_volume(mask)
    if method == Method.CSD:
        rf = None if response is None else build_csd_response(read_mrtrix_response(response), gtab)
        model = CSDModel(rf, lmax=lmax, basis=basis.value)
    else:
        rfs = build_response_functions(image, gtab, mask_volume, wm_eigen_vals=wm_rf,
                                       mean_kurtosis=mk, gm_diffusivity=gm_diffusivity,
                                       csf_diffusivity=csf_rf, gm_noddi=gm_noddi)
        model = SDModel(rfs=rfs, method=method.value, solver=fod_solver,
                        inner_shell_weight=inner_shell_weight, lmax=lmax, basis=basis.value)
    results = model.fit(image, gtab, mask=mask_volume, n_jobs=jobs)
    if merged:
        mode = "majority" if fod_merge == FodMerge.MAJORITY else "ws"
        results = dataclasses.replace(
            results, fods=[merge_fods(results.fods, results.tissue_fractions, mode)])
    _save_fit(image, results, sh_outs, frac_out)
