"""The `lar` command-line driver of the path: `lar fod` (spherical deconvolution to FOD images)
and `lar track` (streamlines from them), run as `python -m larmor_amd.cli`."""

import typer

from larmor_amd.cli.fod import fod
from larmor_amd.cli.track import track

app = typer.Typer(add_completion=False, no_args_is_help=True,
                  help="larmor's FOD fit and tracking on the GPU.")
app.command("fod")(fod)
app.command("track")(track)


def main() -> None:
    app()
