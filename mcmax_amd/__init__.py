"""mcmax_amd -- MI355X-native Monte Carlo photon transport for MCMax disk models.

The transport pass (MCRadiation.f) and the temperature inversion (Main.f:498-558)
run in libmcx.so (HIP, gfx950) behind the C-ABI declared in include/mcx.h,
with Diffuse, the vertical-structure iteration and the ray-traced SED.  This
package builds the model tables on the host, moves them across that ABI, runs
the driver's loop (iterate.py) and writes the reference's output files
(output.py, observe.py); `python -m mcmax_amd input Nphot [obs...] -o outdir`
is the command line.
"""
from . import keywords, model, configs  # noqa: F401

__all__ = ["keywords", "model", "configs"]
