"""mcmax_amd -- MI355X-native Monte Carlo photon transport for MCMax disk models.

The transport pass (MCRadiation.f) and the temperature inversion (Main.f:498-558)
run in libmcx.so (HIP, gfx950) behind the C-ABI declared in include/mcx.h.  This
package builds the model tables on the host and moves them across that ABI.
"""
from . import keywords, model, configs  # noqa: F401

__all__ = ["keywords", "model", "configs"]
