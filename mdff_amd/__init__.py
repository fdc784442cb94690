"""MI355X (gfx950) engforce library for MDFF: C-ABI bindings, synthetic
configurations and the batched force-field-fit binding."""
from . import capi, configs  # noqa: F401

__all__ = ["capi", "configs"]
