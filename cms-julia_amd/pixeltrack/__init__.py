"""pixeltrack: MI355X (gfx950) pixel-track stage - CA doublets, cell connection,
fishbone, ntuplets and broken-line fit - behind a C ABI (include/pixeltrack.h).

Python mirror of the reference producer interface; the compute is in
lib/libpixeltrack_hip.so, the synthetic inputs in lib/libpixeltrack_gen.so."""
from . import _abi  # noqa: F401
from .formats import HitsSoA, TrackSOA  # noqa: F401

__all__ = ["HitsSoA", "TrackSOA", "ca", "generator", "formats"]
