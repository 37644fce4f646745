"""knp_amd - MI355X (gfx950) backend for the KNP spiking-network step.

The product is the C ABI library `libknp_mi355x.so` (include/knp_mi355x.h) built from knp_amd/csrc; this
package holds its Python binding (`backend`), the ABI struct mirrors (`abi`), seeded synthetic workloads
(`workloads`) and the build recipe (`build`).
"""
from . import abi  # noqa: F401

__all__ = ["abi", "backend", "workloads", "build"]
