"""fasttools2d_amd -- MI355X-native doubly periodic 2D Stokes spectral Ewald.

The compute path is ``libse2p.so`` (hand-written HIP kernels for gfx950 behind
the C ABI in ``include/se2p.h``).  This package is the host-side mirror of the
reference's PeriodicEwald interface:

* ``periodic_ewald`` -- the four MEX-equivalent entry points;
* ``params``         -- the drivers' Ewald parameter selection;
* ``workloads``      -- seeded synthetic inputs.
"""
from . import params, workloads  # noqa: F401

__all__ = ["params", "workloads", "periodic_ewald"]


def __getattr__(name):
    # periodic_ewald needs torch and the built library; import it lazily
    if name == "periodic_ewald":
        import importlib

        return importlib.import_module(".periodic_ewald", __name__)
    raise AttributeError(name)
