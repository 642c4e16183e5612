"""trot_amd: the phaseless AFQMC walker step and local energy of `trot`, on MI355X (gfx950).

Host code in Python; compute in hand-written HIP kernels reached through the
C ABI of include/trot_amd.h (trot_amd/lib/libtrot_amd.so).  The public names
mirror the reference package's propagator / trial / measurement / driver API.
"""

from .core import HamChol, MeasOps, System, TrialOps, k_energy, k_force_bias, n_fields  # noqa: F401

__all__ = [
    "HamChol", "MeasOps", "System", "TrialOps", "k_energy", "k_force_bias", "n_fields",
]


def __getattr__(name):
    # GPU-facing modules load lazily so that `import trot_amd` works without the library.
    import importlib

    mods = {"prop": ".prop", "meas": ".meas", "trial": ".trial", "walkers": ".walkers", "blocks": ".blocks",
            "driver": ".driver", "synthetic": ".synthetic", "stats": ".stats", "cisd": ".cisd", "fp": ".fp",
            "graph": ".graph", "batch": ".batch", "diff": ".diff", "staged": ".staged"}
    if name in mods:
        return importlib.import_module(mods[name], __name__)
    raise AttributeError(name)
