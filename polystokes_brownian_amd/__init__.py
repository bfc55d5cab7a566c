"""MI355X-native per-timestep hydrodynamic solve of PolyStokes (HIP kernels behind a C ABI)."""
__all__ = ["PolyStokes", "Box"]


def __getattr__(name):
    if name in ("PolyStokes", "Box"):
        from . import polystokes
        return getattr(polystokes, name)
    raise AttributeError(name)
