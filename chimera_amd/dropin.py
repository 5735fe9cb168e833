"""Register the drop-in modules under the reference's module names.

`install()` makes the reference's own import lines resolve to this package's classes
without editing them:

    exafs_gpu.forward      -> chimera_amd.forward       (ForwardModel, thermal_sigma2)
    exafs_gpu.fourier      -> chimera_amd.fourier       (FourierTransform, r_grid, ...)
    chimera.multift        -> chimera_amd.multift       (MultiEdgeFT)
    methods.method_a_rmc   -> chimera_amd.method_a_rmc  (RMCSampler)

(the reference imports them at chimera/model.py:88-89, chimera/multift.py:22,
chimera/fitrun.py:43).  A parent package that cannot be found (the reference's
`exafs_gpu` or `methods` is absent, or unimportable without CuPy) is registered as an
empty namespace holding the aliased submodules; a parent that exists is imported as it is,
and only the named submodules are replaced.  Call it before anything imports those
modules (a module already imported under a reference name is left alone and reported).
"""
from __future__ import annotations

import importlib
import importlib.util
import sys
import types

ALIASES = {
    "exafs_gpu.forward": "chimera_amd.forward",
    "exafs_gpu.fourier": "chimera_amd.fourier",
    "chimera.multift": "chimera_amd.multift",
    "methods.method_a_rmc": "chimera_amd.method_a_rmc",
}


def _parent_findable(name):
    try:
        return importlib.util.find_spec(name) is not None
    except (ImportError, ValueError):
        return False


def install(aliases=None):
    """Register the aliases; returns {reference name: module registered} and skips (with
    the reason) names already imported from elsewhere."""
    done, skipped = {}, {}
    for ref, ours in (aliases or ALIASES).items():
        mod = importlib.import_module(ours)
        cur = sys.modules.get(ref)
        if cur is not None and cur is not mod:
            skipped[ref] = f"already imported from {getattr(cur, '__file__', cur)}"
            continue
        parent, _, leaf = ref.rpartition(".")
        if parent not in sys.modules:
            if _parent_findable(parent):
                try:
                    importlib.import_module(parent)
                except Exception:          # e.g. the reference package needs CuPy
                    sys.modules.pop(parent, None)
            if parent not in sys.modules:
                pkg = types.ModuleType(parent)
                pkg.__path__ = []          # a namespace: only the registered submodules
                sys.modules[parent] = pkg
        sys.modules[ref] = mod
        setattr(sys.modules[parent], leaf, mod)
        done[ref] = mod
    return {"installed": done, "skipped": skipped}
