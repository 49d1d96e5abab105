"""Import helper: the package directory name contains dashes, so it is loaded
by path under the module name ``louvain_amd``."""
import importlib.util
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))


def load_package():
    if "louvain_amd" in sys.modules:
        return sys.modules["louvain_amd"]
    spec = importlib.util.spec_from_file_location(
        "louvain_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["louvain_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
