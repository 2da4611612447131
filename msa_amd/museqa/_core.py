"""Loads the compiled binding; there is no fallback."""
import glob
import importlib
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

if not glob.glob(os.path.join(_PKG, "_museqa*.so")):
    raise ImportError("msa_amd._museqa is not built: run `python -m msa_amd.build`")

core = importlib.import_module("msa_amd._museqa")
MuseqaError = core.MuseqaError
