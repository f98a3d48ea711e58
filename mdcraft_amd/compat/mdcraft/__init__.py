"""The reference's ``mdcraft`` Python package names over the GPU engine
(implementation: mdcraft_amd.refapi)."""
import os as _os
import sys as _sys

_root = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))))
if _root not in _sys.path:
    _sys.path.append(_root)
