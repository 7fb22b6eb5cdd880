"""CPU oracle for the CTU+CT Godunov step (TEST INFRASTRUCTURE ONLY).

This package is a from-scratch FP64 NumPy restatement of the algorithm in
PythonMHD's ``Source/EvolveFunctions_MHD.py`` and its helper modules.  It exists
to *check* the HIP product path and to time a CPU baseline.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product package ``pythonmhd_amd`` never imports it and has no
CPU fallback.

Parity status of each function is listed in the header of ``oracle/mhd.py``.
"""
from . import mhd  # noqa: F401
