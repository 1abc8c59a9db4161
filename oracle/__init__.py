"""Host oracle and CPU baseline for the spectral front end.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``. The shipped package
``metalmom_amd`` never imports it.
"""
from . import spectral  # noqa: F401
