"""FP64 NumPy restatement of PyPIC3D's per-step PIC hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker against which the
HIP kernels in ``pypic3d_amd`` are compared.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or as the reported CPU baseline -- never
as the thing measured or shipped.  The product path (``pypic3d_amd``) never
imports this package and fails loudly when its HIP library is missing.

Every function cites the reference file:line it restates.  Reference paths
are relative to the PyPIC3D source tree (``PyPIC3D/...``).

Parity pins (what checks each function independently of the kernels):
see the module docstring of ``oracle.pic_oracle``.
"""

from .pic_oracle import *  # noqa: F401,F403
