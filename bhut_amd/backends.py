"""Backend selection of bhut/backends/base.py:102-169, NumPy only.

NumPy arrays are the host carriers of this package; the ``backend`` keyword
keeps the reference's accepted values and errors.  The Dask backend is not
part of the device path: asking for it raises the ImportError the reference
raises when dask is not installed (bhut/backends/dask_.py:231-239).
"""

from __future__ import annotations

import numpy as np


def get_namespace(array, backend: str = "auto"):
    if backend == "auto":
        module = getattr(type(array), "__module__", "") or ""
        backend = "dask" if "dask" in module else "numpy"
    if backend == "numpy":
        return np
    if backend == "dask":
        raise ImportError("Dask backend requires dask[array]. Install with:\n  pip install dask[array]\n"
                          "or\n  conda install dask")
    raise ValueError(f"Unsupported backend: {backend}. Supported: 'numpy', 'dask', 'auto'")
