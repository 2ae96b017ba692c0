"""bhut_amd: bhut's Barnes-Hut force path on AMD Instinct MI355X (gfx950).

Drop-in for ``bhut.accelerations`` and ``bhut.Tree`` (maroba/bhut,
bhut/__init__.py:8-11); the tree build, refit and traversal run in
hand-written HIP kernels (libbhut_hip.so, C ABI in include/bhut_hip.h).
"""

from .api import Tree, accelerations, potential
from .batched import Forest, accelerations_batched

__version__ = "0.1.0"
__all__ = ["Tree", "accelerations", "potential", "Forest", "accelerations_batched"]
