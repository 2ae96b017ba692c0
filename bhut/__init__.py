"""`import bhut` for code written against maroba/bhut.

The package and module names of the reference (bhut/__init__.py:8-11,
bhut/api.py, bhut/tree/{build,node,refit}.py, bhut/traverse/bh.py) resolve
to bhut_amd's modules, whose compute runs in libbhut_hip (HIP, MI355X).
Nothing here computes; it only registers the names.
"""

import sys as _sys

import bhut_amd as _impl
import bhut_amd.api
import bhut_amd.backends
import bhut_amd.traverse
import bhut_amd.traverse.bh
import bhut_amd.tree
import bhut_amd.tree.build
import bhut_amd.tree.node
import bhut_amd.tree.refit

for _name in ("api", "backends", "traverse", "traverse.bh", "tree", "tree.build", "tree.node", "tree.refit"):
    _sys.modules[f"{__name__}.{_name}"] = _sys.modules[f"bhut_amd.{_name}"]

api = bhut_amd.api
backends = bhut_amd.backends
traverse = bhut_amd.traverse
tree = bhut_amd.tree

from bhut_amd import Forest, Tree, accelerations, accelerations_batched, potential  # noqa: E402

__version__ = _impl.__version__
__all__ = ["Tree", "accelerations"]
