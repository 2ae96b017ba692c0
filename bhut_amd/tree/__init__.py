"""Tree construction, data and refit (bhut/tree/)."""

from .build import build_tree
from .node import TreeData
from .refit import refit_tree, should_refit_vs_rebuild

__all__ = ["TreeData", "build_tree", "refit_tree", "should_refit_vs_rebuild"]
