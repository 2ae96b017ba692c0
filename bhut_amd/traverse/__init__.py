"""Tree traversal (bhut/traverse/)."""

from .bh import (HAVE_NUMBA, barnes_hut_accelerations, barnes_hut_accelerations_targets,
                 compute_tree_statistics, interaction_lists, traversal_counters)

__all__ = ["HAVE_NUMBA", "barnes_hut_accelerations", "barnes_hut_accelerations_targets",
           "compute_tree_statistics", "interaction_lists", "traversal_counters"]
