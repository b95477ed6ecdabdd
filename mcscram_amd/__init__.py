"""MI355X-native Monte Carlo fault-tree engine.

The compute path is libmcscram_amd.so (HIP kernels for gfx950 behind a C
ABI declared in include/mcscram_amd.h). This package is the Python host
mirror of that ABI; importing it does not load the library, and every call
that needs it fails loudly if it was not built (python -m mcscram_amd.build).
"""
from .pdag import (AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag,  # noqa: F401
                   batched_trees, large_tree, random_layered_tree, small_tree)
from .engine import (NOALLOC_SCRATCH_BYTES, Batch, LayerManager, Settings, Tally,  # noqa: F401
                     beta_quantile,
                     convergence_update, library_path, load_library,
                     normal_quantile_two_sided, pdag_event_order, pdag_gate_depths,
                     plan_passes, probability_analysis, sample_shape_for)
from . import mef  # noqa: F401  (Open-PSA MEF -> PDAG at compilation level 0)
from . import report  # noqa: F401  (convergence.csv rows)
