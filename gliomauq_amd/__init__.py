"""gliomauq_amd: MI355X-native Fisher-Kolmogorov forward solve + likelihood.

The compute path is libgq.so (HIP kernels for gfx950 behind the C-ABI of
include/gq_api.h); this package is its host-side mirror of the reference's
block-operator API (see solver.py) plus the synthetic-input generator.
"""
from . import _lib  # noqa: F401
from .solver import (BlockFWT, BlockProcessing, Context, Ensemble, FlipLikelihood, Glioma_HG_UQ,  # noqa: F401
                     Glioma_ReactionDiffusionOperator, Grid, HGG_Likelihood, Science, Topology, UpdateTumor,
                     fk_steps, n_steps_until, set_step_timing, step_timing, step_timing_drain, set_autotune,
                     autotune_info, host_stats, set_step_coeffs, set_step_fusion, set_step_persistent,
                     persistent_info, graph_info, set_step_graphs)

__all__ = ["BlockFWT", "BlockProcessing", "Context", "Ensemble", "FlipLikelihood", "Glioma_HG_UQ",
           "Glioma_ReactionDiffusionOperator", "Grid", "HGG_Likelihood", "Science", "Topology", "UpdateTumor",
           "fk_steps", "n_steps_until", "set_step_timing", "step_timing", "step_timing_drain", "set_autotune",
           "autotune_info", "host_stats", "set_step_coeffs", "set_step_fusion", "set_step_persistent", "persistent_info",
           "graph_info", "set_step_graphs"]
