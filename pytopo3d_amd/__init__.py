"""MI355X-native SIMP hot path of PyTopo3D.

``top3d`` is a drop-in for pytopo3d.core.optimizer.top3d whose per-iteration
work runs as hand-written HIP kernels (libpytopo3d_hip.so, built with
``python -m pytopo3d_amd.build``).  There is no CPU fallback.
"""

from .problem import (  # noqa: F401
    build_force_vector,
    edge_load_force_vector,
    lk_H8,
    make_C_matrix,
    support_node_flags,
)
from .optimizer import (  # noqa: F401
    SIMPProblem,
    evaluate_fixed_geometry_compliance,
    evaluate_fixed_geometry_metrics,
    top3d,
)

__all__ = [
    "top3d",
    "evaluate_fixed_geometry_metrics",
    "evaluate_fixed_geometry_compliance",
    "SIMPProblem",
    "lk_H8",
    "make_C_matrix",
    "build_force_vector",
    "edge_load_force_vector",
    "support_node_flags",
]
