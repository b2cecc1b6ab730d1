"""Experiment-runner entry points routed to the HIP path.

``execute_optimization`` has the signature, keyword mapping and return tuple of
pytopo3d.runners.experiment.execute_optimization
(pytopo3d/runners/experiment.py:95-213): the runner the reference's CLI and
pipeline call, which forwards its arguments to ``top3d``
(pytopo3d/runners/experiment.py:173-195) and times the run.  Here the call
goes to :func:`pytopo3d_amd.top3d`, so a runner switched to this module runs
every design iteration on the GPU.  ``save_result`` writes the optimized
design in the reference's result format (``np.save`` of the
``(nely, nelx, nelz)`` float64 array, pytopo3d/utils/results_manager.py:86-104).

The reference's experiment set-up, STL export and visualisation are not part
of the hot path and are not provided.
"""

from __future__ import annotations

import logging
import os
import time
from typing import Any, Dict, MutableMapping, Optional, Sequence, Tuple

import numpy as np

from .optimizer import top3d

_log = logging.getLogger("pytopo3d_amd.runner")


def execute_optimization(
    nelx: int,
    nely: int,
    nelz: int,
    volfrac: float,
    penal: float,
    rmin: float,
    disp_thres: float,
    material_params: Optional[Sequence[float]] = None,
    elem_size: float = 0.01,
    force_field: Optional[np.ndarray] = None,
    support_mask: Optional[np.ndarray] = None,
    tolx: float = 0.01,
    maxloop: int = 2000,
    create_animation: bool = False,
    animation_frequency: int = 10,
    logger: Optional[logging.Logger] = None,
    combined_obstacle_mask: Optional[np.ndarray] = None,
    use_gpu: bool = False,
    protected_zone_mask: Optional[np.ndarray] = None,
    beta_schedule: Sequence[float] = (1.0, 2.0, 4.0, 8.0),
    projection_eta: float = 0.5,
    move: float = 0.2,
    diagnostics_out: Optional[MutableMapping[str, Any]] = None,
    **hip_options: Any,
) -> Tuple[np.ndarray, Optional[Dict], float, Optional[float], float]:
    """Run one optimization and time it.

    Returns ``(xPhys, history, final_compliance, failure_force, run_time)`` like
    the reference runner: history is None unless ``create_animation``
    (snapshots every ``animation_frequency`` iterations), failure_force is the
    ``top3d`` fourth result (None), run_time in seconds.  ``hip_options`` are
    passed to :func:`pytopo3d_amd.top3d` (``precision``, ``solver``,
    ``solver_rtol``, ``force_vector``, ...); ``use_gpu`` is accepted for
    compatibility (the work always runs on the GPU).
    """
    log = logger or _log
    log.info("Starting optimization with %dx%dx%d elements...", nelx, nely, nelz)
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    log.debug("Optimization parameters: tolx=%s, maxloop=%s, save_history=%s, "
              "history_frequency=%s, ndof=%d, use_gpu=%s, beta_schedule=%s, projection_eta=%s",
              tolx, maxloop, create_animation, animation_frequency, ndof, use_gpu,
              tuple(beta_schedule), projection_eta)
    start = time.time()
    xPhys, history, final_compliance, failure_force = top3d(
        nelx,
        nely,
        nelz,
        volfrac,
        penal,
        rmin,
        disp_thres,
        material_params=material_params,
        elem_size=elem_size,
        force_field=force_field,
        support_mask=support_mask,
        obstacle_mask=combined_obstacle_mask,
        tolx=tolx,
        maxloop=maxloop,
        save_history=create_animation,
        history_frequency=animation_frequency,
        use_gpu=use_gpu,
        protected_zone_mask=protected_zone_mask,
        beta_schedule=beta_schedule,
        projection_eta=projection_eta,
        move=move,
        diagnostics_out=diagnostics_out,
        **hip_options,
    )
    run_time = time.time() - start
    if create_animation and history is not None:
        log.info("Optimization history captured with %d frames", len(history["density_history"]))
    log.debug("Optimization finished in %.2f seconds", run_time)
    return xPhys, history, final_compliance, failure_force, run_time


def save_result(result_array: np.ndarray, experiment_dir: str,
                filename: str = "optimized_design.npy") -> str:
    """``np.save`` the design to ``experiment_dir/filename`` and return the
    path (the reference's ResultsManager.save_result,
    pytopo3d/utils/results_manager.py:86-104)."""
    os.makedirs(experiment_dir, exist_ok=True)
    path = os.path.join(experiment_dir, filename)
    np.save(path, result_array)
    return path
