"""GPU test of the training pipeline (range-solver-lib_amd/learned_train.py) at a small size:
pairs recorded from the solver's own flexible-CG runs, the GPU trainer lowers the held-out
loss, and the trained network keeps the solve converged."""
import importlib.util
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _train_module():
    spec = importlib.util.spec_from_file_location(
        "learned_train", os.path.join(ROOT, "range-solver-lib_amd", "learned_train.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_training_pipeline_small(gpu):
    rs = gpu
    lt = _train_module()
    theta, summary = lt.train(rs, cells=16, sigma=2.5, problems=16, seed0=500, holdout=0.25, record=32,
                              batch=16, steps=80, lr=2e-3, out=None, log_every=40)
    print(summary["heldout_loss_start"], summary["heldout_loss_trained"], summary["heldout_iterations"],
          flush=True)
    assert theta.shape == (rs.N_PARAMS,) and np.all(np.isfinite(theta))
    assert summary["pairs"] > 100
    assert summary["heldout_loss_trained"] < 0.9 * summary["heldout_loss_start"]
    its = summary["heldout_iterations"]
    assert its["exact"] <= its["jacobi"]
    assert its["network"] < 2000          # converged (iteration cap of the held-out solves)
