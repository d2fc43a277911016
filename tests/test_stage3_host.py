"""Host-side checks of the stage-3 package (pamo_amd.safe_project) that need
no GPU: the reference's default configuration, the normalisation of
processing.py, the call signatures PaMO uses, and the loud failure without
a device."""
import inspect

import numpy as np
import pytest

from pamo_amd import safe_project as s3
from oracle import stage3 as o3


def test_config_defaults_are_the_references():
    """config.py:6-73: the values the default energies read."""
    c = s3.Stage3Config()
    want = dict(max_blocks=1 << 25, n_newton_iters=10, n_cg_iters=40, n_line_search_iters=10, d_hat=1e-3,
                coll_stiffness=1e2, contact_detection_radius=1e-2, ee_classify_thres=1e-3, ccd_slackness=0.7,
                ccd_thickness=1e-6, ccd_max_iters=100, spd_max_iters=3, elas_young_modulus=1e-1,
                elas_poisson_ratio=0.0, hinge_stiffness=1e-2, mesh2gt_dist_stiffness=1e3, min_angle_thres=0.0)
    for k, v in want.items():
        assert getattr(c, k) == v, k
    # the arrays handed to the library are laid out as the oracle's
    d, i = c.arrays()
    od, oi = o3.Config().arrays()
    assert np.array_equal(d, od) and np.array_equal(i, oi)


def test_normalisation_is_processing_py():
    """processing.py:15-20: scale = 1 / largest extent, t = -mean * scale."""
    rng = np.random.default_rng(3)
    V = rng.normal(size=(500, 3)) * [1.0, 3.0, 0.5] + [4.0, -2.0, 7.0]
    s, t = s3.get_normalization_transform(V)
    assert s == pytest.approx(1.0 / (V[:, 1].max() - V[:, 1].min()))
    assert np.allclose(t, -V.mean(axis=0) * s)
    W = V * s + t
    assert np.allclose(W.mean(axis=0), 0.0) and np.max(W.max(0) - W.min(0)) == pytest.approx(1.0)
    so, to = o3.normalization(V)
    assert s == so and np.array_equal(t, to)


def test_signatures_follow_the_reference():
    """pamo_safe_project.process(gt_V, gt_F, stage2_V, stage2_F, n_iters,
    system=None, config=None) and Stage3System's methods (processing.py:23,
    system.py:17-440)."""
    p = list(inspect.signature(s3.process).parameters)
    assert p[:7] == ["gt_V", "gt_F", "stage2_V", "stage2_F", "n_iters", "system", "config"]
    for m in ("register_mesh", "step", "get_vertices", "clear"):
        assert callable(getattr(s3.Stage3System, m))
    assert list(inspect.signature(s3.Stage3System.register_mesh).parameters)[1:] == ["gt_V", "gt_F", "V", "F"]


def test_system_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        s3.Stage3System()
