"""Shared test set-up: markers, import paths, oracle and scene helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import oit_oracle

    oit_oracle.lib()
    return oit_oracle


def deep_scene(seed=3, width=48, height=40, spheres=160, radius=(0.05, 0.2), alpha=(0.1, 0.6), samples=1,
               cluster_fraction=0.0):
    """A small scene with many layers per pixel, so that every K engages the tail."""
    from vk_order_independent_transparency_amd import scene

    p = scene.SceneParams(seed=seed, width=width, height=height, spheres=spheres, radius=radius, alpha=alpha,
                          samples=samples, cluster_fraction=cluster_fraction, cluster_extent=0.2)
    tg, d, c = scene.generate(p)
    return p, tg, d, c


def per_target_counts(target, T):
    return np.bincount(np.asarray(target, dtype=np.int64), minlength=T)
