"""Training and evaluation loops of the learned E/B field surrogate.

The network, its forward and backward passes and the optimiser run on the
GPU inside libionsolver_hip.so; this module only feeds it data: each training
step writes a random set of LOD sums (scenarios.random_lod), computes the
direct field update_e_b_dynamic as the target (ion_compute_fields) and runs
one training step (ion_surrogate_train_step).
"""
import numpy as np

from .lbm import Field
from .scenarios import random_lod


def lod_buffer(lbm, top):
    """Full LOD buffer with the given top-level entries (others zero)."""
    n_lod = lbm.config.kernel_constants()["n_lod"]
    buf = np.zeros((n_lod, 4), np.float32)
    buf[:top.shape[0]] = top
    return buf.ravel()


def train(lbm, steps, seed=0, log_every=0, **lod_kw):
    """Train the surrogate of `lbm` on `steps` random LOD configurations.
    Returns the list of (rel_e, rel_b) training residuals."""
    rng = np.random.default_rng(seed)
    c = lbm.config
    shape, D = (c.n_x, c.n_y, c.n_z), c.mhd_lod_depth
    hist = []
    for i in range(steps):
        lbm.write_field(Field.QU_LOD, lod_buffer(lbm, random_lod(shape, D, rng, **lod_kw)))
        lbm.compute_fields()
        hist.append(lbm.surrogate_train_step(True))
        if log_every and (i % log_every == 0 or i == steps - 1):
            print(f"surrogate train step {i}: rel_e {hist[-1][0]:.4e} rel_b {hist[-1][1]:.4e}", flush=True)
    return hist


def evaluate(lbm, n, seed=12345, **lod_kw):
    """Mean relative L2 error of the surrogate on n held-out configurations."""
    rng = np.random.default_rng(seed)
    c = lbm.config
    shape, D = (c.n_x, c.n_y, c.n_z), c.mhd_lod_depth
    res = []
    for _ in range(n):
        lbm.write_field(Field.QU_LOD, lod_buffer(lbm, random_lod(shape, D, rng, **lod_kw)))
        lbm.compute_fields()
        res.append(lbm.surrogate_residual())
    return np.array(res)
