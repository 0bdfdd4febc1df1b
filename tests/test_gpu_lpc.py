"""Learned block preconditioner on the GPU: the BF16/MFMA network against its FP64
oracle (oracle/lpc_oracle.py, itself pinned by finite differences), the hand-written
backward pass against the oracle's gradients, the solver coupling (a zero network is
exactly Jacobi), training on residuals recorded from the solver, and the flexible
PCG reaching the FP64 tolerance with it.

Tolerances: the network runs in BF16 with FP32 accumulation.  Against the oracle
with the same BF16 rounding points the outputs agree to 2e-3 (norm-wise relative;
FP32 accumulation order can flip a stored BF16 value by one unit); against exact
FP64 to 3e-2; the gradients (BF16 activations and gradients) to 3e-2 per layer.
"""
import numpy as np
import pytest

import lpc_oracle as LO
import pfflip_amd as G
import pyoracle as O
from _helpers import phase, rel_err, scene

pytestmark = pytest.mark.gpu


def _inputs(nt, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((nt, 18, 18, 18, 4), np.float32)
    X[..., 0] = rng.normal(size=(nt, 18, 18, 18))
    X[..., 1] = rng.uniform(0.0, 1.0, size=(nt, 18, 18, 18))
    X[..., 2] = rng.uniform(0.0, 1.0, size=(nt, 18, 18, 18))
    return X


def test_network_forward_matches_oracle():
    net = G.LearnedPreconditioner(seed=1)
    w = LO.init_weights(3).astype(np.float32)
    net.set_weights(w)
    X = _inputs(3, 7)
    O_gpu = net.forward_tiles(X)
    O_bf, _ = LO.forward(X, w.astype(np.float64), bf16=True)
    O_64, _ = LO.forward(X, w.astype(np.float64), bf16=False)
    assert rel_err(O_gpu, O_bf) < 2e-3
    assert rel_err(O_gpu, O_64) < 3e-2
    # the weights round-trip exactly
    assert np.array_equal(net.weights(), w)


def test_network_gradients_match_oracle():
    net = G.LearnedPreconditioner(seed=1)
    w = LO.init_weights(5).astype(np.float32)
    net.set_weights(w)
    X = _inputs(2, 9)
    gO = np.random.default_rng(4).normal(size=(2, 16, 16, 16)).astype(np.float32)
    gW = net.grad_tiles(X, gO)
    _, cache = LO.forward(X, w.astype(np.float64), bf16=True)
    g_ref = LO.backward(cache, gO, w.astype(np.float64), bf16=True)
    o = 0
    for name, n in zip(("W1", "W2", "W3", "Wh"), LO.SIZES):
        e = rel_err(gW[o:o + n], g_ref[o:o + n])
        assert e < 3e-2, (name, e)
        o += n


def _system(n=32, kind=O.SCENE_RT, rho_g=1.0):
    """A projection-ready context (RT scene, one P2G) and its solve configuration."""
    dx, x, y, z, t = scene(kind, n, seed=21)
    ctx = G.Context(n, n, n, dx)
    ctx.upload(x, y, z, type=t)
    oph = phase(1000.0, rho_g)
    ph = G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)
    ctx.mark_cells()
    ctx.p2g(ph, 1.0)
    ctx.save_velocity()
    ctx.apply_gravity(1e-3, -9.81)
    ctx.apply_walls()
    return ctx, ph


def _saved_faces(ctx):
    return [ctx.field(f) for f in (G.FIELD_U, G.FIELD_V, G.FIELD_W)]


def _restore_faces(ctx, faces):
    for f, a in zip((G.FIELD_U, G.FIELD_V, G.FIELD_W), faces):
        ctx.set_field(f, a)


def test_zero_network_is_jacobi():
    ctx, ph = _system(24)
    faces = _saved_faces(ctx)
    net = G.LearnedPreconditioner(seed=2)
    net.set_weights(np.zeros(G.LearnedPreconditioner.num_weights(), np.float32))
    net.attach(ctx)
    G.lpc_record(ctx, 8, 5)
    st_j = ctx.project(ph, 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-6, 20000))
    assert G.lpc_snapshot_count(ctx) > 0
    a, b = net.loss(ctx, 0)
    assert a == b      # z = r/diag + s * 0 exactly
    _restore_faces(ctx, faces)
    st_l = ctx.project(ph, 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-6, 20000, G.PRECOND_LEARNED))
    assert st_j.converged == 1 and st_l.converged == 1 and st_l.fallback_used == 0
    # flexible beta with a fixed preconditioner is the standard beta up to rounding
    assert abs(st_l.iterations - st_j.iterations) <= max(3, st_j.iterations // 20)


def test_training_lowers_the_loss_and_the_flexible_solve_converges():
    ctx, ph = _system(32)
    faces = _saved_faces(ctx)
    net = G.LearnedPreconditioner(seed=3)
    net.attach(ctx)
    G.lpc_record(ctx, 24, 6)
    st_j = ctx.project(ph, 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-6, 20000))
    ns = G.lpc_snapshot_count(ctx)
    assert ns >= 8
    before = [net.loss(ctx, k) for k in range(ns)]
    net.train(ctx, 150, lr=2e-3, seed=5)
    after = [net.loss(ctx, k) for k in range(ns)]
    l0 = np.mean([a for a, _ in before])
    l1 = np.mean([a for a, _ in after])
    lj = np.mean([b for _, b in after])
    print(f"loss: untrained {l0:.4f} trained {l1:.4f} jacobi {lj:.4f}")
    assert l1 < l0 and l1 < lj
    # the trained network inside flexible PCG reaches the FP64 tolerance
    _restore_faces(ctx, faces)
    st_l = ctx.project(ph, 1e-3, G.default_solve_cfg(G.STOP_REL2, 1e-6, 20000, G.PRECOND_LEARNED))
    print(f"iterations: jacobi {st_j.iterations}, learned {st_l.iterations} (fallback {st_l.fallback_used})")
    assert st_l.converged == 1 and st_l.true_rel_residual <= 1e-6
