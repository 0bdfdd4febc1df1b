"""Learned preconditioner on the GPU: the BF16/MFMA network against its FP64 oracle
(oracle/lpc_oracle.py, itself pinned by finite differences), the hand-written
backward pass against the oracle's gradients, the solver coupling (a network with a
zero head is exactly the plain multigrid cycle), the gradient of the loss through the
cycle against central differences of the loss, and training on residuals recorded
from the solver.

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
from _helpers import phase, rel_err

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


def _mg_system(n=32, kind=O.SCENE_RT, rho_g=1.0, seed=21):
    """A projection-ready context (one P2G of a seeded scene) and its phase."""
    ctx = G.Context(n, n, n, 1.0 / n)
    ctx.scene(kind, seed)
    oph = phase(1000.0, rho_g)
    ph = G.Phase(oph.rho_l, oph.rho_g, oph.alpha_phi, oph.rho_tilde_0)
    _prepare(ctx, ph)
    return ctx, ph


def _prepare(ctx, ph, dt=1e-3):
    ctx.mark_cells()
    ctx.p2g(ph, 1.0)
    ctx.save_velocity()
    ctx.apply_gravity(dt, -9.81)
    ctx.apply_walls()


def _cfg(precond, tol=1e-6):
    cfg = G.default_solve_cfg(G.STOP_REL2, tol, 20000, precond)
    cfg.inner_rel_tol, cfg.check_every = 0.01, 4
    return cfg


def _recorded(n=32, kind=O.SCENE_RT, seed=3, snapshots=16):
    ctx, ph = _mg_system(n, kind, seed=seed)
    net = G.LearnedPreconditioner(seed=seed)
    net.attach(ctx)
    G.lpc_record(ctx, snapshots, 1)
    st = ctx.project(ph, 1e-3, _cfg(G.PRECOND_MULTIGRID))
    assert st.converged == 1
    return ctx, ph, net


def test_zero_head_network_is_the_plain_cycle():
    # a new network has a zero head: every aggregate weight is exp(0) = 1
    ctx, ph, net = _recorded(32, seed=2)
    ns = G.lpc_snapshot_count(ctx)
    assert ns > 0
    a, b = net.loss(ctx, 0)
    assert a == b
    _prepare(ctx, ph)
    st_m = ctx.project(ph, 1e-3, _cfg(G.PRECOND_MULTIGRID))
    _prepare(ctx, ph)
    st_l = ctx.project(ph, 1e-3, _cfg(G.PRECOND_LEARNED))
    assert st_m.converged == 1 and st_l.converged == 1 and st_l.fallback_used == 0
    assert st_l.iterations == st_m.iterations
    assert st_l.true_rel_residual == st_m.true_rel_residual


def _bf16_grid_head(net, rng, scale=4):
    """weights with the head on multiples of 2^-7 (exact in BF16), so that +-2^-7
    steps of a head weight reach the network unrounded"""
    w = net.weights().copy()
    w[-16:] = rng.integers(-scale, scale + 1, 16).astype(np.float32) * 2.0 ** -7
    net.set_weights(w)
    return w


@pytest.mark.parametrize("objective", ["residual", "energy"])
def test_gradient_matches_finite_differences(objective):
    # the adjoint through the cycle (post-smoothing, restriction, the scaled coarse
    # solve and the network) against central differences of the loss itself; the
    # energy form without power steps, so that e0 (the snapshot) does not depend on
    # the weights
    ctx, ph, net = _recorded(32, seed=4)
    net.set_objective(G.LPC_OBJECTIVE_ENERGY if objective == "energy" else G.LPC_OBJECTIVE_RESIDUAL, 0)
    rng = np.random.default_rng(7)
    w = _bf16_grid_head(net, rng)
    for k in (0, G.lpc_snapshot_count(ctx) - 1):
        g, loss = net.grad(ctx, k)
        assert np.isfinite(g).all() and loss > 0
        eps = 2.0 ** -7
        for j in (0, 5, 11, 15):
            idx = len(w) - 16 + j
            wp, wm = w.copy(), w.copy()
            wp[idx] += eps
            wm[idx] -= eps
            net.set_weights(wp)
            lp, _ = net.loss(ctx, k)
            net.set_weights(wm)
            lm, _ = net.loss(ctx, k)
            net.set_weights(w)
            fd = (lp - lm) / (2 * eps)
            print(f"[lpc] {objective} snapshot {k} head {j}: grad {g[idx]:.6e} fd {fd:.6e}", flush=True)
            assert abs(fd - g[idx]) <= 0.03 * max(abs(fd), abs(g[idx])) + 1e-7
        # a direction over the third layer: BF16 rounding of the perturbed weights
        # moves the effective direction by up to ~20% of d, and loss differences below
        # ~1e-6 are rounding noise
        d = np.zeros_like(w)
        o3, n3 = LO.SIZES[0] + LO.SIZES[1], LO.SIZES[2]
        d[o3:o3 + n3] = rng.normal(size=n3).astype(np.float32)
        d *= 0.02 * np.linalg.norm(w[o3:o3 + n3]) / np.linalg.norm(d)
        net.set_weights(w + d)
        lp, _ = net.loss(ctx, k)
        net.set_weights(w - d)
        lm, _ = net.loss(ctx, k)
        net.set_weights(w)
        fd = (lp - lm) / 2
        gd = float(np.dot(g.astype(np.float64), d.astype(np.float64)))
        print(f"[lpc] {objective} snapshot {k} layer-3 direction: grad.d {gd:.6e} fd {fd:.6e}", flush=True)
        assert abs(fd - gd) <= 0.25 * max(abs(fd), abs(gd)) + 1e-6


def test_training_lowers_the_loss():
    ctx, ph, net = _recorded(32, seed=5, snapshots=24)
    ns = G.lpc_snapshot_count(ctx)
    assert ns >= 4
    before = [net.loss(ctx, k) for k in range(ns)]
    net.train(ctx, 60, lr=2e-3, seed=5)
    after = [net.loss(ctx, k) for k in range(ns)]
    l0 = np.mean([a for a, _ in before])
    l1 = np.mean([a for a, _ in after])
    lp = np.mean([b for _, b in after])
    print(f"[lpc] loss: plain cycle {lp:.5f}, untrained {l0:.5f}, trained {l1:.5f}", flush=True)
    assert l0 == lp and l1 < l0
    # the trained network in the solve: same FP64 tolerance, no fallback
    _prepare(ctx, ph)
    st = ctx.project(ph, 1e-3, _cfg(G.PRECOND_LEARNED))
    assert st.converged == 1 and st.true_rel_residual <= 1e-6 and st.fallback_used == 0


def test_learned_solve_against_plain_pcg():
    """The learned preconditioner as shipped (no trained weights ship, so its network
    has the zero head: the multigrid cycle) inside the flexible PCG + FP64 refinement,
    against plain Jacobi PCG + FP64 refinement on the same 128^3 RT system at 1000:1:
    the same FP64 tolerance, no fallback to Jacobi, at most a tenth of the iterations
    and at most half the wall time (best of two solves each)."""
    import time

    n = 128
    ctx, ph = _mg_system(n, O.SCENE_RT, seed=7)
    net = G.LearnedPreconditioner(seed=7)
    net.attach(ctx)
    out = {}
    for name, pre in (("plain", G.PRECOND_JACOBI), ("learned", G.PRECOND_LEARNED)):
        best = None
        for _ in range(2):
            _prepare(ctx, ph)
            ctx.synchronize()
            t0 = time.perf_counter()
            st = ctx.project(ph, 1e-3, _cfg(pre))
            ctx.synchronize()
            ms = (time.perf_counter() - t0) * 1e3
            assert st.converged == 1 and st.true_rel_residual <= 1e-6, (name, st.true_rel_residual)
            if best is None or ms < best[0]:
                best = (ms, st.iterations, st.fallback_used)
        out[name] = best
    print(f"[lpc] 128^3 RT solve: plain PCG {out['plain'][0]:.1f} ms {out['plain'][1]} it, "
          f"learned {out['learned'][0]:.1f} ms {out['learned'][1]} it", flush=True)
    assert out["learned"][2] == 0
    assert out["learned"][1] * 10 <= out["plain"][1]
    assert out["learned"][0] * 2 <= out["plain"][0]
    G.lib().pf_lpc_attach(ctx.h, None)
    net.close()
    ctx.close()


def test_objective_rejects_invalid_values():
    net = G.LearnedPreconditioner(seed=1)
    for obj, p in ((2, 0), (-1, 0), (G.LPC_OBJECTIVE_ENERGY, 65), (G.LPC_OBJECTIVE_ENERGY, -1)):
        with pytest.raises(G.PfError):
            net.set_objective(obj, p)
    net.set_objective(G.LPC_OBJECTIVE_ENERGY, 64)
    net.close()
