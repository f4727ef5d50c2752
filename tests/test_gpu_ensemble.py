"""GPU: the differentiated ensemble (pmpm_batch_step forward, eager or graph
replay, and pmpm_batch_adjoint_step reverse) against each scene run alone
through diff.simulate, the single-scene adjoint whose gradients the adjoint
tests pin against an FP64 restatement and finite differences."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _losses(xs, x0s):
    return [((x - x0) ** 2).mean() for x, x0 in zip(xs, x0s)]


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_matches_scenes_run_alone(graph):
    from pyroclastmpm_amd.diff import simulate
    from pyroclastmpm_amd.ensemble import Ensemble, column_ensemble, simulate_ensemble

    # three columns of different sizes, cell sizes and materials; 10 steps
    # with a checkpoint every 4 (segments of 4, 4 and 2 steps)
    kw = dict(count=3, seed=2, n_min=2000, n_max=9000, dt=1e-4, checkpoint_every=4)
    steps = 10
    print(f"ensemble parity: graph={graph} {kw} steps={steps}", flush=True)
    models, thetas = column_ensemble(**kw)
    ens = Ensemble(models, lanes=3)
    x0s = [m.positions(m.initial) for m in models]

    th = thetas.clone().requires_grad_(True)
    xs = simulate_ensemble(ens, th, steps, graph=graph)
    sum(_losses(xs, x0s)).backward()
    g_ens = th.grad.clone()

    alone, _ = column_ensemble(**kw)
    for i, m in enumerate(alone):
        t = thetas[i].clone().requires_grad_(True)
        x = simulate(m, t, steps)
        assert torch.allclose(xs[i], x, rtol=0, atol=1e-6), (i, float((xs[i] - x).abs().max()))
        assert float((x - x0s[i]).abs().max()) > 1e-6, "the column must move"
        ((x - x0s[i]) ** 2).mean().backward()
        # Every node sum is formed in a fixed order (the wrap-around deposits
        # included), so a scene gives the same bits batched or alone; the
        # cohesion gradient, which collects only from particles at the
        # Drucker-Prager apex (c = 0) and so follows the last bits of the
        # stress, is held to the same bar as phi and E.
        rel = [abs(float(g_ens[i, j]) - float(t.grad[j])) / abs(float(t.grad[j])) for j in range(3)]
        print(f"scene {i}: n={m.n} max |dx|={float((xs[i] - x).abs().max()):.3g} grad rel diff={rel}", flush=True)
        assert max(rel) <= 1e-5, (i, g_ens[i], t.grad)
        assert all(float(v) != 0.0 for v in t.grad), t.grad
