"""torch.autograd forms (BatchedExpval, CircuitExpval) against the oracle and
central finite differences (h = 1e-4, error O(h^2) ~ 1e-8)."""
import numpy as np
import pytest

from _tol import close128
from test_gpu_batch import oracle_run, ring_template

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
N = pytest.importorskip("pennylane_lightning_gpu._native")
from pennylane_lightning_gpu.torch_layer import batched_expval, circuit_expval  # noqa: E402


def test_batched_expval_backward_is_weighted_jacobian():
    rng = np.random.default_rng(21)
    n = 11
    ops, pidx = ring_template(n, layers=2)
    t = N.Template(n, ops, pidx, [("Z", 0)])
    B = 5
    theta = rng.uniform(0, 2 * np.pi, (B, t.num_params))
    w = rng.normal(size=B)
    p = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    ev = batched_expval(p, t)
    (ev * torch.tensor(w)).sum().backward()
    for b in range(B):
        e, g = oracle_run(n, ops, pidx, theta[b], [("Z", 0)])
        close128(ev[b].item(), e)
        close128(p.grad[b].numpy(), w[b] * g)


def test_batched_expval_gradcheck():
    n = 4
    ops, pidx = ring_template(n, layers=1)
    t = N.Template(n, ops, pidx, [("Z", 1)])
    p = torch.tensor(np.random.default_rng(2).uniform(0, 2 * np.pi, (2, t.num_params)), dtype=torch.float64,
                     requires_grad=True)
    assert torch.autograd.gradcheck(lambda x: batched_expval(x, t), (p,), eps=1e-6, atol=1e-6)


@pytest.mark.parametrize("n", [6, 14])
def test_circuit_expval_matches_oracle_and_fd(n):
    rng = np.random.default_rng(n)
    ops, pidx = ring_template(n, layers=2)
    theta = rng.uniform(0, 2 * np.pi, max(pidx) + 1)
    obs = [("Z", 0), ("X", n - 1)]
    p = torch.tensor(theta, requires_grad=True)
    e = circuit_expval(p, n, ops, pidx, obs)
    e.backward()
    e_ref, g_ref = oracle_run(n, ops, pidx, theta, obs)
    close128(e.item(), e_ref)
    close128(p.grad.numpy(), g_ref)
    h = 1e-4
    for j in rng.choice(theta.size, size=3, replace=False):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd = (circuit_expval(torch.tensor(tp), n, ops, pidx, obs).item() -
              circuit_expval(torch.tensor(tm), n, ops, pidx, obs).item()) / (2 * h)
        assert abs(fd - p.grad[j].item()) < 1e-7


def test_circuit_and_batched_forms_agree():
    rng = np.random.default_rng(9)
    n = 12
    ops, pidx = ring_template(n, layers=2)
    t = N.Template(n, ops, pidx, [("Z", 0)])
    theta = rng.uniform(0, 2 * np.pi, t.num_params)
    pb = torch.tensor(theta[None, :], requires_grad=True)
    batched_expval(pb, t).sum().backward()
    pc = torch.tensor(theta, requires_grad=True)
    circuit_expval(pc, n, ops, pidx, [("Z", 0)]).backward()
    close128(pb.grad[0].numpy(), pc.grad.numpy())
