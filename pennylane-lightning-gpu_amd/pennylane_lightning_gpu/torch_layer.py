"""torch.autograd forms of circuit expectation values.

The reference differentiates through PennyLane's interfaces: the device's
``adjoint_jacobian`` gives the Jacobian and ``vjp`` contracts it with the
incoming gradient (pennylane_lightning_gpu/lightning_gpu.py:638-752 and
:771-810).  Here the same contraction is a ``torch.autograd.Function``:

- ``BatchedExpval``: a batch of circuits over one template (a gate structure
  with parameter slots and a Pauli-word observable, up to 20 qubits).  The
  forward is one persistent batched launch (a replay of a cached batch plan,
  lgh_batch_plan_run, per template and batch size) that returns
  <P> for every circuit and, when a gradient is needed, the adjoint
  Jacobian of every circuit in the same launch; the backward is
  grad_out[:, None] * J.
- ``CircuitExpval``: one circuit of any size through a device state
  (lgh_apply, the Pauli-word expectation and lgh_adjoint_jacobian).

Parameters may live on any torch device; the values cross to the host
(B x P doubles), the states never leave the GPU.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np
import torch

from . import _native as N


def _host64(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy()


# Batch plans (captured launch sequence, tuned grid) of recent batch shapes:
# a training loop calls the same template with the same batch size every
# step, so each step replays a graph instead of rebuilding the launch.
_PLANS: "OrderedDict" = OrderedDict()
_PLANS_MAX = 8


def _plan(template: "N.Template", batch: int, grad: bool, dtype) -> "N.BatchPlan":
    key = (template.handle.value, batch, bool(grad), np.dtype(dtype).name)
    plan = _PLANS.get(key)
    if plan is None or not plan.handle.value:
        plan = N.BatchPlan([template] * batch, grad=grad, dtype=dtype)
        _PLANS[key] = plan
        while len(_PLANS) > _PLANS_MAX:
            _PLANS.popitem(last=False)
    _PLANS.move_to_end(key)
    return plan


class BatchedExpval(torch.autograd.Function):
    """expvals[b] = <P> of ``template`` with parameters ``params[b]``."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, template: "N.Template", dtype=np.complex128):
        if params.dim() != 2 or params.shape[1] != template.num_params:
            raise ValueError(f"params must be (batch, {template.num_params})")
        p = _host64(params)
        need = params.requires_grad
        ev, jac = _plan(template, p.shape[0], need, dtype).run(list(p)) if p.shape[0] else \
            N.batch_run([], [], grad=need, dtype=dtype)
        if need:
            J = torch.from_numpy(np.stack(jac) if jac else np.zeros((0, template.num_params)))
            ctx.save_for_backward(J.to(params.device, params.dtype))
        return torch.from_numpy(ev).to(params.device, params.dtype)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (J,) = ctx.saved_tensors
        return grad_out[:, None] * J, None, None


def batched_expval(params: torch.Tensor, template: "N.Template", dtype=np.complex128) -> torch.Tensor:
    return BatchedExpval.apply(params, template, dtype)


class CircuitExpval(torch.autograd.Function):
    """<P> of one circuit on ``num_qubits`` qubits.

    ``ops``: ``(name, wires, params, inverse)`` tuples; ``param_index[i]`` is
    the entry of ``params`` that drives op i (``-1``: fixed).  ``observable``:
    ``[("Z", wire), ...]``.
    """

    @staticmethod
    def forward(ctx, params: torch.Tensor, num_qubits: int, ops, param_index, observable,
                dtype=np.complex128):
        theta = _host64(params).ravel()
        bound = [(o[0], o[1], [theta[p]] if p >= 0 else list(o[2]), o[3]) for o, p in zip(ops, param_index)]
        psi = N.State(num_qubits, dtype=dtype).reset()
        psi.apply(bound)
        word = ["I"] * num_qubits
        for p, w in observable:
            word[w] = p
        e = psi.expval_pauli_words(["".join(word)], [list(range(num_qubits))], [1.0])
        e = float(np.real(e[0] if isinstance(e, (tuple, list)) else e))
        if params.requires_grad:
            bra = N.State(num_qubits, dtype=dtype)
            bra.copy_from(psi)
            pops = [("Pauli" + p, [w], [], False) for p, w in observable]
            if pops:
                bra.apply(pops)
            par_ops = [i for i, o in enumerate(bound) if o[2]]
            tr = [j for j, i in enumerate(par_ops) if param_index[i] >= 0]
            row = N.adjoint_jacobian(psi, [bra], bound, tr)[0]
            g = np.zeros(theta.size)
            for col, j in enumerate(tr):
                g[param_index[par_ops[j]]] += row[col]
            ctx.save_for_backward(torch.from_numpy(g).to(params.device, params.dtype))
        return torch.tensor(e, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None, None, None


def circuit_expval(params, num_qubits, ops, param_index, observable, dtype=np.complex128) -> torch.Tensor:
    return CircuitExpval.apply(params, num_qubits, ops, param_index, observable, dtype)
