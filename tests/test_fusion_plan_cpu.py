"""Host-side plan of the fused passes (lgh_fusion_plan): no device needed.

The plan reports how many passes a gate sequence gets, how many of them are
standalone, and how many diagonal gates are applied inside diagonal runs.
These tests pin the scheduler's grouping of diagonal gates: inside a pass a
diagonal gate may move past a gate that uses its qubits only diagonally (a
control, a phase), never past a dense gate on one of its qubits.
"""
import pytest

N = pytest.importorskip("pennylane_lightning_gpu._native")


def plan(n, ops, adjoint=False):
    return N.fusion_plan(n, ops, adjoint=adjoint)


def test_benchmark_circuit_plan():
    from pennylane_lightning_gpu.workloads import ry_rz_cz_brick

    ops = ry_rz_cz_brick(28, 6, seed=1337)  # 498 gates: 168 RY, 168 RZ, 162 CZ
    ndiag = sum(1 for o in ops if o[0] in ("RZ", "CZ"))
    # default tile heights at 28 qubits: four register bits both ways, up to
    # 96 trainable gates per adjoint pass; the lookahead schedules take 7 + 7
    # passes (the one-ahead greedy schedule alone took 8 + 8)
    adj = plan(28, ops, adjoint=True)
    assert adj["standalone"] == 0 and adj["passes"] <= 7
    # every diagonal gate of the reverse sweep is a member of a diagonal run
    assert adj["run_members"] == ndiag
    fwd = plan(28, ops)
    assert fwd["standalone"] == 0 and fwd["passes"] <= 7
    assert fwd["run_members"] >= ndiag - 2
    # three register bits (complex128 below 24 qubits) at the same size: 10 passes
    N.set_tile_bits(4, 3)
    try:
        assert plan(28, ops, adjoint=True)["passes"] <= 10
    finally:
        N.set_tile_bits(4, 0)


@pytest.mark.parametrize("adjoint", [False, True])
def test_commuting_diagonal_gates_form_one_run(adjoint):
    # RZ / CZ on wires that the RY gates do not touch: one run of four
    ops = [("RZ", [0], [0.1], False), ("RY", [1], [0.2], False), ("RZ", [2], [0.3], False),
           ("RY", [3], [0.4], False), ("RZ", [4], [0.5], False), ("RY", [5], [0.6], False),
           ("CZ", [6, 7], [], False), ("RY", [8], [0.7], False)]
    assert plan(10, ops, adjoint) == {"passes": 1, "standalone": 0, "runs": 1, "run_members": 4}


def test_diagonal_gates_stay_behind_dense_gates_on_their_wire():
    # the two RZ never share a run (one of them may close the pass in a run
    # with the pass's own scale gate, which the plan does not count)
    ops = [("RZ", [0], [0.1], False), ("RY", [0], [0.2], False), ("RZ", [0], [0.3], False), ("RY", [0], [0.4], False)]
    assert plan(10, ops, adjoint=True)["run_members"] <= 1
    assert plan(10, ops)["run_members"] <= 1


@pytest.mark.parametrize("adjoint", [False, True])
def test_controls_commute_with_diagonal_gates_targets_do_not(adjoint):
    # RZ on the control of a CNOT moves past it: the two RZ form a run
    on_control = [("RZ", [0], [0.1], False), ("CNOT", [0, 1], [], False), ("RZ", [0], [0.3], False)]
    assert plan(10, on_control, adjoint)["run_members"] == 2
    # RZ on its target does not
    on_target = [("RZ", [1], [0.1], False), ("CNOT", [0, 1], [], False), ("RZ", [1], [0.3], False)]
    assert plan(10, on_target, adjoint)["run_members"] == 0


def test_standalone_gates_and_small_states():
    ops = [("RY", [0], [0.1], False), ("SWAP", [0, 5], [], False), ("RY", [5], [0.2], False)]
    assert plan(12, ops) == {"passes": 3, "standalone": 1, "runs": 0, "run_members": 0}
    # below 6 lane + 4 register bits no gate fits a fused forward pass
    small = [("RY", [0], [0.1], False), ("RZ", [1], [0.2], False)]
    assert plan(9, small)["standalone"] == 2


def test_plan_errors():
    with pytest.raises(N.LghError, match="adjoint"):
        plan(10, [("Rot", [0], [0.1, 0.2, 0.3], False)], adjoint=True)
    with pytest.raises(N.LghError):
        plan(0, [("RY", [0], [0.1], False)])
    with pytest.raises(N.LghError):
        plan(4, [("RY", [7], [0.1], False)])


def test_pauli_plan_groups_words_by_tile_bits():
    """The Pauli-word expectation groups words whose flip bits fit one tile
    of 2^11 amplitudes into one pass: the 24-qubit Hamiltonian of 256 random
    words (1-4 factors) takes a handful of state reads instead of one per
    distinct flip mask; words flipping more bits than a tile holds fall back
    to the pair passes, one per flip mask (host-only plan, no device)."""
    from pennylane_lightning_gpu.workloads import random_pauli_hamiltonian

    n = 24
    words, wires, _ = random_pauli_hamiltonian(n, 256, 1337)
    masks = {sum(1 << (n - 1 - w) for p, w in zip(wd, ws) if p in "XY") for wd, ws in zip(words, wires)}
    plan = N.pauli_plan(n, words, wires)
    assert plan["pair_launches"] == 0
    assert plan["tile_launches"] <= 24 < len(masks)
    wide = N.pauli_plan(n, ["X" * 12, "XX", "Y" * 12], [list(range(12)), [0, 23], list(range(12))])
    assert wide == {"tile_passes": 1, "tile_launches": 1, "pair_launches": 1}
    small = N.pauli_plan(4, ["XXXX", "YZ", "Z"], [[0, 1, 2, 3], [0, 3], [2]])
    assert small == {"tile_passes": 1, "tile_launches": 1, "pair_launches": 0}


@pytest.mark.parametrize("n,layers", [(21, 1), (27, 2)])
def test_strongly_entangling_plan(n, layers):
    """StronglyEntanglingLayers (Rot expanded, CNOT ranges 1 and 2 with
    wrap-around) fuses completely: no standalone pass, no CNOT and all but at
    most one RZ per layer in a diagonal run (an RZ between two dense gates on
    its wire can be left as a lone diagonal step), few passes each way."""
    from pennylane_lightning_gpu.workloads import strongly_entangling

    ops = strongly_entangling(n, layers, seed=1337)
    assert len(ops) == 4 * n * layers and sum(1 for o in ops if o[2]) == 3 * n * layers
    ranges = {(o[1][1] - o[1][0]) % n for o in ops if o[0] == "CNOT"}
    assert ranges == set(range(1, layers + 1))
    fwd, adj = plan(n, ops), plan(n, ops, adjoint=True)
    assert fwd["standalone"] == 0 and adj["standalone"] == 0
    assert fwd["passes"] <= 3 * layers and adj["passes"] <= 6 * layers
    assert 2 * n * layers - layers <= adj["run_members"] <= 2 * n * layers


def test_lean_pass_coverage():
    """Passes of uncontrolled 2x2 gates, controlled X gates and diagonal runs
    run in the lean kernels (host plan); a controlled rotation or a trainable
    gate without an X/Y generator keeps its pass on the general kernels."""
    from pennylane_lightning_gpu.workloads import ladder, layered_ring, ry_rz_cz_brick, strongly_entangling

    for n, ops in ((28, ry_rz_cz_brick(28, 6)), (27, strongly_entangling(27, 2)), (12, layered_ring(12, 4)),
                   (24, ladder(24, 8))):
        for adjoint in (False, True):
            assert N.fusion_lean_passes(n, ops, adjoint=adjoint) == plan(n, ops, adjoint)["passes"]
    ops = ry_rz_cz_brick(14, 2) + [("CRX", [3, 9], [0.4], False)]
    for adjoint in (False, True):
        p = plan(14, ops, adjoint)["passes"]
        assert N.fusion_lean_passes(14, ops, adjoint=adjoint) == p - 1
    # a trainable PhaseShift (generator not X/Y) on a register bit is lean only forward
    ops = ry_rz_cz_brick(12, 2) + [("PhaseShift", [5], [0.3], False)]
    assert N.fusion_lean_passes(12, ops) == plan(12, ops)["passes"]
