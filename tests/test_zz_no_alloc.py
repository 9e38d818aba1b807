"""lgh_expval_pauli_words_noalloc against lgh_expval_pauli_words.

The no-allocation entry point runs the same passes with its partials in a
fixed device scratch array shared by every call on the device.  Two calls in
a row (the second on a smaller state with fewer words, so stale contents of
the array from the first call would show) must give the results of the
allocating entry point, which the Pauli-word parity tests pin to the oracle;
both paths run the same kernels with the same chunking, so the results are
compared bit for bit.  Both calls go through the C ABI (ctypes).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = pytest.importorskip("pennylane_lightning_gpu._native")

from pennylane_lightning_gpu.workloads import ladder, random_pauli_hamiltonian  # noqa: E402


def note(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_pauli_words_match_allocating_entry(dtype):
    for n, nwords, seed in ((16, 300, 11), (11, 40, 12)):
        note(f"[noalloc] {np.dtype(dtype).name} n={n} words={nwords}")
        s = N.State(n, dtype).reset().apply(ladder(n, 3, seed))
        words, wires, coeffs = random_pauli_hamiltonian(n, nwords, seed)
        # a few words with Y factors and identities only
        words = list(words) + ["YY", "I", "XYZ"]
        wires = list(wires) + [[0, n - 1], [3], [1, 2, n - 2]]
        coeffs = np.concatenate([np.asarray(coeffs, dtype=np.complex128), [0.5, -1.25, 2.0]])
        t_ref, p_ref = s.expval_pauli_words(words, wires, coeffs)
        t_got, p_got = s.expval_pauli_words(words, wires, coeffs, noalloc=True)
        np.testing.assert_array_equal(p_got, p_ref)
        assert t_got == t_ref
        assert np.all(np.abs(p_ref) <= 1.0 + 1e-6)


def test_noalloc_rejects_too_many_words():
    s = N.State(10).reset()
    words = ["Z"] * 1025
    wires = [[w % 10] for w in range(1025)]
    with pytest.raises(N.LghError, match="at most 1024 words"):
        s.expval_pauli_words(words, wires, np.ones(1025), noalloc=True)
    # the allocating entry point has no such limit
    tot, per = s.expval_pauli_words(words, wires, np.ones(1025))
    assert tot == pytest.approx(1025.0)
