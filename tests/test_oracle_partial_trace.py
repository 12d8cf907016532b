"""Pins the oracle's partial-trace restatements (CPU only) against an
independent tensor contraction (numpy reshape + einsum over the traced
axes), known states (Bell pairs give I/2, product states give the product of
the kept kets), and the agreement of the ψ and ρ forms on |ψ><ψ|."""
import numpy as np
import pytest

from oracle import oracle as orc


def rand_psi(N, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << N) + 1j * rng.normal(size=1 << N)
    return v / np.linalg.norm(v)


def einsum_trace(psi, trace_qubits, N):
    """Tensor form: axis a of psi.reshape((2,)*N) is qubit N - a."""
    t = psi.reshape((2,) * N)
    keep = [q for q in range(1, N + 1) if q not in set(trace_qubits)]
    ax = {q: N - q for q in range(1, N + 1)}
    letters = "abcdefghijklmnopqrstuvwxyz"
    src = [letters[i] for i in range(N)]
    src2 = list(src)
    for q in keep:
        src2[ax[q]] = letters[N + ax[q]].upper()
    # output: kept qubits, highest first (row index big-endian over kept)
    rows = "".join(src[ax[q]] for q in reversed(keep))
    cols = "".join(src2[ax[q]] for q in reversed(keep))
    r = np.einsum(f"{''.join(src)},{''.join(src2)}->{rows}{cols}", t, t.conj())
    D = 1 << len(keep)
    return r.reshape(D, D)


@pytest.mark.parametrize("N,trace", [(3, [2]), (4, [1, 3]), (5, [5]), (5, [1, 2, 3, 4]),
                                     (6, [2, 5]), (4, []), (6, [1, 6])])
def test_psi_form_matches_tensor_contraction(N, trace):
    psi = rand_psi(N, N + len(trace))
    got = orc.partial_trace_psi(psi, trace, N)
    assert np.abs(got - einsum_trace(psi, trace, N)).max() < 1e-14
    assert abs(np.trace(got) - 1) < 1e-14
    assert np.abs(got - got.conj().T).max() < 1e-15


@pytest.mark.parametrize("N,trace", [(3, [2]), (4, [1, 3]), (5, [1, 2, 3, 4]), (4, [])])
def test_rho_form_agrees_on_pure_states(N, trace):
    psi = rand_psi(N, 7)
    rho = np.outer(psi, psi.conj())
    assert np.abs(orc.partial_trace_rho(rho, trace, N) -
                  orc.partial_trace_psi(psi, trace, N)).max() < 1e-14
    # and on a mixture, linearity
    psi2 = rand_psi(N, 8)
    mix = 0.3 * rho + 0.7 * np.outer(psi2, psi2.conj())
    want = 0.3 * orc.partial_trace_psi(psi, trace, N) + 0.7 * orc.partial_trace_psi(psi2, trace, N)
    assert np.abs(orc.partial_trace_rho(mix, trace, N) - want).max() < 1e-14


def test_known_states():
    bell = np.array([1, 0, 0, 1], np.complex128) / np.sqrt(2)
    assert np.allclose(orc.partial_trace_psi(bell, [2], 2), np.eye(2) / 2, atol=1e-16)
    assert np.allclose(orc.partial_trace_psi(bell, [1], 2), np.eye(2) / 2, atol=1e-16)
    kets = [np.array([np.cos(a), np.exp(1j * a) * np.sin(a)]) for a in (0.3, 0.9, 1.4)]
    psi = orc.product_ket(kets)
    # keep qubits 1 and 3: rho_A = kron(|k3><k3|, |k1><k1|) (qubit 1 lowest)
    want = np.kron(np.outer(kets[2], kets[2].conj()), np.outer(kets[0], kets[0].conj()))
    assert np.abs(orc.partial_trace_psi(psi, [2], 3) - want).max() < 1e-15
