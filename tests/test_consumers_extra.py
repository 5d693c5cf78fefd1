"""MixedPrecisionFockBuilder and the derivative integral buffer on the host
(include/libaccint/consumers/mixed_precision_fock_builder.hpp:43-147,
include/libaccint/buffers/two_electron_buffer.hpp:41-307).  CPU only."""
import numpy as np
import pytest

import libaccint as lai


def _blocks(seed):
    """A few random integral blocks and their first functions in a 10-function basis."""
    rng = np.random.default_rng(seed)
    shapes = (((3, 1, 3, 1), (0, 3, 4, 7)), ((1, 1, 1, 1), (2, 5, 6, 1)), ((3, 3, 1, 1), (4, 0, 8, 9)))
    return [(rng.standard_normal(dims), offs) for dims, offs in shapes]


def _accumulate(J, K, D, g, offs):
    """J_{mu nu} += g D_{la si}, K_{mu la} += g D_{nu si}
    (src/shared/consumers/mixed_precision_fock_builder.cpp:43-75)."""
    fa, fb, fc, fd = offs
    na, nb, nc, nd = g.shape
    J[fa:fa + na, fb:fb + nb] += np.einsum("abcd,cd->ab", g, D[fc:fc + nc, fd:fd + nd])
    K[fa:fa + na, fc:fc + nc] += np.einsum("abcd,bd->ac", g, D[fb:fb + nb, fd:fd + nd])


def test_mixed_builder_accumulates_promoted_fp32_and_fp64_blocks():
    n = 10
    X = np.random.default_rng(3).standard_normal((n, n))
    D = X + X.T
    mb = lai.MixedPrecisionFockBuilder(n)
    mb.set_density(D)
    J = np.zeros((n, n))
    K = np.zeros((n, n))
    for k, (g, offs) in enumerate(_blocks(5)):
        if k % 2 == 0:  # FP32 block: rounded to float, promoted before accumulation
            b = lai.TwoElectronBufferF32()
            b.from_numpy(g.astype(np.float32))
            gi = g.astype(np.float32).astype(np.float64)
        else:
            b = lai.TwoElectronBuffer()
            b.from_numpy(g)
            gi = g
        mb.accumulate(b, *offs)
        _accumulate(J, K, D, gi, offs)
    assert mb.n_float32_accumulations() == 2 and mb.n_float64_accumulations() == 1
    assert mb.mode() == lai.MixedPrecisionMode.Compute32Accumulate64
    assert np.max(np.abs(mb.get_coulomb_matrix() - J)) < 1e-12
    assert np.max(np.abs(mb.get_exchange_matrix() - K)) < 1e-12
    # the FP32 blocks were rounded: unrounded values give a different J
    J64 = np.zeros((n, n))
    K64 = np.zeros((n, n))
    for g, offs in _blocks(5):
        _accumulate(J64, K64, D, g, offs)
    assert np.max(np.abs(J64 - J)) > 1e-9
    H = np.eye(n)
    assert np.allclose(mb.get_fock_matrix(H, 0.5), H + J - 0.5 * K, atol=1e-12)
    mb.reset()
    assert np.all(mb.get_coulomb_matrix() == 0) and mb.n_float32_accumulations() == 0


def test_mixed_builder_errors_and_threshold():
    mb = lai.MixedPrecisionFockBuilder(4)
    with pytest.raises(ValueError):
        mb.set_density(np.eye(5))
    b = lai.TwoElectronBufferF32()
    b.from_numpy(np.ones((1, 1, 1, 1), dtype=np.float32))
    with pytest.raises(RuntimeError, match="density"):
        mb.accumulate(b, 0, 0, 0, 0)
    assert mb.fp32_threshold() == 1e-7
    mb.set_fp32_threshold(1e-9)
    assert mb.fp32_threshold() == 1e-9


def test_gradient_buffer_layout():
    na, nb, nc, nd = 3, 1, 6, 1
    g = lai.TwoElectronGradientBuffer(na, nb, nc, nd)
    assert lai.TwoElectronGradientBuffer.N_DERIV == 12 and g.size() == 12 * na * nb * nc * nd
    # component = centre * 3 + direction, slowest index
    g.set(2, 0, 5, 0, 7, 1.5)
    g.set(0, 0, 1, 0, 0, -2.0)
    arr = g.to_numpy()
    assert arr.shape == (12, na, nb, nc, nd)
    assert arr[7, 2, 0, 5, 0] == 1.5 and arr[0, 0, 0, 1, 0] == -2.0
    flat = arr.ravel()
    assert flat[7 * na * nb * nc * nd + ((2 * nb + 0) * nc + 5) * nd] == 1.5
    assert g(2, 0, 5, 0, 7) == 1.5 and np.count_nonzero(arr) == 2
