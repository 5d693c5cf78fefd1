"""The Engine routes besides the full-basis Fock build, on the device, against
the oracle:

  * compute(op, span<ShellSetQuartet>, consumer) / compute(op, ShellSetQuartet,
    consumer) (reference engine.hpp:518-537,712-746; cpu_engine.hpp:435-547):
    plain accumulate over every shell quartet of each ShellSetQuartet, for
    FockBuilder, GpuFockBuilder (device ERI batches digested on the device)
    and a generic consumer;
  * GpuFockBuilder::accumulate of host blocks and accumulate_device_eri_batch
    (fock_builder_gpu.cu:475-530,600-621) through the device digestion;
  * compute_and_consume with ScreeningOptions and with a generic consumer;
  * a non-symmetric density: J and K equal the reference's N^4 accumulate
    (fock_builder.cpp:98-117) for any D.

Tolerance: J/K 1e-12 relative (the blocks are the raw-ERI kernels' outputs,
the sums are FP64)."""
import numpy as np
import pytest

import oracle
from libaccint import systems
import libaccint as lai

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def water_ccpvdz():
    atoms = systems.h2o_sto3g_geometry()
    return atoms, lai.basis_set("cc-pvdz", atoms), oracle.RawBasis("cc-pvdz", atoms)


def plain_jk(E, D, quartets, basis):
    """J, K of the plain accumulate over explicit shell quartets (oracle ERIs)."""
    n = D.shape[0]
    J = np.zeros((n, n))
    K = np.zeros((n, n))
    for q in quartets:
        sl = [slice(basis.shell(s).function_index(), basis.shell(s).function_index() + basis.shell(s).n_functions())
              for s in q]
        g = E[sl[0], sl[1], sl[2], sl[3]]
        J[sl[0], sl[1]] += np.einsum("abcd,cd->ab", g, D[sl[2], sl[3]])
        K[sl[0], sl[2]] += np.einsum("abcd,bd->ac", g, D[sl[1], sl[3]])
    return J, K


def set_shells(basis, k):
    s = basis.shell_sets()[k]
    return [s.shell(i).shell_index() for i in range(s.n_shells())]


def expand(basis, set_quartets):
    out = []
    for a, b, c, d in set_quartets:
        for i in set_shells(basis, a):
            for j in set_shells(basis, b):
                for k in set_shells(basis, c):
                    for l in set_shells(basis, d):
                        out.append((i, j, k, l))
    return out


def density(n, seed, symmetric=True):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((n, 4)) / np.sqrt(n)
    if symmetric:
        return C @ C.T
    C2 = rng.standard_normal((n, 4)) / np.sqrt(n)
    return C @ C2.T


@pytest.fixture(scope="module")
def water():
    atoms, b, rb = water_ccpvdz()
    o = oracle.Oracle(rb)
    return b, lai.Engine(b), o, o.eri_tensor()


def test_every_ordered_shell_set_quartet_gives_the_n4_fock_build(water):
    """All (a, b, c, d) over the shell sets cover every ordered shell quartet
    once, so the plain sum is the reference's N^4 J/K (oracle.jk_full)."""
    b, e, o, _ = water
    n = b.n_basis_functions()
    D = density(n, 3)
    ns = len(b.shell_sets())
    allq = [(x, y, z, w) for x in range(ns) for y in range(ns) for z in range(ns) for w in range(ns)]
    Jo, Ko = o.jk_full(D)
    fb = lai.FockBuilder(n)
    fb.set_density(D)
    e.compute_shell_set_quartets(lai.Operator.coulomb(), fb, allq)
    assert rel(fb.get_coulomb_matrix(), Jo) < 1e-12
    assert rel(fb.get_exchange_matrix(), Ko) < 1e-12
    g = lai.GpuFockBuilder(n)
    g.set_density(D)
    e.compute_shell_set_quartets(lai.Operator.coulomb(), g, allq)
    g.synchronize()
    assert rel(g.get_coulomb_matrix(), Jo) < 1e-12
    assert rel(g.get_exchange_matrix(), Ko) < 1e-12


def test_canonical_shell_set_quartets_all_consumers_agree(water):
    """BasisSet::shell_set_quartets() (the reference's canonical set-pair
    list) through FockBuilder, GpuFockBuilder and a generic consumer: each is
    the plain accumulate over the expanded shell quartets."""
    b, e, o, E = water
    n = b.n_basis_functions()
    D = density(n, 4)
    sq = b.shell_set_quartet_indices()
    assert len(sq) == b.n_shell_set_quartets()
    Jx, Kx = plain_jk(E, D, expand(b, sq), b)
    fb = lai.FockBuilder(n)
    fb.set_density(D)
    e.compute_shell_set_quartets(lai.Operator.coulomb(), fb)
    assert rel(fb.get_coulomb_matrix(), Jx) < 1e-12
    assert rel(fb.get_exchange_matrix(), Kx) < 1e-12
    g = lai.GpuFockBuilder(n)
    g.set_density(D)
    e.compute_shell_set_quartets(lai.Operator.coulomb(), g)
    assert rel(g.get_coulomb_matrix(), Jx) < 1e-12
    assert rel(g.get_exchange_matrix(), Kx) < 1e-12
    Jg = np.zeros((n, n))
    Kg = np.zeros((n, n))
    seen = []

    def consumer(blk, fa, fb_, fc, fd):
        na, nb, nc, nd = blk.shape
        Jg[fa:fa + na, fb_:fb_ + nb] += np.einsum("abcd,cd->ab", blk, D[fc:fc + nc, fd:fd + nd])
        Kg[fa:fa + na, fc:fc + nc] += np.einsum("abcd,bd->ac", blk, D[fb_:fb_ + nb, fd:fd + nd])
        seen.append((fa, fb_, fc, fd))

    e.compute_shell_set_quartets(lai.Operator.coulomb(), consumer)
    assert len(seen) == len(expand(b, sq))
    assert rel(Jg, Jx) < 1e-12 and rel(Kg, Kx) < 1e-12


def test_single_shell_set_quartet_device_route(water):
    b, e, o, E = water
    n = b.n_basis_functions()
    D = density(n, 5)
    ns = len(b.shell_sets())
    for q in [(0, 0, 0, 0), (ns - 1, 0, ns - 1, 0), (1, ns - 1, 0, 2 % ns)]:
        Jx, Kx = plain_jk(E, D, expand(b, [q]), b)
        g = lai.GpuFockBuilder(n)
        g.set_density(D)
        e.compute_shell_set_quartet(lai.Operator.coulomb(), g, *q)
        assert rel(g.get_coulomb_matrix(), Jx) < 1e-12, q
        assert rel(g.get_exchange_matrix(), Kx) < 1e-12, q


def test_gpu_fock_builder_host_blocks_match_n4(water):
    """GpuFockBuilder.accumulate over every ordered shell quartet (host
    blocks, digested on the device in batches) equals the N^4 J/K; a second
    set_density digests the pending blocks with the first density."""
    b, e, o, E = water
    n = b.n_basis_functions()
    D1, D2 = density(n, 6), density(n, 7)
    ns = b.n_shells()
    q = np.array([(i, j, k, l) for i in range(ns) for j in range(ns) for k in range(ns) for l in range(ns)],
                 dtype=np.int32)
    vals, offs = e.compute_eri_blocks(q)
    g = lai.GpuFockBuilder(n)
    g.set_density(D1)
    half = len(q) // 2
    for t in range(len(q)):
        if t == half:
            g.set_density(D2)
        i, j, k, l = (int(x) for x in q[t])
        sh = [b.shell(s) for s in (i, j, k, l)]
        buf = lai.TwoElectronBuffer(*[s.n_functions() for s in sh])
        buf.from_numpy(vals[offs[t]:offs[t + 1]].reshape([s.n_functions() for s in sh]))
        g.accumulate(buf, *[s.function_index() for s in sh])
    J1, K1 = plain_jk(E, D1, [tuple(x) for x in q[:half]], b)
    J2, K2 = plain_jk(E, D2, [tuple(x) for x in q[half:]], b)
    assert rel(g.get_coulomb_matrix(), J1 + J2) < 1e-12
    assert rel(g.get_exchange_matrix(), K1 + K2) < 1e-12
    g.reset()
    assert np.max(np.abs(g.get_coulomb_matrix())) == 0.0


def test_device_eri_batch_on_another_stream(water):
    import torch
    b, e, o, E = water
    n = b.n_basis_functions()
    D = density(n, 8)
    ns = b.n_shells()
    rng = np.random.default_rng(9)
    q = rng.integers(0, ns, size=(300, 4)).astype(np.int32)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    batch = e.compute_eri_batch_device(q, s1.cuda_stream)
    assert batch.n_blocks() == 300
    g = lai.GpuFockBuilder(n, s2.cuda_stream)
    g.set_density(D)
    g.accumulate_device_eri_batch(batch)
    g.synchronize()
    Jx, Kx = plain_jk(E, D, [tuple(x) for x in q], b)
    assert rel(g.get_coulomb_matrix(), Jx) < 1e-12
    assert rel(g.get_exchange_matrix(), Kx) < 1e-12


def test_non_symmetric_density_matches_n4(water):
    """For a non-symmetric D the J/K routes return the N^4 accumulate's J and
    K (K of the antisymmetric part included), not K of (D + D^T)/2."""
    b, e, o, _ = water
    n = b.n_basis_functions()
    D = density(n, 10, symmetric=False)
    Jo, Ko = o.jk_full(D)
    Ks = o.jk_full(0.5 * (D + D.T))[1]
    assert rel(Ks, Ko) > 1e-3  # the antisymmetric part matters
    fb = lai.FockBuilder(n)
    fb.set_density(D)
    e.compute_and_consume(lai.Operator.coulomb(), fb)
    assert rel(fb.get_coulomb_matrix(), Jo) < 1e-10
    assert rel(fb.get_exchange_matrix(), Ko) < 1e-10
    g = lai.GpuFockBuilder(n)
    g.set_density(D)
    e.compute_and_consume(lai.Operator.coulomb(), g)
    assert rel(g.get_coulomb_matrix(), Jo) < 1e-10
    assert rel(g.get_exchange_matrix(), Ko) < 1e-10


def test_screening_options_overload_and_generic_consumer(water):
    """compute_and_consume(op, consumer, ScreeningOptions): the FockBuilder's
    screened build equals the oracle's canonical screened J/K from the
    device's Q table; a generic consumer gets N^4 blocks (unscreened) or only
    the quartets passing Q_ij Q_kl >= thr."""
    b, e, o, E = water
    n = b.n_basis_functions()
    D = density(n, 11)
    thr = 1e-6
    e.precompute_schwarz_bounds()
    Q = e.schwarz_bounds().packed()
    Jo, Ko = o.jk_canonical(D, Q, thr)
    fb = lai.FockBuilder(n)
    fb.set_density(D)
    e.compute_and_consume_screened(lai.Operator.coulomb(), fb, lai.ScreeningOptions(threshold=thr))
    assert rel(fb.get_coulomb_matrix(), Jo) < 1e-12
    assert rel(fb.get_exchange_matrix(), Ko) < 1e-12

    ns = b.n_shells()

    def q_of(i, j):
        i, j = min(i, j), max(i, j)
        return Q[j * (j + 1) // 2 + i]

    for opts in (None, lai.ScreeningOptions(threshold=thr)):
        got = []
        Jg = np.zeros((n, n))
        Kg = np.zeros((n, n))

        def consumer(blk, fa, fb_, fc, fd):
            na, nb, nc, nd = blk.shape
            Jg[fa:fa + na, fb_:fb_ + nb] += np.einsum("abcd,cd->ab", blk, D[fc:fc + nc, fd:fd + nd])
            Kg[fa:fa + na, fc:fc + nc] += np.einsum("abcd,bd->ac", blk, D[fb_:fb_ + nb, fd:fd + nd])
            got.append(1)

        e.compute_and_consume(lai.Operator.coulomb(), consumer, opts)
        keep = [(i, j, k, l) for i in range(ns) for j in range(ns) for k in range(ns) for l in range(ns)
                if opts is None or q_of(i, j) * q_of(k, l) >= thr]
        assert len(got) == len(keep)
        Jx, Kx = plain_jk(E, D, keep, b)
        assert rel(Jg, Jx) < 1e-12 and rel(Kg, Kx) < 1e-12
