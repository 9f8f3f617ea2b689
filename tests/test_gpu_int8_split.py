"""The int8 split-product form of the hopping wraps (ALFQ_PREC_INT8_SPLIT), through the C ABI.

* The product itself: for real x complex, complex x real and complex x complex operands of
  ragged and full shapes, every entry of A B lies within 16 K 2^-56 max_k |A_ik| max_k |B_kj| of
  the exact product (the bound of 8 base-128 digits per operand), and the result
  is not the FP64 MFMA one (the integer path ran).
* Sweeps: with the wraps of G on the split products, the chain takes exactly the FP64
  drop-in's decisions on the same random tape (fields and counters bit-exact), G after
  every sweep equals the drop-in's within 1e-10, and at every stabilisation point the wrapped
  G stays within 1e-8 of the G the FP64 UDV chain rebuilds (Precision Green), on the 16 x 16
  U = 4 beta = 10 workload and the flux honeycomb (complex exp(-dtau T)).  Since the two
  runs rebuild the same G from the same fields, |G_split - G_FP64| at a stabilisation point
  is at most the sum of the two runs' Precision Green there.
* A batch of 16 chains of the 16 x 16 workload on the split products: every chain's mean
  Precision Green below 1e-8 (the reference's criterion), single values within 5 x
  max(1e-8, the FP64 batch's largest).
"""
import numpy as np
import pytest

from helpers import fields_for, hubbard, rel, tape

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rt():
    from alf_amd import runtime
    runtime.load_library()
    return runtime


def _bound(A, B):
    K = A.shape[1]
    ra = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1) if np.iscomplexobj(A) else np.abs(A).max(axis=1)
    cb = np.maximum(np.abs(B.real), np.abs(B.imag)).max(axis=0) if np.iscomplexobj(B) else np.abs(B).max(axis=0)
    return 8.0 * K * 2.0 ** -56 * np.outer(ra, cb) * 2.0 + 4.0 * 2.0 ** -53 * np.abs(A) @ np.abs(B)


@pytest.mark.parametrize("shape", [(16, 16, 16), (37, 50, 29), (100, 36, 64), (256, 256, 256), (300, 17, 5),
                                   (130, 200, 300)])
def test_split_product_accuracy(rt, shape):
    M, N, K = shape
    r = np.random.default_rng(M + 3 * N + 7 * K)
    cplx = lambda m, n: r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))
    # rows of very different magnitudes: the exponents are per row / column
    Ar = r.standard_normal((M, K)) * np.exp(r.uniform(-20, 20, (M, 1)))
    Ac = cplx(M, K)
    Br = r.standard_normal((K, N))
    Bc = cplx(K, N) * np.exp(r.uniform(-5, 5, (1, N)))
    for A, B in ((Ar, Bc), (Ac, Br), (Ac, Bc)):
        got = rt.gemm(A, B, prec=rt.PREC_INT8_SPLIT)
        want = A @ B
        assert np.all(np.abs(got - want) <= _bound(A, B)), np.abs(got - want).max()
        assert not np.array_equal(got, rt.gemm(A, B))
    # alpha, beta
    C0 = cplx(M, N)
    got = rt.gemm(Ar, Bc, alpha=0.5 - 2j, beta=1.5 + 0.25j, Cin=C0, prec=rt.PREC_INT8_SPLIT)
    want = (0.5 - 2j) * (Ar @ Bc) + (1.5 + 0.25j) * C0
    assert rel(got, want) < 1e-12


def _paired(rt, ham, f, u, sweeps, delay=32):
    engines = []
    for prec in (rt.PREC_FP64, rt.PREC_INT8_SPLIT):
        e = rt.Engine(ham, delay=delay)
        e.set_precision(prec)
        e.set_fields(f)
        e.set_tape(u)
        e.init()
        engines.append(e)
    for k in range(sweeps):
        for e in engines:
            e.sweep(1)
        a, b = engines
        assert np.array_equal(a.get_fields(), b.get_fields()), f"sweep {k}: decisions differ"
        assert rel(b.get_green(), a.get_green()) < 1e-10
        assert abs(b.get_phase() - a.get_phase()) < 1e-10
    ca, cb = engines[0].control(), engines[1].control()
    assert ca["ACC_up"] == cb["ACC_up"] and ca["NCG"] == cb["NCG"]
    la, lb = engines[0].precision_log(), engines[1].precision_log()
    for e in engines:
        e.close()
    return la, lb


@pytest.mark.parametrize("lattice,L,beta,U,nwrap,n_phi,sweeps", [("Square", 16, 10.0, 4.0, 10, 0, 2),
                                                                  ("Honeycomb", 12, 3.0, 3.0, 8, 18, 1)])
def test_split_wraps_match_fp64_at_every_stabilisation(rt, lattice, L, beta, U, nwrap, n_phi, sweeps):
    ham = hubbard(L1=L, L2=L, beta=beta, U=U, nwrap=nwrap, lattice=lattice, n_phi=n_phi)
    f = fields_for(ham, 3)
    u = tape(3 * 2 * ham.Ndim * ham.Ltrot * sweeps + 64, seed=9)
    la, lb = _paired(rt, ham, f, u, sweeps)
    worst = float((la[:, 0] + lb[:, 0]).max())
    print(f"\n{lattice} {L}x{L} beta={beta}: Precision Green FP64 max {la[:, 0].max():.2e}, split max "
          f"{lb[:, 0].max():.2e} mean {lb[:, 1].mean():.2e}; |G_split - G_FP64| <= {worst:.2e} over "
          f"{len(lb)} stabilisations", flush=True)
    assert len(la) == len(lb) > 0
    assert lb[:, 0].max() <= 1e-8
    assert worst <= 1e-8


def _batch_precision(rt, ham, prec, nchain, sweeps):
    b = rt.BatchEngine([ham] * nchain, delay=32)
    b.set_precision(prec)
    for z in range(nchain):
        b.chain(z).set_fields(fields_for(ham, z))
        b.chain(z).set_philox(seed=z, stream=0)
    b.init()
    b.sweep(sweeps)
    c = [b.chain(z).control() for z in range(nchain)]
    b.close()
    return np.array([x["XMAXG"] for x in c]), np.array([x["precision_green_mean"] for x in c])


def test_split_wraps_batch_meets_the_bar(rt):
    """16 chains of the 16 x 16 workload, 2 sweeps from the random start, on the split products:
    the mean Precision Green over the stabilisations (the reference's criterion,
    Documentation/stabilization.tex:225) stays below 1e-8 for every chain, and no single value
    exceeds 5 x max(1e-8, the FP64 batch's largest value on the same seeds).  Single values of
    both forms reach ~1e-8 here (FP64 7.5e-9 and split 9.7e-9 on one build, 2.1e-9 and 1.04e-8
    on another: the chains' trajectories differ after the first differing decision); 7
    truncated digits gave 1.1e-7 and fail the bound."""
    ham = hubbard(L1=16, L2=16, beta=10.0, U=4.0, nwrap=10)
    x64, m64 = _batch_precision(rt, ham, rt.PREC_FP64, 16, 2)
    xs, ms = _batch_precision(rt, ham, rt.PREC_INT8_SPLIT, 16, 2)
    print(f"\nbatch of 16: Precision Green max FP64 {x64.max():.2e}, split {xs.max():.2e}; mean FP64 "
          f"{m64.max():.2e}, split {ms.max():.2e}", flush=True)
    assert ms.max() <= 1e-8
    assert xs.max() <= 5.0 * max(1e-8, x64.max())


@pytest.mark.parametrize("mixed", [False, True])
def test_split_batch_chains_match_single_chains(rt, mixed):
    """Batched split products equal single-chain ones: a batch whose chains share one
    exp(-dtau T) (its digit planes split once and read by every chain) and a batch of two
    hoppings (dtau 0.1 and 0.05 at the same Ltrot: planes per chain).  Every chain takes the
    decisions it takes alone, with G within 1e-10."""
    ha = hubbard(L1=8, L2=8, beta=4.0, U=4.0, nwrap=10)
    hb = hubbard(L1=8, L2=8, beta=2.0, U=6.0, nwrap=10, dtau=0.05) if mixed else ha
    assert ha.Ltrot == hb.Ltrot
    hams = [ha, ha, hb, hb]
    b = rt.BatchEngine(hams, delay=32)
    b.set_precision(rt.PREC_INT8_SPLIT)
    for z, h in enumerate(hams):
        b.chain(z).set_fields(fields_for(h, 40 + z))
        b.chain(z).set_philox(seed=40 + z, stream=0)
    b.init()
    b.sweep(2)
    for z, h in enumerate(hams):
        e = rt.Engine(h, delay=32)
        e.set_precision(rt.PREC_INT8_SPLIT)
        e.set_fields(fields_for(h, 40 + z))
        e.set_philox(seed=40 + z, stream=0)
        e.init()
        e.sweep(2)
        c = b.chain(z)
        assert np.array_equal(c.get_fields(), e.get_fields())
        assert rel(c.get_green(), e.get_green()) < 1e-10
        assert abs(c.get_phase() - e.get_phase()) < 1e-10
        e.close()
    b.close()
