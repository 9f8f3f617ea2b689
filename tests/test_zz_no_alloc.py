"""alfq_gemm_noalloc (staging through the library's fixed per-device array) against alfq_gemm
(staging through buffers it allocates): the same kernels on the same operands, so the
products agree bit for bit.  Two calls in a row, the second on smaller operands than the
first, so the second reads nothing the first left behind; an operand above the staging
limit is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lib, fn, A, B, opa, opb, alpha, beta, Cin, prec=0):
    a_real, b_real = not np.iscomplexobj(A), not np.iscomplexobj(B)
    Aa = np.asfortranarray(A, dtype=np.float64 if a_real else np.complex128)
    Ba = np.asfortranarray(B, dtype=np.float64 if b_real else np.complex128)
    M = Aa.shape[0] if opa == "N" else Aa.shape[1]
    K = Aa.shape[1] if opa == "N" else Aa.shape[0]
    N = Ba.shape[1] if opb == "N" else Ba.shape[0]
    Cc = np.array(Cin, dtype=np.complex128, order="F")
    al = np.array([complex(alpha).real, complex(alpha).imag])
    be = np.array([complex(beta).real, complex(beta).imag])
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    st = getattr(lib, fn)(M, N, K, opa.encode(), int(a_real), dp(Aa), Aa.shape[0], opb.encode(), int(b_real),
                          dp(Ba), Ba.shape[0], dp(al), dp(be), dp(Cc), M, prec)
    return st, Cc


def _operands(rng, m, n, k, a_real):
    A = rng.standard_normal((m, k)) if a_real else rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k))
    B = rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))
    Cin = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
    return A, B, Cin


def test_noalloc_gemm_matches_allocating_gemm():
    from alf_amd import runtime
    lib = runtime.load_library()
    rng = np.random.default_rng(5)
    # first call: complex x complex, 300 x 280 x 260 with beta != 0; second: real x complex, smaller
    for (m, n, k, a_real, opa, alpha, beta) in [(300, 280, 260, False, "N", 1.0 - 0.5j, 0.25 + 1j),
                                                (70, 50, 90, True, "T", 2.0, 0.0)]:
        A, B, Cin = _operands(rng, m, n, k, a_real)
        if opa == "T":
            A = np.ascontiguousarray(A.T)
        print(f"gemm_noalloc case m={m} n={n} k={k} a_real={a_real} op={opa}", flush=True)
        st1, C1 = _call(lib, "alfq_gemm", A, B, opa, "N", alpha, beta, Cin)
        st2, C2 = _call(lib, "alfq_gemm_noalloc", A, B, opa, "N", alpha, beta, Cin)
        assert st1 == 0 and st2 == 0, lib.alfq_last_error()
        assert np.array_equal(C1, C2)
        opA = A.T if opa == "T" else A
        ref = alpha * (opA @ B) + beta * Cin
        assert np.abs(C2 - ref).max() <= 1e-12 * np.abs(ref).max()


def test_noalloc_gemm_refuses_operands_above_the_limit():
    from alf_amd import runtime
    lib = runtime.load_library()
    n = 1024                      # 1024 x 1024 complex = 16 MiB > 8 MiB per operand
    A = np.zeros((n, n), dtype=np.complex128)
    st, _ = _call(lib, "alfq_gemm_noalloc", A, A, "N", "N", 1.0, 0.0, np.zeros((n, n)))
    assert st != 0 and b"staging limit" in lib.alfq_last_error()
