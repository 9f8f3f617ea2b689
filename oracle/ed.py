"""Exact diagonalisation of tiny Hubbard clusters: the exact value of the sweep's estimators.

Test infrastructure only.  Many-body Hamiltonian of the Hubbard plugin with N_SUN = 2
spin flavours sharing the hopping matrix O (H_T = sum_s c^+_s O c_s) and the
density-channel interaction of Predefined_Int_U_SUN (Prog/Predefined_Int_mod.F90:59-76),
H_V = U / N_SUN sum_i (sum_s n_is - N_SUN / 2)^2 = U/2 sum_i (n_i - 1)^2.

The sweep measures at the boundaries of Trotter slices whose propagator is
B = exp(-dtau H_V) exp(-dtau H_T) (hopping first: WRAPGRUP, Prog/Wrapgr_mod.F90:110-147); after
the sum over the auxiliary fields every slice is the same B, so the expectation of an
equal-time estimator is Tr[O B^L] / Tr[B^L] at every slice (``trotter``).  ``thermal`` is
the exact Tr[O exp(-beta H)] / Tr[exp(-beta H)].  The only difference left between
``trotter`` and the Monte Carlo is the O(dtau^4) error of the 4-state Hubbard-Stratonovich
decoupling per slice (Fields_mod.F90:270-301).

Pinned (tests/test_oracle_pins.py): U = 0 against free-fermion closed forms, the atomic
limit t = 0 against the single-site partition function, and trotter -> thermal as dtau -> 0.
``trotter_tau`` (tests/test_tau_host.py): U = 0 against e^{-tau T} (1 + e^{-beta T})^-1, and at
t = 0 the equal-time correlators of ``observables``.
"""
from __future__ import annotations

import numpy as np


class HubbardED:
    def __init__(self, hop, U):
        self.hop = np.asarray(hop, dtype=complex)
        n = self.hop.shape[0]
        self.n = n
        m = 2 * n                                     # modes (site i, spin s) -> i + n s
        dim = 1 << m
        # Jordan-Wigner annihilators in the occupation basis (bit k = mode k)
        idx = np.arange(dim)
        self.c = []
        for k in range(m):
            a = np.zeros((dim, dim))
            occ = (idx >> k) & 1
            sign = (-1.0) ** np.array([bin(v & ((1 << k) - 1)).count("1") for v in idx])
            src = idx[occ == 1]
            a[src ^ (1 << k), src] = sign[src]
            self.c.append(a)
        self.num = [np.diag(((idx >> k) & 1).astype(float)) for k in range(m)]
        HT = np.zeros((dim, dim), dtype=complex)
        for s in range(2):
            for i in range(n):
                for j in range(n):
                    if self.hop[i, j] != 0:
                        HT += self.hop[i, j] * self.c[i + n * s].T @ self.c[j + n * s]
        nd = [self.num[i] + self.num[i + n] for i in range(n)]
        HV = 0.5 * U * sum((d - np.eye(dim)) @ (d - np.eye(dim)) for d in nd)
        self.HT, self.HV, self.U = HT, HV, float(U)
        self.nup = [self.num[i] for i in range(n)]
        self.ndn = [self.num[i + n] for i in range(n)]

    def _expm(self, H, t):
        e, V = np.linalg.eigh(H)
        return (V * np.exp(-t * e)) @ V.conj().T

    def trotter_rho(self, dtau, L):
        B = self._expm(self.HV, dtau) @ self._expm(self.HT, dtau)
        M = np.linalg.matrix_power(B, L)
        return M / np.trace(M)

    def thermal_rho(self, beta):
        M = self._expm(self.HT + self.HV, beta)
        return M / np.trace(M)

    def trotter_tau(self, dtau, L):
        """Time-displaced correlators of the Trotterised chain for t = 0..L slices,
        <A(t) B> = Tr[M^(L - t) A M^t B] / Tr[M^L], M = exp(-dtau H_V) exp(-dtau H_T):
        green[t, i, j] = <c_i(t) c_j^+> of one flavour (the estimator G(tau, 0)),
        spin[t, i, j] = <S_i(t) S_j> with S = n_up - n_dn, den[t, i, j] = <n_i(t) n_j>."""
        M = self._expm(self.HV, dtau) @ self._expm(self.HT, dtau)
        pw = [np.eye(M.shape[0], dtype=complex)]
        for _ in range(L):
            pw.append(pw[-1] @ M)
        Z = np.trace(pw[L])
        n = self.n
        sz = [self.nup[i] - self.ndn[i] for i in range(n)]
        nt = [self.nup[i] + self.ndn[i] for i in range(n)]
        green = np.zeros((L + 1, n, n), dtype=complex)
        spin = np.zeros_like(green)
        den = np.zeros_like(green)
        for t in range(L + 1):
            X, Y = pw[L - t], pw[t]
            for i in range(n):
                XcY = X @ self.c[i] @ Y
                XsY = X @ sz[i] @ Y
                XnY = X @ nt[i] @ Y
                for j in range(n):
                    green[t, i, j] = np.sum(XcY * self.c[j].conj()) / Z        # Tr[X c_i Y c_j^+]
                    spin[t, i, j] = np.sum(XsY * sz[j].T) / Z
                    den[t, i, j] = np.sum(XnY * nt[j].T) / Z
        return {"Green": green, "SpinZ": spin, "Den": den}

    def observables(self, rho):
        """Kin, Pot (U sum <n_up n_dn>), Part, Ener and the pair correlations
        green[i, j] = sum_s <c^+_is c_js>, spin[i, j] = <(n_iup - n_idn)(n_jup - n_jdn)>,
        den[i, j] = <n_i n_j>."""
        ex = lambda O: np.trace(O @ rho)
        n = self.n
        kin = ex(self.HT)
        pot = self.U * sum(ex(self.nup[i] @ self.ndn[i]) for i in range(n))
        part = sum(ex(self.nup[i] + self.ndn[i]) for i in range(n))
        green = np.array([[sum(ex(self.c[i + n * s].T @ self.c[j + n * s]) for s in range(2)) for j in range(n)]
                          for i in range(n)])
        sz = [self.nup[i] - self.ndn[i] for i in range(n)]
        nt = [self.nup[i] + self.ndn[i] for i in range(n)]
        spin = np.array([[ex(sz[i] @ sz[j]) for j in range(n)] for i in range(n)])
        den = np.array([[ex(nt[i] @ nt[j]) for j in range(n)] for i in range(n)])
        return {"Kin": kin, "Pot": pot, "Part": part, "Ener": kin + pot, "Green": green, "SpinZ": spin, "Den": den}
