"""CPU restatement of the Hubbard plugin's equal-time measurements and of their bins.

Test infrastructure only: the checker of the device measurement kernel, never part of
the product path.  N_FL = 1 (the density-channel Hubbard model), Mc_step_weight = 1.

Follows
  * Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90:695-801  Obser (scalars Kin, Pot, Part,
    Ener; correlations Green, SpinZ, Den)
  * Prog/Predefined_Hop_mod.F90:1850-1959                    Predefined_Hoppings_Compute_Kin
  * Prog/Predefined_Obs_mod.F90:77-120, 236-325              Green, SpinSUN, Den measures
  * Prog/observables_mod.F90:355-480, 772-860                bin averages (Print_bin_Latt,
    Print_bin_Vec)
  * Libraries/Modules/lattices_v3_mod.F90:847-875            FT_R_to_K_C
  * Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90:826-857   ObserT (Green, SpinZ, Den of
    tau), Prog/Predefined_Obs_mod.F90:337-446, 533-594

Pinned (tests/test_oracle_pins.py) by free fermions, where every observable has a closed
form in the eigenvalues of the hopping matrix, and by sum rules at k = 0.
"""
from __future__ import annotations

import numpy as np


class HubbardObs:
    """Accumulators of one chain: scal[4] = (Kin, Pot, Part, Ener), latt[3] = (Green, SpinZ,
    Den)[d, no_I, no_J], den0[no], with the counts and sign sums of the reference."""

    def __init__(self, hop, U, n_sun, imj, site_cell, site_orb, norb):
        self.hop = np.asarray(hop, dtype=complex)       # hopping matrix O of H_T = c^+ O c
        self.U = float(U)
        self.n_sun = int(n_sun)
        self.imj = np.asarray(imj)
        self.cell = np.asarray(site_cell)
        self.orb = np.asarray(site_orb)
        self.norb = int(norb)
        self.ncell = self.imj.shape[0]
        self.reset()

    def reset(self):
        self.scal = np.zeros(4, dtype=complex)
        self.n_scal = 0
        self.sign_scal = 0.0
        self.latt = np.zeros((3, self.ncell, self.norb, self.norb), dtype=complex)
        self.den0 = np.zeros(self.norb, dtype=complex)
        self.n_eq = 0
        self.sign_eq = 0.0

    def measure(self, G, phase):
        """Obser(GR, PHASE, Ntau, 1.0) for N_FL = 1 (Hamiltonian_Hubbard_smod.F90:718-799)."""
        G = np.asarray(G, dtype=complex)
        n = G.shape[0]
        zp = phase / phase.real
        zs = np.sign(phase.real)
        w = zp * zs
        grc = np.eye(n) - G.T                  # GRC(i, j) = <c^+_i c_j> = delta_ij - G(j, i)
        kin = self.n_sun * np.sum(self.hop * grc)
        d = np.diag(grc)
        pot = self.U * np.sum(d * d)
        part = self.n_sun * np.sum(d)
        self.scal += np.array([kin, pot, part, kin + pot]) * w
        self.n_scal += 1
        self.sign_scal += zs
        dist = self.imj[self.cell[:, None], self.cell[None, :]]
        oi = np.broadcast_to(self.orb[:, None], (n, n))
        oj = np.broadcast_to(self.orb[None, :], (n, n))
        zi = self.n_sun * d
        green = self.n_sun * grc
        spin = self.n_sun * grc * G
        den = zi[:, None] * zi[None, :] + spin
        for q, v in enumerate((green, spin, den)):
            np.add.at(self.latt[q], (dist, oi, oj), v * w)
        np.add.at(self.den0, self.orb, zi * w)
        self.n_eq += 1
        self.sign_eq += zs

    def bins(self):
        """Bin averages as the writers form them: scalars and sign / N; correlations / N;
        the Den background / (N Ns) (Ntau = 1)."""
        return {"scal": self.scal / self.n_scal, "sign_scal": self.sign_scal / self.n_scal,
                "latt": self.latt / self.n_eq, "den0": self.den0 / (self.n_eq * self.ncell),
                "sign_eq": self.sign_eq / self.n_eq}


class HubbardTauObs:
    """Time-displaced accumulators of one chain (ObserT of the Hubbard plugin, N_FL = 1,
    Hamiltonian_Hubbard_smod.F90:826-857): latt[3] = (Green, SpinZ, Den)[tau, d, no_I, no_J]
    for tau = 0..Ltrot, den0[no], and the count N and sign sum, which grow once per TAU_M
    call (at tau = 0).

    With G(tau, 0) = GT0, G(0, tau) = G0T, G(0, 0) = G00, G(tau, tau) = GTT and the pair
    I1 = (cell I, orbital a), J1 = (cell J, orbital b), d = imj(I, J), each term times ZP ZS
    (Predefined_Obs_mod.F90:337-388, 400-446, 533-594):
      Green += GT0(I1, J1)
      SpinZ += -N_SUN G0T(J1, I1) GT0(I1, J1)
      Den   += Z_I Z_J - N_SUN G0T(J1, I1) GT0(I1, J1),  Z_I = N_SUN (1 - GTT(I1, I1)),
               Z_J = N_SUN (1 - G00(J1, J1));  Den0(a) += Z_I
    """

    def __init__(self, ltrot, n_sun, imj, site_cell, site_orb, norb):
        self.ntau = int(ltrot) + 1
        self.n_sun = int(n_sun)
        self.imj = np.asarray(imj)
        self.cell = np.asarray(site_cell)
        self.orb = np.asarray(site_orb)
        self.norb = int(norb)
        self.ncell = self.imj.shape[0]
        self.reset()

    def reset(self):
        self.latt = np.zeros((3, self.ntau, self.ncell, self.norb, self.norb), dtype=complex)
        self.den0 = np.zeros(self.norb, dtype=complex)
        self.n = 0
        self.sign = 0.0

    def measure(self, nt, GT0, G0T, G00, GTT, phase):
        n = GT0.shape[0]
        zp = phase / phase.real
        zs = np.sign(phase.real)
        w = zp * zs
        if nt == 0:
            self.n += 1
            self.sign += zs
        dist = self.imj[self.cell[:, None], self.cell[None, :]]
        oi = np.broadcast_to(self.orb[:, None], (n, n))
        oj = np.broadcast_to(self.orb[None, :], (n, n))
        green = np.asarray(GT0)
        spin = -self.n_sun * np.asarray(G0T).T * green          # G0T(J1, I1) GT0(I1, J1)
        zi = self.n_sun * (1.0 - np.diag(GTT))
        zj = self.n_sun * (1.0 - np.diag(G00))
        den = zi[:, None] * zj[None, :] + spin
        for q, v in enumerate((green, spin, den)):
            np.add.at(self.latt[q, nt], (dist, oi, oj), v * w)
        np.add.at(self.den0, self.orb, zi * w)

    def bins(self):
        """Print_bin_Latt's averages: correlations / N, the Den background / (N Ns Ntau)."""
        return {"latt": self.latt / self.n, "den0": self.den0 / (self.n * self.ncell * self.ntau),
                "sign": self.sign / self.n}


def fourier_r_to_k(x_r, r_vecs, k_vecs):
    """FT_R_to_K_C: X(k) = sum_r exp(-i k.r) X(r) / N over the first axis."""
    ph = np.exp(-1j * (k_vecs @ r_vecs.T))             # (Nk, Nr)
    return np.tensordot(ph, x_r, axes=(1, 0)) / r_vecs.shape[0]
