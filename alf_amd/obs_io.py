"""Bins of the equal-time measurements in the reference's text format, and their analysis.

A bin is what the device accumulated between two ``alfq_obs_reset`` calls (``Engine.obs_get``).
``BinWriter.write`` appends it the way the reference's writers do with -DOBS_LEGACY:

* ``Kin_scal``, ``Pot_scal``, ``Part_scal``, ``Ener_scal`` (Print_bin_Vec,
  Prog/observables_mod.F90:772-860): the bin averages Obs/N and sign/N, one line per bin,
  ``I10`` count, `` (E25.17E3,E25.17E3)`` per value, ``E26.17E3`` sign; an ``*_info`` file
  with the analysis mode.
* ``Green_eq``, ``SpinZ_eq``, ``Den_eq`` (Print_bin_Latt, observables_mod.F90:355-500): the
  averages Obs/N and Obs0/(N Ns), Fourier transformed per orbital pair to the cluster's
  momenta (FT_R_to_K_C, Libraries/Modules/lattices_v3_mod.F90:847-875); a header line
  (sign, Norb, N), Norb background values, then per momentum its coordinates and the
  Norb x Norb values; an ``*_info`` file with the lattice.

``write_tau`` appends a bin of the time-displaced sums to ``Green_tau``, ``SpinZ_tau``,
``Den_tau`` the same way with Ntau = Ltrot + 1 time slices (channel P for Green, PH for the
others, Hamiltonian_Hubbard_smod.F90:652-669).

``read_scal`` and ``jackknife_scal`` read the scalar files back and give the sign-weighted
averages with the reference's jackknife errors (jack and errcalc,
Analysis/ana_mod.F90:862-933, Libraries/Modules/errors_mod.F90:98-126).
"""
from __future__ import annotations

import os

import numpy as np

from alf_amd.host import lattices

SCALARS = ("Kin", "Pot", "Part", "Ener")
CORRELATIONS = ("Green", "SpinZ", "Den")


def fortran_e(x: float, w: int = 25, d: int = 17, e: int = 3) -> str:
    """Fortran Ew.dEe output of x: 0.dddd...E+xxx, right-justified in w columns."""
    x = float(x)
    if x == 0.0 or not np.isfinite(x):
        body = "0." + "0" * d + "E+" + "0" * e if x == 0.0 else str(x)
        return body.rjust(w)
    m, ex = f"{abs(x):.{d - 1}e}".split("e")
    digits = m.replace(".", "")
    p = int(ex) + 1
    body = f"{'-' if x < 0 else ''}0.{digits}E{'+' if p >= 0 else '-'}{abs(p):0{e}d}"
    return body.rjust(w)


def _cplx(z, w=25, lead=" ("):
    return f"{lead}{fortran_e(z.real, w)},{fortran_e(z.imag, w)})"


class BinWriter:
    """Appends bins of one chain's measurements to the reference's text files in `directory`."""

    def __init__(self, directory: str, model):
        self.dir = directory
        os.makedirs(directory, exist_ok=True)
        self.m = model
        latt = model.latt
        self.latt, self.unit = latt, model.unit
        self.R = np.array([latt.list[c, 0] * latt.a1_p + latt.list[c, 1] * latt.a2_p for c in range(latt.N)])
        k = lattices.momentum_list(latt)
        self.K = np.array([m * latt.b1_p + n * latt.b2_p for m, n in k])
        self.ft = np.exp(-1j * (self.K @ self.R.T)) / latt.N          # FT_R_to_K_C

    def _path(self, name):
        return os.path.join(self.dir, name)

    def write(self, sums: dict) -> None:
        n_s, n_e = sums["n_scal"], sums["n_eq"]
        if n_s <= 0 or n_e <= 0:
            raise ValueError("empty bin")
        sign = sums["sign_scal"] / n_s
        for q, name in enumerate(SCALARS):
            f = self._path(name + "_scal")
            if not os.path.exists(f + "_info"):
                with open(f + "_info", "w") as fh:
                    fh.write("====== Analysis Mode ======\nidentity\n")
            with open(f, "a") as fh:
                fh.write(f"{2:10d}" + _cplx(sums["scal"][q] / n_s) + fortran_e(sign, 26) + "\n")
        norb, N = self.unit.Norb, self.latt.N
        sign_eq = sums["sign_eq"] / n_e
        den0 = sums["den0"] / (n_e * N)
        for q, name in enumerate(CORRELATIONS):
            x = sums["latt"][q] / n_e                                     # [d, a, b]
            xk = np.tensordot(self.ft, x, axes=(1, 0))                     # [k, a, b]
            f = self._path(name + "_eq")
            if not os.path.exists(f + "_info"):
                self._latt_info(f + "_info", name + "_eq", "--", 1)
            with open(f, "a") as fh:
                fh.write(fortran_e(sign_eq) + f"{norb:11d}{N:11d}\n")
                for a in range(norb):
                    v = den0[a] if name == "Den" else 0j
                    fh.write(_cplx(v, lead="(") + "\n")
                for ik in range(N):
                    fh.write(fortran_e(self.K[ik, 0]) + " " + fortran_e(self.K[ik, 1]) + "\n")
                    for a in range(norb):
                        for b in range(norb):
                            fh.write(_cplx(xk[ik, a, b], lead="(") + "\n")

    def write_tau(self, sums: dict) -> None:
        """Appends a bin of the time-displaced sums (``Engine.tau_get``) to ``Green_tau``,
        ``SpinZ_tau``, ``Den_tau`` (Print_bin_Latt with Ntau = Ltrot + 1): header (sign, Norb,
        N, Ntau, dtau), Norb background values, then per momentum its coordinates and the
        Ntau x Norb x Norb values."""
        n = sums["n"]
        if n <= 0:
            raise ValueError("empty bin")
        norb, N = self.unit.Norb, self.latt.N
        ntau = sums["latt"].shape[1]
        sign = sums["sign"] / n
        den0 = sums["den0"] / (n * N * ntau)
        dtau = self.m.params.Dtau
        for q, (name, channel) in enumerate(zip(CORRELATIONS, ("P", "PH", "PH"))):
            x = sums["latt"][q] / n                                       # [tau, d, a, b]
            xk = np.einsum("kd,tdab->ktab", self.ft, x)
            f = self._path(name + "_tau")
            if not os.path.exists(f + "_info"):
                self._latt_info(f + "_info", name + "_tau", channel, ntau)
            with open(f, "a") as fh:
                fh.write(fortran_e(sign) + f"{norb:11d}{N:11d}{ntau:11d}" + fortran_e(dtau, 26) + "\n")
                for a in range(norb):
                    v = den0[a] if name == "Den" else 0j
                    fh.write(_cplx(v, lead="(") + "\n")
                for ik in range(N):
                    fh.write(fortran_e(self.K[ik, 0]) + " " + fortran_e(self.K[ik, 1]) + "\n")
                    for t in range(ntau):
                        for a in range(norb):
                            for b in range(norb):
                                fh.write(_cplx(xk[ik, t, a, b], lead="(") + "\n")

    def _latt_info(self, path, observable, channel, ntau):
        L, U = self.latt, self.unit
        row = lambda k, v: f"{k:>20s}: {v}\n"
        vec = lambda a: "".join(fortran_e(t, 26) for t in a)
        with open(path, "w") as fh:
            fh.write(row("Observable", observable))
            fh.write(row("Channel", channel))
            fh.write(f"{'Ntau':>20s}: {ntau:10d}\n")
            fh.write(f"{'dtau':>20s}: {fortran_e(self.m.params.Dtau, 26)}\n")
            fh.write("       ====== Bravais Lattice ======\n")
            fh.write(f"{'Unit cells':>20s}: {L.N:10d}\n")
            for k, a in (("L1", L.L1_p), ("L2", L.L2_p), ("a1", L.a1_p), ("a2", L.a2_p)):
                fh.write(f"{k:>20s}: {vec(a)}\n")
            fh.write("       ========= Unit cell =========\n")
            fh.write(f"{'Coordination number':>20s}: {U.N_coord:10d}\n")
            fh.write(f"{'Number of orbitals':>20s}: {U.Norb:10d}\n")
            fh.write(f"{'Ndim':>20s}: {U.Orb_pos_p.shape[1]:10d}\n")
            for o in range(U.Norb):
                fh.write(f"{'Orbital ' + str(o + 1):>20s}: {vec(U.Orb_pos_p[o])}\n")


def read_scal(path: str):
    """Bins of a *_scal file: (values[bins, nobs] complex, sign[bins])."""
    vals, signs = [], []
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line:
                continue
            head, rest = line.split(None, 1)
            nobs = int(head) - 1
            parts = rest.replace("(", " ").replace(")", " ").replace(",", " ").split()
            nums = [float(p.replace("E", "e")) for p in parts]
            vals.append([complex(nums[2 * i], nums[2 * i + 1]) for i in range(nobs)])
            signs.append(nums[2 * nobs])
    return np.array(vals), np.array(signs)


def jackknife(func, data: np.ndarray, n_skip: int = 0, n_rebin: int = 1):
    """jack (ana_mod.F90:862-933): data[nobs, bins]; skip and rebin, jackknife values
    func(mean of the other bins), then per real and imaginary part errcalc's mean and
    sqrt(sum (j - mean)^2)."""
    nb = (data.shape[1] - n_skip) // n_rebin
    if nb < 2:
        raise ValueError("need at least two bins")
    r = data[:, n_skip:n_skip + nb * n_rebin].reshape(data.shape[0], nb, n_rebin).mean(axis=2)
    tot = r.sum(axis=1)
    j = np.array([func((tot - r[:, i]) / (nb - 1)) for i in range(nb)])
    best = j.real.mean() + 1j * j.imag.mean()
    err = np.sqrt(np.sum((j.real - j.real.mean()) ** 2)) + 1j * np.sqrt(np.sum((j.imag - j.imag.mean()) ** 2))
    return best, err


def jackknife_scal(path: str, n_skip: int = 0, n_rebin: int = 1):
    """Sign-weighted averages <O> = <O s> / <s> of every value of a *_scal file, with errors."""
    vals, signs = read_scal(path)
    data = np.vstack([vals.T, signs[None, :].astype(complex)])
    out = []
    for i in range(vals.shape[1]):
        out.append(jackknife(lambda X, i=i: X[i] / X[-1], data, n_skip, n_rebin))
    return out
