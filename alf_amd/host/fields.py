"""Auxiliary (Hubbard-Stratonovich) fields.

Host-side mirror of ALF's ``Type(Fields)`` for field types 1 (Ising) and 2
(four-state discrete HS), the two kinds the device sweep updates.

Follows
  * ``Prog/Fields_mod.F90:79-99``    (Type Fields: f(N_op, Ltrot), t(N_op))
  * ``Prog/Fields_mod.F90:112-163``  (Phi, Gama)
  * ``Prog/Fields_mod.F90:173-217``  (Flip)
  * ``Prog/Fields_mod.F90:258-303``  (Fields_init: Phi_st, Gama_st, Flip_st)
  * ``Prog/Fields_mod.F90:588-610``  (Fields_set: random +-1 start)
  * ``Prog/Fields_mod.F90:631-662``  (Fields_read_conf: seed line, then one integer per
    (op, slice) with the slice index outer)
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_S6 = math.sqrt(6.0)
_PHI2 = {-2: -math.sqrt(2.0 * (3.0 + _S6)), -1: -math.sqrt(2.0 * (3.0 - _S6)),
         1: math.sqrt(2.0 * (3.0 - _S6)), 2: math.sqrt(2.0 * (3.0 + _S6))}
_GAMA2 = {-2: 1.0 - _S6 / 3.0, 2: 1.0 - _S6 / 3.0, -1: 1.0 + _S6 / 3.0, 1: 1.0 + _S6 / 3.0}
# FLIP_st(s, k): the k-th (k = 1, 2, 3) alternative to field value s
FLIP_ST = {-2: (-1, 1, 2), -1: (1, 2, -2), 1: (2, -2, -1), 2: (-2, -1, 1)}


def phi_st(s: int, typ: int) -> float:
    """Phi_st(s, type) (Fields_mod.F90:270-275)."""
    if typ == 1:
        return float(s)
    if typ == 2:
        return _PHI2[int(s)]
    raise ValueError(f"field type {typ} is not a discrete type")


def gama_st(s: int, typ: int) -> float:
    """Gama_st(s, type) (Fields_mod.F90:277-282)."""
    if typ == 1:
        return 1.0
    if typ == 2:
        return _GAMA2[int(s)]
    raise ValueError(f"field type {typ} is not a discrete type")


def nranf(u: float, n: int) -> int:
    """nranf(N) = clamp(nint(u*N + 0.5), 1, N) (random_wrap_mod.F90:159-168);
    Fortran nint rounds halves away from zero."""
    x = u * n + 0.5
    k = int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))
    return min(max(k, 1), n)


@dataclass
class Fields:
    f: np.ndarray     # (N_op, Ltrot) int field values
    t: np.ndarray     # (N_op,) field type per operator

    @classmethod
    def make(cls, n_op: int, ltrot: int) -> "Fields":
        return cls(np.zeros((n_op, ltrot), dtype=np.int64), np.zeros(n_op, dtype=np.int64))

    def phi(self, n: int, nt: int) -> float:
        return phi_st(int(self.f[n, nt]), int(self.t[n]))

    def gama(self, n: int, nt: int) -> float:
        return gama_st(int(self.f[n, nt]), int(self.t[n]))

    def set_random_pm1(self, uniforms) -> None:
        """Fields_set for types < 4: f = -1 if u > 0.5 else +1, slice outer, op inner."""
        it = iter(uniforms)
        for nt in range(self.f.shape[1]):
            for i in range(self.f.shape[0]):
                self.f[i, nt] = -1 if next(it) > 0.5 else 1

    def set_uniform_four_state(self, seed: int) -> None:
        """Uniform i.i.d. start over {-2,-1,+1,+2} from numpy's default_rng(seed).

        This is the start the workloads of this engine use ('discrete 4-state HS
        field drawn uniformly from seed'); the reference's own random start is +-1
        only, so such a start enters it through a confin file.
        """
        rng = np.random.default_rng(seed)
        vals = np.array([-2, -1, 1, 2], dtype=np.int64)
        self.f[:, :] = vals[rng.integers(0, 4, size=self.f.shape)]

    def write_conf(self, path: str, seeds) -> None:
        """confout/confin text format: seed integers on the first line, then one integer
        per (op, slice), slice outer (Fields_mod.F90:750-778)."""
        with open(path, "w") as fh:
            fh.write(" ".join(str(int(s)) for s in seeds) + "\n")
            for nt in range(self.f.shape[1]):
                for i in range(self.f.shape[0]):
                    fh.write(f"{int(self.f[i, nt])}\n")

    def read_conf(self, path: str, n_seeds: int):
        """Fields_read_conf (Fields_mod.F90:631-662); returns the seed vector."""
        with open(path) as fh:
            tokens = fh.read().split()
        seeds = [int(x) for x in tokens[:n_seeds]]
        vals = iter(tokens[n_seeds:])
        for nt in range(self.f.shape[1]):
            for i in range(self.f.shape[0]):
                self.f[i, nt] = int(next(vals))
        return seeds
