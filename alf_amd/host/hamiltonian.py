"""Model plugin API and the Hubbard plugin.

``HamBase`` mirrors the part of ALF's ``ham_base`` the sweep calls
(``Prog/Hamiltonian_main_mod.F90:153-197``): after ``ham_set`` the plugin exposes
``Op_T``, ``Op_V``, ``nsigma``, ``Ndim``, ``N_FL``, ``N_SUN``, ``Ltrot`` and ``S0``.

``Hubbard`` restates ``Prog/Hamiltonians/Hamiltonian_Hubbard_smod.F90`` for the
density-channel discrete HS coupling (``Mz=.F.``, ``Continuous=.F.``) on the
square and honeycomb lattices, with the dense hopping exponential or the checkerboard
decomposition (``Checkerboard``), and the symmetric Trotter form (``Symm``: symmetrised
checkerboard families, and measurements on exp(-dtau T/2) G exp(dtau T/2)):
  * Ham_Set 207-337 (Ltrot = nint(beta/dtau), N_FL handling)
  * Ham_Hop 362-448 -> Predefined_Hoppings_set_OPT (dense and checkerboard branches)
  * Ham_V 477-542 -> Predefined_Int_U_SUN (Prog/Predefined_Int_mod.F90:59-76)
  * S0 870-885 (== 1 for discrete fields)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import hoppings, lattices
from .fields import Fields
from .operator import ExpOpT, Operator, exp_opt_init, op_make, op_set


@dataclass
class HubbardParams:
    """Namelist values (VAR_lattice, VAR_Model_Generic, VAR_Hubbard, VAR_QMC subset)."""
    Lattice_type: str = "Square"
    L1: int = 4
    L2: int = 4
    N_SUN: int = 2
    N_FL: int = 1
    Phi_X: float = 0.0
    Phi_Y: float = 0.0
    Bulk: bool = True
    N_Phi: int = 0
    Dtau: float = 0.1
    Beta: float = 5.0
    Checkerboard: bool = False
    Symm: bool = False
    ham_T: float = 1.0
    Ham_chem: float = 0.0
    Ham_U: float = 4.0
    Mz: bool = False
    Continuous: bool = False
    Nwrap: int = 10


@dataclass
class HamBase:
    Ndim: int
    N_FL: int
    N_SUN: int
    Ltrot: int
    Nwrap: int
    Op_T: ExpOpT
    Op_V: list
    nsigma: Fields
    params: object = None
    List: np.ndarray = None
    Invlist: np.ndarray = None
    latt: object = None
    unit: object = None
    hop_matrix: np.ndarray = None
    extra: dict = field(default_factory=dict)

    def S0(self, n: int, nt: int, hs_new) -> float:
        """Bosonic weight ratio; 1 for the discrete Hubbard coupling
        (Hamiltonian_Hubbard_smod.F90:870-885)."""
        return 1.0

    @property
    def n_op(self) -> int:
        return len(self.Op_V)


def predefined_int_u_sun(I: int, n_sun: int, dtau: float, U: float) -> Operator:
    """Predefined_Int_U_SUN (Predefined_Int_mod.F90:59-76): U/N_SUN [sum_s (n_s - 1/2)]^2."""
    op = op_make(1)
    op.P[0] = I
    op.O[0, 0] = 1.0
    op.alpha = -0.5 + 0j
    op.g = np.sqrt(complex(-dtau * U / float(n_sun), 0.0))
    op.type = 2
    return op_set(op)


def hubbard_ham_set(p: HubbardParams) -> HamBase:
    """Ham_Set of the Hubbard plugin (Hamiltonian_Hubbard_smod.F90:207-337)."""
    if p.Mz or p.Continuous:
        raise NotImplementedError("this build covers the density-channel discrete HS coupling (Mz=.F.)")
    if p.N_FL != 1:
        raise NotImplementedError("this build covers N_FL = 1")
    ltrot = int(np.rint(p.Beta / p.Dtau))
    ndim, List, Invlist, latt, unit = lattices.predefined_latt(p.Lattice_type, p.L1, p.L2)
    t = p.Lattice_type.upper()
    if t == "SQUARE":
        hop = hoppings.default_hopping_square(p.ham_T, p.Ham_chem, p.Phi_X, p.Phi_Y, p.Bulk, p.N_Phi, latt, unit)
    elif t == "HONEYCOMB":
        hop = hoppings.default_hopping_honeycomb(p.ham_T, p.Ham_chem, p.Phi_X, p.Phi_Y, p.Bulk, p.N_Phi, latt, unit)
    else:
        raise NotImplementedError(t)
    O = hoppings.dense_hopping_matrix(hop, Invlist, latt, unit)
    op_t = op_make(ndim)
    op_t.O[:, :] = O
    op_t.P[:] = np.arange(ndim)
    op_t.g = -p.Dtau + 0j
    op_t.alpha = 0j
    op_set(op_t)
    cb = None
    if p.Checkerboard:
        # bond operators (Predefined_Hop_mod.F90:1773-1834); Op_T holds their dense products
        # P = E_1 ... E_N and P^-1, the matrices Hop_mod_mmthr / Hop_mod_mmthr_m1 apply
        cb = hoppings.checkerboard_opt(hop, Invlist, latt, unit, p.Lattice_type, p.Dtau, symm=p.Symm)
        exp_t = ExpOpT(cb.is_real, np.arange(ndim), cb.dense(ndim, "P"), cb.dense(ndim, "Pinv"),
                       cb.dense(ndim, "Q") if p.Symm else None, cb.dense(ndim, "Qinv") if p.Symm else None,
                       -p.Dtau + 0j)
    else:
        exp_t = exp_opt_init(op_t)

    op_v = []
    for I1 in range(latt.N):
        for no in range(unit.Norb):
            I = int(Invlist[I1, no])
            if abs(p.Ham_U) > lattices.EPS_SMALL:
                op_v.append(predefined_int_u_sun(I, p.N_FL * p.N_SUN, p.Dtau, p.Ham_U))
    nsigma = Fields.make(len(op_v), ltrot)
    nsigma.t[:] = [op.type for op in op_v]
    ham = HamBase(ndim, p.N_FL, p.N_SUN, ltrot, p.Nwrap, exp_t, op_v, nsigma, p,
                  List, Invlist, latt, unit, O)
    ham.extra["op_t"] = op_t
    ham.extra["symm"] = bool(p.Symm)
    if cb is not None:
        ham.extra["checkerboard"] = cb
    return ham
