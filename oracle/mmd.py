"""ctypes wrapper of the C oracle (test infrastructure only).

Each wrapper cites the reference function its C routine restates.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborc.so")
BOYS_REFERENCE = 0
BOYS_ACCURATE = 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_boys_ref.restype = ctypes.c_double
        _lib.orc_boys_ref.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib.orc_boys_acc.restype = ctypes.c_double
        _lib.orc_boys_acc.argtypes = [ctypes.c_int, ctypes.c_double]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def boys(n: int, t, mode: int = BOYS_ACCURATE) -> np.ndarray:
    """jax_qc/integrals/boys.py:86-109 (mode 0) or the accurate series."""
    t = np.ascontiguousarray(np.atleast_1d(np.asarray(t, dtype=np.float64)))
    out = np.empty_like(t)
    lib().orc_boys_array(ctypes.c_int(mode), ctypes.c_int(n), _p(t), _p(out),
                         ctypes.c_int64(t.size))
    return out


def hermite_e(la, lb, Xpa, Xpb, p) -> np.ndarray:
    """jax_qc/integrals/mcmurchie_davidson.py:60-108."""
    E = np.zeros((la + 1, lb + 1, la + lb + 1))
    lib().orc_hermite_e(ctypes.c_int(la), ctypes.c_int(lb), ctypes.c_double(Xpa),
                        ctypes.c_double(Xpb), ctypes.c_double(p), _p(E))
    return E


def hermite_r_aux(t_max, u_max, v_max, n_max, p, PC, mode=BOYS_ACCURATE) -> np.ndarray:
    """jax_qc/integrals/mcmurchie_davidson.py:113-172."""
    R = np.zeros((t_max + 1, u_max + 1, v_max + 1, n_max + 1))
    PC = np.ascontiguousarray(PC, dtype=np.float64)
    lib().orc_hermite_r(ctypes.c_int(mode), ctypes.c_int(t_max), ctypes.c_int(u_max),
                        ctypes.c_int(v_max), ctypes.c_int(n_max), ctypes.c_double(p), _p(PC), _p(R))
    return R


def _nc(l):
    return (l + 1) * (l + 2) // 2


def _vec(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def primitive_eri_block(a, A, la, b, B, lb, c, C, lc, d, D, ld, mode=BOYS_ACCURATE):
    """jax_qc/integrals/obara_saika.py:227-332."""
    out = np.zeros((_nc(la), _nc(lb), _nc(lc), _nc(ld)))
    A, B, C, D = map(_vec, (A, B, C, D))
    dbl, i = ctypes.c_double, ctypes.c_int
    lib().orc_prim_eri(i(mode), dbl(a), _p(A), i(la), dbl(b), _p(B), i(lb), dbl(c), _p(C), i(lc),
                       dbl(d), _p(D), i(ld), _p(out))
    return out


def primitive_overlap_block(a, A, la, b, B, lb):
    """obara_saika.py:90-112."""
    out = np.zeros((_nc(la), _nc(lb)))
    A, B = _vec(A), _vec(B)
    lib().orc_prim_overlap(ctypes.c_double(a), _p(A), ctypes.c_int(la), ctypes.c_double(b),
                           _p(B), ctypes.c_int(lb), _p(out))
    return out


def primitive_kinetic_block(a, A, la, b, B, lb):
    """obara_saika.py:141-174."""
    out = np.zeros((_nc(la), _nc(lb)))
    A, B = _vec(A), _vec(B)
    lib().orc_prim_kinetic(ctypes.c_double(a), _p(A), ctypes.c_int(la), ctypes.c_double(b),
                           _p(B), ctypes.c_int(lb), _p(out))
    return out


def primitive_nuclear_block(a, A, la, b, B, lb, C, Z, mode=BOYS_ACCURATE):
    """obara_saika.py:177-222."""
    out = np.zeros((_nc(la), _nc(lb)))
    A, B = _vec(A), _vec(B)
    C = _vec(np.asarray(C).reshape(-1, 3))
    Z = _vec(np.asarray(Z).reshape(-1))
    lib().orc_prim_nuclear(ctypes.c_int(mode), ctypes.c_double(a), _p(A), ctypes.c_int(la),
                           ctypes.c_double(b), _p(B), ctypes.c_int(lb), ctypes.c_int(len(Z)),
                           _p(C), _p(Z), _p(out))
    return out


class OracleBasis:
    """Flat arrays of a BasisSet (l, nprim, prim_off, exps, coefs, centers)."""

    def __init__(self, basis):
        sh = basis.shells
        self.spherical = bool(basis.spherical)
        self.n_basis = int(basis.n_basis)
        self.l = np.ascontiguousarray([s.angular_momentum for s in sh], dtype=np.int32)
        self.nprim = np.ascontiguousarray([len(s.exponents) for s in sh], dtype=np.int32)
        self.prim_off = np.ascontiguousarray(
            np.concatenate([[0], np.cumsum(self.nprim)[:-1]]), dtype=np.int32)
        self.exps = _vec(np.concatenate([s.exponents for s in sh]))
        self.coefs = _vec(np.concatenate([s.coefficients for s in sh]))
        self.centers = _vec(np.array([s.center for s in sh]).reshape(-1, 3))
        self.nshell = len(sh)

    def args(self):
        return (ctypes.c_int(self.nshell), _p(self.l), _p(self.nprim), _p(self.prim_off),
                _p(self.exps), _p(self.coefs), _p(self.centers))


def eri_tensor(basis, mode=BOYS_ACCURATE) -> np.ndarray:
    """jax_qc/integrals/eri.py:69-105 (dense (n,n,n,n), 8-fold scatter)."""
    ob = OracleBasis(basis)
    n = ob.n_basis
    out = np.zeros((n, n, n, n))
    lib().orc_eri_tensor(ctypes.c_int(mode), *ob.args(), ctypes.c_int(int(ob.spherical)), _p(out))
    return out


def eri_tensor_touching(basis, atom: int, mode=BOYS_ACCURATE) -> np.ndarray:
    """eri_tensor with only the canonical quartets that hold a shell on
    ``atom`` computed (zeros elsewhere): the elements that move when that
    atom is displaced."""
    ob = OracleBasis(basis)
    n = ob.n_basis
    sat = np.ascontiguousarray([s.atom_index for s in basis.shells], dtype=np.int32)
    out = np.zeros((n, n, n, n))
    lib().orc_eri_tensor_atom(ctypes.c_int(mode), *ob.args(), ctypes.c_int(int(ob.spherical)),
                              _p(out), _p(sat), ctypes.c_int(int(atom)))
    return out


def eri_quartets_checksum(basis, quartets, mode=BOYS_ACCURATE) -> float:
    """Contracted Cartesian blocks of the (nq, 4) shell quartets (OpenMP
    over quartets); returns the sum of all elements (for timing runs)."""
    ob = OracleBasis(basis)
    q = np.ascontiguousarray(quartets, dtype=np.int32)
    f = lib().orc_eri_quartets
    f.restype = ctypes.c_double
    return float(f(ctypes.c_int(mode), *ob.args(), ctypes.c_int64(len(q)), _p(q)))


def contracted_eri_block(basis, i, j, k, l, mode=BOYS_ACCURATE) -> np.ndarray:
    """jax_qc/integrals/obara_saika.py:394-426 (Cartesian block)."""
    ob = OracleBasis(basis)
    L = ob.l
    out = np.zeros((_nc(L[i]), _nc(L[j]), _nc(L[k]), _nc(L[l])))
    lib().orc_contracted_eri(ctypes.c_int(mode), *ob.args(), ctypes.c_int(i), ctypes.c_int(j),
                             ctypes.c_int(k), ctypes.c_int(l), _p(out))
    return out


def shell_pair_bounds(basis, mode=BOYS_ACCURATE) -> np.ndarray:
    """jax_qc/integrals/screening.py:32-58."""
    ob = OracleBasis(basis)
    Q = np.zeros((ob.nshell, ob.nshell))
    lib().orc_schwarz(ctypes.c_int(mode), *ob.args(), _p(Q))
    return Q


def one_electron(basis, which: str, mol=None, mode=BOYS_ACCURATE) -> np.ndarray:
    """S (overlap.py:67-85), T (kinetic.py:42-59) or V (nuclear.py:62-81)."""
    ob = OracleBasis(basis)
    n = ob.n_basis
    M = np.zeros((n, n))
    code = {"S": 0, "T": 1, "V": 2}[which]
    if mol is not None:
        C = _vec(np.asarray(mol.coords).reshape(-1, 3))
        Z = _vec(np.asarray(mol.atomic_numbers, dtype=np.float64))
        nat = len(Z)
    else:
        C = Z = np.zeros(3)
        nat = 0
    lib().orc_one_electron(ctypes.c_int(mode), ctypes.c_int(code), *ob.args(),
                           ctypes.c_int(int(ob.spherical)), ctypes.c_int(nat), _p(C), _p(Z), _p(M))
    return M


def nuclear_repulsion_energy(mol) -> float:
    """jax_qc/integrals/nuclear.py:84-98."""
    c = np.asarray(mol.coords, dtype=np.float64)
    Z = np.asarray(mol.atomic_numbers, dtype=np.float64)
    tot = 0.0
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            r = float(np.linalg.norm(c[i] - c[j]))
            if r == 0.0:
                raise ValueError(f"Atoms {i} and {j} are coincident; cannot compute E_nuc.")
            tot += float(Z[i] * Z[j]) / r
    return tot


def compute_integrals(mol, basis, mode=BOYS_ACCURATE):
    """jax_qc/integrals/interface.py:26-67 -> dict of numpy arrays."""
    S = one_electron(basis, "S")
    T = one_electron(basis, "T")
    V = one_electron(basis, "V", mol, mode)
    return dict(S=S, T=T, V=V, H_core=T + V, ERI=eri_tensor(basis, mode),
                E_nuc=nuclear_repulsion_energy(mol))
