"""NumPy FP64 restatement of the reference SCF and gradient drivers
(test infrastructure only).

Each function cites the reference function it restates.  ``eigh`` is
numpy.linalg.eigh where the reference calls jnp.linalg.eigh.
"""

from __future__ import annotations

import numpy as np

from oracle import mmd


# --- jax_qc/scf/fock.py:31-79 ------------------------------------------------

def coulomb(D, ERI):
    return np.einsum("ls,mnls->mn", D, ERI)


def exchange(D, ERI):
    return np.einsum("ls,msln->mn", D, ERI)


def build_fock_rhf(H, D, ERI):
    return H + coulomb(D, ERI) - 0.5 * exchange(D, ERI)


def build_fock_uhf(H, Da, Db, ERI):
    J = coulomb(Da + Db, ERI)
    return H + J - exchange(Da, ERI), H + J - exchange(Db, ERI)


# --- density.py:23-58, energy.py:17-38, orthogonalize.py:22-26 -----------------

def density_rhf(C, n_occ):
    Co = C[:, :n_occ]
    return 2.0 * Co @ Co.T


def density_uhf(Ca, Cb, na, nb):
    return Ca[:, :na] @ Ca[:, :na].T, Cb[:, :nb] @ Cb[:, :nb].T


def electronic_energy_rhf(D, H, F):
    return 0.5 * float(np.sum(D * (H + F)))


def electronic_energy_uhf(Da, Db, H, Fa, Fb):
    return 0.5 * float(np.sum(Da * (H + Fa))) + 0.5 * float(np.sum(Db * (H + Fb)))


def symmetric_orthogonalization(S):
    w, U = np.linalg.eigh(S)
    return U @ np.diag(1.0 / np.sqrt(w)) @ U.T


def _diag(F, X):
    e, Cp = np.linalg.eigh(X.T @ F @ X)
    return e, X @ Cp


# --- diis.py:62-199 --------------------------------------------------------

class DIIS:
    def __init__(self, max_size):
        self.max_size = max_size
        self.focks, self.errors = [], []

    def push(self, F, e):
        if len(self.focks) == self.max_size:
            self.focks.pop(0)
            self.errors.pop(0)
        self.focks.append(F)
        self.errors.append(e)

    def solve(self):
        m = len(self.errors)
        B = np.empty((m + 1, m + 1))
        for i in range(m):
            for j in range(m):
                B[i, j] = float(np.sum(self.errors[i] * self.errors[j]))
        B[-1, :-1] = -1.0
        B[:-1, -1] = -1.0
        B[-1, -1] = 0.0
        rhs = np.zeros(m + 1)
        rhs[-1] = -1.0
        try:
            return np.linalg.solve(B, rhs)[:-1]
        except np.linalg.LinAlgError:
            return None


def diis_error(F, D, S, X):
    return X.T @ (F @ D @ S - S @ D @ F) @ X


# --- rhf.py:95-169 ---------------------------------------------------------

def run_rhf(mol, ints, max_iter=128, conv=1e-10, diis_space=8, damping=0.0):
    H, S, ERI = ints["H_core"], ints["S"], ints["ERI"]
    n_occ = mol.n_electrons // 2
    X = symmetric_orthogonalization(S)
    _, C = _diag(H, X)
    D = density_rhf(C, n_occ)
    F = build_fock_rhf(H, D, ERI)
    e, C = _diag(F, X)
    E = electronic_energy_rhf(D, H, F)
    diis = DIIS(diis_space)
    converged, last, it = False, E, 0
    for _ in range(max_iter):
        Fn = build_fock_rhf(H, D, ERI)
        if damping > 0.0:
            Fn = (1.0 - damping) * Fn + damping * F
        err = diis_error(Fn, D, S, X)
        diis.push(Fn, err)
        Fe = Fn
        if len(diis.focks) >= 2:
            c = diis.solve()
            if c is not None:
                Fe = np.einsum("i,imn->mn", c, np.array(diis.focks))
        e, C = _diag(Fe, X)
        D = density_rhf(C, n_occ)
        E = electronic_energy_rhf(D, H, Fe)
        F = Fe
        it += 1
        dE = abs(E - last)
        last = E
        if dE < conv and float(np.max(np.abs(err))) < max(conv, 1e-6):
            converged = True
            break
    return dict(converged=converged, energy=E + ints["E_nuc"], E_elec=E, D=D, F=F, C=C, eps=e,
                n_iter=it, n_occ=n_occ)


# --- guess.py:42-115, uhf.py:187-287 ----------------------------------------

def core_guess_uhf(H, X, na, nb, ERI):
    _, Ca = _diag(H, X)
    if na != nb:
        Da0 = Ca[:, :na] @ Ca[:, :na].T
        Fb = H + coulomb(Da0, ERI) - exchange(Da0, ERI)
        spread = np.ptp(np.linalg.eigvalsh(X.T @ Fb @ X))
        scale = 0.01 * float(spread) + 0.1
        idx = np.arange(H.shape[0], dtype=np.float64)
        noise = scale * np.outer(np.sin(idx * 1.23 + 0.45), np.cos(idx * 0.67 + 0.89))
        Fb = Fb + (noise + noise.T) / 2.0
        _, Cb = _diag(Fb, X)
    else:
        Cb = Ca
    Da, Db = density_uhf(Ca, Cb, na, nb)
    return Da, Db


def run_uhf(mol, ints, max_iter=128, conv=1e-10, diis_space=8, damping=0.0):
    H, S, ERI = ints["H_core"], ints["S"], ints["ERI"]
    na, nb = mol.n_alpha, mol.n_beta
    n = H.shape[0]
    X = symmetric_orthogonalization(S)
    Da, Db = core_guess_uhf(H, X, na, nb, ERI)
    Fa, Fb = build_fock_uhf(H, Da, Db, ERI)
    ea, Ca = _diag(Fa, X)
    eb, Cb = _diag(Fb, X)
    Da, Db = density_uhf(Ca, Cb, na, nb)
    E = electronic_energy_uhf(Da, Db, H, Fa, Fb)
    diis = DIIS(diis_space)
    init_damp = 0.3 if na != nb else 0.0
    converged, last, it = False, E, 0
    for k in range(max_iter):
        dv = init_damp if k < 5 else damping
        Fan, Fbn = build_fock_uhf(H, Da, Db, ERI)
        if dv > 0.0:
            Fan = (1 - dv) * Fan + dv * Fa
            Fbn = (1 - dv) * Fbn + dv * Fb
        eA = diis_error(Fan, Da, S, X)
        eB = diis_error(Fbn, Db, S, X)
        diis.push(np.concatenate([Fan, Fbn]), np.concatenate([eA, eB]))
        Fa_e, Fb_e = Fan, Fbn
        if len(diis.focks) >= 2:
            c = diis.solve()
            if c is not None:
                Fx = np.einsum("i,imn->mn", c, np.array(diis.focks))
                Fa_e, Fb_e = Fx[:n], Fx[n:]
        ea, Ca = _diag(Fa_e, X)
        eb, Cb = _diag(Fb_e, X)
        Da, Db = density_uhf(Ca, Cb, na, nb)
        E = electronic_energy_uhf(Da, Db, H, Fa_e, Fb_e)
        Fa, Fb = Fa_e, Fb_e
        it += 1
        err = max(float(np.max(np.abs(eA))), float(np.max(np.abs(eB))))
        dE = abs(E - last)
        last = E
        if dE < conv and err < max(conv, 1e-6):
            converged = True
            break
    return dict(converged=converged, energy=E + ints["E_nuc"], E_elec=E, Da=Da, Db=Db, Fa=Fa,
                Fb=Fb, Ca=Ca, Cb=Cb, ea=ea, eb=eb, n_iter=it)


# --- grad/rhf_grad.py:92-222, grad/numerical_grad.py:51-94 -----------------

def energy_weighted_density(C, eps, n_occ):
    Co = C[:, :n_occ]
    return 2.0 * (Co * eps[None, :n_occ]) @ Co.T


def _nuc_rep(coords, Z):
    tot = 0.0
    for i in range(len(Z)):
        for j in range(i + 1, len(Z)):
            tot += Z[i] * Z[j] / float(np.linalg.norm(coords[i] - coords[j]))
    return tot


def rhf_gradient(mol, res, basis_fn, displace_fn, step=1e-4, mode=mmd.BOYS_ACCURATE,
                 touching_only=False, atoms=None):
    """FD-of-integrals Hellmann-Feynman + Pulay gradient at fixed D, W
    (rhf_grad.py:106-211).  ``basis_fn(mol)`` rebuilds the basis,
    ``displace_fn(mol, coords)`` returns the molecule at new coordinates.
    touching_only: the two ERI rebuilds of a displaced atom compute only the
    quartets with a shell on that atom (every other element is computed from
    identical inputs at both geometries, so its difference is exactly zero);
    the result is the full-rebuild one up to summation order.  atoms: only
    these atoms' rows are computed (the rest stay zero)."""
    coords = np.asarray(mol.coords, dtype=np.float64)
    Z = np.asarray(mol.atomic_numbers, dtype=np.float64)
    D = res["D"]
    W = energy_weighted_density(res["C"], res["eps"], mol.n_electrons // 2)
    g = np.zeros_like(coords)
    for a in (range(len(Z)) if atoms is None else atoms):
        for k in range(3):
            cp, cm = coords.copy(), coords.copy()
            cp[a, k] += step
            cm[a, k] -= step
            dEn = (_nuc_rep(cp, Z) - _nuc_rep(cm, Z)) / (2 * step)
            mp, mm = displace_fn(mol, cp), displace_fn(mol, cm)
            if touching_only:
                bp, bm = basis_fn(mp), basis_fn(mm)
                ip = dict(S=mmd.one_electron(bp, "S"),
                          H_core=mmd.one_electron(bp, "T") + mmd.one_electron(bp, "V", mp, mode),
                          ERI=mmd.eri_tensor_touching(bp, a, mode))
                im = dict(S=mmd.one_electron(bm, "S"),
                          H_core=mmd.one_electron(bm, "T") + mmd.one_electron(bm, "V", mm, mode),
                          ERI=mmd.eri_tensor_touching(bm, a, mode))
            else:
                ip = mmd.compute_integrals(mp, basis_fn(mp), mode)
                im = mmd.compute_integrals(mm, basis_fn(mm), mode)
            dH = (ip["H_core"] - im["H_core"]) / (2 * step)
            dS = (ip["S"] - im["S"]) / (2 * step)
            dERI = (ip["ERI"] - im["ERI"]) / (2 * step)
            dJ = np.einsum("ls,mnls->mn", D, dERI)
            dK = np.einsum("ls,msln->mn", D, dERI)
            g[a, k] = dEn + np.sum(D * dH) + 0.5 * np.sum(D * (dJ - 0.5 * dK)) - np.sum(W * dS)
    return g


def numerical_gradient(mol, energy_fn, displace_fn, step=1e-4):
    """Central differences of full-SCF energies (numerical_grad.py:51-94)."""
    coords = np.asarray(mol.coords, dtype=np.float64)
    g = np.zeros_like(coords)
    for a in range(coords.shape[0]):
        for k in range(3):
            cp, cm = coords.copy(), coords.copy()
            cp[a, k] += step
            cm[a, k] -= step
            g[a, k] = (energy_fn(displace_fn(mol, cp)) - energy_fn(displace_fn(mol, cm))) / (2 * step)
    return g
