"""Contracted solid-harmonic Gaussian basis in the layout of include/pgxc.h.

A shell is one (atom, l) block of ParaGauss's unique_atom_basis_type
(modules/unique_atom_module.f90:143-168): n_exp exponents, the first n_unc
used uncontracted, then n_contr contractions over all n_exp normalised
primitives.  Functions are ordered atom -> l -> m -> radial function
(modules/orbitalprojection_module.f90:30-50, C1 symmetry).
"""
from dataclasses import dataclass, field

import numpy as np

PI_REF = 3.14159265358979324  # the literal used by unique_atom_basis_norm_ob


def dfac(k):
    """1 * 3 * ... * (2k - 1)  (modules/unique_atom_module.f90:465-486)."""
    n = 1
    for i in range(1, k + 1):
        n *= 2 * i - 1
    return n


def legacy_norm(L, alpha):
    """(2a/pi)^(3/4) (2 sqrt a)^L / sqrt((2L-1)!!)
    (unique_atom_basis_norm_ob, modules/unique_atom_module.f90:426-444)."""
    alpha = np.asarray(alpha, dtype=np.float64)
    return (2 * alpha / PI_REF) ** 0.75 * (2 * np.sqrt(alpha)) ** L / np.sqrt(float(dfac(L)))


@dataclass
class Shell:
    atom: int
    l: int
    exps: np.ndarray
    n_unc: int
    contr: np.ndarray  # (n_exp, n_contr)
    norms: np.ndarray = None

    def __post_init__(self):
        self.exps = np.asarray(self.exps, dtype=np.float64)
        self.contr = np.asarray(self.contr, dtype=np.float64).reshape(len(self.exps), -1)
        if self.norms is None:
            self.norms = legacy_norm(self.l, self.exps)

    @property
    def n_contr(self):
        return self.contr.shape[1]

    @property
    def nfunc(self):
        return (2 * self.l + 1) * (self.n_unc + self.n_contr)


@dataclass
class Basis:
    atom_xyz: np.ndarray           # (n_atoms, 3) bohr
    shells: list = field(default_factory=list)

    def __post_init__(self):
        self.atom_xyz = np.ascontiguousarray(self.atom_xyz, dtype=np.float64).reshape(-1, 3)

    @property
    def n_atoms(self):
        return self.atom_xyz.shape[0]

    @property
    def n_shells(self):
        return len(self.shells)

    @property
    def nbas(self):
        return sum(s.nfunc for s in self.shells)

    @property
    def lmax(self):
        return max(s.l for s in self.shells)

    @property
    def shell_info(self):
        return np.array([[s.atom, s.l, len(s.exps), s.n_unc, s.n_contr] for s in self.shells],
                        dtype=np.int32).reshape(-1)

    @property
    def exps(self):
        return np.concatenate([s.exps for s in self.shells])

    @property
    def norms(self):
        return np.concatenate([s.norms for s in self.shells])

    @property
    def contr(self):
        parts = [s.contr.reshape(-1, order="F") for s in self.shells if s.contr.size]
        return np.concatenate(parts) if parts else np.zeros(1)

    def function_atoms(self):
        """Atom index of every basis function."""
        out = []
        for s in self.shells:
            out += [s.atom] * s.nfunc
        return np.array(out, dtype=np.int64)


# Shell templates (l, n_exp, n_unc, n_contr) per element.  "3s2p1d" counts
# contracted + uncontracted radial functions per l.
SHELL_SETS = {
    # 24-function H2O basis: O 3s2p1d, H 2s1p
    "small": {
        "O": [(0, 6, 1, 2), (1, 4, 1, 1), (2, 1, 1, 0)],
        "H": [(0, 3, 1, 1), (1, 1, 1, 0)],
        "C": [(0, 6, 1, 2), (1, 4, 1, 1), (2, 1, 1, 0)],
        "N": [(0, 6, 1, 2), (1, 4, 1, 1), (2, 1, 1, 0)],
    },
    # angular momenta up to g: O 2s2p1d1f1g (43 functions), H 2s1p1d (10)
    "spdfg": {
        "O": [(0, 4, 1, 1), (1, 3, 1, 1), (2, 1, 1, 0), (3, 1, 1, 0), (4, 1, 1, 0)],
        "H": [(0, 3, 1, 1), (1, 1, 1, 0), (2, 1, 1, 0)],
    },
    # 33 functions per water: O 4s3p2d, H 2s1p
    "medium": {
        "O": [(0, 8, 1, 3), (1, 5, 1, 2), (2, 2, 2, 0)],
        "H": [(0, 3, 1, 1), (1, 1, 1, 0)],
        "C": [(0, 8, 1, 3), (1, 5, 1, 2), (2, 2, 2, 0)],
        "N": [(0, 8, 1, 3), (1, 5, 1, 2), (2, 2, 2, 0)],
    },
}


def seeded_basis(elements, atom_xyz, seed, shell_set="small", exp_range=(0.1, 5000.0)):
    """Synthetic basis: exponents log-uniform in exp_range (ascending, so the
    uncontracted ones are the most diffuse), contraction coefficients N(0,1)
    from the seed.  One set of shells per element, shared by its atoms."""
    rng = np.random.default_rng(seed)
    per_elem = {}
    for el in sorted(set(elements)):
        shells = []
        for (l, ne, nu, nc) in SHELL_SETS[shell_set][el]:
            ex = np.sort(np.exp(rng.uniform(np.log(exp_range[0]), np.log(exp_range[1]), ne)))
            ct = rng.normal(size=(ne, nc)) if nc else np.zeros((ne, 0))
            shells.append((l, ex, nu, ct))
        per_elem[el] = shells
    b = Basis(atom_xyz)
    for a, el in enumerate(elements):
        for (l, ex, nu, ct) in per_elem[el]:
            b.shells.append(Shell(a, l, ex.copy(), nu, ct.copy()))
    return b
