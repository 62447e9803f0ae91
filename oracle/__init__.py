"""ctypes view of the CPU FP64 oracle (oracle/liboracle.so) and of the
reference build (oracle/_ref/libpgref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
Reference lines each oracle function follows are listed in pgxc_oracle.c.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

XC_XALPHA, XC_PWLDAC, XC_PBEX, XC_PBEC = 1, 3, 10, 18
MASK_PBE = (1 << XC_XALPHA) | (1 << XC_PWLDAC) | (1 << XC_PBEX) | (1 << XC_PBEC)
MASK_LDA = (1 << XC_XALPHA) | (1 << XC_PWLDAC)


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -f oracle/Makefile")
        L = ctypes.CDLL(path)
        L.or_sh_point.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _dp, _dp, _dp, _dp]
        L.or_solid_harmonics.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.or_prim_norm.argtypes = [ctypes.c_int, ctypes.c_double]
        L.or_prim_norm.restype = ctypes.c_double
        L.or_nbas.argtypes = [ctypes.c_int, _ip]
        L.or_nbas.restype = ctypes.c_int
        L.or_basis_eval.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _dp,
                                    ctypes.c_int, _dp, ctypes.c_long, _dp, _dp]
        for name in ("or_xalpha", "or_pw92"):
            getattr(L, name).argtypes = [ctypes.c_int, _dp, _dp, _dp]
        for name in ("or_pbex", "or_pbec"):
            getattr(L, name).argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_functional.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.or_functional.restype = ctypes.c_int
        L.or_functional_spin.argtypes = [ctypes.c_int, ctypes.c_int] + [_dp] * 8
        L.or_functional_spin.restype = ctypes.c_int
        L.or_pw92_spin.argtypes = [ctypes.c_int] + [_dp] * 5
        L.or_pbec_spin.argtypes = [ctypes.c_int] + [_dp] * 9
        for name in ("or_xc_build", "or_xc_build_spin"):
            fn = getattr(L, name)
            fn.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, _dp,
                           _dp, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                           ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp,
                           ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
            fn.restype = ctypes.c_int
        L.or_slater_radius.argtypes = [ctypes.c_int]
        L.or_slater_radius.restype = ctypes.c_double
        L.or_radial_grid.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        L.or_radial_grid.restype = ctypes.c_int
        L.or_lebedev.argtypes = [ctypes.c_int, _dp]
        L.or_lebedev.restype = ctypes.c_int
        L.or_atomic_grid.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp,
                                     ctypes.c_long, _dp]
        L.or_atomic_grid.restype = ctypes.c_long
        L.or_becke0.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_long, ctypes.c_long,
                                _dp, _dp, ctypes.c_int]
        _LIB = L
    return _LIB


def ref_lib():
    """The compiled reference modules, or None where they were not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libpgref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path, mode=os.RTLD_LAZY)
        L.pgref_pw92.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.pgref_pbec.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.pgref_pw92_spin.argtypes = [ctypes.c_int] + [_dp] * 5
        L.pgref_pbec_spin.argtypes = [ctypes.c_int] + [_dp] * 11
        L.pgref_solid_harmonics.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        _REF = L
    return _REF


# ---------------------------------------------------------------- functionals
def _term1(name, rho):
    rho = _c(rho)
    f, fr = np.zeros_like(rho), np.zeros_like(rho)
    getattr(lib(), name)(rho.size, _d(rho), _d(f), _d(fr))
    return f, fr


def _term2(name, rho, gam):
    rho, gam = _c(rho), _c(gam)
    f, fr, fg = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    getattr(lib(), name)(rho.size, _d(rho), _d(gam), _d(f), _d(fr), _d(fg))
    return f, fr, fg


def xalpha(rho):
    return _term1("or_xalpha", rho)


def pw92(rho):
    return _term1("or_pw92", rho)


def pbex(rho, gam):
    return _term2("or_pbex", rho, gam)


def pbec(rho, gam):
    return _term2("or_pbec", rho, gam)


def functional(mask, rho, gam):
    rho, gam = _c(rho), _c(gam)
    f, fr, fg = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    rc = lib().or_functional(mask, rho.size, _d(rho), _d(gam), _d(f), _d(fr), _d(fg))
    if rc != 0:
        raise ValueError("or_functional: unsupported functional mask %#x (%d)" % (mask, rc))
    return f, fr, fg


def functional_spin(mask, na, nb, gaa, gbb, gab):
    """ispin = 2: f, fr (n, 2) = dF/d(na, nb), fg (n, 3) = dF/d(gaa, gbb, gab)."""
    na, nb, gaa, gbb, gab = (_c(a) for a in (na, nb, gaa, gbb, gab))
    n = na.size
    f, fr, fg = np.zeros(n), np.zeros((n, 2)), np.zeros((n, 3))
    rc = lib().or_functional_spin(mask, n, _d(na), _d(nb), _d(gaa), _d(gbb), _d(gab), _d(f),
                                  _d(fr), _d(fg))
    if rc != 0:
        raise ValueError("or_functional_spin: unsupported functional mask %#x (%d)" % (mask, rc))
    return f, fr, fg


def pw92_spin(na, nb):
    na, nb = _c(na), _c(nb)
    f, fa, fb = np.zeros_like(na), np.zeros_like(na), np.zeros_like(na)
    lib().or_pw92_spin(na.size, _d(na), _d(nb), _d(f), _d(fa), _d(fb))
    return f, fa, fb


def pbec_spin(na, nb, gaa, gbb, gab):
    """f, dF/dna, dF/dnb, dF/dg2 (g2 = gaa + gbb + 2 gab)."""
    na, nb, gaa, gbb, gab = (_c(a) for a in (na, nb, gaa, gbb, gab))
    out = [np.zeros_like(na) for _ in range(4)]
    lib().or_pbec_spin(na.size, _d(na), _d(nb), _d(gaa), _d(gbb), _d(gab), *(_d(o) for o in out))
    return tuple(out)


def ref_pw92_spin(na, nb):
    na, nb = _c(na), _c(nb)
    f, fa, fb = np.zeros_like(na), np.zeros_like(na), np.zeros_like(na)
    ref_lib().pgref_pw92_spin(na.size, _d(na), _d(nb), _d(f), _d(fa), _d(fb))
    return f, fa, fb


def ref_pbec_spin(na, nb, gaa, gbb, gab):
    """f, dF/dna, dF/dnb, dF/dgaa, dF/dgbb, dF/dgab from the reference module."""
    na, nb, gaa, gbb, gab = (_c(a) for a in (na, nb, gaa, gbb, gab))
    out = [np.zeros_like(na) for _ in range(6)]
    ref_lib().pgref_pbec_spin(na.size, _d(na), _d(nb), _d(gaa), _d(gbb), _d(gab),
                              *(_d(o) for o in out))
    return tuple(out)


def ref_pw92(rho):
    rho = _c(rho)
    f, fr = np.zeros_like(rho), np.zeros_like(rho)
    ref_lib().pgref_pw92(rho.size, _d(rho), _d(f), _d(fr))
    return f, fr


def ref_pbec(rho, gam):
    rho, gam = _c(rho), _c(gam)
    f, fr, fg = np.zeros_like(rho), np.zeros_like(rho), np.zeros_like(rho)
    ref_lib().pgref_pbec(rho.size, _d(rho), _d(gam), _d(f), _d(fr), _d(fg))
    return f, fr, fg


# ------------------------------------------------------------ solid harmonics
def solid_harmonics(lmax, rel_xyz):
    """rel_xyz (npts,3) -> C (nlm,npts), G (nlm,3,npts)."""
    xyz = np.asfortranarray(rel_xyz, dtype=np.float64)
    npts = xyz.shape[0]
    nlm = (lmax + 1) ** 2
    C = np.zeros((nlm, npts))
    G = np.zeros((nlm, 3, npts))
    lib().or_solid_harmonics(lmax, npts, xyz.ctypes.data_as(_dp), _d(C), _d(G))
    return C, G


def ref_solid_harmonics(lmax, pts, center):
    pts = np.asfortranarray(pts, dtype=np.float64)
    npts = pts.shape[0]
    nlm = (lmax + 1) ** 2
    C = np.zeros((nlm, npts))          # Fortran c(npts, nlm)
    G = np.zeros((nlm, 3, npts))       # Fortran g(npts, 3, nlm)
    ctr = _c(center)
    ref_lib().pgref_solid_harmonics(npts, lmax, pts.ctypes.data_as(_dp), _d(ctr), _d(C), _d(G))
    return C, G


def prim_norm(L, alpha):
    return lib().or_prim_norm(L, alpha)


# ----------------------------------------------------------------- basis, XC
def _basis_args(basis):
    info = _c(basis.shell_info, np.int32)
    return (basis.n_atoms, _c(basis.atom_xyz), basis.n_shells, info,
            _c(basis.exps), _c(basis.norms), _c(basis.contr))


def basis_eval(basis, pts):
    """pts (npts,3) -> phi (nbas,npts), grad (nbas,3,npts)."""
    na, axyz, ns, info, ex, nm, ct = _basis_args(basis)
    P = np.asfortranarray(pts, dtype=np.float64)
    npts = P.shape[0]
    n = basis.nbas
    phi = np.zeros((n, npts))
    grad = np.zeros((n, 3, npts))
    lib().or_basis_eval(na, _d(axyz), ns, info.ctypes.data_as(_ip), _d(ex), _d(nm), _d(ct),
                        npts, P.ctypes.data_as(_dp), npts, _d(phi), _d(grad))
    return phi, grad


def xc_build(basis, grid, mask, P=None, C=None, occ=None, veclen=128,
             batch_range=(0, -1), nthreads=1, ispin=1):
    """Reference-order FP64 XC build.  grid: (npts,4) x,y,z,w.
    ispin = 2: P is (n, n, 2) or C (n, nocc, 2) with occ (nocc, 2), alpha
    then beta, and ham holds the two spins interleaved (alpha at even, beta
    at odd positions).  Returns (ham_packed, exc, charge)."""
    na, axyz, ns, info, ex, nm, ct = _basis_args(basis)
    G = np.asfortranarray(grid, dtype=np.float64)
    npts = G.shape[0]
    n = basis.nbas
    if ispin not in (1, 2):
        raise ValueError("ispin must be 1 or 2")
    if P is not None:
        P = np.asarray(P, dtype=np.float64)
        if P.shape != ((n, n) if ispin == 1 else (n, n, 2)):
            raise ValueError("P has shape %s" % (P.shape,))
        dform, dens, nocc, oc = 0, np.asfortranarray(P), 0, np.zeros(2)
    else:
        C = np.asarray(C, dtype=np.float64)
        if C.ndim != (2 if ispin == 1 else 3) or C.shape[0] != n:
            raise ValueError("C has shape %s" % (C.shape,))
        dform, dens = 1, np.asfortranarray(C)
        nocc = C.shape[1]
        oc = np.asfortranarray(np.asarray(occ, dtype=np.float64).reshape(C.shape[1:]))
    ham = np.zeros(ispin * n * (n + 1) // 2)
    e = np.zeros(1)
    q = np.zeros(1)
    fn = lib().or_xc_build if ispin == 1 else lib().or_xc_build_spin
    rc = fn(na, _d(axyz), ns, info.ctypes.data_as(_ip), _d(ex), _d(nm), _d(ct),
            G.ctypes.data_as(_dp), npts, veclen, mask, dform,
            dens.ctypes.data_as(_dp), nocc, oc.ctypes.data_as(_dp), _d(ham), _d(e), _d(q),
            batch_range[0], batch_range[1], nthreads)
    if rc != 0:
        raise ValueError("or_xc_build failed: %d" % rc)
    return ham, float(e[0]), float(q[0])


# ----------------------------------------------------------------------- grid
ZNUM = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
ICGRD = {194: 231, 302: 291}  # NANG of the reference's Lebedev rules


def slater_radius(el):
    return lib().or_slater_radius(ZNUM[el])


def radial_grid(nradpts, R):
    r, w = np.zeros(nradpts), np.zeros(nradpts)
    n = lib().or_radial_grid(nradpts, R, _d(r), _d(w))
    return r[:n], w[:n]


def lebedev(nang):
    """(points (n,3), weights (n,)) in INIGRD's order."""
    p = np.zeros(4 * nang)
    n = lib().or_lebedev(ICGRD[nang], _d(p))
    if n != nang:
        raise ValueError("or_lebedev: %d points" % n)
    p = p.reshape(4, nang).T
    return np.ascontiguousarray(p[:, :3]), np.ascontiguousarray(p[:, 3])


def atomic_grid(nradpts, R, nang, centre):
    cap = nradpts * nang
    out = np.zeros(4 * cap)
    n = lib().or_atomic_grid(nradpts, R, ICGRD[nang], _d(_c(centre)), cap, _d(out))
    if n < 0:
        raise ValueError("or_atomic_grid failed")
    g = out.reshape(4, cap).T[:n]
    return np.ascontiguousarray(g[:, :3]), np.ascontiguousarray(g[:, 3])


def becke0(atom_xyz, host, pts, nthreads=1):
    """Partition weight of atom host at the points (npts, 3)."""
    A = _c(atom_xyz)
    X = np.asfortranarray(pts, dtype=np.float64)
    w = np.zeros(X.shape[0])
    lib().or_becke0(A.shape[0], _d(A), host, X.shape[0], X.shape[0], X.ctypes.data_as(_dp),
                    _d(w), nthreads)
    return w


def molecular_grid(elements, atom_xyz, nrad, nang, nthreads=1):
    """The reference's linear grid grdpts(npts, 4) (mkgrid,
    modules/grid_module.f90:3045-3164): per atom in order, Slater radius R,
    int(nrad R) Pople shells times the Lebedev rule, weights times becke0."""
    parts = []
    for a, (el, c) in enumerate(zip(elements, np.asarray(atom_xyz))):
        R = slater_radius(el)
        x, w = atomic_grid(int(nrad * R), R, nang, c)
        w = w * becke0(atom_xyz, a, x, nthreads) * 1  # times n_equal_atoms = 1 (C1)
        parts.append(np.column_stack([x, w]))
    return np.asfortranarray(np.vstack(parts))
