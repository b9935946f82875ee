"""Initial conditions: Elements, ElementsIC, CartesianIC and init_nbody.

Host-side mirror of NbodyGradient.jl src/ics/ (InitialConditions.jl,
init_nbody.jl, kepler_init.jl, kepler.jl, setup_hierarchy.jl).  It runs once
per system on the CPU, producing the Cartesian state and jac_init (the
derivative of (x, v, m) with respect to (P, t0, e cos w, e sin w, I, Omega, m))
that the GPU integration starts from.  Arrays use Julia's column-major layout
(numpy order "F"), indices are 0-based.

Scalar arithmetic follows the reference's operation order; elementary
functions come from the C library (math.*, and cbrt through ctypes because
Python 3.10's math module has none), so results are reproducible bit for bit.
"""
import ctypes
import ctypes.util
import math

import numpy as np

YEAR = 365.242
GNEWT = 39.4845 / (YEAR * YEAR)  # NbodyGradient.jl:15
THIRD = 1.0 / 3.0                 # NbodyGradient.jl:16
PI = math.pi

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.cbrt.argtypes = [ctypes.c_double]
_libm.cbrt.restype = ctypes.c_double


def _cbrt(x):
    return _libm.cbrt(x)


def _sign(x):
    return -1.0 if x < 0 else (1.0 if x > 0 else x)


# --------------------------------------------------------------------- Elements
class Elements:
    """Orbital elements of a binary and mass of its outer body
    (InitialConditions.jl:21-107).  Keyword construction follows
    Elements(; m, P, t0, ecosw, esinw, I, Omega, a, e, omega, tp)."""

    FIELDS = ("m", "P", "t0", "ecosw", "esinw", "I", "Omega")

    def __init__(self, m, P=math.nan, t0=math.nan, ecosw=math.nan, esinw=math.nan, I=math.nan, Omega=math.nan,
                 a=math.nan, e=math.nan, omega=math.nan, tp=math.nan):
        vals = [P, t0, ecosw, esinw, I, Omega, a, omega, e, tp]
        if all(math.isnan(x) for x in vals):
            self.m, self.P, self.t0, self.ecosw, self.esinw, self.I, self.Omega = float(m), 0.0, 0.0, 0.0, 0.0, 0.0, 0.0
            self.a = self.e = self.omega = self.tp = 0.0
            return
        if math.isnan(P) and math.isnan(a):
            raise ValueError("Must specify either P or a.")
        if P > 0 and a > 0:
            raise ValueError("Only one of P (period) or a (semi-major axis) can be defined")
        if P < 0:
            raise ValueError("Period must be positive")
        if a < 0:
            raise ValueError("Orbital semi-major axis must be positive")
        if not math.isnan(e) and not (0.0 <= e < 1.0):
            raise ValueError("Eccentricity must be in [0,1), e=%r" % e)
        if not (math.isnan(ecosw) and math.isnan(esinw)) and not math.isnan(e):
            raise ValueError("Must specify either e and ecosw/esinw, but not both.")
        if not (math.isnan(tp) or math.isnan(t0)):
            raise ValueError("Must specify either initial transit time or time of periastron passage, but not both.")
        omega = 0.0 if math.isnan(omega) else omega
        if math.isnan(e) and math.isnan(ecosw) and math.isnan(esinw):
            e = ecosw = esinw = 0.0
        if math.isnan(e):
            ecosw = 0.0 if math.isnan(ecosw) else ecosw
            esinw = 0.0 if math.isnan(esinw) else esinw
            e = math.sqrt(ecosw * ecosw + esinw * esinw)
            if not (0.0 <= e < 1.0):
                raise ValueError("Eccentricity must be in [0,1)")
            omega = math.atan2(esinw, ecosw)
        else:
            esinw, ecosw = e * math.sin(omega), e * math.cos(omega)
            esinw = 0.0 if abs(esinw) < 2.220446049250313e-16 else esinw
            ecosw = 0.0 if abs(ecosw) < 2.220446049250313e-16 else ecosw
        t0 = 0.0 if math.isnan(t0) else t0
        n = 2 * PI / P
        if math.isnan(tp) and e != 0:
            tp = math.fmod(t0 - math.sqrt(1.0 - e * e) * ecosw / (n * (1.0 - esinw))
                           - (2.0 / n) * math.atan2(math.sqrt(1.0 - e) * (esinw + ecosw + e),
                                                    math.sqrt(1.0 + e) * (esinw - ecosw - e)), P)
        elif math.isnan(tp):
            tp = (PI / 2 + omega) / n
        self.m, self.P, self.t0, self.ecosw, self.esinw = float(m), float(P), float(t0), float(ecosw), float(esinw)
        self.I = 0.0 if math.isnan(I) else float(I)
        self.Omega = 0.0 if math.isnan(Omega) else float(Omega)
        self.a = 0.0 if math.isnan(a) else float(a)
        self.e, self.omega, self.tp = float(e), float(omega), float(tp)

    def row(self):
        return [getattr(self, f) for f in self.FIELDS]


# --------------------------------------------------------------------- hierarchy
def hierarchy(ic_vec):
    """Hierarchy matrix from [n, 1, 1, ..., 1] (setup_hierarchy.jl:9-145, the
    fully nested case that an integer H also selects)."""
    n = int(ic_vec[0])
    bins = [int(b) for b in ic_vec[1:]]
    if bins != [1] * (n - 1):
        raise NotImplementedError("only fully nested hierarchies [n, 1, ..., 1] are built here; "
                                  "pass the hierarchy matrix itself for other shapes")
    h = np.zeros((n, n), order="F")
    h[0, 0] = -1.0
    h[0, 1] = 1.0
    for r in range(1, n - 1):
        h[r, :r + 1] = -1.0
        h[r, r + 1] = 1.0
    h[n - 1, :] = -1.0
    return h


def _sigma_m(m, i, j, eps):  # init_nbody.jl:203-210
    s = 0.0
    n = len(m)
    for l in range(n):
        s += m[l] * (1.0 if eps[i, j] == eps[i, l] else 0.0)
    return s


def amatrix(eps, m):  # init_nbody.jl:176-184
    n = len(m)
    A = np.zeros((n, n), order="F")
    for i in range(n):
        for j in range(n):
            A[i, j] = (eps[i, j] * m[j]) / _sigma_m(m, i, j, eps)
    return A


def _invert(A0):
    """Gauss-Jordan elimination with partial pivoting (the reference calls
    LAPACK's inv; agreement is to rounding)."""
    n = A0.shape[0]
    a = [[float(A0[r, c]) for c in range(n)] for r in range(n)]
    b = [[1.0 if r == c else 0.0 for c in range(n)] for r in range(n)]
    for c in range(n):
        piv = c
        best = abs(a[c][c])
        for r in range(c + 1, n):
            if abs(a[r][c]) > best:
                best = abs(a[r][c])
                piv = r
        if piv != c:
            a[c], a[piv] = a[piv], a[c]
            b[c], b[piv] = b[piv], b[c]
        inv = 1.0 / a[c][c]
        a[c] = [q * inv for q in a[c]]
        b[c] = [q * inv for q in b[c]]
        for r in range(n):
            if r == c:
                continue
            f = a[r][c]
            if f == 0.0:
                continue
            a[r] = [a[r][q] - f * a[c][q] for q in range(n)]
            b[r] = [b[r][q] - f * b[c][q] for q in range(n)]
    out = np.zeros((n, n), order="F")
    for r in range(n):
        for c in range(n):
            out[r, c] = b[r][c]
    return out


# --------------------------------------------------------------------- Kepler
def ekepler(m, ecc):  # kepler.jl:1-41
    if m == 0.0:
        return 0.0
    pi2 = 2.0 * PI
    ms = math.fmod(m, pi2)
    if ms != 0.0 and ((ms < 0.0) != (pi2 < 0.0)):
        ms += pi2
    de0 = (ecc * 0.85) * _sign(ms)
    de1 = 2.0 * de0
    de2 = 3.0 * de0
    it = 0
    while True:
        de2 = de1
        de1 = de0
        f3 = ecc * math.cos(de0 + ms)
        f2 = ecc * math.sin(de0 + ms)
        de0 = (f2 - de1 * f3) / (1.0 - f3)
        it += 1
        if it >= 20 or de0 == de1 or de0 == de2:
            break
    return de0 + m


def _mm(A, B):  # 3x3 product, column-major lists
    return [(A[r] * B[3 * c] + A[r + 3] * B[1 + 3 * c]) + A[r + 6] * B[2 + 3 * c] for c in range(3) for r in range(3)]


def _mv(A, b):
    return [(A[r] * b[0] + A[r + 3] * b[1]) + A[r + 6] * b[2] for r in range(3)]


def kepler_init(time, mass, el):
    """Position, velocity and the 7x7 Jacobian of one Keplerian
    (kepler_init.jl:66-210)."""
    period, t0, ecosomega, esinomega, inc, capomega = [float(q) for q in el[:6]]
    n = (2.0 * PI) / period
    semi = _cbrt(((((GNEWT * mass) * (period * period)) / 4.0) / (PI * PI)))
    dsemidp = ((2.0 * THIRD) * semi) / period
    dsemidm = (THIRD * semi) / mass
    ecc = math.sqrt(esinomega * esinomega + ecosomega * ecosomega)
    deccdecos = ecosomega / ecc if ecc != 0.0 else 0.0
    deccdesin = esinomega / ecc if ecc != 0.0 else 0.0
    sqrt1mecc2 = math.sqrt(1.0 - ecc * ecc)
    den1 = (esinomega - ecosomega) - ecc
    if ecc == 0.0:
        tp = t0 - (3.0 * period) / 4.0
    else:
        tp = (t0 - ((sqrt1mecc2 / n) * ecosomega) / (1.0 - esinomega)) - \
            (2.0 / n) * math.atan2(math.sqrt(1.0 - ecc) * ((esinomega + ecosomega) + ecc), math.sqrt(1.0 + ecc) * den1)
    dtpdp = (tp - t0) / period
    fac = math.sqrt((1.0 - ecc) / (1.0 + ecc))
    den2 = 1.0 / (den1 * den1)
    theta = (fac * ((esinomega + ecosomega) + ecc)) / den1
    dthetadecc = (((((ecc + ecosomega) * (ecc + ecosomega)) + ((2.0 * (1.0 - ecc * ecc)) * esinomega)) -
                   esinomega * esinomega) / (sqrt1mecc2 * (1.0 + ecc))) * den2
    dthetadecos = ((2.0 * fac) * esinomega) * den2
    dthetadesin = ((-(2.0 * fac)) * (ecosomega + ecc)) * den2
    dtpdecc = ((((ecc / sqrt1mecc2) / n) * ecosomega) / (1.0 - esinomega)) - (((2.0 / n) / (1.0 + theta * theta)) * dthetadecc)
    dtpdecos = (dtpdecc * deccdecos - ((sqrt1mecc2 / n) / (1.0 - esinomega))) - \
        (((2.0 / n) / (1.0 + theta * theta)) * dthetadecos)
    dtpdesin = (dtpdecc * deccdesin - (((sqrt1mecc2 / n) * ecosomega) / ((1.0 - esinomega) * (1.0 - esinomega)))) - \
        (((2.0 / n) / (1.0 + theta * theta)) * dthetadesin)
    dtpdt0 = 1.0
    mean = n * (time - tp)
    dmdp = (-mean) / period
    dmdtp = -n
    ekep = ekepler(mean, ecc)
    cosekep, sinekep = math.cos(ekep), math.sin(ekep)
    r = semi * (1.0 - ecc * cosekep)
    denom = semi / r
    dekepdecos = (sinekep * denom) * deccdecos
    dekepdesin = (sinekep * denom) * deccdesin
    dekepdm = denom
    coscapomega, sincapomega = math.cos(capomega), math.sin(capomega)
    cosomega = ecosomega / ecc if ecc != 0.0 else 1.0
    sinomega = esinomega / ecc if ecc != 0.0 else 0.0
    cosinc, sininc = math.cos(inc), math.sin(inc)
    P1 = [cosomega, sinomega, 0.0, -sinomega, cosomega, 0.0, 0.0, 0.0, 1.0]
    P2 = [1.0, 0.0, 0.0, 0.0, cosinc, sininc, 0.0, -sininc, cosinc]
    P3 = [coscapomega, sincapomega, 0.0, -sincapomega, coscapomega, 0.0, 0.0, 0.0, 1.0]
    P32 = _mm(P3, P2)
    P321 = _mm(P32, P1)
    xplane = [semi * (cosekep - ecc), semi * (sqrt1mecc2 * sinekep), semi * 0.0]
    vplane = [-sinekep, sqrt1mecc2 * cosekep, 0.0]
    x = _mv(P321, xplane)
    dxda = [q / semi for q in x]
    dxdekep = _mv([q * semi for q in P321], vplane)
    dxdecc = _mv([((-q) * semi) / ecc for q in P321], [cosekep, sinekep / sqrt1mecc2, 0.0])
    Ps = [q / ecc for q in P32]
    tmp = _mv(Ps, xplane)
    dxdecos = [dxdecc[q] * deccdecos + tmp[q] for q in range(3)]
    tmp = _mv(Ps, [-xplane[1], xplane[0], 0.0])
    dxdesin = [dxdecc[q] * deccdesin + tmp[q] for q in range(3)]
    Dinc = [0.0, 0.0, 0.0, 0.0, -sininc, cosinc, 0.0, -cosinc, -sininc]
    Dcom = [-sincapomega, coscapomega, 0.0, -coscapomega, -sincapomega, 0.0, 0.0, 0.0, 0.0]
    Ainc = _mm(_mm(P3, Dinc), P1)
    Acom = _mm(_mm(Dcom, P2), P1)
    dxdinc = _mv(Ainc, xplane)
    dxdcom = _mv(Acom, xplane)
    nsd = (n * semi) * denom
    Pn = [((q * n) * semi) * denom for q in P321]
    v = _mv(Pn, vplane)
    tmp2 = _mv(Pn, [-cosekep, (-sqrt1mecc2) * sinekep, 0.0])
    dvda = [q / semi for q in v]
    dvdp = [(-q) / period for q in v]
    dvdekep = [((((-v[q]) * ecc) * sinekep) * denom) + tmp2[q] for q in range(3)]
    tmp = _mv(Pn, [0.0, ((-ecc) / sqrt1mecc2) * cosekep, 0.0])
    dvdecc = [(((-v[q]) / ecc) + ((v[q] * cosekep) * denom)) + tmp[q] for q in range(3)]
    Ps = [(((q * n) * semi) * denom) / ecc for q in P32]
    tmp = _mv(Ps, vplane)
    dvdecos = [dvdecc[q] * deccdecos + tmp[q] for q in range(3)]
    tmp = _mv(Ps, [-vplane[1], vplane[0], 0.0])
    dvdesin = [dvdecc[q] * deccdesin + tmp[q] for q in range(3)]
    w = [q * nsd for q in vplane]
    dvdinc = _mv(Ainc, w)
    dvdcom = _mv(Acom, w)
    jac = np.zeros((7, 7), order="F")
    for q in range(3):
        jac[q, 0] = dxda[q] * dsemidp + ((dxdekep[q] * dekepdm) * (dmdp + dmdtp * dtpdp))
        jac[q, 1] = ((dxdekep[q] * dekepdm) * dmdtp) * dtpdt0
        jac[q, 2] = dxdecos[q] + dxdekep[q] * ((dekepdm * dmdtp) * dtpdecos + dekepdecos) if ecc != 0.0 else 0.0
        jac[q, 3] = dxdesin[q] + dxdekep[q] * ((dekepdm * dmdtp) * dtpdesin + dekepdesin) if ecc != 0.0 else 0.0
        jac[q, 4] = dxdinc[q]
        jac[q, 5] = dxdcom[q]
        jac[q, 6] = dxda[q] * dsemidm
        jac[3 + q, 0] = (dvdp[q] + dvda[q] * dsemidp) + ((dvdekep[q] * dekepdm) * (dmdp + dmdtp * dtpdp))
        jac[3 + q, 1] = ((dvdekep[q] * dekepdm) * dmdtp) * dtpdt0
        jac[3 + q, 2] = dvdecos[q] + dvdekep[q] * ((dekepdm * dmdtp) * dtpdecos + dekepdecos) if ecc != 0.0 else 0.0
        jac[3 + q, 3] = dvdesin[q] + dvdekep[q] * ((dekepdm * dmdtp) * dtpdesin + dekepdesin) if ecc != 0.0 else 0.0
        jac[3 + q, 4] = dvdinc[q]
        jac[3 + q, 5] = dvdcom[q]
        jac[3 + q, 6] = dvda[q] * dsemidm
    jac[6, 6] = 1.0
    return np.array(x), np.array(v), jac


# --------------------------------------------------------------------- ICs
class ElementsIC:
    """Initial conditions from a hierarchy and orbital elements
    (InitialConditions.jl:142-241).  elements: n x 7 rows
    (m, P, t0, ecosw, esinw, I, Omega); H: int n (fully nested), a hierarchy
    vector [n, 1, ..., 1], or an n x n hierarchy matrix."""

    def __init__(self, t0, H, *elems, der=True):
        if len(elems) == 1 and isinstance(elems[0], str):
            elements = np.loadtxt(elems[0], delimiter=",", comments="#", ndmin=2)
        elif len(elems) == 1 and not isinstance(elems[0], Elements) and np.ndim(elems[0]) == 2:
            elements = np.asarray(elems[0], dtype=np.float64)
        else:
            es = elems[0] if len(elems) == 1 and isinstance(elems[0], (list, tuple)) else elems
            elements = np.array([e.row() for e in es], dtype=np.float64)
        if isinstance(H, (int, np.integer)):
            Hm = hierarchy([int(H)] + [1] * (int(H) - 1))
        elif np.ndim(H) == 1:
            Hm = hierarchy(list(H))
        else:
            Hm = np.asfortranarray(np.asarray(H, dtype=np.float64))
        n = Hm.shape[0]
        if elements.shape[0] < n:
            elements = np.vstack([elements, np.zeros((n - elements.shape[0], 7))])
        self.elements = np.asfortranarray(elements[:n, :7].copy())
        self.H = Hm
        self.nbody = n
        self.m = self.elements[:, 0].copy()
        self.amat = amatrix(self.H, self.m)
        self.t0 = float(t0)
        self.der = der

    def init_nbody(self):
        return init_nbody(self)


class CartesianIC:
    """Initial conditions from Cartesian coordinates (InitialConditions.jl:259-282).
    coords: N x 7 rows (m, x, y, z, vx, vy, vz) or a file name."""

    def __init__(self, t0, N, coords):
        if isinstance(coords, str):
            coords = np.loadtxt(coords, delimiter=",", ndmin=2)
        coords = np.asarray(coords, dtype=np.float64)
        self.m = coords[:N, 0].copy()
        self.x = np.asfortranarray(coords[:N, 1:4].T.copy())
        self.v = np.asfortranarray(coords[:N, 4:7].T.copy())
        self.nbody = int(N)
        self.t0 = float(t0)

    def init_nbody(self):
        return init_nbody(self)


def init_nbody(ic):
    """x (3 x n), v (3 x n), jac_init (7n x 7n): init_nbody.jl:13-36 with
    kepcalc (:50-105) and d_dm (:120-162)."""
    if isinstance(ic, CartesianIC):
        n = ic.nbody
        return ic.x.copy(order="F"), ic.v.copy(order="F"), np.asfortranarray(np.eye(7 * n))
    n = ic.nbody
    eps = ic.H
    m = ic.m
    A = ic.amat
    sevn = 7 * n
    rk = np.zeros((n, 3), order="F")
    rdk = np.zeros((n, 3), order="F")
    jk = np.zeros((6 * n, sevn), order="F")
    i, b = 0, 0
    while i < n - 1:
        mu = 0.0
        for l in range(n):
            if eps[i, l] != 0.0:
                mu += m[l]
        if eps[i, 0] == 0.0:
            b += 1
        el = ic.elements[i + 1 + b, 1:7]
        r, rd, jac21 = kepler_init(ic.t0, mu, el)
        rk[i, :] = r
        rdk[i, :] = rd
        for j in range(6):
            for k in range(6):
                jk[6 * i + j, 7 * (i + 1) + k] = jac21[j, k]
        for j in range(n):
            if eps[i, j] != 0.0:
                for k in range(6):
                    jk[6 * i + k, 7 * j + 6] = jac21[k, 6]
        if b > 0:
            b -= 2
        elif b < 0:
            b = 0
        i += 1
    Ainv = _invert(A)
    jac_init = np.zeros((sevn, sevn), order="F")
    dAdm = np.zeros((n, n, n), order="F")
    for k in range(n):
        for ii in range(n):
            for j in range(n):
                eij = eps[ii, j]
                sm = _sigma_m(m, ii, j, eps)
                t1 = ((1.0 if k == j else 0.0) * eij) / sm
                t2 = (((1.0 if eij == eps[ii, k] else 0.0) * eij) * m[j]) / (sm * sm)
                dAdm[ii, j, k] = t1 - t2
    dAinvdm = np.zeros((n, n, n), order="F")
    for k in range(n):
        tmpm = np.zeros((n, n))
        for r in range(n):
            for c in range(n):
                s = 0.0
                for q in range(n):
                    s += (-Ainv[r, q]) * dAdm[q, c, k]
                tmpm[r, c] = s
        for r in range(n):
            for c in range(n):
                s = 0.0
                for q in range(n):
                    s += tmpm[r, q] * Ainv[q, c]
                dAinvdm[r, c, k] = s
    for ii in range(n):
        for k in range(n):
            # row-wise: the same multiply and add per element, k ascending
            aik = Ainv[ii, k]
            jac_init[7 * ii:7 * ii + 6, :] += aik * jk[6 * k:6 * k + 6, :]
        for k in range(n):
            for j in range(3):
                sx = 0.0
                sv = 0.0
                for q in range(n):
                    sx += dAinvdm[ii, q, k] * rk[q, j]
                    sv += dAinvdm[ii, q, k] * rdk[q, j]
                jac_init[7 * ii + j, 7 * k + 6] += sx
                jac_init[7 * ii + 3 + j, 7 * k + 6] += sv
        jac_init[7 * ii + 6, 7 * ii + 6] = 1.0
    x = np.zeros((3, n), order="F")
    v = np.zeros((3, n), order="F")
    for ii in range(n):
        for q in range(3):
            sx = 0.0
            sv = 0.0
            for k in range(n):
                sx += Ainv[ii, k] * rk[k, q]
                sv += Ainv[ii, k] * rdk[k, q]
            x[q, ii] = sx
            v[q, ii] = sv
    return x, v, jac_init
