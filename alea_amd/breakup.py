"""Optimized breakup of a radial pair function over a periodic 2-D cell
(Natoli & Ceperley, J. Comput. Phys. 117, 171 (1995)): the set-up tables of
the 2-D Ewald sum and of the image-summed Padé Jastrow, built on the host.

A radial f(r) is split as f = f_short + f_long.  Inside the cutoff r_c (the
Wigner-Seitz inradius, so the short part sums by minimum image) f_long is an
expansion in the locally piecewise quintic Hermite (LPQHI) basis on a uniform
knot mesh; at and beyond r_c it is f itself, so f_short has compact support.
Its coefficients t_n make the Fourier transform of f_long negligible above
k_cut, so the reciprocal sum over |k| <= k_cut carries the rest.  That is the
weighted least-squares problem

    minimise  sum_{|k| > k_cut} w_k |X(k) - sum_n t_n c_n(k)|^2,
    X(k) = -2 pi int_{r_c}^inf r J0(kr) f(r) dr,
    c_n(k) = 2 pi int_0^{r_c} r J0(kr) h_n(r) dr,

over the lattice shells up to k_continuum (weights = their multiplicities)
and, beyond, log-spaced samples of the continuum density of states A k dk /
(2 pi) up to k_max; the last knot's three coefficients are the value, slope
and curvature of f at r_c (assigned, not fitted).  The reference builds the
same tables in numerics::build_long_range_breakup
(src/numerics/long_range_breakup.cpp:96-354), numerics::build_ewald
(src/numerics/ewald_breakup.cpp:205-276) and PeriodicPairSum
(src/wavefunction/periodic_pair_sum.cpp:16-64); the formats here are the ones
the C-ABI takes (alea_add_ewald, alea_add_pade_jastrow).

How each piece is computed here:
  * c_n(k): Gauss-Legendre on every knot panel, subdivided so that each
    sub-panel spans at most about one radian of J0's phase (the LPQHI
    functions are polynomial on a panel, kinked at the knots).
  * Coulomb, X(k) = -(2 pi / (eps k)) int_{k r_c}^inf J0(t) dt, from the
    integral of J0 in closed form.
  * A rational u = P/Q, f = u - u_inf, split by partial fractions over the
    simple roots rho_i = -c_i of Q, f = sum_i A_i / (r + c_i); for Re c > 0
        2 pi int_0^inf r J0(kr) / (r + c) dr = 2 pi c G(kc),
        G(x) = int_0^inf e^{-xt} (1 - 1/sqrt(1 + t^2)) dt,
    from int_0^inf J0(kr) e^{-sr} dr = 1/sqrt(k^2 + s^2); G is a positive,
    smooth Laplace integral (no cancellation at large x), and the part
    inside r_c is subtracted by quadrature.
  * The least-squares solve is a QR factorization of the weighted design
    matrix (never the normal equations, which square its condition number).
"""
from __future__ import annotations

import functools
import math

import numpy as np
from scipy import linalg as sla
from scipy import special

TWO_PI = 2.0 * math.pi

# Gauss-Legendre nodes and weights on [0, 1], 20 points
_GX, _GW = np.polynomial.legendre.leggauss(20)
_GX = 0.5 * (_GX + 1.0)
_GW = 0.5 * _GW


# ── cell ────────────────────────────────────────────────────────────────────
def reduced_cell(lattice):
    """The reduced basis the cell works in (axis-aligned rows in axis order,
    else Gauss-reduced; simulation_cell.cpp:16-36,100-132), its reciprocal
    basis b_i . a_j = 2 pi delta_ij, the area and the Wigner-Seitz inradius
    (half the shortest nonzero lattice vector, simulation_cell.cpp:202-224)."""
    from .config import cell_vectors
    a = cell_vectors(lattice)
    det = a[0] * a[3] - a[1] * a[2]
    if det == 0.0:
        raise ValueError("breakup: singular cell")
    inv = 1.0 / det
    ainv = [a[3] * inv, -a[2] * inv, -a[1] * inv, a[0] * inv]
    b = [TWO_PI * x for x in ainv]
    shortest2 = 0.0
    for n0 in range(-2, 3):
        for n1 in range(-2, 3):
            if n0 == 0 and n1 == 0:
                continue
            s0 = n0 * a[0] + n1 * a[2]
            s1 = n0 * a[1] + n1 * a[3]
            r2 = s0 * s0 + s1 * s1
            if shortest2 == 0.0 or r2 < shortest2:
                shortest2 = r2
    return dict(a=a, b=b, area=abs(det), r_ws=0.5 * math.sqrt(shortest2))


# ── LPQHI basis ─────────────────────────────────────────────────────────────
# The quintic Hermite shape functions on x in [0, 1]: S_a has unit a-th
# derivative at x = 0 (value for a = 0) and zero other derivatives up to the
# second at both ends.  Coefficients of x^0..x^5.
_S = np.array([[1.0, 0.0, 0.0, -10.0, 15.0, -6.0],
               [0.0, 1.0, 0.0, -6.0, 8.0, -3.0],
               [0.0, 0.0, 0.5, -1.5, 1.5, -0.5]])
# reflection about the knot: value and curvature even, slope odd
_MIRROR = np.array([1.0, -1.0, 1.0])


def _shape(x):
    """S_a(x) for a = 0, 1, 2: [3, len(x)]."""
    p = np.vstack([x ** m for m in range(6)])
    return _S @ p


def _panel_counts(ks, width):
    return np.clip(1 + (np.asarray(ks) * width).astype(np.int64), 1, 4000)


def basis_transforms(n_knots, r_cut, ks):
    """c_n(k) = 2 pi int r J0(kr) h_n(r) dr for every LPQHI function n = 3 i + a
    (a = value, slope, curvature at knot i, scaled by delta^a) and every k:
    [len(ks), 3 n_knots]."""
    ks = np.asarray(ks, dtype=float)
    nk = n_knots
    delta = r_cut / (nk - 1)
    out = np.zeros((len(ks), 3 * nk))
    scale = delta ** np.arange(3)
    m_all = _panel_counts(ks, delta)
    panels = np.arange(nk - 1, dtype=float)
    for m in np.unique(m_all):
        sel = np.nonzero(m_all == m)[0]
        # m equal sub-panels of each knot panel, 20 points each, in x
        x = ((np.arange(m)[:, None] + _GX[None, :]) / m).ravel()
        wq = np.tile(_GW, m) / m * delta            # dr = delta dx
        right = _shape(x) * scale[:, None]         # knot j's functions on panel j
        left = _shape(1.0 - x) * (scale * _MIRROR)[:, None]   # knot j+1's on panel j
        r = (panels[:, None] + x[None, :]) * delta   # [panel, q]
        for i0 in range(0, len(sel), 64):
            kk = ks[sel[i0:i0 + 64]]
            g = r[None, :, :] * special.j0(kk[:, None, None] * r[None, :, :]) * wq[None, None, :]
            cr = np.einsum("kpq,aq->kpa", g, right)
            cl = np.einsum("kpq,aq->kpa", g, left)
            o = np.zeros((len(kk), nk, 3))
            o[:, :-1, :] += cr
            o[:, 1:, :] += cl
            out[sel[i0:i0 + 64]] = o.reshape(len(kk), 3 * nk)
    return TWO_PI * out


def _inner_transform(f, r_cut, n_panels, ks):
    """2 pi int_0^{r_c} r J0(kr) f(r) dr for a smooth f, on n_panels equal
    panels subdivided like the basis transforms."""
    ks = np.asarray(ks, dtype=float)
    width = r_cut / n_panels
    out = np.zeros(len(ks))
    m_all = _panel_counts(ks, width)
    for m in np.unique(m_all):
        sel = np.nonzero(m_all == m)[0]
        x = ((np.arange(n_panels * m)[:, None] + _GX[None, :]) / m).ravel()
        r = x * width
        wq = np.tile(_GW, n_panels * m) / m * width
        fr = r * f(r) * wq
        for i0 in range(0, len(sel), 64):
            kk = ks[sel[i0:i0 + 64]]
            out[sel[i0:i0 + 64]] = special.j0(kk[:, None] * r[None, :]) @ fr
    return TWO_PI * out


def integral_j0(x):
    """int_0^x J0(t) dt = x J0(x) + (pi x / 2) (J1(x) H0(x) - J0(x) H1(x)),
    H the Struve functions (Abramowitz & Stegun 11.1.7)."""
    x = np.asarray(x, dtype=float)
    return x * special.j0(x) + 0.5 * math.pi * x * (special.j1(x) * special.struve(0, x)
                                                    - special.j0(x) * special.struve(1, x))


def _laplace_g(x):
    """G(x) = int_0^inf e^{-xt} (1 - 1/sqrt(1 + t^2)) dt for complex x with
    Re x > 0, by Gauss-Legendre on geometric panels in s = x t (the integrand
    turns over at s ~ |x| and decays as e^{-s})."""
    x = np.atleast_1d(np.asarray(x, dtype=complex))
    out = np.zeros(x.shape, dtype=complex)
    for i, xv in enumerate(x):
        lo = 1e-3 * min(abs(xv), 1.0)
        edges = [0.0, lo]
        while edges[-1] < 80.0:
            edges.append(2.0 * edges[-1])
        e = np.array(edges)
        a, b = e[:-1], e[1:]
        s = (a[:, None] + (b - a)[:, None] * _GX[None, :]).ravel()
        w = ((b - a)[:, None] * _GW[None, :]).ravel()
        # along the ray s = x t (t = s / x): e^{-s} (1 - (1 + t^2)^{-1/2}) dt
        t = s / xv
        q = t * t
        root = np.sqrt(1.0 + q)
        h = q / (root * (1.0 + root))            # 1 - 1/sqrt(1+q), no cancellation
        out[i] = np.sum(w * np.exp(-s) * h) / xv
    return out


def _laplace_g2(x):
    """int_0^inf e^{-xt} (1 - 1/sqrt(1 + t^2)) (1 - x t) dt = G(x) + x G'(x)
    for complex x with Re x > 0, on the quadrature of _laplace_g: the
    transform of 1/(r + c)^2 below is minus the c-derivative of that of
    1/(r + c)."""
    x = np.atleast_1d(np.asarray(x, dtype=complex))
    out = np.zeros(x.shape, dtype=complex)
    for i, xv in enumerate(x):
        lo = 1e-3 * min(abs(xv), 1.0)
        edges = [0.0, lo]
        while edges[-1] < 80.0:
            edges.append(2.0 * edges[-1])
        e = np.array(edges)
        a, b = e[:-1], e[1:]
        s = (a[:, None] + (b - a)[:, None] * _GX[None, :]).ravel()
        w = ((b - a)[:, None] * _GW[None, :]).ravel()
        t = s / xv
        q = t * t
        root = np.sqrt(1.0 + q)
        h = q / (root * (1.0 + root))
        out[i] = np.sum(w * np.exp(-s) * h * (1.0 - s)) / xv
    return out


# ── the breakup ─────────────────────────────────────────────────────────────
def _fit_rows(cell, k_cut, k_continuum, k_max, n_continuum):
    """|k| of the fitted rows and their weights: the lattice shells in
    (k_cut, k_continuum] with their multiplicities, then n_continuum
    log-spaced samples of the continuum up to k_max, each weighted by the
    number of lattice points it stands for, A k dk / (2 pi)."""
    b1x, b1y, b2x, b2y = cell["b"]
    b_min = min(math.hypot(b1x, b1y), math.hypot(b2x, b2y))
    nmax = int(math.ceil(k_continuum / b_min)) + 1
    i, j = np.meshgrid(np.arange(-nmax, nmax + 1), np.arange(-nmax, nmax + 1), indexing="ij")
    kx = i * b1x + j * b2x
    ky = i * b1y + j * b2y
    kk = np.hypot(kx, ky)[(i != 0) | (j != 0)]
    kk = np.sort(kk[(kk > k_cut) & (kk <= k_continuum)])
    ks, ws = [], []
    a = 0
    while a < len(kk):
        z = a
        while z < len(kk) and kk[z] - kk[a] < 1e-9 * b_min:
            z += 1
        ks.append(kk[a])
        ws.append(float(z - a))
        a = z
    lo, hi = math.log(k_continuum), math.log(k_max)
    dl = (hi - lo) / n_continuum
    for n in range(n_continuum):
        k = math.exp(lo + (n + 0.5) * dl)
        ks.append(k)
        ws.append(cell["area"] * k * (k * dl) / TWO_PI)
    return np.array(ks), np.array(ws), len(ks) - n_continuum


def _retained_k(cell, k_cut):
    """The half space of reciprocal vectors with 0 < |k| <= k_cut (one of each
    +-k pair: n1 > 0, or n1 = 0 and n2 > 0), in the order the C-ABI's tables
    list them (n1 outer, n2 inner, both ascending)."""
    b1x, b1y, b2x, b2y = cell["b"]
    b_min = min(math.hypot(b1x, b1y), math.hypot(b2x, b2y))
    nmax = int(math.ceil(k_cut / b_min)) + 1
    out = []
    for i in range(0, nmax + 1):
        for j in range(-nmax, nmax + 1):
            if i == 0 and j <= 0:
                continue
            kx = i * b1x + j * b2x
            ky = i * b1y + j * b2y
            k = math.hypot(kx, ky)
            if k > k_cut or k == 0.0:
                continue
            out.append((i, j, kx, ky, k))
    return out


def breakup(cell, f, n_knots, target, k_cut, k_continuum, k_max, n_continuum):
    """The optimized breakup of f over the cell.

    f: (value, d1, d2) callables of r; target(ks) -> X(k).  Returns the
    coefficients t, the retained reciprocal vectors with weights
    f~_long(k) / A, and the largest |residual| of the fit over the lattice
    shells above k_cut (the part the truncated reciprocal sum drops)."""
    r_cut = cell["r_ws"]
    ks, ws, n_shell = _fit_rows(cell, k_cut, k_continuum, k_max, n_continuum)
    nb = 3 * n_knots
    n_free = nb - 3
    delta = r_cut / (n_knots - 1)
    t = np.zeros(nb)
    # f_long meets f at r_c with two derivatives (the knot's value, slope and
    # curvature coefficients)
    t[nb - 3] = f[0](r_cut)
    t[nb - 2] = f[1](r_cut)
    t[nb - 1] = f[2](r_cut)
    C = basis_transforms(n_knots, r_cut, ks)
    X = target(ks) - C[:, n_free:] @ t[n_free:]
    sw = np.sqrt(ws)
    q, rr = sla.qr(sw[:, None] * C[:, :n_free], mode="economic")
    t[:n_free] = sla.solve_triangular(rr, q.T @ (sw * X))
    resid = X - C[:, :n_free] @ t[:n_free]
    residual = float(np.max(np.abs(resid[:n_shell]))) if n_shell else 0.0
    kept = _retained_k(cell, k_cut)
    kr = np.array([k[4] for k in kept])
    f_long = -target(kr) + basis_transforms(n_knots, r_cut, kr) @ t
    return dict(t=t, n_knots=n_knots, r_cut=r_cut, delta=delta,
                n1=[k[0] for k in kept], n2=[k[1] for k in kept],
                kx=[k[2] for k in kept], ky=[k[3] for k in kept],
                w=list(f_long / cell["area"]), residual=residual)


def _spline(br):
    """LongRangeSpline::scaled (long_range_spline.h): t_n delta^a, 3 per knot."""
    a = np.arange(len(br["t"])) % 3
    return list(br["t"] * br["delta"] ** a)


def _cell_b_min(cell):
    b1x, b1y, b2x, b2y = cell["b"]
    return min(math.hypot(b1x, b1y), math.hypot(b2x, b2y))


# ── Ewald (Coulomb) ─────────────────────────────────────────────────────────
@functools.lru_cache(maxsize=256)
def _ewald(lattice: tuple, tol: float, epsilon: float):
    cell = reduced_cell(lattice)
    r_cut = cell["r_ws"]
    inv_eps = 1.0 / epsilon
    f = (lambda r: inv_eps / r, lambda r: -inv_eps / (r * r), lambda r: 2.0 * inv_eps / (r * r * r))

    def target(ks):
        # -f~(k) + 2 pi int_0^{r_c} r J0(kr) f dr = -(2 pi/(eps k)) (1 - int_0^{k r_c} J0)
        ks = np.asarray(ks, dtype=float)
        return -(TWO_PI * inv_eps / ks) * (1.0 - integral_j0(ks * r_cut))

    k_cut = 8.0 * _cell_b_min(cell)
    br = None
    # the fit's accuracy bottoms out near 20 knots and degrades beyond as the
    # design matrix's condition number grows: the first count that reaches
    # the tolerance is taken
    for n_knots in (12, 16, 20, 24):
        br = breakup(cell, f, n_knots, target, k_cut, 60.0 * math.pi / r_cut, 6000.0 / r_cut, 400)
        if br["residual"] < tol:
            break
    else:
        raise ValueError(f"Ewald breakup: residual {br['residual']:.3g} above the tolerance {tol:g}; "
                         "loosen the tolerance or use a larger cell")
    # (1/A) int f_short d^2r: the k = 0 component of the short-range half
    c0 = basis_transforms(br["n_knots"], r_cut, [0.0])[0]
    sr_k0 = (TWO_PI * r_cut * inv_eps - float(c0 @ br["t"])) / cell["area"]
    return dict(n1=br["n1"], n2=br["n2"], kx=br["kx"], ky=br["ky"], w=br["w"], spline=_spline(br),
                n_knots=br["n_knots"], r_cut=r_cut, inv_delta=1.0 / br["delta"], v_long_zero=float(br["t"][0]),
                sr_k0=sr_k0, k_cut=k_cut, area=cell["area"], tol=tol, residual=br["residual"])


def ewald_tables(lattice, tol=1e-6, epsilon=1.0) -> dict:
    """EwaldSetup of a periodic cell for the Coulomb kernel 1/(eps r)
    (numerics::build_ewald, src/numerics/ewald_breakup.cpp:205-276): the
    half-space k set with weights, the long-range spline, r_cut, 1/delta,
    v_long_zero = f_long(0) and sr_k0 = (1/A) int f_short d^2r."""
    return dict(_ewald(tuple(float(x) for x in lattice), float(tol), float(epsilon)))


# ── image-summed Padé ───────────────────────────────────────────────────────
def pade_limit(a, b) -> float:
    """u_inf = lim u of u = (sum a_k r^k) / (1 + sum b_k r^k): a_N / b_M for
    N = M, 0 for N < M (pair_radial.h:50-55); refuses N > M."""
    a, b = _trim(a), _trim(b)
    if len(a) > len(b):
        raise ValueError(f"padé jastrow: numerator of order {len(a)} over denominator of order {len(b)} "
                         "grows without bound, so its sum over periodic images does not exist")
    return a[-1] / b[-1] if a and len(a) == len(b) else 0.0


def _trim(c):
    c = [float(x) for x in c]
    while c and c[-1] == 0.0:
        c.pop()
    return c


def _rational(a, b):
    """u, u', u'' of the rational at r (vectorized)."""
    pa = np.polynomial.Polynomial([0.0] + list(a))
    pq = np.polynomial.Polynomial([1.0] + list(b))
    dpa, d2pa, dpq, d2pq = pa.deriv(), pa.deriv(2), pq.deriv(), pq.deriv(2)

    def ev(r):
        p, q = pa(r), pq(r)
        u = p / q
        du = (dpa(r) - u * dpq(r)) / q
        d2u = (d2pa(r) - 2.0 * du * dpq(r) - u * d2pq(r)) / q
        return u, du, d2u
    return ev


@functools.lru_cache(maxsize=256)
def _pade(lattice: tuple, a: tuple, b: tuple):
    cell = reduced_cell(lattice)
    r_cut = cell["r_ws"]
    a, b = _trim(a), _trim(b)
    u_inf = pade_limit(a, b)
    if not b:
        raise ValueError("padé jastrow: a periodic image sum needs a denominator")
    # f = u - u_inf = R / Q with deg R < deg Q, by partial fractions over the
    # roots rho = -c of Q: a simple root gives A / (r + c), A = R(rho)/Q'(rho);
    # a double one (the squared denominators of a Padé's parameter
    # derivatives, pade_jastrow.cpp:227-244) A / (r + c) + B / (r + c)^2 with
    # Q = (r - rho)^2 S: B = R(rho)/S(rho), A = (R'(rho) S(rho) - R(rho) S'(rho)) / S(rho)^2
    pq = np.polynomial.Polynomial([1.0] + b)
    pr = np.polynomial.Polynomial([0.0] + a) - u_inf * pq
    terms = []   # (c, A, B)
    for rho, mult in _root_groups(pq):
        if rho.real >= 0.0:
            raise ValueError("padé jastrow: the denominator has a root with a non-negative real part; the "
                             "image-sum split needs every root in the left half plane")
        if mult == 1:
            terms.append((-rho, pr(rho) / pq.deriv()(rho), 0.0))
        else:
            sq, rem = divmod(pq, np.polynomial.Polynomial([rho * rho, -2.0 * rho, 1.0]))
            sv, dsv = sq(rho), sq.deriv()(rho)
            terms.append((-rho, (pr.deriv()(rho) * sv - pr(rho) * dsv) / (sv * sv), pr(rho) / sv))
    ev = _rational(a, b)
    f = (lambda r: ev(r)[0] - u_inf, lambda r: ev(r)[1], lambda r: ev(r)[2])

    def target(ks):
        # X(k) = -2 pi int_{r_c}^inf r J0(kr) f dr = -(full transform - inside r_c);
        # 2 pi int_0^inf r J0(kr)/(r + c) dr = 2 pi c G(kc), and for 1/(r + c)^2
        # minus its c-derivative, -2 pi (G(kc) + kc G'(kc))
        ks = np.asarray(ks, dtype=float)
        full = np.zeros(len(ks))
        for ci, ai, bi in terms:
            full += np.real(TWO_PI * ai * ci * _laplace_g(ks * ci))
            if bi != 0.0:
                full += np.real(-TWO_PI * bi * _laplace_g2(ks * ci))
        inner = _inner_transform(lambda r: ev(r)[0] - u_inf, r_cut, 19, ks)
        return -(full - inner)

    b_min = _cell_b_min(cell)
    # a Jastrow's transform falls much faster than Coulomb's, so the fit needs
    # a far shorter k range; the retained k reach 5 |b_min| (periodic_pair_sum.cpp:
    # 30-49 measures that choice)
    br = breakup(cell, f, 20, target, 5.0 * b_min, 20.0 * math.pi / r_cut, 300.0 / r_cut, 150)
    return dict(n1=br["n1"], n2=br["n2"], kx=br["kx"], ky=br["ky"], w=br["w"], spline=_spline(br),
                n_knots=br["n_knots"], r_cut=r_cut, inv_delta=1.0 / br["delta"], u_inf=u_inf,
                residual=br["residual"])


def _root_groups(pq):
    """The roots of the polynomial pq with their multiplicities (1 or 2):
    roots closer than 1e-6 of their size pair up, polished as a root of pq'
    by Newton; a triple root is refused."""
    roots = list(pq.roots())
    out = []
    while roots:
        r0 = roots.pop(0)
        near = [i for i, r in enumerate(roots) if abs(r - r0) < 1e-6 * max(abs(r0), 1e-300)]
        if not near:
            out.append((r0, 1))
            continue
        if len(near) > 1:
            raise ValueError("padé jastrow: the denominator has a root of multiplicity above two")
        r1 = roots.pop(near[0])
        rho = 0.5 * (r0 + r1)
        d1, d2 = pq.deriv(), pq.deriv(2)
        for _ in range(3):
            rho = rho - d1(rho) / d2(rho)
        out.append((rho, 2))
    return out


def pade_parameter_tables(lattice, a, b, q) -> dict:
    """The image-sum tables of du/dp for parameter q of the Padé u = P/Q,
    P = sum a_k r^k, Q = 1 + sum b_k r^k (q < len(a): a_{q+1}; else
    b_{q-len(a)+1}), as PadeJastrow builds them for PeriodicPairSum
    (pade_jastrow.cpp:227-245): du/da_k = r^k / Q, du/db_k = -P r^k / Q^2,
    each broken up like u itself on the same k set (its weights, spline and
    limit)."""
    a = [float(x) for x in a]
    b = [float(x) for x in b]
    if q < len(a):
        num = [0.0] * q + [1.0]
        den = b
    else:
        k = q - len(a) + 1
        num = [0.0] * (len(a) + k)
        for j in range(len(a)):
            num[j + k] = -a[j]
        pq = np.polynomial.Polynomial([1.0] + b)
        den = list((pq * pq).coef[1:])
    return pade_tables(lattice, num, den)


def pade_tables(lattice, a, b) -> dict:
    """PeriodicPairSum's tables for the Padé u = (sum a_k r^k)/(1 + sum b_k
    r^k) (periodic_pair_sum.cpp:16-64, pade_jastrow.cpp:203-245): the
    half-space k set (|k| <= 5 |b_min|) with the weights u~_long(k)/A, the
    long-range spline, r_cut, 1/delta and u_inf."""
    return dict(_pade(tuple(float(x) for x in lattice), tuple(float(x) for x in a), tuple(float(x) for x in b)))
