"""Host-side system description: cell, particle groups, trial-wavefunction
factors and Hamiltonian terms, built the way the reference's builders build
them for the 2-D electron-gas inputs.

What is mirrored (reference file:line):
  cell from rs            src/qmc/app.cpp:274-341   (area = N*pi*rs^2)
  Kato cusp a1            src/builders/jastrow.cpp:48,159-175
  image-summed Padé       src/builders/jastrow.cpp:366-368
  closed-shell plane waves src/wavefunction/orbital_set.h:566-691
  twisted plane waves     src/wavefunction/complex_orbital_set.h:278-341
  Ewald                   src/builders/ewald.cpp:40-66 (tolerance 1e-6)

The Natoli-Ceperley breakup tables (Ewald and the image-summed Padé) are
built on the host by alea_amd.breakup for any cell; a caller may pass tables
of its own instead (`tables=`: {"ewald": ..., "pade": {pade_key: ...}}), as a
drop-in caller passes the tables its builders computed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import breakup


def pade_key(a, b) -> str:
    return "a=" + ",".join(f"{x:.15g}" for x in a) + ";b=" + ",".join(f"{x:.15g}" for x in b)


def ewald_table(lattice, tables=None, tol=1e-6):
    """The Ewald set-up of a cell: from `tables` when given, else the
    product's own breakup at tolerance tol (the builders' default 1e-6)."""
    if tables is not None:
        return tables["ewald"]
    return breakup.ewald_tables(lattice, tol)


def pade_table(lattice, a, b, tables=None):
    """The image-sum split of the Padé (a, b) in a cell: from `tables` when
    given, else the product's own breakup."""
    if tables is not None:
        return tables["pade"][pade_key(a, b)]
    return breakup.pade_tables(lattice, a, b)


def ewald_table_tightest(lattice, tables=None, tols=(1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 3e-2)):
    """The Ewald set-up at the tightest tolerance of `tols` the cell's
    breakup reaches (an elongated cell, whose Wigner-Seitz inradius leaves
    the fit little room, does not reach 1e-6)."""
    if tables is not None:
        return tables["ewald"]
    last = None
    for tol in tols:
        try:
            return breakup.ewald_tables(lattice, tol)
        except ValueError as e:
            last = e
    raise last


def square_lattice(n_particles: int, rs: float):
    L = math.sqrt(math.pi * n_particles) * rs
    return [L, 0.0, 0.0, L]


def hexagonal_lattice(n_particles: int, rs: float):
    area = n_particles * math.pi * rs * rs
    L = math.sqrt(2.0 * area / math.sqrt(3.0))
    return [L, 0.0, 0.5 * L, L * math.sqrt(3.0) / 2.0]


def reciprocal(lattice):
    """2*pi*inverse-transpose of the (reduced) basis, as SimulationCell does
    for axis-aligned cells (simulation_cell.cpp:146-201)."""
    a = np.array(lattice, dtype=float).reshape(2, 2)
    det = a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]
    ainv = np.array([[a[1, 1], -a[1, 0]], [-a[0, 1], a[0, 0]]]) / det
    return 2.0 * math.pi * ainv


def plane_wave_shells(lattice, n_orb: int):
    """Closed-shell real plane waves (orbital_set.h:566-691): returns k [n,2]
    and is_sin [n]; refuses an open shell."""
    b = reciprocal(lattice)
    area = abs(b[0, 0] * b[1, 1] - b[0, 1] * b[1, 0])
    h_min = area / max(math.hypot(*b[0]), math.hypot(*b[1]))
    M = 4
    while M <= 4096:
        r_safe2 = (M * h_min) ** 2 * (1.0 + 1e-12)
        kv = []
        for i in range(-M, M + 1):
            for j in range(-M, M + 1):
                k = i * b[0] + j * b[1]
                k2 = k[0] * k[0] + k[1] * k[1]
                if k2 <= r_safe2:
                    kv.append((k2, k[0], k[1]))
        kv.sort()
        ks, sins, closed = [], [], []
        s = 0
        while s < len(kv):
            e = s + 1
            while e < len(kv) and abs(kv[e][0] - kv[s][0]) <= 1e-9 * (1.0 + kv[s][0]):
                e += 1
            if kv[s][0] <= 1e-12 * (1.0 + kv[s][0]):
                ks.append((0.0, 0.0))
                sins.append(0)
            else:
                for q in range(s, e):
                    kx, ky = kv[q][1], kv[q][2]
                    positive = kx > 1e-12 or (not kx < -1e-12 and ky > 1e-12)
                    if not positive:
                        continue
                    for which in (0, 1):
                        ks.append((kx, ky))
                        sins.append(which)
            closed.append(len(sins))
            s = e
        if len(sins) >= n_orb:
            if n_orb not in closed:
                raise ValueError(f"plane-wave orbitals: {n_orb} leaves a shell open; closed: {closed[:12]}")
            return np.array(ks[:n_orb]), np.array(sins[:n_orb], dtype=np.uint8)
        M *= 2
    raise RuntimeError("shell enumeration did not converge")


def twisted_wavevectors(lattice, n: int, twist):
    """q = k + theta for the n lowest |k + theta| (complex_orbital_set.h:278-341);
    twist is in units of the reciprocal vectors as given."""
    b = reciprocal(lattice)
    theta = twist[0] * b[0] + twist[1] * b[1]
    tn = math.hypot(*theta)
    area = abs(b[0, 0] * b[1, 1] - b[0, 1] * b[1, 0])
    h_min = area / max(math.hypot(*b[0]), math.hypot(*b[1]))
    M = 4
    while M <= 4096:
        qs = []
        for i in range(-M, M + 1):
            for j in range(-M, M + 1):
                q = i * b[0] + theta + j * b[1]
                qs.append((q[0] * q[0] + q[1] * q[1], q[0], q[1]))
        if len(qs) > n:
            qs.sort(key=lambda t: t[0])
            r_safe = M * h_min - tn
            if r_safe > 0 and qs[n][0] <= r_safe * r_safe:
                if n > 0 and abs(qs[n][0] - qs[n - 1][0]) <= 1e-9 * (1.0 + qs[n][0]):
                    raise ValueError("twist splits a degenerate |k+theta| shell")
                return np.array([(t[1], t[2]) for t in qs[:n]])
        M *= 2
    raise RuntimeError("twisted enumeration did not converge")


@dataclass
class Spec:
    """Everything the C-ABI needs to build one system."""
    n_particles: int
    lattice: list | None
    groups: list = field(default_factory=list)          # (first, last, charge, inv_mass)
    pade: list = field(default_factory=list)            # dicts
    drummond: list = field(default_factory=list)
    nosanow: list = field(default_factory=list)
    slater_pw: list = field(default_factory=list)       # (first, last, k, is_sin)
    slater_twisted: list = field(default_factory=list)  # (first, last, q)
    slater_gauss: list = field(default_factory=list)    # (first, last, centers, alphas)
    ewald: dict | None = None
    coulomb: list = field(default_factory=list)         # (ga, gb, eps)
    inverse_power: list = field(default_factory=list)   # (ga, gb, strength, exponent)
    mcmillan: list = field(default_factory=list)        # (ga, gb, b, n)
    sites: object = None                                # crystal sites [N][2], if any
    name: str = ""


def heg_spec(n_up: int, n_dn: int, rs: float, twist=None, slater=True, jastrow=True,
             ewald=True, tables=None, lattice=None, drummond=None) -> Spec:
    """The 2-D paramagnetic (or polarized) electron gas of the HEG inputs;
    `tables` (one cell's {"ewald", "pade"}) replaces the product's breakup."""
    N = n_up + n_dn
    lat = lattice or square_lattice(N, rs)
    s = Spec(n_particles=N, lattice=lat, name=f"heg N={N} rs={rs}")
    s.groups.append((0, n_up, -1.0, 1.0))
    if n_dn:
        s.groups.append((n_up, N, -1.0, 1.0))
    b = [0.5 / rs]
    if jastrow:
        pairs = [(0, 0, 1.0 / 3.0)]
        if n_dn:
            pairs += [(1, 1, 1.0 / 3.0), (0, 1, 1.0)]
        for ga, gb, a1 in pairs:
            tab = pade_table(lat, [a1], b, tables)
            s.pade.append(dict(ga=ga, gb=gb, num=1, den=1, coeffs=[a1] + b, image=tab))
    if drummond is not None:
        s.drummond.extend(drummond)
    if slater:
        for gi, (first, last, _, _) in enumerate(s.groups):
            n = last - first
            if twist is None:
                k, sn = plane_wave_shells(lat, n)
                s.slater_pw.append((first, last, k, sn))
            else:
                s.slater_twisted.append((first, last, twisted_wavevectors(lat, n, twist)))
    if ewald:
        s.ewald = ewald_table(lat, tables)
    return s


def heg_drummond_spec(n_up: int, n_dn: int, rs: float, order: int = 8, sigma: float = 1e-3, seed: int = 2,
                      tables=None) -> Spec:
    """The paramagnetic HEG of the published DMC case (N = 74, rs = 40 there):
    plane-wave Slater determinants, the image-summed Padé pair Jastrow with
    the Kato cusps, plus a Drummond polynomial (order `order`, C = 3, cutoff
    L/2) on each spin pair with no cusp of its own, so the Padé factor alone
    sets the cusp; 2-D Ewald.  The Drummond polynomial is in r itself
    (drummond_jastrow.cpp:47-64), so coefficients of size 1e-3 would make
    u ~ 1e20 at r ~ 300: they are drawn N(0, sigma) from `seed` in units of
    the cutoff, alpha_l = xi_l / L^(C + l) for the term in r^l, so each term
    is ~sigma at r ~ L."""
    n = n_up + n_dn
    half = square_lattice(n, rs)[0] / 2.0
    rng = np.random.default_rng(seed)
    C = 3
    # coeffs[0] multiplies r^0, coeffs[k >= 1] multiplies r^(k+1)
    powers = [0] + list(range(2, order + 1))
    dr = []
    for ga, gb in ((0, 0), (1, 1), (0, 1)):
        xi = rng.normal(0.0, sigma, order)
        dr.append(dict(ga=ga, gb=gb, C=C, gamma=0.0, L=half,
                       coeffs=[float(x) / half ** (C + l) for x, l in zip(xi, powers)]))
    s = heg_spec(n_up, n_dn, rs, drummond=dr, tables=tables)
    s.name = f"heg N={n} rs={rs:g} pade+drummond{order}"
    return s


def wigner_crystal_spec(rs: float = 60.0, rows: int = 8, cols: int = 16, nosanow_alpha: float | None = None,
                        drummond_order: int = 8, drummond_sigma: float = 1e-3, drummond_seed: int = 2,
                        tables=None) -> Spec:
    """The polarized 2-D Wigner crystal: N = rows x cols electrons of one spin
    on a triangular lattice in the rectangular cell of area N pi rs^2 (cols
    sites along x, rows along y, every other row shifted by half a spacing).
    Trial: one real determinant of Gaussians exp(-c |r - R_j|^2) (minimum
    image) centred on the sites with c = 0.1 / rs^1.5, a Nosanow one-body
    term -alpha |r_i - R_i|^2 on each particle's own site (alpha defaults to
    c), the image-summed Pade u = r/3 / (1 + r/(2 rs)) and a Drummond
    polynomial (C = 3, cutoff the Wigner-Seitz radius, coefficients drawn
    N(0, sigma) in units of the cutoff, as for the paramagnetic DMC case);
    2-D Ewald.  Particle i sits on site i."""
    n = rows * cols
    a0 = rs * math.sqrt(2.0 * math.pi / math.sqrt(3.0))
    lat = [cols * a0, 0.0, 0.0, rows * a0 * math.sqrt(3.0) / 2.0]
    sites = np.array([((c + 0.5 * (r % 2)) * a0, r * a0 * math.sqrt(3.0) / 2.0)
                      for r in range(rows) for c in range(cols)])
    sites -= 0.5 * np.array([lat[0], lat[3]])   # centred on the origin, like the cell
    c = 0.1 / rs ** 1.5
    s = Spec(n_particles=n, lattice=lat, name=f"wigner crystal N={n} rs={rs:g}")
    s.groups.append((0, n, -1.0, 1.0))
    b = [0.5 / rs]
    s.pade.append(dict(ga=0, gb=0, num=1, den=1, coeffs=[1.0 / 3.0] + b,
                       image=pade_table(lat, [1.0 / 3.0], b, tables)))
    half = 0.5 * min(lat[0], lat[3])
    rng = np.random.default_rng(drummond_seed)
    C = 3
    powers = [0] + list(range(2, drummond_order + 1))
    xi = rng.normal(0.0, drummond_sigma, drummond_order)
    s.drummond.append(dict(ga=0, gb=0, C=C, gamma=0.0, L=half,
                           coeffs=[float(x) / half ** (C + l) for x, l in zip(xi, powers)]))
    s.nosanow.append(dict(group=0, alpha=c if nosanow_alpha is None else nosanow_alpha,
                          sites=sites.reshape(-1).copy()))
    s.slater_gauss.append((0, n, sites.copy(), np.full(n, c)))
    s.ewald = ewald_table_tightest(lat, tables)
    s.sites = sites
    return s


def site_positions(spec: Spec, nw: int, rng: np.random.Generator, spread: float):
    """Walkers near the crystal sites: site + N(0, spread^2) per coordinate,
    SoA [p*2*nw + d*nw + w]."""
    sites = np.asarray(spec.sites)
    pos = sites[:, :, None] + rng.normal(0.0, spread, size=(spec.n_particles, 2, nw))
    return np.ascontiguousarray(pos.reshape(-1))


def dipole_spec(n: int, density: float = 1.0, b: float | None = None, n_exp: int = 5,
                lattice: str = "hexagonal") -> Spec:
    """2-D bosonic dipoles perpendicular to the plane (configs/
    dipoles_2d_gas_n256.yaml): one group of n bosons of mass 1, H = -1/2 sum
    nabla^2 + sum_{i<j} 1/r^3 in units of r0, density n r0^2 = `density`
    (rs = 1/sqrt(pi density)), and a Jastrow-only trial: McMillan
    u = -(b/r)^5 (b defaults to the mean spacing rs)."""
    rs = 1.0 / math.sqrt(math.pi * density)
    lat = hexagonal_lattice(n, rs) if lattice == "hexagonal" else square_lattice(n, rs)
    s = Spec(n_particles=n, lattice=lat, name=f"dipoles N={n} n r0^2={density:g}")
    s.groups.append((0, n, 0.0, 1.0))
    s.mcmillan.append((0, 0, rs if b is None else b, n_exp))
    s.inverse_power.append((0, 0, 1.0, 3.0))
    return s


def exciton_spec() -> Spec:
    """Open-cell 2-D exciton with the exact u = -r (tests/configs/exciton_2d_cusp_m1.yaml):
    E_L = -1 for every configuration."""
    s = Spec(n_particles=2, lattice=None, name="exciton")
    s.groups = [(0, 1, -1.0, 1.0), (1, 2, 1.0, 1.0)]
    s.pade.append(dict(ga=0, gb=1, num=1, den=0, coeffs=[-1.0], image=None))
    s.coulomb.append((0, 1, 1.0))
    return s


def uniform_positions(spec: Spec, nw: int, rng: np.random.Generator, box=1.0):
    """Uniform fractional coordinates in [-1/2, 1/2) mapped through the cell
    (app.cpp:56-116), SoA [p*2*nw + d*nw + w]."""
    N = spec.n_particles
    f = rng.uniform(-0.5, 0.5, size=(N, 2, nw))
    if spec.lattice is None:
        return np.ascontiguousarray((f * box).reshape(-1))
    a = np.array(spec.lattice).reshape(2, 2)
    pos = np.einsum("pin,id->pdn", f, a)
    return np.ascontiguousarray(pos.reshape(-1))


def free_parameters(spec: Spec, free, param_tables=None) -> list:
    """The free parameters of `spec`'s wavefunction for
    Walkers.parameter_derivatives: `free` lists (factor, index) pairs, factors
    counted in the order Walkers registers them (Padé factors, Drummond,
    Nosanow, then the determinants and McMillan), indices as
    alea_parameter_derivatives takes them.  A Padé summed over images carries
    the breakup of its du/dp (breakup.pade_parameter_tables, or
    param_tables(factor, index) when given, e.g. another fit's tables)."""
    out = []
    n_pade = len(spec.pade)
    for f, i in free:
        q = {"factor": int(f), "index": int(i)}
        if f < n_pade and spec.pade[f].get("image"):
            j = spec.pade[f]
            if param_tables is not None:
                q["tables"] = param_tables(f, i)
            else:
                c = list(j["coeffs"])
                q["tables"] = breakup.pade_parameter_tables(spec.lattice, c[:j["num"]], c[j["num"]:], i)
        out.append(q)
    return out
