"""Run configurations in the reference's YAML schema, built into a Spec for the
GPU path, and the reference's initial walkers for them.

The reference reads a config (src/io/parser.cpp) and App::build_drivers
(src/qmc/app.cpp:722-927) turns it into a WalkerEnsemble, a TrialWavefunction
and a Hamiltonian.  This module is that seam for `device: gpu`: the same keys,
defaults and derived values, ending in a Spec that alea_amd.walkers registers
through the C-ABI.

  cell from the lattice    src/qmc/app.cpp:273-341 (build_cell)
  particle groups          src/qmc/app.cpp:129-153 (one group per entry)
  initial walkers          src/qmc/app.cpp:56-116 (initialize_positions) with
                           numerics::Random, src/numerics/random.h:11-27
  stage seeds              src/qmc/app.cpp:764-766, 815-826
  Kato cusp                src/builders/jastrow.cpp:47-96
  Padé / Drummond          src/builders/jastrow.cpp:101-176, 325-386
  Slater spin blocks       src/builders/slater.cpp:85-95, 205-300
  Coulomb / Ewald          src/builders/ewald.cpp:40-74 (a periodic run's
                           Coulomb terms are one cell-wide Ewald sum)

What the GPU path does not hold (3-D, PIMC, backflow, pairing, crystal sites,
Keldysh and other potentials, layer offsets, minimum-image truncation) is
refused with a ValueError that names the item, never approximated.

The initial walkers need the reference's generator bit for bit: MT19937-64
(Matsumoto and Nishimura's 64-bit Mersenne Twister, the C++ standard's
std::mt19937_64) and std::uniform_real_distribution<double>, which for a
64-bit engine is one draw x mapped to x / 2^64 (the standard's
generate_canonical with one term), written here in numpy.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .system import Spec, ewald_table, pade_table, plane_wave_shells, twisted_wavevectors

_U64 = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister, vectorised over each 312-word twist."""
    N, M = 312, 156
    MATRIX_A = np.uint64(0xB5026F5AA96619E9)
    UPPER = np.uint64(0xFFFFFFFF80000000)
    LOWER = np.uint64(0x7FFFFFFF)

    def __init__(self, seed: int):
        mt = [seed & _U64]
        for i in range(1, self.N):
            prev = mt[-1]
            mt.append((6364136223846793005 * (prev ^ (prev >> 62)) + i) & _U64)
        self.mt = np.array(mt, dtype=np.uint64)
        self.idx = self.N

    def _mix(self, a, b, c):
        x = (a & self.UPPER) | (b & self.LOWER)
        xa = x >> np.uint64(1)
        xa ^= np.where((x & np.uint64(1)).astype(bool), self.MATRIX_A, np.uint64(0))
        return c ^ xa

    def _twist(self):
        mt, M = self.mt, self.M
        # word i needs words i+1 (not yet twisted) and i+M mod N: the first
        # M words read only old words, the next N-M-1 the new words i-M,
        # and the last the new words 0 and M-1
        mt[:M] = self._mix(mt[:M], mt[1:M + 1], mt[M:])
        mt[M:-1] = self._mix(mt[M:-1], mt[M + 1:], mt[:M - 1])
        mt[-1:] = self._mix(mt[-1:], mt[:1], mt[M - 1:M])
        self.idx = 0

    def next_u64(self, n: int) -> np.ndarray:
        out = np.empty(n, dtype=np.uint64)
        got = 0
        while got < n:
            if self.idx >= self.N:
                self._twist()
            take = min(n - got, self.N - self.idx)
            y = self.mt[self.idx:self.idx + take].copy()
            y ^= (y >> np.uint64(29)) & np.uint64(0x5555555555555555)
            y ^= (y << np.uint64(17)) & np.uint64(0x71D67FFFEDA60000)
            y ^= (y << np.uint64(37)) & np.uint64(0xFFF7EEE000000000)
            y ^= y >> np.uint64(43)
            out[got:got + take] = y
            got += take
            self.idx += take
        return out

    def uniform(self, n: int) -> np.ndarray:
        """std::uniform_real_distribution<double>(0, 1): x / 2^64 in [0, 1)."""
        u = self.next_u64(n).astype(np.float64) / 18446744073709551616.0
        return np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)


# ── config file ──────────────────────────────────────────────────────────────
def load(path: str) -> dict:
    import yaml
    with open(path, encoding="utf-8") as f:
        cfg = yaml.safe_load(f)
    cfg.setdefault("_path", os.path.abspath(path))
    return cfg


def _refuse(what: str):
    raise ValueError(f"config: {what} is not on the GPU path")


def _lattice(cfg) -> dict:
    lat = dict(cfg.get("system", {}).get("lattice", {}))
    lat.setdefault("type", "open")
    lat.setdefault("dimension", 2)
    return lat


def _particles(cfg) -> list:
    return list(cfg.get("system", {}).get("particles", []))


def _items(cfg, kind: str) -> list:
    return [it[kind] for it in cfg.get("wavefunction", []) or [] if kind in it]


def _stages(cfg) -> list:
    return list(cfg.get("qmc", []) or [])


def build_groups(cfg) -> list[dict]:
    """One group per `particles:` entry, in order (app.cpp:129-153)."""
    groups, off = [], 0
    for p in _particles(cfg):
        if p.get("flavours") is not None:
            _refuse("`flavours` counts")
        if float(p.get("z", 0.0)) != 0.0:
            _refuse(f"the layer offset z of group {p.get('group')}")
        mass = p.get("mass", 1.0)
        if isinstance(mass, (list, tuple)):
            if len(set(float(m) for m in mass)) != 1:
                _refuse(f"the anisotropic mass of group {p.get('group')}")
            mass = mass[0]
        up, dn = int(p.get("up", 0)), int(p.get("down", 0))
        groups.append(dict(name=p.get("group", ""), first=off, last=off + up + dn, up=up, down=dn,
                           charge=float(p.get("charge", 0.0)), mass=float(mass),
                           init=p.get("initialization", ""), type=p.get("type", "")))
        off += up + dn
    return groups


def build_cell(cfg, n_particles: int):
    """The cell vectors [a1x, a1y, a2x, a2y], or None for an open system
    (app.cpp:273-341)."""
    lat = _lattice(cfg)
    if int(lat["dimension"]) != 2:
        _refuse(f"dimension {lat['dimension']}")
    if lat["type"] != "periodic":
        return None
    if lat.get("vectors") is not None:
        v = [float(x) for row in lat["vectors"] for x in (row if isinstance(row, (list, tuple)) else [row])]
        if len(v) != 4:
            raise ValueError("config: lattice.vectors needs 4 components in 2-D")
        return v
    shape = lat.get("shape", "square")
    L = lat.get("L", 0.0)
    if isinstance(L, (list, tuple)):
        if len(L) != 2:
            raise ValueError("config: lattice.L as a list needs one length per dimension")
        return [float(L[0]), 0.0, 0.0, float(L[1])]
    L = float(L)
    rs = float(lat.get("rs", 0.0))
    if rs > 0.0:
        if n_particles == 0:
            raise ValueError("config: lattice.rs needs a nonzero particle count")
        area = n_particles * math.pi * rs * rs
        L = math.sqrt(2.0 * area / math.sqrt(3.0)) if shape == "hexagonal" else math.sqrt(area)
    if L <= 0.0:
        raise ValueError("config: a periodic lattice needs lattice.vectors, lattice.rs or lattice.L")
    if shape == "hexagonal":
        return [L, 0.0, 0.5 * L, L * math.sqrt(3.0) / 2.0]
    if shape in ("square", "rectangular"):
        return [L, 0.0, 0.0, L]
    raise ValueError(f"config: unknown lattice shape '{shape}'")


def cell_vectors(lattice):
    """The basis SimulationCell keeps (vectors(), simulation_cell.cpp:16-36,
    100-132): rows each along one axis are put in axis order; any other 2-D
    basis is Gauss-reduced.  A hexagonal basis sits on the reduction's tie
    (|a1| = |a2|), so rounding decides whether the rows swap."""
    a = [float(x) for x in lattice]
    axis, ok = [0, 0], True
    for i in range(2):
        nz = [d for d in range(2) if a[i * 2 + d] != 0.0]
        if len(nz) != 1 or nz[0] in axis[:i]:
            ok = False
            break
        axis[i] = nz[0]
    if ok:
        re = [0.0] * 4
        for i in range(2):
            for d in range(2):
                re[axis[i] * 2 + d] = a[i * 2 + d]
        return re
    for _ in range(64):
        n0 = a[0] * a[0] + a[1] * a[1]
        n1 = a[2] * a[2] + a[3] * a[3]
        if n1 < n0:
            a[0], a[2] = a[2], a[0]
            a[1], a[3] = a[3], a[1]
            continue
        if n0 == 0.0:
            break
        mu = float(np.rint((a[0] * a[2] + a[1] * a[3]) / n0))
        if mu == 0.0:
            break
        a[2] -= mu * a[0]
        a[3] -= mu * a[1]
    return a


def ws_inradius(lattice) -> float:
    """Half the shortest nonzero lattice vector."""
    a = np.array(lattice, dtype=float).reshape(2, 2)
    best = math.inf
    for i in range(-3, 4):
        for j in range(-3, 4):
            if i or j:
                v = i * a[0] + j * a[1]
                best = min(best, math.hypot(v[0], v[1]))
    return 0.5 * best


def _is_gamma(twist) -> bool:
    return not twist or all(float(t) == round(float(t)) for t in twist)


def _group(groups, name):
    for g in groups:
        if g["name"] == name:
            return g
    raise ValueError(f"config: no particle group '{name}'")


def _coulomb_terms(cfg):
    terms = (cfg.get("hamiltonian", {}) or {}).get("terms", []) or []
    out = []
    for t in terms:
        kind = t.get("type", "")
        if kind == "coulomb":
            out.append(("coulomb", t.get("pair"), float(t.get("epsilon", 1.0))))
        elif kind == "inverse_power":
            for k in ("pair", "strength", "exponent"):
                if k not in t:
                    raise ValueError(f"config: hamiltonian term 'inverse_power' needs `{k}`")
            out.append(("inverse_power", t["pair"], float(t["strength"]), float(t["exponent"])))
        else:
            _refuse(f"hamiltonian term '{kind}'")
    return out


def _cusp_residue(terms, ga, gb, periodic: bool) -> float:
    """Sum of lim r V(r) over the pair potentials (HamiltonianComponent::
    coulomb_cusp_residue): q_a q_b / eps from each Coulomb term on the pair;
    in a periodic cell the Coulomb terms are one Ewald sum over every pair
    (ewald_term.cpp:310-328); an inverse power C/r on the pair adds C
    (inverse_power.cpp:170-173), steeper ones nothing."""
    def own(pair):
        return sorted(pair) == sorted((ga["name"], gb["name"]))
    R = 0.0
    coul = [t for t in terms if t[0] == "coulomb"]
    if periodic and coul:
        R += ga["charge"] * gb["charge"] / coul[0][2]
    for t in terms:
        if t[0] == "coulomb" and not periodic and own(t[1]):
            R += ga["charge"] * gb["charge"] / t[2]
        elif t[0] == "inverse_power" and own(t[1]) and t[3] == 1.0:
            R += t[2]
    return R


def _coalescence_ell(cfg, groups, a: str, b: str) -> int:
    """1 for a same-group pair of a fully polarised group that carries a
    determinant, else 0 (jastrow.cpp:71-96)."""
    if a != b:
        return 0
    if not any(s.get("group") == a for s in _items(cfg, "slater")):
        return 0
    g = _group(groups, a)
    return 1 if (g["up"] == 0 or g["down"] == 0) else 0


def kato_cusp(cfg, groups, terms, a: str, b: str, periodic: bool) -> float:
    """Gamma = 2 mu R / (2 l + d - 1) (jastrow.cpp:47-63)."""
    ga, gb = _group(groups, a), _group(groups, b)
    R = _cusp_residue(terms, ga, gb, periodic)
    mu = ga["mass"] * gb["mass"] / (ga["mass"] + gb["mass"])
    ell = _coalescence_ell(cfg, groups, a, b)
    return 2.0 * mu * R / float(2 * ell + 2 - 1)


def build_spec(cfg, tables=None, requests: list | None = None) -> Spec:
    """The Spec of a config: groups, cell, Jastrow and determinant factors and
    the Hamiltonian, with every derived value (cusps, k-points, breakup
    tables) the reference's builders derive.

    The breakup tables come from alea_amd.breakup, or from `tables` (one
    cell's {"ewald", "pade"}) when given.  With `requests` (a list), the
    tables the config needs are appended to it as ("pade", lattice, a, b) and
    ("ewald", lattice, tol) instead; the Spec then holds no tables."""
    if any(s.get("method") == "pimc" for s in _stages(cfg)) or cfg.get("action"):
        _refuse("a PIMC run")
    groups = build_groups(cfg)
    n = groups[-1]["last"] if groups else 0
    lattice = build_cell(cfg, n)
    periodic = lattice is not None
    lat = _lattice(cfg)
    if lat.get("truncation") not in (None, "", "image_sum", "none"):
        _refuse(f"lattice.truncation '{lat.get('truncation')}'")
    if lat.get("capacitor") not in (None, "include"):
        _refuse(f"lattice.capacitor '{lat.get('capacitor')}'")
    twist = lat.get("twist") or []
    terms = _coulomb_terms(cfg)
    name = os.path.basename(cfg.get("_path", "")) or (cfg.get("project", {}) or {}).get("name", "")
    s = Spec(n_particles=n, lattice=lattice, name=name)
    gindex = {g["name"]: i for i, g in enumerate(groups)}
    for g in groups:
        s.groups.append((g["first"], g["last"], g["charge"], 1.0 / g["mass"]))
    ewald = periodic and any(t[0] == "coulomb" for t in terms)

    # ── wavefunction ──
    for item in cfg.get("wavefunction", []) or []:
        (kind, body), = item.items()
        if kind == "jastrow":
            jt = body.get("type", "")
            if body.get("pairs"):
                if jt != "drummond":
                    raise ValueError("config: `pairs:` is for drummond jastrows only")
                pairs = [tuple(p) for p in body["pairs"]]
            else:
                pairs = [tuple(body.get("pair", ("", "")))]
            if jt in ("pade", "padé"):
                a_name, b_name = pairs[0]
                a1 = kato_cusp(cfg, groups, terms, a_name, b_name, periodic)
                num = [a1] + [float(x) for x in body.get("numerator", []) or []]
                den = [float(x) for x in body.get("denominator", []) or []]
                image = None
                if periodic:
                    if not ewald:
                        raise ValueError(f"config: padé jastrow on [{a_name}, {b_name}] in a periodic "
                                         "cell needs an Ewald term to be summed over images")
                    if requests is not None:
                        requests.append(("pade", lattice, num, den))
                    else:
                        image = pade_table(lattice, num, den, tables)
                s.pade.append(dict(ga=gindex[a_name], gb=gindex[b_name], num=len(num), den=len(den),
                                   coeffs=num + den, image=image))
            elif jt == "drummond":
                params = [float(x) for x in body.get("params", []) or []]
                Lc = float(body.get("cutoff", 0.0))
                C = int(body.get("truncation", 3))
                if periodic and Lc > ws_inradius(lattice):
                    raise ValueError(f"config: drummond cutoff {Lc} exceeds the Wigner-Seitz inradius")
                gamma = None
                for a_name, b_name in pairs:
                    g = kato_cusp(cfg, groups, terms, a_name, b_name, periodic)
                    if gamma is None:
                        gamma = g
                    elif abs(g - gamma) > 1e-12 * max(1.0, abs(gamma)):
                        raise ValueError("config: one drummond u(r) carries one cusp; split the pairs")
                for a_name, b_name in pairs:
                    s.drummond.append(dict(ga=gindex[a_name], gb=gindex[b_name], C=C, gamma=gamma, L=Lc,
                                           coeffs=params))
            else:
                _refuse(f"jastrow type '{jt}'")
        elif kind == "slater":
            if body.get("backflow"):
                _refuse("a backflow determinant")
            g = _group(groups, body.get("group", ""))
            orb = body.get("orbitals", {}) or {}
            otype = orb.get("type", "gaussian")
            blocks = [(g["first"], g["first"] + g["up"], g["up"])] if g["up"] else []
            if g["down"]:
                blocks.append((g["first"] + g["up"], g["last"], g["down"]))
            if otype == "plane_wave":
                if orb.get("k") is not None:
                    _refuse("an explicit plane-wave occupation (`k:`)")
                if not periodic:
                    raise ValueError("config: plane-wave orbitals need a periodic cell")
                for first, last, cnt in blocks:
                    if _is_gamma(twist):
                        k, sn = plane_wave_shells(lattice, cnt)
                        s.slater_pw.append((first, last, k, sn))
                    else:
                        s.slater_twisted.append((first, last, twisted_wavevectors(lattice, cnt, twist)))
            elif otype == "gaussian":
                if orb.get("sites") or (orb.get("centers") and orb.get("alpha") is not None and not orb.get("alphas")):
                    _refuse("a crystal determinant (one site per particle)")
                centers = [list(map(float, c)) for c in orb.get("centers", []) or []]
                alphas = [float(x) for x in orb.get("alphas", []) or []]
                if len(centers) != len(alphas):
                    raise ValueError(f"config: slater centers/alphas length mismatch for group {g['name']}")
                if len(alphas) < max(g["up"], g["down"]):
                    raise ValueError(f"config: slater group {g['name']} needs more orbitals")
                if any(len(c) != 2 for c in centers):
                    raise ValueError("config: slater orbital centers need 2 components")
                for first, last, cnt in blocks:   # the shared pool's first cnt orbitals
                    s.slater_gauss.append((first, last, np.array(centers[:cnt]), np.array(alphas[:cnt])))
            else:
                _refuse(f"slater orbitals '{otype}'")
        else:
            _refuse(f"wavefunction item '{kind}'")

    # ── hamiltonian (the kinetic term is always there) ──
    if ewald:
        eps = {t[2] for t in terms if t[0] == "coulomb"}
        if eps != {1.0}:
            _refuse("a periodic Coulomb term with epsilon != 1")
        tol = float(lat.get("ewald_tol", 1e-6))
        if requests is not None:
            requests.append(("ewald", lattice, tol))
        else:
            e = ewald_table(lattice, tables, tol)
            if abs(e["tol"] - tol) > 1e-12 * tol:
                raise ValueError(f"the given Ewald table has tolerance {e['tol']}, the config asks {tol}")
            s.ewald = e
    for t in terms:
        if t[0] == "coulomb" and not periodic:
            a_name, b_name = t[1]
            s.coulomb.append((gindex[a_name], gindex[b_name], t[2]))
        elif t[0] == "inverse_power":
            a_name, b_name = t[1]
            s.inverse_power.append((gindex[a_name], gindex[b_name], t[2], t[3]))
    return s


# ── seeds and initial walkers ────────────────────────────────────────────────
def first_ensemble_seed(cfg, seed: int | None = None) -> int:
    """Seed of the first ensemble build_drivers creates: an optimizer's uses the
    project seed, a VMC/DMC stage's the stage seed seed + stage + 1 (with no
    PIMC stage, app.cpp:815-817)."""
    base = int(seed if seed is not None else (cfg.get("project", {}) or {}).get("seed", 0))
    stages = _stages(cfg)
    if any(st.get("method") == "optimizer" for st in stages):
        return base
    return base + 1


def initial_positions(cfg, nw: int, seed: int, groups=None, lattice=None) -> np.ndarray:
    """The reference's initial walkers for `seed` (app.cpp:56-116), as the
    walker-fastest SoA [p*2*nw + d*nw + w] the C-ABI takes.  A periodic cell
    draws fractional coordinates in [-1/2, 1/2) per particle and maps them
    through the cell's reduced basis (cell_vectors); an open one draws each
    coordinate in
    [-L/2, L/2).  Particles of an `initialization: fixed` group sit at the
    origin and draw nothing."""
    groups = groups or build_groups(cfg)
    n = groups[-1]["last"] if groups else 0
    if lattice is None:
        lattice = build_cell(cfg, n)
    if any(s.get("sites") or (s.get("orbitals", {}) or {}).get("sites")
           or ((s.get("orbitals", {}) or {}).get("alpha") is not None
               and not (s.get("orbitals", {}) or {}).get("alphas"))
           for s in _items(cfg, "slater")) or _items(cfg, "nosanow") or _items(cfg, "pairing"):
        _refuse("crystal-site or pairing-partner initial walkers")
    fixed = np.zeros(n, dtype=bool)
    for g in groups:
        if g["init"] == "fixed":
            fixed[g["first"]:g["last"]] = True
    n_free = int((~fixed).sum())
    u = MT19937_64(seed).uniform(nw * n_free * 2).reshape(nw, n_free, 2)
    pos = np.zeros((nw, n, 2))
    if lattice is not None:
        a = np.array(cell_vectors(lattice), dtype=float).reshape(2, 2)
        f = -0.5 + 1.0 * u
        free = f[..., 0, None] * a[0] + f[..., 1, None] * a[1]
    else:
        L = float(_lattice(cfg).get("L", 0.0))
        free = -0.5 * L + L * u
    pos[:, ~fixed, :] = free
    return np.ascontiguousarray(pos.transpose(1, 2, 0).reshape(-1))
