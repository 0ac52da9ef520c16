"""Per-system molecular topology records and AMBER prmtop / inpcrd intake and output.

A :class:`Topology` carries exactly what the batched energy path consumes for one molecule:
charges, Lennard-Jones type tables, valence terms, the 1:1/1:2/1:3/1:4 exclusion map and the
Generalized Born radii.  It is the host-side counterpart of the reference's ``AtomGraph``
abstracts (``NonbondedKit``, ``ValenceKit``, ``ImplicitSolventKit``, src/Topology/
atomgraph_abstracts.h) restricted to the terms on this path.

The prmtop reader follows ``AtomGraph::buildFromPrmtop`` (src/Topology/atomgraph_detailers.cpp:
607-1300): charges are scaled by 1 / 18.2223 (src/Topology/amber_prmtop_util.h:22-23) and then
passed through ``smoothCharges`` (src/Topology/atomgraph_refinement.cpp:1519-1620); exclusions
are re-derived from the bond graph by shortest path (``mapExclusions``, 1:2, 1:3 and 1:4 pairs),
not read from EXCLUDED_ATOMS_LIST.
"""
from __future__ import annotations

import dataclasses
import math
import re
from collections import deque
from typing import Dict, List, Optional, Tuple

import numpy as np

# Coulomb constant, src/Constants/symbol_values.h:33.
AMBER_COULOMB = 332.0522172900000
# prmtop charge scaling, src/Topology/amber_prmtop_util.h:22-23.
AMBER_CHARGE_SCALING = 18.2223
# Default 1:4 attenuation divisors, src/Topology/atomgraph_constants.h.
AMBER_SCEE = 1.2
AMBER_SCNB = 2.0
# Charge smoothing defaults, src/Topology/atomgraph_constants.h:12-18.
DEFAULT_CHARGE_ROUNDING_TOL = 0.001
DEFAULT_CHARGE_DISCRETIZATION = 1.1920928955078125e-7

# Implicit solvent models (AMBER igb numbers) and their alpha / beta / gamma,
# src/Constants/generalized_born.h:56-66 and src/Topology/atomgraph_setters.cpp:955-985.
IGB_NONE, IGB_HCT, IGB_OBC, IGB_OBC2 = 0, 1, 2, 5
GB_ABG = {
    IGB_NONE: (0.0, 0.0, 0.0),
    IGB_HCT: (0.0, 0.0, 0.0),
    IGB_OBC: (0.8, 0.0, 2.909125),
    IGB_OBC2: (1.0, 0.8, 4.85),
}
GB_RADII_OFFSET = 0.09        # src/Constants/generalized_born.h:112
SALT_KAPPA_DEPENDENCE = 0.10806  # src/Constants/generalized_born.h:23


@dataclasses.dataclass
class Topology:
    """One molecule's parameters for the valence + non-bonded + GB energy path."""

    natom: int
    charge: np.ndarray                  # (N,) e
    lj_idx: np.ndarray                  # (N,) int32
    n_lj_types: int
    bonds: np.ndarray                   # (nb, 2) int32
    bond_k: np.ndarray
    bond_r0: np.ndarray
    angles: np.ndarray                  # (na, 3) int32
    angle_k: np.ndarray
    angle_theta: np.ndarray
    dihedrals: np.ndarray               # (nd, 4) int32
    dihe_amp: np.ndarray
    dihe_phase: np.ndarray
    dihe_freq: np.ndarray
    dihe_improper: np.ndarray           # (nd,) int32
    dihe_scee: np.ndarray               # (nd,) 1:4 electrostatic divisor (0 = no 1:4 pair)
    dihe_scnb: np.ndarray               # (nd,) 1:4 Lennard-Jones divisor
    pb_radii: np.ndarray                # (N,) intrinsic (Poisson-Boltzmann) radii, A
    gb_screen: np.ndarray               # (N,) screening factors
    igb: int = IGB_NONE
    dielectric: float = 78.5
    salt: float = 0.0
    coulomb: float = AMBER_COULOMB
    infr14: Optional[np.ndarray] = None  # (ni, 2) inferred 1:4 pairs
    infr14_scee: Optional[np.ndarray] = None
    infr14_scnb: Optional[np.ndarray] = None
    name: str = ""
    # Lennard-Jones parameters: either per-type sigma / epsilon under Lorentz-Berthelot combining
    # (the tables below are then built on demand), or the stored T x T tables (index ti * T + tj)
    # for the ordinary and the 1:4 interactions, as an AMBER topology holds them.
    lj_sigma: Optional[np.ndarray] = None
    lj_epsilon: Optional[np.ndarray] = None
    lja: Optional[np.ndarray] = None
    ljb: Optional[np.ndarray] = None
    lja14: Optional[np.ndarray] = None
    ljb14: Optional[np.ndarray] = None
    _excl_cache: Optional[np.ndarray] = dataclasses.field(default=None, repr=False)

    # ------------------------------------------------------------------------------------------
    @property
    def kappa(self) -> float:
        """Debye screening parameter, src/Potential/energy_abstracts.tpp:15-16."""
        return math.sqrt(SALT_KAPPA_DEPENDENCE * self.salt)

    def gb_alpha_beta_gamma(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        a, b, g = GB_ABG[self.igb]
        n = self.natom
        return (np.full(n, a), np.full(n, b), np.full(n, g))

    def exclusion_pairs(self) -> np.ndarray:
        """All (i < j) pairs excluded from Coulomb / LJ: 1:2, 1:3 and 1:4 by shortest bond path.

        Restates ``mapExclusions`` (src/Topology/atomgraph_refinement.cpp) for systems without
        virtual sites.
        """
        if self._excl_cache is not None:
            return self._excl_cache
        n = self.natom
        adj: List[List[int]] = [[] for _ in range(n)]
        for i, j in self.bonds:
            adj[int(i)].append(int(j))
            adj[int(j)].append(int(i))
        pairs = []
        for src in range(n):
            dist = {src: 0}
            dq = deque([src])
            while dq:
                u = dq.popleft()
                if dist[u] == 3:
                    continue
                for v in adj[u]:
                    if v not in dist:
                        dist[v] = dist[u] + 1
                        dq.append(v)
            for v, d in dist.items():
                if v > src and 1 <= d <= 3:
                    pairs.append((src, v))
        arr = np.array(sorted(pairs), dtype=np.int32).reshape(-1, 2)
        self._excl_cache = arr
        return arr

    def exclusion_matrix(self) -> np.ndarray:
        n = self.natom
        m = np.zeros((n, n), dtype=np.uint8)
        p = self.exclusion_pairs()
        if len(p):
            m[p[:, 0], p[:, 1]] = 1
            m[p[:, 1], p[:, 0]] = 1
        return m

    def lj_tables(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """(A, B, A14, B14), each (T*T,): the stored tables, or the Lorentz-Berthelot tables
        A_ij = 4 eps_ij sigma_ij^12, B_ij = 4 eps_ij sigma_ij^6 of the per-type parameters."""
        if self.lja is None:
            if self.lj_sigma is None:
                raise ValueError("topology has neither LJ tables nor per-type parameters")
            sig = np.asarray(self.lj_sigma, dtype=np.float64)
            eps = np.asarray(self.lj_epsilon, dtype=np.float64)
            sij = 0.5 * (sig[:, None] + sig[None, :])
            eij = np.sqrt(eps[:, None] * eps[None, :])
            a = (4.0 * eij * sij ** 12).reshape(-1)
            b = (4.0 * eij * sij ** 6).reshape(-1)
            return a, b, a, b
        return self.lja, self.ljb, self.lja14, self.ljb14

    def lorentz_berthelot(self, rtol: float = 1.0e-12) -> Optional[Tuple[np.ndarray, np.ndarray]]:
        """Per-type (sigma, epsilon) if every table entry follows Lorentz-Berthelot combining.

        A_ij = 4 eps_ij sigma_ij^12, B_ij = 4 eps_ij sigma_ij^6 with sigma_ij = (s_i + s_j) / 2
        and eps_ij = sqrt(e_i e_j).  Returns None when any entry departs by more than rtol or a
        type has a zero coefficient paired with a nonzero one.
        """
        if self.lj_sigma is not None:
            return (np.asarray(self.lj_sigma, dtype=np.float64),
                    np.asarray(self.lj_epsilon, dtype=np.float64))
        t = self.n_lj_types
        A = self.lja.reshape(t, t)
        B = self.ljb.reshape(t, t)
        A14 = self.lja14.reshape(t, t)
        B14 = self.ljb14.reshape(t, t)
        if not (np.array_equal(A, A14) and np.array_equal(B, B14)):
            return None
        d_a = np.diag(A)
        d_b = np.diag(B)
        sig = np.zeros(t)
        eps = np.zeros(t)
        nz = (d_a > 0) & (d_b > 0)
        sig[nz] = (d_a[nz] / d_b[nz]) ** (1.0 / 6.0)
        eps[nz] = d_b[nz] ** 2 / (4.0 * d_a[nz])
        sij = 0.5 * (sig[:, None] + sig[None, :])
        eij = np.sqrt(eps[:, None] * eps[None, :])
        A_lb = 4.0 * eij * sij ** 12
        B_lb = 4.0 * eij * sij ** 6
        scale_a = np.maximum(np.abs(A), 1.0e-300)
        scale_b = np.maximum(np.abs(B), 1.0e-300)
        if np.any(np.abs(A_lb - A) > rtol * scale_a) or np.any(np.abs(B_lb - B) > rtol * scale_b):
            return None
        return sig, eps

    def valence_term_count(self) -> int:
        return len(self.bonds) + len(self.angles) + len(self.dihedrals)


# ----------------------------------------------------------------------------------------------
# Charge smoothing
# ----------------------------------------------------------------------------------------------
def smooth_charges(q: np.ndarray, rounding_tol: float = DEFAULT_CHARGE_ROUNDING_TOL,
                   discretization: float = DEFAULT_CHARGE_DISCRETIZATION) -> np.ndarray:
    """Restatement of ``smoothCharges`` (src/Topology/atomgraph_refinement.cpp:1519-1608).

    If the net charge is within rounding_tol of an integer, spread the difference evenly, round
    every charge to the discretization grid, then sprinkle single grid increments over the atoms
    with the largest charge magnitudes (ties broken by atom index) until the sum is integral.
    """
    q = np.array(q, dtype=np.float64)
    n = len(q)
    qsum = float(np.sum(q))
    if abs(qsum - round(qsum)) >= rounding_tol:
        return q
    dq = (round(qsum) - qsum) / n
    q = q + dq
    if discretization < 1.0e-8:  # constants::tiny
        return q
    q = np.array([math.floor(v / discretization + 0.5) if v >= 0 else
                  -math.floor(-v / discretization + 0.5) for v in q], dtype=np.float64)
    q *= discretization
    disc_qsum = float(np.sum(q))
    ddq = discretization if (round(disc_qsum) - disc_qsum > 0.0) else -discretization
    increments = int(round(abs((round(disc_qsum) - disc_qsum) / discretization)))
    if increments > 0 and q.max() < 21.4 and q.min() > -21.4:
        keys = [(int(v * 100000000), i) for i, v in enumerate(q)]
        keys.sort(key=lambda t: -abs(t[0]))
        # Order runs of equal charge by atom index, as the reference does after its sort.
        seci = 0
        while seci < n:
            secj = seci + 1
            while secj < n and abs(q[keys[seci][1]] - q[keys[secj][1]]) < discretization:
                secj += 1
            keys[seci:secj] = sorted(keys[seci:secj], key=lambda t: t[1])
            seci = secj
        j = 0
        for _ in range(increments):
            q[keys[j][1]] += ddq
            j += 1
            if j == n:
                j = 0
    return q


# ----------------------------------------------------------------------------------------------
# AMBER prmtop / inpcrd
# ----------------------------------------------------------------------------------------------
_FMT_RE = re.compile(r"\(\s*(\d+)\s*([aAiIeEfF])\s*(\d+)(?:\.(\d+))?\s*\)")


def _read_prmtop_sections(path: str) -> Dict[str, List[str]]:
    sections: Dict[str, Tuple[str, List[str]]] = {}
    flag = None
    fmt = None
    lines: List[str] = []
    with open(path, "r") as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            if line.startswith("%FLAG"):
                if flag is not None:
                    sections[flag] = (fmt, lines)
                flag = line[5:].strip()
                fmt = None
                lines = []
            elif line.startswith("%FORMAT"):
                fmt = line[7:].strip()
            elif line.startswith("%VERSION") or line.startswith("%COMMENT"):
                continue
            elif flag is not None:
                lines.append(line)
        if flag is not None:
            sections[flag] = (fmt, lines)
    out = {}
    for name, (fmt, body) in sections.items():
        m = _FMT_RE.search(fmt or "")
        if m is None:
            out[name] = []
            continue
        count, kind, width = int(m.group(1)), m.group(2).lower(), int(m.group(3))
        vals: List[str] = []
        for ln in body:
            for k in range(count):
                field = ln[k * width:(k + 1) * width]
                if field.strip() == "" and kind != "a":
                    continue
                if kind == "a" and field == "":
                    continue
                vals.append(field)
        out[name] = (kind, vals)
    return out


def _ints(sec) -> np.ndarray:
    return np.array([int(v) for v in sec[1]], dtype=np.int64)


def _floats(sec) -> np.ndarray:
    return np.array([float(v) for v in sec[1]], dtype=np.float64)


def read_prmtop(path: str, igb: int = IGB_NONE, dielectric: float = 78.5, salt: float = 0.0,
                coulomb: float = AMBER_COULOMB, scee_default: float = AMBER_SCEE,
                scnb_default: float = AMBER_SCNB,
                charge_rounding_tol: float = DEFAULT_CHARGE_ROUNDING_TOL,
                charge_discretization: float = DEFAULT_CHARGE_DISCRETIZATION) -> Topology:
    """Read an AMBER prmtop into a :class:`Topology` (atomgraph_detailers.cpp:607-1300)."""
    s = _read_prmtop_sections(path)
    ptr = _ints(s["POINTERS"])
    natom, ntypes = int(ptr[0]), int(ptr[1])
    q = _floats(s["CHARGE"]) * (1.0 / AMBER_CHARGE_SCALING)
    q = smooth_charges(q, charge_rounding_tol, charge_discretization)
    lj_idx = _ints(s["ATOM_TYPE_INDEX"]).astype(np.int32) - 1
    nbidx = _ints(s["NONBONDED_PARM_INDEX"]).reshape(ntypes, ntypes)
    acoef = _floats(s["LENNARD_JONES_ACOEF"])
    bcoef = _floats(s["LENNARD_JONES_BCOEF"])
    a14 = _floats(s["LENNARD_JONES_14_ACOEF"]) if "LENNARD_JONES_14_ACOEF" in s else acoef
    b14 = _floats(s["LENNARD_JONES_14_BCOEF"]) if "LENNARD_JONES_14_BCOEF" in s else bcoef

    def expand(table):
        out = np.zeros((ntypes, ntypes))
        for ti in range(ntypes):
            for tj in range(ntypes):
                k = int(nbidx[ti, tj])
                out[ti, tj] = table[k - 1] if k > 0 else 0.0
        return out.reshape(-1)

    bond_k_t = _floats(s["BOND_FORCE_CONSTANT"])
    bond_r_t = _floats(s["BOND_EQUIL_VALUE"])
    angl_k_t = _floats(s["ANGLE_FORCE_CONSTANT"])
    angl_t_t = _floats(s["ANGLE_EQUIL_VALUE"])
    dihe_k_t = _floats(s["DIHEDRAL_FORCE_CONSTANT"])
    dihe_n_t = _floats(s["DIHEDRAL_PERIODICITY"])
    dihe_p_t = _floats(s["DIHEDRAL_PHASE"])
    nparm = len(dihe_k_t)
    scee_t = _floats(s["SCEE_SCALE_FACTOR"]) if "SCEE_SCALE_FACTOR" in s else \
        np.full(nparm, scee_default)
    scnb_t = _floats(s["SCNB_SCALE_FACTOR"]) if "SCNB_SCALE_FACTOR" in s else \
        np.full(nparm, scnb_default)

    b = np.concatenate([_ints(s["BONDS_INC_HYDROGEN"]), _ints(s["BONDS_WITHOUT_HYDROGEN"])])
    b = b.reshape(-1, 3)
    a = np.concatenate([_ints(s["ANGLES_INC_HYDROGEN"]), _ints(s["ANGLES_WITHOUT_HYDROGEN"])])
    a = a.reshape(-1, 4)
    d = np.concatenate([_ints(s["DIHEDRALS_INC_HYDROGEN"]),
                        _ints(s["DIHEDRALS_WITHOUT_HYDROGEN"])]).reshape(-1, 5)
    bonds = (np.abs(b[:, :2]) // 3).astype(np.int32)
    angles = (np.abs(a[:, :3]) // 3).astype(np.int32)
    dihedrals = (np.abs(d[:, :4]) // 3).astype(np.int32)
    bpar, apar, dpar = b[:, 2] - 1, a[:, 3] - 1, d[:, 4] - 1
    no14 = d[:, 2] < 0
    improper = (d[:, 3] < 0).astype(np.int32)
    scee = np.where(no14, 0.0, scee_t[dpar])
    scnb = np.where(no14, 0.0, scnb_t[dpar])
    # A 1:4 divisor of zero in the file means "no attenuated pair" as well.
    scee = np.where(scee_t[dpar] == 0.0, 0.0, scee)

    radii = _floats(s["RADII"]) if "RADII" in s else np.zeros(natom)
    screen = _floats(s["SCREEN"]) if "SCREEN" in s else np.zeros(natom)
    topo = Topology(
        natom=natom, charge=q, lj_idx=lj_idx, n_lj_types=ntypes,
        lja=expand(acoef), ljb=expand(bcoef), lja14=expand(a14), ljb14=expand(b14),
        bonds=bonds, bond_k=bond_k_t[bpar], bond_r0=bond_r_t[bpar],
        angles=angles, angle_k=angl_k_t[apar], angle_theta=angl_t_t[apar],
        dihedrals=dihedrals, dihe_amp=dihe_k_t[dpar], dihe_phase=dihe_p_t[dpar],
        dihe_freq=dihe_n_t[dpar], dihe_improper=improper, dihe_scee=scee, dihe_scnb=scnb,
        pb_radii=radii, gb_screen=screen, igb=igb, dielectric=dielectric, salt=salt,
        coulomb=coulomb, name=path)
    return topo


def read_inpcrd(path: str) -> np.ndarray:
    """Read AMBER inpcrd coordinates as an (N, 3) float64 array (12.7 fixed-width fields)."""
    with open(path, "r") as fh:
        lines = fh.read().split("\n")
    natom = int(lines[1].split()[0])
    vals: List[float] = []
    for ln in lines[2:]:
        k = 0
        while k + 12 <= len(ln) and len(vals) < 3 * natom:
            vals.append(float(ln[k:k + 12]))
            k += 12
        if len(vals) >= 3 * natom:
            break
    return np.array(vals[:3 * natom], dtype=np.float64).reshape(natom, 3)


def _write_section(fh, flag: str, fmt: str, values, kind: str, per_line: int, width: int):
    fh.write(f"%FLAG {flag:<74s}\n%FORMAT({fmt})\n")
    if len(values) == 0:
        fh.write("\n")
        return
    for start in range(0, len(values), per_line):
        chunk = values[start:start + per_line]
        if kind == "i":
            fh.write("".join(f"{int(v):{width}d}" for v in chunk) + "\n")
        elif kind == "e":
            fh.write("".join(f"{float(v):{width}.8E}" for v in chunk) + "\n")
        else:
            fh.write("".join(f"{str(v):<{width}s}" for v in chunk) + "\n")


def write_prmtop(topo: Topology, path: str, title: str = "synthetic") -> None:
    """Write an AMBER-style prmtop holding the sections this path reads back.

    Each unique bond / angle / dihedral parameter set becomes its own parameter type.  Charges
    are written multiplied by 18.2223.  The excluded-atom list is emitted from the shortest-path
    1:2/1:3/1:4 map (amber_prmtop_util / atomgraph_writing.cpp:82- is the reference writer).
    """
    n = topo.natom
    t = topo.n_lj_types
    lja, ljb, _, _ = topo.lj_tables()
    A = lja.reshape(t, t)
    B = ljb.reshape(t, t)
    nb_index = np.zeros((t, t), dtype=np.int64)
    acoef, bcoef = [], []
    for ti in range(t):
        for tj in range(ti + 1):
            acoef.append(A[ti, tj])
            bcoef.append(B[ti, tj])
            k = len(acoef)
            nb_index[ti, tj] = k
            nb_index[tj, ti] = k
    excl = topo.exclusion_pairs()
    per_atom: List[List[int]] = [[] for _ in range(n)]
    for i, j in excl:
        per_atom[int(i)].append(int(j) + 1)
    excl_counts, excl_list = [], []
    for i in range(n):
        lst = sorted(per_atom[i])
        if not lst:
            lst = [0]
        excl_counts.append(len(lst))
        excl_list.extend(lst)
    nb, na, nd = len(topo.bonds), len(topo.angles), len(topo.dihedrals)
    bond_words = []
    for k, (i, j) in enumerate(topo.bonds):
        bond_words += [3 * int(i), 3 * int(j), k + 1]
    angl_words = []
    for k, (i, j, kk) in enumerate(topo.angles):
        angl_words += [3 * int(i), 3 * int(j), 3 * int(kk), k + 1]
    dihe_words = []
    for k, (i, j, kk, l) in enumerate(topo.dihedrals):
        kw = 3 * int(kk) if topo.dihe_scee[k] > 0 else -3 * int(kk)
        lw = -3 * int(l) if topo.dihe_improper[k] else 3 * int(l)
        dihe_words += [3 * int(i), 3 * int(j), kw, lw, k + 1]
    ptrs = [n, t, 0, nb, 0, na, 0, nd, 0, 0, len(excl_list), 1, nb, na, nd, nb, na, nd, 0, 0,
            0, 0, 0, 0, 0, 0, 0, 0, n, 0, 0, 0]
    with open(path, "w") as fh:
        fh.write("%VERSION  VERSION_STAMP = V0001.000  DATE = 01/01/26  00:00:00\n")
        _write_section(fh, "TITLE", "20a4", [title[:80]], "a", 1, 80)
        _write_section(fh, "POINTERS", "10I8", ptrs, "i", 10, 8)
        _write_section(fh, "ATOM_NAME", "20a4", [f"A{i % 1000:<3d}"[:4] for i in range(n)],
                       "a", 20, 4)
        _write_section(fh, "CHARGE", "5E16.8", topo.charge * AMBER_CHARGE_SCALING, "e", 5, 16)
        _write_section(fh, "MASS", "5E16.8", np.full(n, 12.011), "e", 5, 16)
        _write_section(fh, "ATOM_TYPE_INDEX", "10I8", topo.lj_idx + 1, "i", 10, 8)
        _write_section(fh, "NUMBER_EXCLUDED_ATOMS", "10I8", excl_counts, "i", 10, 8)
        _write_section(fh, "NONBONDED_PARM_INDEX", "10I8", nb_index.reshape(-1), "i", 10, 8)
        _write_section(fh, "RESIDUE_LABEL", "20a4", ["MOL "], "a", 20, 4)
        _write_section(fh, "RESIDUE_POINTER", "10I8", [1], "i", 10, 8)
        _write_section(fh, "BOND_FORCE_CONSTANT", "5E16.8", topo.bond_k, "e", 5, 16)
        _write_section(fh, "BOND_EQUIL_VALUE", "5E16.8", topo.bond_r0, "e", 5, 16)
        _write_section(fh, "ANGLE_FORCE_CONSTANT", "5E16.8", topo.angle_k, "e", 5, 16)
        _write_section(fh, "ANGLE_EQUIL_VALUE", "5E16.8", topo.angle_theta, "e", 5, 16)
        _write_section(fh, "DIHEDRAL_FORCE_CONSTANT", "5E16.8", topo.dihe_amp, "e", 5, 16)
        _write_section(fh, "DIHEDRAL_PERIODICITY", "5E16.8", topo.dihe_freq, "e", 5, 16)
        _write_section(fh, "DIHEDRAL_PHASE", "5E16.8", topo.dihe_phase, "e", 5, 16)
        scee_w = np.where(topo.dihe_scee > 0, topo.dihe_scee, AMBER_SCEE)
        scnb_w = np.where(topo.dihe_scnb > 0, topo.dihe_scnb, AMBER_SCNB)
        _write_section(fh, "SCEE_SCALE_FACTOR", "5E16.8", scee_w, "e", 5, 16)
        _write_section(fh, "SCNB_SCALE_FACTOR", "5E16.8", scnb_w, "e", 5, 16)
        _write_section(fh, "SOLTY", "5E16.8", [0.0], "e", 5, 16)
        _write_section(fh, "LENNARD_JONES_ACOEF", "5E16.8", acoef, "e", 5, 16)
        _write_section(fh, "LENNARD_JONES_BCOEF", "5E16.8", bcoef, "e", 5, 16)
        _write_section(fh, "BONDS_INC_HYDROGEN", "10I8", [], "i", 10, 8)
        _write_section(fh, "BONDS_WITHOUT_HYDROGEN", "10I8", bond_words, "i", 10, 8)
        _write_section(fh, "ANGLES_INC_HYDROGEN", "10I8", [], "i", 10, 8)
        _write_section(fh, "ANGLES_WITHOUT_HYDROGEN", "10I8", angl_words, "i", 10, 8)
        _write_section(fh, "DIHEDRALS_INC_HYDROGEN", "10I8", [], "i", 10, 8)
        _write_section(fh, "DIHEDRALS_WITHOUT_HYDROGEN", "10I8", dihe_words, "i", 10, 8)
        _write_section(fh, "EXCLUDED_ATOMS_LIST", "10I8", excl_list, "i", 10, 8)
        _write_section(fh, "AMBER_ATOM_TYPE", "20a4", [f"T{k % 1000:<3d}"[:4] for k in
                                                       topo.lj_idx], "a", 20, 4)
        _write_section(fh, "RADIUS_SET", "1a80", ["synthetic radii"], "a", 1, 80)
        _write_section(fh, "RADII", "5E16.8", topo.pb_radii, "e", 5, 16)
        _write_section(fh, "SCREEN", "5E16.8", topo.gb_screen, "e", 5, 16)


def write_inpcrd(xyz: np.ndarray, path: str, title: str = "synthetic") -> None:
    xyz = np.asarray(xyz, dtype=np.float64).reshape(-1)
    with open(path, "w") as fh:
        fh.write(f"{title}\n{len(xyz) // 3:6d}\n")
        for start in range(0, len(xyz), 6):
            fh.write("".join(f"{v:12.7f}" for v in xyz[start:start + 6]) + "\n")
