"""CPU oracle for reading Wannier90 files (Wannier90::from_hr, NoRMatrix).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product path (``rustb_amd``) never imports, calls or links it.

A plain-Python restatement of ``src/wannier90.rs:49-1228`` of the reference
crate (LiuyichenYanwushang/Rustb v0.7.2) for ``Model<SPIN, 3, NoRMatrix>``:
``_hr.dat``, ``.win``, the optional ``_centres.xyz`` and ``_wsvec.dat``, and the
checks of ``Model::validate`` (``model.rs:564-667``).  It reads line by line, as
the reference does, so it is meant for small files.  Text follows Rust's
``str`` rules: ``lines()`` splits at ``\\n`` and drops one ``\\r`` before it;
``split_whitespace``/``trim`` use Unicode White_Space (not Python's wider
set); integers and floats follow Rust's ``FromStr`` grammar, and floats are
rounded correctly (Python's ``float`` on a string Rust accepts).  Where the
reference panics (index beyond the file or beyond nsta), this raises
``W90Error("FileParse", ...)``.

Pinning status (``tests/test_wannier90.py``):

* the reference's own regression tests (``wannier90.rs:1230-1518``: minimal
  dataset, dropped atoms without orbitals, species mismatch, unknown
  projection species, comment stripping, atoms_cart fallback, multi-atom
  and multi-line projections, length units and atoms_frac, malformed atom
  rows, truncated centres) with their expected counts, positions and errors;
* closed forms: storage of seeded non-Hermitian blocks written in the file
  layout (element ``<m,0|H|n,R>`` on line (n outer, m inner) lands at
  ``ham[R][m][n]`` divided by its degeneracy, bit for bit), R = 0 moved to
  row 0 wherever it sits, the Fermi level on the R = 0 diagonal only; the
  simple-cubic s band ``eps - 2 t (cos 2pi k1 + cos 2pi k2 + cos 2pi k3)``
  from a written model; Wigner-Seitz shifts that split a hopping over
  equivalent R leave H(k) at shift-invariant k unchanged and reproduce the
  hand-computed storage of a two-orbital case;
* number grammar: tokens Rust accepts or rejects (signs, ``.5``, ``5.``,
  exponents, ``inf``/``nan``, Fortran ``D`` exponents, hex) against their
  known values.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

BOHR_TO_ANGSTROM = 0.529_177_210_67  # wannier90.rs:9

# AtomType::from_str (atom_struct.rs:485-580): H .. Ra, case-sensitive.
ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra"
).split()

# OrbProj (atom_struct.rs:15-77), in declaration order.
ORB_PROJ = (
    "s px py pz dxy dyz dxz dz2 dx2y2 fz3 fxz2 fyz2 fzx2y2 fxyz fxx23y2 fy3x2y2 sp_1 sp_2 sp2_1 sp2_2 sp2_3 "
    "sp3_1 sp3_2 sp3_3 sp3_4 sp3d_1 sp3d_2 sp3d_3 sp3d_4 sp3d_5 sp3d2_1 sp3d2_2 sp3d2_3 sp3d2_4 sp3d2_5 sp3d2_6"
).split()

# the projection items of wannier90.rs:355-420 and their orbitals
PROJECTIONS = {
    "s": ["s"],
    "p": ["pz", "px", "py"],
    "d": ["dz2", "dxz", "dyz", "dx2y2", "dxy"],
    "f": ["fz3", "fxz2", "fyz2", "fzx2y2", "fxyz", "fxx23y2", "fy3x2y2"],
    "sp3": ["sp3_1", "sp3_2", "sp3_3", "sp3_4"],
    "sp2": ["sp2_1", "sp2_2", "sp2_3"],
    "sp": ["sp_1", "sp_2"],
    "sp3d": ["sp3d_1", "sp3d_2", "sp3d_3", "sp3d_4", "sp3d_5"],
    "sp3d2": ["sp3d2_1", "sp3d2_2", "sp3d2_3", "sp3d2_4", "sp3d2_5", "sp3d2_6"],
    "px": ["px"],
    "py": ["py"],
    "pz": ["pz"],
    "dxy": ["dxy"],
    "dxz": ["dxz"],
    "dyz": ["dyz"],
    "dz2": ["dz2"],
    "dx2-y2": ["dx2y2"],
}

# Rust char::is_whitespace (Unicode White_Space)
_WS = ("\t\n\x0b\x0c\r \x85\xa0\u1680\u2000\u2001\u2002\u2003\u2004\u2005\u2006\u2007\u2008"
       "\u2009\u200a\u2028\u2029\u202f\u205f\u3000")
_WS_RE = re.compile("[" + re.escape(_WS) + "]+")
_FLOAT_RE = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)\Z", re.ASCII)
_INFNAN_RE = re.compile(r"[+-]?(?:inf|infinity|nan)\Z", re.IGNORECASE)
_INT_RE = re.compile(r"[+-]?\d+\Z", re.ASCII)


class W90Error(Exception):
    """kind: the TbError variant (FileParse, FileCreation, Io, ...)."""

    def __init__(self, kind: str, message: str):
        self.kind = kind
        super().__init__(f"{kind}: {message}")


def _split(s):
    return [t for t in _WS_RE.split(s) if t]


def _trim(s):
    return s.strip(_WS)


def _parse_usize(tok):
    if tok == "":
        return None, "cannot parse integer from empty string"
    if not _INT_RE.match(tok) or tok[0] == "-":
        return None, "invalid digit found in string"
    v = int(tok)
    if v >= 1 << 64:
        return None, "number too large to fit in target type"
    return v, ""


def _parse_isize(tok):
    if tok == "":
        return None, "cannot parse integer from empty string"
    if not _INT_RE.match(tok):
        return None, "invalid digit found in string"
    v = int(tok)
    if v >= 1 << 63:
        return None, "number too large to fit in target type"
    if v < -(1 << 63):
        return None, "number too small to fit in target type"
    return v, ""


def parse_f64(tok):
    """<f64 as FromStr>; None where Rust reports "invalid float literal"."""
    if _INFNAN_RE.match(tok):
        return float(tok)
    if not _FLOAT_RE.match(tok):
        return None
    return float(tok)


def _read_lines(path):
    """BufRead::lines over the file (UnicodeDecodeError for non-UTF-8 text)."""
    with open(path, "rb") as f:
        text = f.read().decode("utf-8")
    parts = text.split("\n")
    if parts[-1] == "":
        parts.pop()
    # one '\r' before a '\n' is dropped; a last line without '\n' keeps it
    return [p[:-1] if p.endswith("\r") and (i < len(parts) - 1 or text.endswith("\n")) else p
            for i, p in enumerate(parts)]


def win_data_line(line):
    """wannier90.rs:11-17."""
    line = line.split("!")[0]
    return _trim(line.split("#")[0])


def _length_unit_scale(tok):
    """wannier90.rs:19-25."""
    t = tok.lower()
    if t in ("ang", "angstrom", "angstroms"):
        return 1.0
    if t == "bohr":
        return BOHR_TO_ANGSTROM
    return None


@dataclass
class W90Model:
    spin: bool
    lat: np.ndarray
    orb: np.ndarray
    orb_projection: list
    atoms: list  # (species, fractional position, [orbital indices])
    hamR: np.ndarray
    ham: np.ndarray
    nsta: int = field(default=0)
    rmatrix: np.ndarray = None  # [N_R, 3, n, n] with _r.dat

    @property
    def norb(self):
        return self.orb.shape[0]


def from_hr(path, file_name, zero_energy, spin=None, rmatrix=False):
    """Wannier90::from_hr (wannier90.rs:49-1228); spin None takes the file's;
    rmatrix=True is R = HasRMatrix (from_hr_with_rmatrix, wannier90.rs:1208)."""
    prefix = path + file_name
    hr_path = prefix + "_hr.dat"

    def fp(file, msg):
        raise W90Error("FileParse", f"Failed to parse data from file '{file}': {msg}")

    try:
        reads = _read_lines(hr_path)
    except UnicodeDecodeError:
        raise W90Error("Io", "I/O error: stream did not contain valid UTF-8")
    except OSError as e:
        raise W90Error("FileCreation", f"Failed to create file '{hr_path}': Unable to open HR file: {e}")

    def line(i):
        if i >= len(reads):
            fp(hr_path, "Truncated _hr.dat")
        return reads[i]

    # wannier90.rs:91-125
    nsta, e = _parse_usize(_trim(line(1)))
    if e:
        fp(hr_path, f"Failed to parse nsta: {e}")
    n_R, e = _parse_usize(_trim(line(2)))
    if e:
        fp(hr_path, f"Failed to parse n_R: {e}")
    weights = []
    n_line = 0
    for i in range(3, len(reads)):
        if "." in reads[i]:
            n_line = i
            break
        for t in _split(_trim(reads[i])):
            v, e = _parse_usize(t)
            if e:
                fp(hr_path, f"Failed to parse weight: {e}")
            weights.append(v)

    # wannier90.rs:127-245: R = 0 overwrites row 0, other R appended
    hamR = [(0, 0, 0)]
    ham = [np.zeros((nsta, nsta), complex)]
    for i in range(n_R):
        toks = _split(_trim(line(i * nsta * nsta + n_line)))
        Rv = []
        for d in range(3):
            if d >= len(toks):
                fp(hr_path, "Missing R vector component")
            v, e = _parse_isize(toks[d])
            if e:
                fp(hr_path, f"Failed to parse R vector: {e}")
            Rv.append(v)
        mat = np.zeros((nsta, nsta), complex)
        for ind_i in range(nsta):
            for ind_j in range(nsta):
                toks = _split(_trim(line(i * nsta * nsta + ind_i * nsta + ind_j + n_line)))
                if len(toks) < 6:
                    fp(hr_path, "Missing Hamiltonian real part")
                re_ = parse_f64(toks[5])
                if re_ is None:
                    fp(hr_path, "Failed to parse Hamiltonian real part: invalid float literal")
                if len(toks) < 7:
                    fp(hr_path, "Missing Hamiltonian imaginary part")
                im_ = parse_f64(toks[6])
                if im_ is None:
                    fp(hr_path, "Failed to parse Hamiltonian imaginary part: invalid float literal")
                if i >= len(weights):
                    fp(hr_path, "Truncated _hr.dat: missing degeneracy weight")
                w = float(weights[i])
                mat[ind_j, ind_i] = complex(_div(re_, w), _div(im_, w))
        if Rv == [0, 0, 0]:
            ham[0] = mat
        else:
            ham.append(mat)
            hamR.append(tuple(Rv))
    for i in range(nsta):  # wannier90.rs:247-249
        ham[0][i, i] = complex(ham[0][i, i].real - zero_energy, ham[0][i, i].imag - 0.0)

    # .win (wannier90.rs:250-559)
    win_path = prefix + ".win"
    try:
        wl = _read_lines(win_path)
    except UnicodeDecodeError:
        raise W90Error("Io", "I/O error: stream did not contain valid UTF-8")
    except OSError as e:
        raise W90Error("FileCreation", f"Failed to create file '{win_path}': Unable to open win file: {e}")
    it = iter(wl)
    lat = np.zeros((3, 3))
    file_spin = False
    proj_name, proj_list, atom_proj = [], [], []
    atom_name, atom_pos = [], []

    def nxt(msg="Unexpected end of file"):
        try:
            return next(it)
        except StopIteration:
            fp(win_path, msg)

    def coord(tok, ctx):
        v = parse_f64(tok)
        if v is None:
            fp(win_path, f"Failed to parse {ctx}: invalid float literal")
        return v

    for a in it:
        key = win_data_line(a).lower()
        if "begin unit_cell_cart" in key:
            unit, rows = 1.0, []
            while len(rows) < 3:
                ln = win_data_line(nxt("Missing lattice vector line"))
                if not ln:
                    continue
                toks = _split(ln)
                if not rows and len(toks) == 1:
                    unit = _length_unit_scale(toks[0])
                    if unit is None:
                        fp(win_path, f"Unknown unit_cell_cart unit '{toks[0]}'")
                    continue
                if len(toks) != 3:
                    fp(win_path, f"A unit_cell_cart row must contain 3 numbers, found {len(toks)} in '{ln}'")
                rows.append([coord(t, "lattice vector") * unit for t in toks])
            lat = np.array(rows, dtype=float)
        elif "spinors" in key and ("t" in key or "true" in key):
            file_spin = True
        elif "begin projections" in key:
            while True:
                s = win_data_line(nxt())
                if not s:
                    continue
                if "end projections" in s.lower():
                    break
                prj = [_trim(x) for x in re.split("[,;:]", s)]
                if len(prj) < 2 or prj[0] == "":
                    fp(win_path, f"Malformed projection line '{s}': expected species:orbital")
                orbs = []
                for item in prj[1:]:
                    if item not in PROJECTIONS:
                        raise W90Error("InvalidOrbitalProjection",
                                       f"Invalid orbital projection string: 'Unrecognized projection '{item}' "
                                       "in seedname.win'")
                    orbs += PROJECTIONS[item]
                proj_list.append(len(orbs))
                atom_proj.append(orbs)
                if prj[0] not in ELEMENTS:
                    fp(win_path, f"Unknown atomic species '{prj[0]}' in begin projections")
                proj_name.append(prj[0])
        elif "begin atoms_cart" in key:
            unit, first = 1.0, True
            while True:
                s = win_data_line(nxt())
                if not s:
                    continue
                if "end atoms_cart" in s.lower():
                    break
                f = _split(s)
                if first and len(f) == 1:
                    unit = _length_unit_scale(f[0])
                    if unit is None:
                        fp(win_path, f"Unknown atoms_cart unit '{f[0]}'")
                    first = False
                    continue
                first = False
                if len(f) != 4:
                    fp(win_path, "An atoms_cart row must contain a species and 3 coordinates, found "
                                 f"{len(f)} fields in '{s}'")
                atom_name.append(f[0])
                atom_pos.append([coord(t, "Cartesian atom position") * unit for t in f[1:]])
        elif "begin atoms_frac" in key:
            while True:
                s = win_data_line(nxt())
                if not s:
                    continue
                if "end atoms_frac" in s.lower():
                    break
                f = _split(s)
                if len(f) != 4:
                    fp(win_path, "An atoms_frac row must contain a species and 3 coordinates, found "
                                 f"{len(f)} fields in '{s}'")
                atom_name.append(f[0])
                frac = np.array([coord(t, "fractional atom position") for t in f[1:]])
                atom_pos.append(list(frac @ lat))
    if spin is not None and bool(spin) != file_spin:  # wannier90.rs:559-564
        raise W90Error("Other", f"Unexpected error: Spin mismatch: Wannier90 .win file has spin="
                                f"{str(file_spin).lower()} but Model was constructed with SPIN={str(bool(spin)).lower()}")

    def inv_lat():  # LAPACK getrf/getri, singular on an exact zero pivot
        try:
            return np.linalg.inv(lat)
        except np.linalg.LinAlgError:
            raise W90Error("Linalg", "Linear algebra operation failed: the lattice matrix is singular")

    # _centres.xyz (wannier90.rs:565-711) or the .win atoms (wannier90.rs:712-782)
    xyz_path = prefix + "_centres.xyz"
    orb_proj, atoms = [], []
    try:
        xl = _read_lines(xyz_path)
        have_xyz = True
    except UnicodeDecodeError:
        fp(xyz_path, "Failed to read line: stream did not contain valid UTF-8")
    except OSError:
        have_xyz = False
    if have_xyz:
        if len(xl) < 2:
            fp(xyz_path, "_centres.xyz must contain a count and comment line")
        declared, e = _parse_usize(_trim(xl[0]))
        if e:
            fp(xyz_path, f"Invalid _centres.xyz entry count: {e}")
        available = len(xl) - 2
        if declared < nsta or available < declared:
            fp(xyz_path, f"Truncated _centres.xyz: header declares {declared} entries, HR requires {nsta} "
                         f"Wannier-centre entries, but only {available} data lines are present")
        norb = nsta // 2 if file_spin else nsta
        cart = np.zeros((norb, 3))
        for i in range(norb):
            f = _split(xl[i + 2])
            if len(f) < 4:
                fp(xyz_path, f"Malformed Wannier-centre row {i + 1}: expected a label and 3 coordinates")
            for d in range(3):
                v = parse_f64(f[d + 1])
                if v is None:
                    fp(xyz_path, "Failed to parse Wannier-centre position: invalid float literal")
                cart[i, d] = v
        inv = inv_lat()
        orb = cart @ inv
        natom = declared - nsta
        new_pos, new_name = [], []
        for i in range(natom):
            f = _split(xl[i + 2 + nsta])
            if len(f) < 4:
                fp(xyz_path, f"Malformed atom row {i + 1} in _centres.xyz: expected a species and 3 coordinates")
            pos = []
            for d in range(3):
                v = parse_f64(f[d + 1])
                if v is None:
                    fp(xyz_path, "Failed to parse atom position: invalid float literal")
                pos.append(v)
            if f[0] not in ELEMENTS:
                fp(xyz_path, f"Unknown atomic species '{f[0]}' in _centres.xyz; species must be listed in the "
                             "win file's begin projections block")
            new_pos.append(pos)
            new_name.append(f[0])
        for i, name in enumerate(new_name):
            idx = []
            for j, jn in enumerate(proj_name):
                if jn == name:
                    idx += list(range(len(orb_proj), len(orb_proj) + proj_list[j]))
                    orb_proj += atom_proj[j]
            if not idx:
                continue
            atoms.append((name, np.array(new_pos[i]) @ inv, idx))
        if len(orb_proj) != norb:
            got = len(orb_proj)
            detail = (f"{norb} orbitals declared in the projections block, but only {got} could be assigned to atoms"
                      if got < norb else
                      f"{norb} orbitals declared in the projections block, but {got} were assigned to atoms")
            fp(xyz_path, f"species mismatch between _centres.xyz and the win projections block: {detail}")
    else:
        inv = inv_lat()
        pos_frac = np.array(atom_pos).reshape(-1, 3) @ inv
        rows = []
        for i, name in enumerate(atom_name):
            if name not in ELEMENTS:
                fp(win_path, f"Unknown atomic species '{name}' in begin atoms_frac block")
            idx = []
            for j, jn in enumerate(proj_name):
                if jn == name:
                    idx += list(range(len(orb_proj), len(orb_proj) + proj_list[j]))
                    orb_proj += atom_proj[j]
                    rows += [pos_frac[i]] * proj_list[j]
            if not idx:
                continue
            atoms.append((name, pos_frac[i].copy(), idx))
        expected = nsta // 2 if file_spin else nsta
        if len(orb_proj) != expected:
            got = len(orb_proj)
            detail = (f"{expected} orbitals declared in the HR file, but only {got} could be assigned to atoms"
                      if got < expected else
                      f"{expected} orbitals declared in the HR file, but {got} were assigned to atoms")
            fp(win_path, f"species mismatch between begin atoms_frac and the projections block: {detail}")
        orb = np.array(rows, dtype=float).reshape(-1, 3)

    # _r.dat (wannier90.rs:805-902)
    rmat = None
    if rmatrix:
        r_path = prefix + "_r.dat"
        try:
            rl = _read_lines(r_path)
        except UnicodeDecodeError:
            fp(xyz_path, "Failed to read line: stream did not contain valid UTF-8")
        except OSError:
            raise W90Error("FileCreation", f"Failed to create file '{r_path}': R::HAS_RMATRIX=true but _r.dat "
                                           "file not found")
        if len(rl) < 3:
            fp(r_path, "Truncated _r.dat: missing the R-vector count line")
        nRr, e = _parse_usize(_trim(rl[2]))
        if e:
            fp(r_path, f"Failed to parse n_R: {e}")
        rmat = np.zeros((len(hamR), 3, nsta, nsta), complex)
        index_of = {}
        for r, Rv in enumerate(hamR):
            index_of.setdefault(Rv, r)

        def rline(i):
            if i >= len(rl):
                fp(r_path, "Truncated _r.dat")
            return rl[i]

        for i in range(nRr if nsta else 0):
            t = _split(_trim(rline(i * nsta * nsta + 3)))
            Rv = []
            for d in range(3):
                if d >= len(t):
                    fp(r_path, "Missing R vector component")
                v, e = _parse_isize(t[d])
                if e:
                    fp(r_path, f"Failed to parse R vector: {e}")
                Rv.append(v)
            if tuple(Rv) not in index_of:
                fp(r_path, f"R vector [{Rv[0]}, {Rv[1]}, {Rv[2]}] not found in Hamiltonian")
            idx = index_of[tuple(Rv)]
            if i >= len(weights):
                fp(r_path, "Truncated _r.dat: missing degeneracy weight")
            w = float(weights[i])
            for ind_i in range(nsta):
                for ind_j in range(nsta):
                    t = _split(_trim(rline(i * nsta * nsta + ind_i * nsta + ind_j + 3)))
                    for d in range(3):
                        vals = []
                        for c, part in ((0, "real"), (1, "imaginary")):
                            pos = 5 + 2 * d + c
                            if pos >= len(t):
                                fp(r_path, f"Missing R matrix {part} part")
                            x = parse_f64(t[pos])
                            if x is None:
                                fp(r_path, f"Failed to parse R matrix {part} part: invalid float literal")
                            vals.append(x)
                        rmat[idx, d, ind_j, ind_i] = complex(_div(vals[0], w), _div(vals[1], w))

    # _wsvec.dat (wannier90.rs:904-1133; with an r-matrix it moves with the hopping)
    ws_path = prefix + "_wsvec.dat"
    try:
        ws = _read_lines(ws_path)
    except UnicodeDecodeError:  # the reference names the _centres.xyz path here
        fp(xyz_path, "Failed to read line: stream did not contain valid UTF-8")
    except OSError:
        ws = None
    if ws is not None:
        if not ws:
            fp(ws_path, "Truncated _wsvec.dat: empty file")
        new_R = [(0, 0, 0)]
        new_ham = [np.zeros((nsta, nsta), complex)]
        new_rm = [np.zeros((3, nsta, nsta), complex)] if rmat is not None else None
        index_of = {}
        for r, Rv in enumerate(hamR):
            index_of.setdefault(Rv, r)  # find_R: first match
        i = 0

        def wline(k):
            if k >= len(ws):
                fp(ws_path, f"Truncated _wsvec.dat: missing line {k + 1}")
            return ws[k]

        while i < len(ws) - 1:
            i += 1
            t = _split(_trim(wline(i)))
            Rv = []
            for d in range(3):
                if d >= len(t):
                    fp(ws_path, "Missing R vector component")
                v, e = _parse_isize(t[d])
                if e:
                    fp(ws_path, f"Failed to parse R vector: {e}")
                Rv.append(v)
            oi = []
            for d in range(2):
                if 3 + d >= len(t):
                    fp(ws_path, "Missing orbital index")
                v, e = _parse_usize(t[3 + d])
                if e:
                    fp(ws_path, f"Failed to parse orbital index: {e}")
                if v == 0 or v > nsta:
                    fp(ws_path, f"orbital index {v} out of range")
                oi.append(v - 1)
            i += 1
            t = _split(_trim(wline(i)))
            if not t:
                fp(ws_path, "Missing weight value")
            weight, e = _parse_usize(t[0])
            if e:
                fp(ws_path, f"Failed to parse weight: {e}")
            Rt = tuple(Rv)
            if Rt not in index_of:
                fp(ws_path, f"R vector [{Rv[0]}, {Rv[1]}, {Rv[2]}] not found in Hamiltonian")
            h = ham[index_of[Rt]][oi[0], oi[1]]
            hop = complex(_div(h.real, float(weight)), _div(h.imag, float(weight)))
            if rmat is not None:
                hr = [rmat[index_of[Rt], d, oi[0], oi[1]] for d in range(3)]
                hop_r = [complex(_div(x.real, float(weight)), _div(x.imag, float(weight))) for x in hr]
            for _ in range(weight):
                i += 1
                t = _split(_trim(wline(i)))
                sh = []
                for d in range(3):
                    v, e = _parse_isize(t[d]) if d < len(t) else (None, "missing")
                    if e:
                        fp(ws_path, f"malformed shift line {i + 1}")
                    sh.append(v)
                nR = (Rv[0] + sh[0], Rv[1] + sh[1], Rv[2] + sh[2])
                if nR in new_R:
                    k = new_R.index(nR)
                    new_ham[k][oi[0], oi[1]] += hop
                    if rmat is not None:
                        for d in range(3):
                            new_rm[k][d, oi[0], oi[1]] += hop_r[d]
                else:
                    blk = np.zeros((nsta, nsta), complex)
                    if rmat is not None:  # this branch adds to zero (wannier90.rs:1026-1031)
                        blk[oi[0], oi[1]] += hop
                        rb = np.zeros((3, nsta, nsta), complex)
                        for d in range(3):
                            rb[d, oi[0], oi[1]] += hop_r[d]
                        new_rm.append(rb)
                    else:
                        blk[oi[0], oi[1]] = hop
                    new_R.append(nR)
                    new_ham.append(blk)
        hamR, ham = new_R, new_ham
        if rmat is not None:
            rmat = np.array(new_rm)

    if rmat is not None:  # wannier90.rs:1134-1160, rows 0 .. nR - 2 only
        pos = {}
        for r, Rv in enumerate(hamR):
            pos.setdefault(tuple(Rv), r)
        for r in range(len(hamR) - 1):
            Rv = tuple(hamR[r])
            inv = tuple(-x for x in Rv)
            if inv not in pos:
                raise W90Error("MissingHermitianConjugate", f"Missing Hermitian conjugate for R vector: {list(Rv)}")
            q = pos[inv]
            for i in range(nsta):
                for j in range(nsta):
                    for d in range(3):
                        a, b = rmat[r, d, i, j], rmat[q, d, j, i]
                        a = complex((a.real + b.real) / 2.0, (a.imag + -b.imag) / 2.0)
                        rmat[r, d, i, j] = a
                        rmat[q, d, j, i] = complex(a.real, -a.imag)

    m = W90Model(spin=file_spin, lat=lat, orb=np.asarray(orb, float).reshape(-1, 3), orb_projection=orb_proj,
                 atoms=atoms, hamR=np.array(hamR, dtype=np.int64).reshape(-1, 3),
                 ham=np.array(ham, dtype=complex).reshape(-1, nsta, nsta), nsta=nsta, rmatrix=rmat)
    _validate(m)
    return m


def _div(a, b):
    """IEEE a / b, including b == 0 (inf/nan as Rust's f64 division)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))


def _validate(m):
    """Model::validate (model.rs:564-667) for the loaded arrays."""
    if m.norb == 0:
        raise W90Error("NoOrbitals", "a model must contain at least one orbital")
    if not (np.all(np.isfinite(m.lat)) and np.all(np.isfinite(m.orb))):
        raise W90Error("InvalidModelInvariant", "finite_geometry")
    for a, (_, pos, _) in enumerate(m.atoms):
        if not np.all(np.isfinite(pos)):
            raise W90Error("InvalidModelInvariant", f"atomic_position: atom {a}")
    for a, (_, pos, idx) in enumerate(m.atoms):
        for o in idx:
            for d in range(3):
                delta = abs(m.orb[o, d] - pos[d]) % 1.0
                if min(delta, 1.0 - delta) > 0.1:
                    raise W90Error("InvalidModelInvariant", f"orbital_atom_position: orbital {o}, atom {a}")
    nsta_model = 2 * m.norb if m.spin else m.norb
    if nsta_model != m.nsta:
        raise W90Error("InvalidModelInvariant", "hamiltonian_shape")
