"""Physical data for the reacting-flow path: nondimensional constants, the
flow solver's thermodynamic/transport tables (``thd_type``) and the CHEMKIN
mechanism tables used by the kinetics kernels.

Everything here is host-side table construction; the numbers are handed to
the HIP library as flat arrays.  Each constant is computed with the same
expression order as the reference so that the doubles are identical.

Reference (saras13015/Phd, devel-izem/src):
  constants()            module_tools.f90:453-562, suth :573-581
  thd_init               thd/thd_tools_perfectgas.f90:12-248
  therm.dat reader       thd/thd_tools_perfectgas.f90:367-473
  tranfit.out reader     thd/thd_tools_perfectgas.f90:478-638
  chem.inp interpreter   chemkin/ckinterp39.f (unit conversion :2415-2430)
  CHEMKIN constants      chemkin/vode.f:462-479
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field

import numpy as np

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
NRV = 9
NIV = 5
NV = NIV + NRV

# CHEMKIN-II atomic weights as the interpreter stores them: the DATA table of
# ckinterp39.f holds default-REAL literals, so each weight is the float32
# value widened to double.
_AWT = {"H": 1.00797, "O": 15.9994, "N": 14.0067, "C": 12.01115, "AR": 39.948, "HE": 4.00260}


def _f32(x: float) -> float:
    return float(np.float32(x))


# --------------------------------------------------------------------------
# nondimensionalisation
# --------------------------------------------------------------------------
@dataclass
class Adim:
    """adi_type as filled by constants() (module_tools.f90:453-562)."""

    gamma_inf: float = 1.40
    re: float = 65000.0
    pr: float = 0.72
    le: float = 1.0
    ma: float = 2.25
    T_inf: float = 169.44
    L_ref: float = 0.00004839025546795050
    R_ref: float = 8.31451
    r_m_ref: float = 287.15

    def __post_init__(self):
        g = self.gamma_inf
        self.gamma = g
        self.gm1 = g - 1.0
        self.gm1_i = 1.0 / self.gm1
        self.ttrd = 2.0 / 3.0
        self.ggmopr = g * self.gm1_i / self.pr
        self.sqgmr = math.sqrt(g) * self.ma / self.re
        self.Sc = self.le * self.pr
        self.T_ref = self.T_inf
        self.W_ref = self.R_ref / self.r_m_ref
        self.cp_ref = g * (self.R_ref / self.W_ref) / (g - 1.0)
        T = self.T_ref
        if T < 110.4:
            raise NotImplementedError("vand() branch of constants() not needed for T_ref=169.44")
        self.mu_ref = 1.4580e-6 * T ** (1.5) / (T + 110.4)
        self.c_inf = math.sqrt(g * (self.R_ref / self.W_ref) * self.T_inf)
        self.u_ref = self.c_inf / math.sqrt(g)
        self.u_inf = self.c_inf * self.ma
        self.time_ref = self.L_ref / self.u_ref
        self.rho_ref = self.re * self.mu_ref / (self.L_ref * self.u_inf)
        self.p_ref = self.rho_ref * (self.u_ref * self.u_ref)
        self.lbda_ref = self.cp_ref * self.mu_ref / self.pr
        self.D_ref = self.lbda_ref / (self.rho_ref * self.cp_ref * self.le)
        # energy scale of the nondimensional variables: R T_ref / W_ref = u_ref^2
        self.e_ref = self.u_ref * self.u_ref


# --------------------------------------------------------------------------
# therm.dat
# --------------------------------------------------------------------------
@dataclass
class NasaEntry:
    name: str
    tlo: float
    thi: float
    tmid: float       # 0.0 when the species line leaves columns 66-75 blank
    high: list        # 7 coefficients, T > Tmid
    low: list         # 7 coefficients


def read_therm(path: str | None = None) -> tuple[dict, tuple[float, float, float]]:
    """Parse a CHEMKIN therm.dat.  Field positions follow the fixed format the
    reference reader uses (thd_tools_perfectgas.f90:418-447: a18,a6,a20,a1,
    3e10.0,I5 then 5E15 / 5E15 / 4E15)."""
    path = path or os.path.join(DATA_DIR, "therm.dat")
    with open(path) as f:
        lines = f.read().splitlines()
    if not lines[0].strip().upper().startswith("THERMO"):
        raise ValueError("therm.dat must start with THERMO")
    defaults = tuple(float(lines[1][10 * q:10 * q + 10]) for q in range(3))
    out = {}
    i = 2
    while i < len(lines):
        ln = lines[i]
        if ln.strip().upper().startswith("END"):
            break
        if len(ln) < 80 or ln[79] != "1":
            i += 1
            continue
        name = ln[0:18].split()[0]

        def fnum(s):
            s = s.strip()
            return float(s) if s else 0.0

        tlo, thi, tmid = fnum(ln[45:55]), fnum(ln[55:65]), fnum(ln[65:75])
        c = []
        for q in range(1, 4):
            row = lines[i + q]
            n = 5 if q < 3 else 4
            c += [float(row[15 * p:15 * p + 15].replace("D", "E")) for p in range(n)]
        # the first entry of a species is the one read (the reference's reader
        # stops at the first match, thd_tools_perfectgas.f90:409-440)
        out.setdefault(name.upper(), NasaEntry(name, tlo, thi, tmid, c[0:7], c[7:14]))
        i += 4
    return out, defaults


# --------------------------------------------------------------------------
# tranfit.out
# --------------------------------------------------------------------------
@dataclass
class TransportFits:
    names: list
    cofd: np.ndarray    # [l][m][n]  = CofD(n,l,m)
    coflam: np.ndarray  # [l][n]
    cofeta: np.ndarray  # [l][n]
    W: np.ndarray       # kg/mol
    awt: np.ndarray     # kg/mol
    kncf: np.ndarray    # [element][species]


def read_tranfit(path: str | None = None) -> TransportFits:
    """Reader with the semantics of thd_cofd_read (thd_tools_perfectgas.f90:478-638)."""
    path = path or os.path.join(DATA_DIR, "tranfit.out")
    tok = open(path).read().split()
    p = 0

    def nxt():
        nonlocal p
        p += 1
        return tok[p - 1]

    nreac, nspec, npoly, nele = int(nxt()), int(nxt()), int(nxt()), int(nxt())
    cofd_lu = np.zeros((nspec, nspec, npoly))
    names = [None] * nspec
    for j in range(nspec):
        if int(nxt()) != j + 1:
            raise ValueError("tranfit.out: bad species index")
        names[j] = nxt()
        for k in range(j + 1):
            nxt()
            for n in range(npoly):
                cofd_lu[j, k, n] = float(nxt())
    if int(nxt()) != -2:
        raise ValueError("tranfit.out: missing separator")
    cofd = np.zeros((nspec, nspec, npoly))
    for j in range(nspec):
        for k in range(nspec):
            cofd[j, k] = cofd_lu[j, k] if k <= j else cofd_lu[k, j]
    coflam = np.zeros((nspec, npoly))
    for k in range(nspec):
        nxt()
        for n in range(npoly):
            coflam[k, n] = float(nxt())
    if int(nxt()) != -2:
        raise ValueError("tranfit.out: missing separator")
    cofeta = np.zeros((nspec, npoly))
    for k in range(nspec):
        nxt()
        for n in range(npoly):
            cofeta[k, n] = float(nxt())
    if int(nxt()) != -2:
        raise ValueError("tranfit.out: missing separator")
    W = np.zeros(nspec)
    for k in range(nspec):
        nxt()
        W[k] = float(nxt())
    W = W / 1000.0
    nxt()
    awt = np.array([float(nxt()) for _ in range(nele)]) / 1000.0
    if int(nxt()) != -2:
        raise ValueError("tranfit.out: missing separator")
    kncf = np.zeros((nele, nspec), dtype=np.int64)
    for k in range(nspec):
        for j in range(nele):
            kncf[j, k] = int(nxt())
    if int(nxt()) != -1:
        raise ValueError("tranfit.out: missing end mark")
    return TransportFits(names, cofd, coflam, cofeta, W, awt, kncf)


# --------------------------------------------------------------------------
# thd tables for the flow kernels
# --------------------------------------------------------------------------
@dataclass
class Thd:
    names: list
    W: np.ndarray
    Wc: np.ndarray
    Wc_i: np.ndarray
    heatcap: np.ndarray      # [species][range 0 low, 1 high][7]
    tpolymin: np.ndarray
    tpolylim: np.ndarray
    tpolymax: np.ndarray
    cofd: np.ndarray         # [l][m][4]
    coflam: np.ndarray       # [l][4]
    cofeta: np.ndarray
    wilke_c1: np.ndarray     # [m][l] = (W(l)/W(m))**0.25
    wilke_c2: np.ndarray     # [m][l] = 1/sqrt((1+W(m)/W(l))*8)
    gam2: float
    Tinf: float
    p0: float
    pinf: float
    Dinf: float
    laminf: float
    muinf: float
    kncf: np.ndarray
    awt: np.ndarray

    def flat(self) -> np.ndarray:
        """Flat float64 image in the order of the C struct ``izem_thd``."""
        parts = [self.W, self.Wc, self.Wc_i, self.heatcap.ravel(), self.tpolymin, self.tpolylim,
                 self.tpolymax, self.cofd.ravel(), self.coflam.ravel(), self.cofeta.ravel(),
                 self.wilke_c1.ravel(), self.wilke_c2.ravel(),
                 np.array([self.gam2, self.Tinf, self.p0, self.pinf, self.Dinf, self.laminf,
                           self.muinf])]
        return np.ascontiguousarray(np.concatenate(parts), dtype=np.float64)


def build_thd(adi: Adim, therm_path=None, tranfit_path=None,
              fits: "TransportFits | None" = None) -> Thd:
    """thd_init (thd_tools_perfectgas.f90:12-248).  The transport fits come
    from a tranfit.out file (the bundled one by default), or are given
    (phd_amd.tranfit.fit_transport)."""
    tr = fits if fits is not None else read_tranfit(tranfit_path)
    nasa, _ = read_therm(therm_path)
    n = len(tr.names)
    heat = np.zeros((n, 2, 7))
    tmin, tlim, tmax = np.zeros(n), np.zeros(n), np.zeros(n)
    for k, nm in enumerate(tr.names):
        e = nasa[nm.upper()]
        heat[k, 0] = e.low
        heat[k, 1] = e.high
        tmin[k], tmax[k], tlim[k] = e.tlo, e.thi, e.tmid
    Winf = adi.W_ref
    Wc = tr.W / Winf
    Wc_i = 1.0 / Wc
    gam = adi.gamma_inf
    gam2 = (gam - 1.0) / gam
    c1 = np.zeros((n, n))
    c2 = np.zeros((n, n))
    for m in range(n):
        for l in range(n):
            c1[m, l] = (tr.W[l] / tr.W[m]) ** 0.25
            c2[m, l] = 1.0 / math.sqrt((1.0 + tr.W[m] / tr.W[l]) * 8.0)
    return Thd(tr.names, tr.W, Wc, Wc_i, heat, tmin, tlim, tmax, tr.cofd, tr.coflam, tr.cofeta,
               c1, c2, gam2, adi.T_ref, 101325.0, adi.p_ref, adi.D_ref, adi.lbda_ref, adi.mu_ref,
               tr.kncf, tr.awt)


# --------------------------------------------------------------------------
# CHEMKIN mechanism
# --------------------------------------------------------------------------
KMAX = 16
IMAX = 48
TBMAX = 12
RU = 8.314510e7          # erg/(mol K), vode.f:462
PATM = 1.01325e6         # dyn/cm^2
RU_JOUL = 8.314510


@dataclass
class Mechanism:
    species: list
    elements: list
    wt: np.ndarray                      # g/mol
    ncf: np.ndarray                     # [k][element]
    a: np.ndarray                       # [k][range][7], range 0: T <= Tmid
    tmid: np.ndarray
    reactions: list = field(default_factory=list)

    @property
    def kk(self):
        return len(self.species)

    @property
    def ii(self):
        return len(self.reactions)

    def tables(self) -> dict:
        """Fixed-size arrays in the layout of the C struct ``izem_mech``."""
        kk, ii = self.kk, self.ii
        if kk > KMAX or ii > IMAX:
            raise ValueError("mechanism too large for the kernel tables")
        t = dict(
            kk=kk, ii=ii,
            wt=np.zeros(KMAX), a=np.zeros((KMAX, 2, 7)), tmid=np.zeros(KMAX),
            A=np.zeros(IMAX), b=np.zeros(IMAX), EoR=np.zeros(IMAX),
            nu=np.zeros((IMAX, 6), np.int32), nunk=np.zeros((IMAX, 6), np.int32),
            reversible=np.zeros(IMAX, np.int32), ntb=-np.ones(IMAX, np.int32),
            tb_k=np.zeros((IMAX, TBMAX), np.int32), tb_eff=np.zeros((IMAX, TBMAX)),
            fal=np.zeros(IMAX, np.int32), fpar=np.zeros((IMAX, 8)),
            kfal=np.zeros(IMAX, np.int32), rpar=np.zeros((IMAX, 3)),
        )
        t["wt"][:kk] = self.wt
        t["a"][:kk] = self.a
        t["tmid"][:kk] = self.tmid
        for i, r in enumerate(self.reactions):
            t["A"][i], t["b"][i], t["EoR"][i] = r["A"], r["b"], r["EoR"]
            for s, (k, nu) in enumerate(r["reac"]):
                t["nunk"][i, s], t["nu"][i, s] = k + 1, -nu
            for s, (k, nu) in enumerate(r["prod"]):
                t["nunk"][i, 3 + s], t["nu"][i, 3 + s] = k + 1, nu
            # 0 irreversible, 1 reverse from the equilibrium constant,
            # 2 explicit reverse Arrhenius parameters (REV)
            t["reversible"][i] = (2 if r.get("rpar") else 1) if r["rev"] else 0
            if r.get("rpar"):
                t["rpar"][i] = r["rpar"]
            t["kfal"][i] = r.get("kfal", 0)
            if r["tb"]:
                eff = r["eff"]
                t["ntb"][i] = len(eff)
                for s, (k, e) in enumerate(eff):
                    t["tb_k"][i, s], t["tb_eff"][i, s] = k + 1, e
            if r["fal"]:
                t["fal"][i] = r["fal"]
                t["fpar"][i, :len(r["fpar"])] = r["fpar"]
        return t


def _efac(units: str) -> float:
    # ckinterp39.f:2415-2426; the literals are default-REAL constants
    if units == "KELV":
        return 1.0
    if units == "CAL/":
        return _f32(4.184) / RU_JOUL
    if units == "KCAL":
        return _f32(4184.0) / RU_JOUL
    if units == "JOUL":
        return 1.00 / RU_JOUL
    if units == "KJOU":
        return _f32(4000.0) / RU_JOUL
    raise ValueError(units)


def _parse_side(s: str, spidx: dict):
    """Species and coefficients of one side; whether it has +M; the falloff
    collider: None, "M" or the name of the species in (+name)."""
    s = s.strip()
    third = False
    fall = None
    m = re.search(r"\(\+([A-Za-z0-9]+)\)", s)
    if m:
        fall = m.group(1).upper()
        s = s[:m.start()] + s[m.end():]
    terms = [x for x in re.split(r"\+(?![^()]*\))", s) if x]
    out = []
    for term in terms:
        term = term.strip()
        if term.upper() == "M":
            third = True
            continue
        mm = re.match(r"^(\d*)(.*)$", term)
        coef = int(mm.group(1)) if mm.group(1) else 1
        name = mm.group(2).strip().upper()
        if name not in spidx:
            raise ValueError(f"unknown species {name}")
        k = spidx[name]
        for q, (kk, nu) in enumerate(out):
            if kk == k:
                out[q] = (kk, nu + coef)
                break
        else:
            out.append((k, coef))
    return out, third, fall


def read_mechanism(inp_path: str | None = None, therm_path: str | None = None) -> Mechanism:
    """Parse a CHEMKIN-II mechanism (ELEMENTS / SPECIES / REACTIONS with +M,
    (+M) and (+species) falloff, LOW, TROE (3 or 4 values), SRI (3 or 5
    values), REV, irreversible =>, DUPLICATE and third-body efficiencies)
    into the tables of the kinetics kernels, with the unit conversions and
    defaults of ckinterp39.f (SRI with 3 values: d = 1, e = 0, :1831-1834;
    E of LOW and REV converted like the forward E, :2431-2439)."""
    inp_path = inp_path or os.path.join(DATA_DIR, "h2_air_li2004.inp")
    lines = []
    for raw in open(inp_path).read().splitlines():
        ln = raw.split("!")[0].rstrip()
        if ln.strip():
            lines.append(ln)
    sec = None
    elements, species, rlines = [], [], []
    eunits = "CAL/"
    for ln in lines:
        u = ln.strip().upper()
        if u.startswith("ELEM"):
            sec = "E"; rest = u.split()[1:]; elements += [x for x in rest if x != "END"]; continue
        if u.startswith("SPEC"):
            sec = "S"; rest = u.split()[1:]; species += [x for x in rest if x != "END"]; continue
        if u.startswith("REAC"):
            sec = "R"
            for w in u.split()[1:]:
                if w.startswith("MOLEC"):
                    raise ValueError("MOLECULES units for the pre-exponential factors are not read")
                if w.startswith("KELV"): eunits = "KELV"
                elif w.startswith("KCAL"): eunits = "KCAL"
                elif w.startswith("KJOU"): eunits = "KJOU"
                elif w.startswith("JOUL"): eunits = "JOUL"
            continue
        if u == "END":
            sec = None
            continue
        if sec == "E":
            elements += u.split()
        elif sec == "S":
            species += u.split()
        elif sec == "R":
            rlines.append(ln.strip())
    spidx = {s: i for i, s in enumerate(species)}
    nasa, defaults = read_therm(therm_path)
    kk = len(species)
    wt = np.zeros(kk)
    ncf = np.zeros((kk, len(elements)), dtype=np.int64)
    a = np.zeros((kk, 2, 7))
    tmid = np.zeros(kk)
    for k, s in enumerate(species):
        e = nasa[s]
        a[k, 0] = e.low
        a[k, 1] = e.high
        tmid[k] = e.tmid if e.tmid > 0 else defaults[1]
        comp = _composition(s)
        w = 0.0
        for el, cnt in comp.items():
            ncf[k, elements.index(el)] = cnt
        for j, el in enumerate(elements):
            w = w + ncf[k, j] * _f32(_AWT[el])
        wt[k] = w
    mech = Mechanism(species, elements, wt, ncf, a, tmid)
    efac = _efac(eunits)
    cur = None
    for ln in rlines:
        u = ln.upper()
        if "=" in u and not u.startswith(("LOW", "TROE", "SRI", "REV", "DUP")) and "/" not in u.split()[0]:
            parts = ln.split()
            eq = parts[0]
            A, b, E = float(parts[1]), float(parts[2]), float(parts[3])
            rev = "=>" not in eq or "<=>" in eq
            lhs, rhs = re.split(r"<=>|=>|=", eq, maxsplit=1)
            reac, tb1, f1 = _parse_side(lhs, spidx)
            prod, tb2, f2 = _parse_side(rhs, spidx)
            if f1 != f2 or tb1 != tb2:
                raise ValueError(f"third body or falloff collider differs across {eq}")
            # (+M): third-body falloff; (+species): that species is [M]
            # and the reaction has no third-body efficiencies (KFAL)
            kfal = 0
            if f1 is not None and f1 != "M":
                if f1 not in spidx:
                    raise ValueError(f"unknown falloff collider {f1} in {eq}")
                kfal = spidx[f1] + 1
            cur = dict(eq=eq, A=A, b=b, EoR=E * efac, reac=reac, prod=prod, rev=rev,
                       tb=tb1 or f1 == "M", fal=1 if f1 else 0, kfal=kfal, fpar=[],
                       eff=[], dup=False, rpar=None)
            mech.reactions.append(cur)
            continue
        for kw, body in re.findall(r"([A-Z0-9()]+)\s*/([^/]*)/", u):
            if kw == "LOW":
                v = [float(x) for x in body.split()]
                cur["fpar"] = [v[0], v[1], v[2] * efac]
            elif kw == "TROE":
                v = [float(x) for x in body.split()]
                if len(v) not in (3, 4) or not cur["fal"]:
                    raise ValueError(f"bad TROE data for {cur['eq']}")
                cur["fal"] = 3 if len(v) == 3 else 4
                cur["fpar"] = cur["fpar"][:3] + v
            elif kw == "SRI":
                v = [float(x) for x in body.split()]
                if len(v) == 3:
                    v = v + [1.0, 0.0]
                if len(v) != 5 or not cur["fal"]:
                    raise ValueError(f"bad SRI data for {cur['eq']}")
                cur["fal"] = 2
                cur["fpar"] = cur["fpar"][:3] + v
            elif kw == "REV":
                v = [float(x) for x in body.split()]
                if len(v) != 3 or not cur["rev"] or cur["fal"]:
                    raise ValueError(f"REV not allowed for {cur['eq']}")
                cur["rpar"] = [v[0], v[1], v[2] * efac]
            elif kw in spidx:
                cur["eff"].append((spidx[kw], float(body)))
            else:
                raise ValueError(f"unsupported auxiliary keyword {kw}")
        if "DUP" in u:
            cur["dup"] = True
    return mech


def _composition(name: str) -> dict:
    comp = {}
    for el, cnt in re.findall(r"(AR|HE|[A-Z])(\d*)", name.upper()):
        comp[el] = comp.get(el, 0) + (int(cnt) if cnt else 1)
    return comp


def mass_fractions_premix(thd_names, phi: float, products: float = 0.0) -> np.ndarray:
    """H2/air premix at equivalence ratio phi (air = 0.21 O2 + 0.79 N2 by
    mole), optionally blended with a mass fraction ``products`` of the
    complete-combustion H2O/OH mixture (90/10 by mole) of the same elements."""
    idx = {n.upper(): i for i, n in enumerate(thd_names)}
    X = np.zeros(len(thd_names))
    X[idx["H2"]] = 2.0 * phi
    X[idx["O2"]] = 1.0
    X[idx["N2"]] = 0.79 / 0.21
    tr = read_tranfit()
    W = tr.W
    Y = X * W
    Y = Y / Y.sum()
    if products > 0.0:
        Xp = np.zeros(len(thd_names))
        Xp[idx["H2O"]] = 0.9
        Xp[idx["OH"]] = 0.1
        Xp[idx["N2"]] = 0.79 / 0.21 * 0.5
        Yp = Xp * W
        Yp = Yp / Yp.sum()
        Y = (1.0 - products) * Y + products * Yp
    return Y
