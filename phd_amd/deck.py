"""The reference's input deck: its INI-style configuration files and the
keys of the hot path.

File syntax (the configuration reader of devel-izem/src/module_parsing.f90,
CFG_read_file / parse_line / trim_comment, :227-445):

  * ``#`` and ``;`` start a comment, except inside '...' or "..." strings;
  * a line without ``=`` must be a category, ``[NAME]``;
  * ``KEY = value`` sets NAME%KEY when the line is indented by at least two
    blanks or a tab, and the bare KEY otherwise; ``KEY += value`` appends to
    the stored text;
  * values are read as the key's type: logicals as Fortran reads them
    (T, F, .TRUE., .false., ...: the first letter after an optional '.'),
    integers, reals (D exponents allowed), strings (quotes and outer
    separators dropped), and arrays split on blanks, commas, quotes and tabs;
  * every key must be a known one (CFG_check); an unknown key is an error.

The known keys, their types and defaults are those of read_inp
(devel-izem/src/module_input.f90:266-400).  options() maps the hot-path keys
onto the solver's Options and refuses the settings that path does not
build (WENO7, LES, EGLIB, passive/variance variables, fixed or limited dt,
non-periodic boundaries, MPI decompositions, the simplified DVODE, MIL).
"""
from __future__ import annotations

import re
from dataclasses import dataclass

# name: (type, default); arrays are lists
KEYS = {
    "FILENAME": (str, "shearlayer.ini"),
    "GENERAL%DIM": (int, 1), "GENERAL%NRV": (int, 2), "GENERAL%NPV": (int, 0),
    "GENERAL%NVV": (int, 0), "GENERAL%EGLIB": (bool, False),
    "GENERAL%EGLIB_PRECISION": (int, 1), "GENERAL%VIS": (bool, True),
    "GENERAL%INIT": (str, "AMBIENT"), "GENERAL%FORCESUM": (bool, True),
    "MONITORING%IND_FILES": (bool, False), "MONITORING%INI_SOL": (bool, False),
    "MONITORING%FILTER": (bool, False), "MONITORING%FIL_XINI": (float, 0.5),
    "MONITORING%FIL_XEND": (float, 1.0), "MONITORING%ITSHOW": (int, 1),
    "MONITORING%WALLTIME": (int, 1000), "MONITORING%ITLIM": (bool, False),
    "MONITORING%ITMAX": (int, 10), "MONITORING%READ_RESTART": (bool, False),
    "MONITORING%NAME_RESTART": (str, "RESTART"), "MONITORING%NUMBER_RESTART": (int, 1),
    "MONITORING%FREQ_RESTART": (int, 1), "MONITORING%INITIAL_TIME": (float, 0.0),
    "MONITORING%TIME_OFFSET": (float, 1.0e-10), "MONITORING%DTIME": (float, 1.0e-7),
    "MONITORING%STAT": (bool, False), "MONITORING%NSTAT": (int, 1000),
    "MONITORING%VOLUME": ([float], [0.0, 0.0, 0.0, 1.0, 1.0, 1.0]),
    "MONITORING%Z_XYSTAT": ([float], [1.0]), "MONITORING%Y_XZSTAT": ([float], [1.0]),
    "MONITORING%X_YZSTAT": ([float], [1.0]),
    "MONITORING%PROBE_NAME": ([str], ["PROBE1", "PROBE2"]),
    "MONITORING%X-PROBE_COORD": ([float], [0.0, 1.0]),
    "MONITORING%Y-PROBE_COORD": ([float], [0.0, 1.0]),
    "MONITORING%Z-PROBE_COORD": ([float], [0.0, 1.0]),
    "MONITORING%NVOLUME": (int, 1),
    "TIMESTEPPING%VIS_DT": (bool, True), "TIMESTEPPING%FIXEDDT": (bool, False),
    "TIMESTEPPING%DT": (float, 1.0e-10), "TIMESTEPPING%CFLMIN": (float, 0.1),
    "TIMESTEPPING%CFL": (float, 0.5), "TIMESTEPPING%FO": (float, 0.1),
    "TIMESTEPPING%CPR": (float, 20.0), "TIMESTEPPING%DTLIMIT": (bool, False),
    "TIMESTEPPING%DTMIN_ALLOWED": (float, 1.0e-10),
    "IFLES%LES": (bool, False), "IFLES%MU_SGS_MODEL": (str, "SMAGORINSKY"),
    "IFLES%MU_SGS_FACTOR": (float, 0.18), "IFLES%TAU_ISO_SGS_SWITCH": (bool, False),
    "IFLES%TAU_ISO_SGS_MODEL": (str, "YOSHIZAWA"), "IFLES%TAU_ISO_SGS_FACTOR": (float, 1.0),
    "IFLES%CTE_DARLYHARLOW_SGS": (float, 0.08), "IFLES%PERCENT_WEIGHT_SGS": (float, 70.0),
    "IFLES%SC_SGS": (float, 0.7), "IFLES%PR_SGS": (float, 0.7),
    "REACTING%REACTION": (bool, False), "REACTING%REAC_SELEC": (str, "DVODE"),
    "REACTING%DVODESIMP": (bool, False), "REACTING%INDEX_N2": (int, 1),
    "REACTING%INDEX_H2": (int, 1), "REACTING%INDEX_O2": (int, 1),
    "REACTING%INDEX_H2O": (int, 1), "REACTING%YST": (float, 0.41),
    "REACTING%DIS_RATE_THERM": (bool, True), "REACTING%NREAC": (int, 21),
    "REACTING%INDEX_FUEL": (int, 2), "REACTING%INDEX_OXIDIZER": (int, 4),
    "REACTING%INDEX_DA": (int, 1), "REACTING%INDEX_SCALAR": (int, 1),
    "WENO%ORDER_WENO": (int, 7), "WENO%WEIGHTS": (bool, True),
    "WENO%PERCENT_WEIGHT": (float, 15.0), "WENO%OPT": (bool, False), "WENO%OPTORD": (int, 1),
    "WENO%DUCROS_TREDUC": (float, 0.2), "WENO%SHOCK_SENSOR_TYPE": (str, "ADAMSSHARIFF"),
    "GRID%DIM_LENGTH_COORD": (float, 1.0), "GRID%NONDIM_GRID": (bool, False),
    "GRID%LENGTH_REF": (float, 1.44e-4),
    "GRID%BCXMIN": (str, "PERIODIC"), "GRID%BCXMAX": (str, "PERIODIC"),
    "GRID%BCYMIN": (str, "PERIODIC"), "GRID%BCYMAX": (str, "PERIODIC"),
    "GRID%BCZMIN": (str, "PERIODIC"), "GRID%BCZMAX": (str, "PERIODIC"),
    "GRID%NUMPROCX": (int, 1), "GRID%NUMPROCY": (int, 1), "GRID%NUMPROCZ": (int, 1),
    "POSTPARALLEL%NPROCX": (int, 1), "POSTPARALLEL%NPROCY": (int, 1),
    "POSTPARALLEL%NPROCZ": (int, 1), "POSTPARALLEL%GHOSTPTX": (int, 0),
    "POSTPARALLEL%GHOSTPTY": (int, 0), "POSTPARALLEL%GHOSTPTZ": (int, 0),
    "DUMP%READ_BE": (bool, False), "DUMP%WRITE_BE": (bool, False), "DUMP%TRAFO": (bool, False),
    "POSTPRCESSING%NBINS": (int, 200), "POSTPRCESSING%PLANE_TYPE": (str, "RESTART"),
    "POSTPRCESSING%PLANE_NUMBER": (int, 0), "POSTPRCESSING%TEMP_AVG": (bool, False),
    "POSTPRCESSING%SPAT_AVG": (bool, False), "POSTPRCESSING%COND_AVG": (bool, False),
    "POSTPRCESSING%PDFS": (bool, False), "POSTPRCESSING%SPECTRAS": (bool, False),
    "POSTPRCESSING%SIMILARITY_X": (bool, False), "POSTPRCESSING%SECOND_LOOP": (bool, False),
    "POSTPRCESSING%READ_STAT": (bool, False), "POSTPRCESSING%START_FILE": (int, 1),
    "POSTPRCESSING%END_FILE": (int, 1), "POSTPRCESSING%SKIP_FILE": (int, 1),
    "SHEARLAYER%Y1F": ([float], [0.0, 0.05, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.95]),
    "SHEARLAYER%Y0O": ([float], [5.6e-7, 0.0, 1.55e-4, 0.278, 1.83e-3, 0.17, 5.1e-6, 2.5e-7,
                                 0.55]),
    "SHEARLAYER%PERTURB": (bool, False), "SHEARLAYER%SKEWANGLE": (float, 0.0),
    "SHEARLAYER%UMAX": (float, 1634.0), "SHEARLAYER%UMIN": (float, 973.3),
    "SHEARLAYER%UC": (float, 1208.18), "SHEARLAYER%RHOM": (float, 0.2785),
    "SHEARLAYER%MUM": (float, 4.1455e-5),
}
_FIXED_SIZE = {"MONITORING%VOLUME": 6}
_SEPARATORS = " ,'\"\t"


class DeckError(ValueError):
    pass


def _strip_comment(line: str) -> str:
    quote = ""
    for n, ch in enumerate(line):
        if not quote:
            if ch in ("'", '"'):
                quote = ch
            elif ch in "#;":
                return line[:n]
        elif ch == quote:
            quote = ""
    return line


def _fields(text: str) -> list[str]:
    return [f for f in re.split("[" + re.escape(_SEPARATORS) + "]+", text) if f]


def _logical(s: str) -> bool:
    t = s.lstrip(".")
    if t[:1] in ("T", "t"):
        return True
    if t[:1] in ("F", "f"):
        return False
    raise DeckError(f"not a logical: {s!r}")


def _real(s: str) -> float:
    return float(s.replace("D", "E").replace("d", "e"))


def _scalar(kind, s: str):
    try:
        if kind is bool:
            return _logical(s)
        if kind is int:
            return int(s)
        if kind is float:
            return _real(s)
        return s
    except ValueError as e:
        raise DeckError(str(e)) from None


@dataclass
class Deck:
    raw: dict  # NAME -> stored text of the keys the file set

    @classmethod
    def parse(cls, text: str) -> "Deck":
        raw: dict = {}
        category = ""
        for number, line in enumerate(text.splitlines(), 1):
            line = _strip_comment(line)
            if not line.strip():
                continue
            eq = line.find("=")
            if eq < 0:
                s = line.strip()
                close = s.find("]")
                if not s.startswith("[") or close < 0:
                    raise DeckError(f"line {number}: not a category or a KEY = value")
                category = s[1:close]
                continue
            append = eq > 0 and line[eq - 1] == "+"
            name = line[:eq - 1] if append else line[:eq]
            # only an indented key belongs to the current category
            cat = category if (name[:2] == "  " or name[:1] == "\t") else ""
            name = name.strip()
            if not name:
                raise DeckError(f"line {number}: no key before '='")
            full = f"{cat}%{name}" if cat else name
            value = line[eq + 1:]
            raw[full] = raw.get(full, "").rstrip() + value.rstrip() if append else value
            if full not in KEYS:
                raise DeckError(f"line {number}: unknown variable {full} specified")
        deck = cls(raw)
        for name in raw:
            deck.get(name)  # type check every value as the reader does
        return deck

    @classmethod
    def read(cls, path: str) -> "Deck":
        with open(path) as f:
            return cls.parse(f.read())

    def get(self, name: str):
        kind, default = KEYS[name]
        if name not in self.raw:
            return default
        fields = _fields(self.raw[name])
        if isinstance(kind, list):
            vals = [_scalar(kind[0], f) for f in fields]
            want = _FIXED_SIZE.get(name)
            if want is not None and len(vals) != want:
                raise DeckError(f"variable [{name}] has the wrong size")
            return vals
        if kind is str:
            # a string of length 1 takes everything between its first and
            # last field
            text = self.raw[name].strip(_SEPARATORS)
            return text
        if len(fields) != 1:
            raise DeckError(f"variable [{name}] has the wrong size")
        return _scalar(kind, fields[0])


SENSORS = {"ADAMSSHARIFF": 0, "DUCROS": 1}


BC_KEYS = ("GRID%BCXMIN", "GRID%BCXMAX", "GRID%BCYMIN", "GRID%BCYMAX", "GRID%BCZMIN",
           "GRID%BCZMAX")
BCS = {"PERIODIC": 0, "EXTRAPOLATION": 1, "POSTSHOCK_SLIPWALL": 2, "NSCBC_OUT": 3}


def face_conditions(deck: Deck) -> tuple:
    """The six face conditions W, E, S, N, B, F for izem_set_bc (0 periodic,
    1 extrapolation, 2 post-shock slip wall, 3 characteristic outflow);
    faces beyond GENERAL%DIM are periodic."""
    dim = deck.get("GENERAL%DIM")
    return tuple(BCS[deck.get(k).upper()] if f // 2 < dim else 0 for f, k in enumerate(BC_KEYS))


def ghost_layers(deck: Deck) -> int:
    """ng = (ORDER_WENO + 1) / 2 (module_input.f90:712): 3 for WENO5, 4 for
    WENO7."""
    return (deck.get("WENO%ORDER_WENO") + 1) // 2


def library_key(deck: Deck) -> tuple[int, int]:
    """(species, ghost layers) of the library a deck runs on: the full
    library (9 species, WENO5, chemistry) or a flow-only build
    (phd_amd/build.py FLOW_VARIANTS)."""
    return deck.get("GENERAL%NRV"), ghost_layers(deck)


def decomposition(deck: Deck) -> tuple[int, int, int]:
    """GRID%NUMPROCX/Y/Z: the ranks of the reference's decomposed run, which
    the GPU run reproduces on one device (izem_set_decomposition)."""
    return tuple(deck.get(k) for k in ("GRID%NUMPROCX", "GRID%NUMPROCY", "GRID%NUMPROCZ"))


def options(deck: Deck, **overrides):
    """The solver Options of a deck (module_input.f90:673-705), refusing
    what the hot path does not build."""
    from .build import FLOW_VARIANTS
    from .solver import Options
    g = deck.get
    refuse = []
    if g("WENO%ORDER_WENO") not in (5, 7):
        refuse.append(f"ORDER_WENO = {g('WENO%ORDER_WENO')} (WENO5 or WENO7)")
    key = library_key(deck)
    if key != (9, 3) and key not in FLOW_VARIANTS:
        refuse.append(f"NRV = {key[0]} with ORDER_WENO = {g('WENO%ORDER_WENO')} (built: 9 "
                      f"species WENO5, and flow-only {', '.join(f'{n} species ng {q}' for n, q in FLOW_VARIANTS)})")
    elif key != (9, 3) and g("REACTING%REACTION"):
        refuse.append(f"REACTION = T with NRV = {key[0]}, ORDER_WENO = "
                      f"{g('WENO%ORDER_WENO')} (chemistry: the 9-species WENO5 library)")
    for key in ("GENERAL%NPV", "GENERAL%NVV"):
        if g(key) != 0:
            refuse.append(f"{key} = {g(key)}")
    for key in ("GENERAL%EGLIB", "IFLES%LES", "TIMESTEPPING%FIXEDDT", "TIMESTEPPING%DTLIMIT",
                "REACTING%DVODESIMP", "WENO%OPT"):
        if g(key):
            refuse.append(f"{key} = T")
    for key in BC_KEYS:
        bc = g(key).upper()
        if bc not in BCS:
            refuse.append(f"{key} = {g(key)} (periodic, extrapolated, post-shock slip-wall "
                          f"on BCYMIN or characteristic-outflow on BCXMAX faces only)")
        elif bc == "NSCBC_OUT" and key != "GRID%BCXMAX":
            refuse.append(f"{key} = {g(key)} (declared for BCXMAX only, module_BCs.f90:2511-2514)")
        elif bc == "POSTSHOCK_SLIPWALL" and key != "GRID%BCYMIN":
            refuse.append(f"{key} = {g(key)} (declared for BCYMIN only, module_BCs.f90:1163-1168)")
        elif bc == "POSTSHOCK_SLIPWALL" and g("GENERAL%INIT").upper() != "DMR":
            refuse.append(f"{key} = {g(key)} (its state is the double Mach reflection's: "
                          f"INIT = 'DMR')")
    for q, key in enumerate(("GRID%NUMPROCX", "GRID%NUMPROCY", "GRID%NUMPROCZ")):
        if g(key) < 1 or (q >= g("GENERAL%DIM") and g(key) != 1):
            refuse.append(f"{key} = {g(key)}")
    sensor = g("WENO%SHOCK_SENSOR_TYPE").upper()
    if sensor not in SENSORS:
        refuse.append(f"SHOCK_SENSOR_TYPE = {sensor}")
    selec = g("REACTING%REAC_SELEC").upper()
    if g("REACTING%REACTION") and selec != "DVODE":
        refuse.append(f"REAC_SELEC = {selec} (DVODE only)")
    if refuse:
        raise DeckError("settings the MI355X path does not build: " + "; ".join(refuse))
    kw = dict(weno_avg=int(g("WENO%WEIGHTS")), max_rel_weight=g("WENO%PERCENT_WEIGHT"),
              vis=int(g("GENERAL%VIS")), vis_dt=int(g("TIMESTEPPING%VIS_DT")),
              reaction=int(g("REACTING%REACTION")), forcesum=int(g("GENERAL%FORCESUM")),
              CFL=g("TIMESTEPPING%CFL"), CFLmin=g("TIMESTEPPING%CFLMIN"),
              Fo=g("TIMESTEPPING%FO"), CPR=g("TIMESTEPPING%CPR"),
              sensor_type=SENSORS[sensor], ducros_treduc=g("WENO%DUCROS_TREDUC"))
    kw.update(overrides)
    return Options(**kw)
