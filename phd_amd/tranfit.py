"""Transport property fits from Lennard-Jones parameters: the project's own
restatement of CHEMKIN's TRANFIT as the reference builds it
(devel-izem/src/tranfit/tranfit.f: TRANFIT :3-378, LAMFIT36 :379-500,
DIFFIT36 :501-612, D1236 :734-788, PARKER36 :789-812, INTERP36 :813-839,
OMEG1236 :840-1005, OMEGXX36 :1006-1039, CKCPML36 :1978-2038).

For every species the viscosity and the thermal conductivity, and for every
pair the binary diffusion coefficient at 1 atm, are evaluated at 50
temperatures evenly spaced over the range that the thermodynamic data of all
species cover, and ln(property) is fitted by a cubic in ln T.  The result
has the layout of the reference's tranfit.out (phd_amd.chemdata.TransportFits),
so that build_thd can take fits made here or read from a file.

* Viscosity: Chapman-Enskog with the reduced collision integral
  Omega(2,2)*(T*, delta*) of the Stockmayer potential; self-diffusion with
  Omega(1,1)*.
* Conductivity: translational, rotational and vibrational parts (Warnatz),
  with the rotational relaxation number scaled from 298 K by Parker's
  temperature dependence (Brau and Jonkman 1970).
* Binary diffusion: combination rules with the polar/non-polar induction
  correction, and the Marrero-Mason second-order correction factor (1972).
* Collision integrals: the Monchick-Mason tables (J. Chem. Phys. 35, 1676,
  1961) of Omega(1,1)* and Omega(2,2)* over 37 reduced temperatures and 8
  reduced dipole moments, with TRANFIT's three-point interpolation in delta*
  and then in T*, and its closed forms above T* = 75.

TRANFIT writes its constants as default-REAL (single precision) literals:
pi, the unit factors, the viscosity and diffusion prefactors, Avogadro's
number, the table entries and the fit constants of the A*, B*, C* ratios
and of Parker's function.  They are rounded to single precision here in the
same places, so that the fits agree with the reference program's to
rounding.  The cubic is an ordinary least-squares fit (TRANFIT calls
SLATEC's DPOLFT with uniform weights and DPCOEF to expand it in powers of
ln T); the two agree to the conditioning of the cubic in ln T.
"""
from __future__ import annotations

import os

import numpy as np

from . import chemdata as cd
from .chemdata import _f32

NT = 50     # fit temperatures (tranfit.f:46)
NO = 4      # polynomial coefficients (cubic)

# constants of TRANFIT (tranfit.f:71-76, 140) and of the CHEMKIN-II 3.6
# linking file it reads (RU, PATM: tranfit.f:1153), single-precision literals
PI = _f32(3.1415926535)
FDTCGS = _f32(1.0e-18)     # Debye -> cgs
FATCM = _f32(1.0e8)        # Angstrom -> cm
DIPMIN = _f32(1.0e-20)
RU = _f32(8.314e7)         # erg/(mol K)
PATM = _f32(1.01325e6)     # dyn/cm^2
BOLTZ = RU / _f32(6.02217e23)

_TSTERN = [_f32(x) for x in (.1, .2, .3, .4, .5, .6, .7, .8, .9, 1., 1.2, 1.4, 1.6, 1.8, 2., 2.5,
                             3., 3.5, 4., 5., 6., 7., 8., 9., 10., 12., 14., 16., 18., 20., 25.,
                             30., 35., 40., 50., 75., 100.)]
_DELTA = [_f32(x) for x in (0., .25, .5, .75, 1., 1.5, 2., 2.5)]

# Monchick-Mason reduced collision integrals, rows T* (above), columns
# delta* (above): Omega(1,1)* and Omega(2,2)*
_OMEGA11 = """
  4.008   4.002   4.655    5.52   6.454   8.214   9.824   11.31
  3.130   3.164   3.355   3.721   4.198    5.23   6.225   7.160
  2.649   2.657    2.77   3.002   3.319   4.054   4.785   5.483
  2.314    2.32   2.402   2.572   2.812   3.386   3.972   4.539
  2.066   2.073    2.14   2.278   2.472   2.946   3.437   3.918
  1.877   1.885   1.944    2.06   2.225   2.628   3.054   3.747
  1.729   1.738    1.79   1.893   2.036   2.388   2.763   3.137
 1.6122   1.622    1.67    1.76   1.886   2.198   2.535   2.872
  1.517   1.527   1.572   1.653   1.765   2.044    2.35   2.657
   1.44    1.45    1.49   1.564   1.665   1.917   2.196  2.4780
 1.3204    1.33   1.364   1.425    1.51    1.72   1.956   2.199
  1.234    1.24   1.272   1.324   1.394   1.573   1.777    1.99
  1.168   1.176   1.202   1.246   1.306    1.46    1.64   1.827
 1.1166   1.124   1.146   1.185   1.237   1.372    1.53     1.7
  1.075   1.082   1.102   1.135   1.181     1.3   1.441   1.592
 1.0006   1.005    1.02   1.046    1.08    1.17   1.278   1.397
    .95   .9538   .9656   .9852   1.012   1.082   1.168   1.265
  .9131   .9162   .9256   .9413   .9626   1.019    1.09    1.17
  .8845   .8871   .8948   .9076   .9252    .972    1.03   1.098
  .8428   .8446    .850    .859   .8716   .9053   .9483   .9984
   .813   .8142   .8183    .825   .8344   .8598   .8927   .9316
  .7898    .791    .794   .7993   .8066   .8265   .8526   .8836
  .7711    .772   .7745   .7788   .7846   .8007    .822   .8474
  .7555   .7562   .7584   .7619   .7667     .78   .7976   .8189
  .7422    .743   .7446   .7475   .7515   .7627   .7776    .796
 .72022   .7206    .722   .7241   .7271   .7354   .7464     .76
  .7025    .703    .704   .7055   .7078   .7142   .7228   .7334
 .68776    .688   .6888   .6901   .6919    .697    .704   .7125
  .6751   .6753    .676    .677   .6785   .6827   .6884   .6955
   .664   .6642   .6648   .6657   .6669   .6704   .6752    .681
  .6414   .6415   .6418   .6425   .6433   .6457    .649    .653
  .6235   .6236   .6239   .6243   .6249   .6267    .629    .632
 .60882   .6089   .6091   .6094     .61   .6112    .613   .6154
  .5964   .5964   .5966    .597   .5972   .5983    .600   .6017
  .5763   .5763   .5764   .5766   .5768   .5775   .5785     .58
  .5415   .5415   .5416   .5416   .5418    .542   .5424    .543
   .518    .518   .5182   .5184   .5184   .5185   .5186   .5187
"""
_OMEGA22 = """
    4.1   4.266   4.833   5.742   6.729   8.624   10.34  11.890
  3.263   3.305   3.516   3.914   4.433    5.57   6.637   7.618
   2.84   2.836   2.936   3.168   3.511   4.329   5.126   5.874
  2.531   2.522   2.586   2.749   3.004    3.64   4.282   4.895
  2.284   2.277   2.329    2.46   2.665   3.187   3.727   4.249
  2.084   2.081    2.13   2.243   2.417   2.862   3.329   3.786
  1.922   1.924    1.97   2.072   2.225   2.641   3.028   3.435
 1.7902   1.795    1.84   1.934    2.07   2.417   2.788   3.156
  1.682   1.689   1.733    1.82   1.944   2.258   2.596   2.933
  1.593    1.60   1.644   1.725    1.84   2.124   2.435   2.746
  1.455   1.465   1.504   1.574    1.67   1.913   2.181    2.45
  1.355   1.365     1.4   1.461   1.544   1.754   1.989   2.228
   1.28   1.289   1.321   1.374   1.447    1.63   1.838   2.053
  1.222   1.231    1.26   1.306    1.37   1.532   1.718   1.912
  1.176   1.184   1.209    1.25   1.307    1.45   1.618   1.795
 1.0933     1.1   1.119    1.15   1.193   1.304   1.435   1.578
  1.039   1.044    1.06   1.083   1.117   1.204    1.31   1.428
  .9996   1.004   1.016   1.035   1.062   1.133    1.22    1.32
  .9699   .9732    .983   .9991   1.021    1.08   1.153   1.236
  .9268   .9291    .936   .9473   .9628   1.005   1.058    1.12
  .8962   .8979    .903   .9114    .923   .9545   .9955   1.044
  .8727   .8741    .878   .8845   .8935    .918   .9505   .9893
  .8538   .8549    .858   .8632   .8703    .890   .9164   .9482
  .8379   .8388   .8414   .8456   .8515    .868   .8895    .916
  .8243   .8251   .8273   .8308   .8356   .8493   .8676     .89
  .8018   .8024   .8039   .8065    .810    .820   .8337   .8504
  .7836    .784   .7852   .7872   .7899   .7976    .808   .8212
  .7683   .7687   .7696    .771   .7733   .7794    .788   .7983
  .7552   .7554   .7562   .7575   .7592    .764    .771   .7797
  .7436   .7438   .7445   .7455    .747   .7512    .757   .7642
 .71982     .72   .7204   .7211   .7221    .725   .7289   .7339
   .701   .7011   .7014    .702   .7026   .7047   .7076   .7112
 .68545   .6855    .686    .686   .6867   .6883   .6905    .693
  .6723   .6724   .6726    .673   .6733   .6745    .676   .6784
   .651    .651   .6512   .6513   .6516   .6524   .6534   .6546
   .614    .614   .6143   .6145   .6147   .6148   .6148   .6147
  .5887   .5889   .5894     .59   .5903   .5901   .5895   .5885
"""


def _table(text):
    return [[_f32(float(x)) for x in row.split()] for row in text.strip().splitlines()]


_O11 = _table(_OMEGA11)
_O22 = _table(_OMEGA22)


def _interp3(arg, x, y):
    """TRANFIT's quadratic interpolation through three points (INTERP36)."""
    val1 = y[0] + (arg - x[0]) * (y[1] - y[0]) / (x[1] - x[0])
    val2 = y[1] + (arg - x[1]) * (y[2] - y[1]) / (x[2] - x[1])
    fac1 = (arg - x[0]) / (x[1] - x[0]) / 2.0
    fac2 = (x[2] - arg) / (x[2] - x[1]) / 2.0
    if arg >= x[1]:
        return (val1 * fac2 + val2) / (1.0 + fac2)
    return (val1 + val2 * fac1) / (1.0 + fac1)


def omega(n: int, tr: float, dr: float) -> float:
    """Reduced collision integral Omega(1,1)* (n = 1) or Omega(2,2)* (n = 2)
    at reduced temperature tr and reduced dipole moment dr (OMEG1236)."""
    if dr < -1.0e-5 or dr > 2.5 or tr < 0.09 or tr > 500.0 or (abs(dr) > 1.0e-5 and tr > 75.0):
        raise ValueError(f"collision integral undefined at T* = {tr}, delta* = {dr}")
    if tr > 75.0:
        if n == 1:
            return (_f32(0.623) - _f32(0.136e-2) * tr + _f32(0.346e-5) * tr * tr
                    - _f32(0.343e-8) * tr * tr * tr)
        return (_f32(0.703) - _f32(0.146e-2) * tr + _f32(0.357e-5) * tr * tr
                - _f32(0.343e-8) * tr * tr * tr)
    tab = _O11 if n == 1 else _O22
    if tr <= 0.2:
        ii = 2
    else:
        ii = 37
        for i in range(2, 37):
            if _TSTERN[i - 2] < tr and _TSTERN[i - 1] >= tr:
                ii = i
    targ = [_TSTERN[ii - 2 + i] for i in range(3)]
    if abs(dr) >= 1.0e-5:
        if dr <= 0.25:
            kk = 2
        else:
            kk = 7
            for k in range(2, 8):
                if _DELTA[k - 2] < dr and _DELTA[k - 1] >= dr:
                    kk = k
        darg = [_DELTA[kk - 2 + k] for k in range(3)]
        vert = [_interp3(dr, darg, [tab[ii - 2 + i][kk - 2 + k] for k in range(3)])
                for i in range(3)]
        return _interp3(tr, targ, vert)
    return _interp3(tr, targ, [tab[ii - 2 + i][0] for i in range(3)])


def _omegxx(n: int, tr: float, dst: float) -> float:
    """Omega(1,1)*, Omega(2,2)* and the ratios B*, C* (OMEGXX36)."""
    if n in (1, 2):
        return omega(n, tr, dst)
    if n == 3:
        if tr <= 5.0:
            return _f32(1.36) - _f32(0.223) * tr + _f32(0.0613) * tr ** 2 - _f32(0.00554) * tr ** 3
        return _f32(1.095)
    if tr <= 5.0:
        return _f32(0.823) + _f32(0.0128) * tr + _f32(0.0112) * tr ** 2 - _f32(0.00193) * tr ** 3
    return _f32(0.9483)


def _d12_correction(w1, w2, t, eps1, eps2, sig1, sig2, dst):
    """Second-order correction of the binary diffusion coefficient (Marrero
    and Mason, J. Phys. Chem. Ref. Data 1, 3, 1972; D1236)."""
    sumw = w1 + w2
    sig12 = 0.5 * (sig1 + sig2)
    tr11, tr22, tr12 = t / eps1, t / eps2, t / np.sqrt(eps1 * eps2)
    om2f11 = _omegxx(2, tr11, dst)
    om2f22 = _omegxx(2, tr22, dst)
    om1f12 = _omegxx(1, tr12, dst)
    om2f12 = _omegxx(2, tr12, dst)
    bst12 = _omegxx(3, tr12, dst)
    cst12 = _omegxx(4, tr12, dst)
    ast12 = om2f12 / om1f12
    c12, c16, c24 = _f32(1.2), _f32(1.6), _f32(2.4)
    h = (sig1 / sig2) ** 2 * np.sqrt(2.0 * w2 / sumw) * 2.0 * w1 ** 2 / (sumw * w2)
    p1 = h * om2f11 / om1f12
    h = (sig2 / sig1) ** 2 * np.sqrt(2.0 * w1 / sumw) * 2.0 * w2 ** 2 / (sumw * w1)
    p2 = h * om2f22 / om1f12
    p12 = 15.0 * ((w1 - w2) / sumw) ** 2 + 8.0 * w1 * w2 * ast12 / sumw ** 2
    h = 2.0 / (w2 * sumw) * np.sqrt(2.0 * w2 / sumw) * om2f11 / om1f12 * (sig1 / sig12) ** 2
    q1 = ((2.5 - c12 * bst12) * w1 ** 2 + 3.0 * w2 ** 2 + c16 * w1 * w2 * ast12) * h
    h = 2.0 / (w1 * sumw) * np.sqrt(2.0 * w1 / sumw) * om2f22 / om1f12 * (sig2 / sig12) ** 2
    q2 = ((2.5 - c12 * bst12) * w2 ** 2 + 3.0 * w1 ** 2 + c16 * w1 * w2 * ast12) * h
    h = (((w1 - w2) / sumw) ** 2 * (2.5 - c12 * bst12) * 15.0
         + 4.0 * w1 * w2 * ast12 / sumw ** 2 * (11.0 - c24 * bst12))
    q12 = h + c16 * sumw * om2f11 * om2f22 / (np.sqrt(w1 * w2) * om1f12 ** 2) * \
        sig1 ** 2 * sig2 ** 2 / sig12 ** 4
    return ((6.0 * cst12 - 5.0) ** 2 / 10.0) * (p1 + p2 + p12) / (q1 + q2 + q12) + 1.0


def _parker(t, eps):
    """Temperature dependence of the rotational collision number (PARKER36)."""
    a, b, c = _f32(2.7842), _f32(4.4674), _f32(5.5683)
    d, dr = eps / t, eps / 298.0
    return ((1.0 + a * np.sqrt(dr) + b * dr + c * dr ** 1.5)
            / (1.0 + a * np.sqrt(d) + b * d + c * d ** 1.5))


def read_tran_dat(species, path: str | None = None) -> dict:
    """Lennard-Jones parameters of the given species from a CHEMKIN transport
    database: name, geometry (0 atom, 1 linear, 2 nonlinear), eps/k [K],
    sigma [A], dipole [Debye], polarizability [A^3], Z_rot(298 K)."""
    path = path or os.path.join(cd.DATA_DIR, "tran.dat")
    want = {s.upper() for s in species}
    out = {}
    for raw in open(path).read().splitlines():
        ln = raw.split("!")[0].split()
        if len(ln) < 7 or ln[0].upper() not in want:
            continue
        v = [float(x) for x in ln[1:7]]
        out[ln[0].upper()] = dict(nlin=int(v[0]), eps=v[1], sig=v[2], dip=v[3], pol=v[4],
                                  zrot=v[5])
    missing = [s for s in species if s.upper() not in out]
    if missing:
        raise ValueError(f"transport data not given for {missing}")
    return out


def _cv(mech, k, t):
    """Molar heat capacity at constant volume (CKCPML36 minus R)."""
    a = mech.a[k][1] if t > mech.tmid[k] else mech.a[k][0]
    tn = (1.0, t, t ** 2, t ** 3, t ** 4)
    cp = 0.0
    for n in range(5):
        cp = cp + RU * tn[n] * a[n]
    return cp - RU


def _cubic(x, y):
    # coefficients of ln T^0..3, least squares
    return np.polynomial.polynomial.polyfit(np.asarray(x), np.asarray(y), NO - 1)


def fit_transport(mech: "cd.Mechanism", tran_path: str | None = None,
                  therm_path: str | None = None, p: float = PATM) -> "cd.TransportFits":
    """Viscosity, conductivity and binary diffusion fits of the mechanism's
    species (TRANFIT): the content of the reference's tranfit.out."""
    kk = mech.kk
    tp = read_tran_dat(mech.species, tran_path)
    lj = [tp[s.upper()] for s in mech.species]
    nasa, defaults = cd.read_therm(therm_path)
    tlow = max((nasa[s.upper()].tlo or defaults[0]) for s in mech.species)
    thigh = min((nasa[s.upper()].thi or defaults[2]) for s in mech.species)
    dt = (thigh - tlow) / (NT - 1)
    temps = [tlow + n * dt for n in range(NT)]
    alogt = [np.log(t) for t in temps]
    wt = [float(w) for w in mech.wt]
    fcon = 0.5 * FDTCGS ** 2 * FATCM ** 3 / BOLTZ
    pfac = PATM / p
    coflam = np.zeros((kk, NO))
    cofeta = np.zeros((kk, NO))
    for k in range(kk):
        s = lj[k]
        dst = fcon * s["dip"] ** 2 / (s["eps"] * s["sig"] ** 3)
        helpe = _f32(2.6693e-5) * np.sqrt(wt[k]) / s["sig"] ** 2
        helpd = _f32(2.6280e-3) * pfac / (np.sqrt(wt[k]) * s["sig"] ** 2)
        xla, xeta = [], []
        for t in temps:
            prut = p / (RU * t)
            tr = t / s["eps"]
            dii = np.sqrt(t ** 3) * helpd / omega(1, tr, dst)
            eta = np.sqrt(t) * helpe / omega(2, tr, dst)
            rodet = dii * wt[k] * prut / eta
            aa = 2.5 - rodet
            park = _parker(t, s["eps"])
            if s["nlin"] == 2:
                bb = s["zrot"] * park + 2.0 * (5.0 / 2.0 + rodet) / PI
                croct = 1.0
            else:
                bb = s["zrot"] * park + 2.0 * (_f32(5.0 / 3.0) + rodet) / PI
                croct = _f32(2.0 / 3.0)
            ftra = 2.5 * (1.0 - 2.0 * croct * aa / (bb * PI))
            frot = rodet * (1.0 + 2.0 * aa / (bb * PI))
            fvib = rodet
            if s["nlin"] == 0:
                faktor = 5.0 / 2.0 * 1.5 * RU
            elif s["nlin"] == 1:
                faktor = (ftra * 1.5 + frot) * RU + fvib * (_cv(mech, k, t) - 2.5 * RU)
            else:
                faktor = (ftra + frot) * 1.5 * RU + fvib * (_cv(mech, k, t) - 3.0 * RU)
            xla.append(np.log(eta / wt[k] * faktor))
            xeta.append(np.log(eta))
        coflam[k] = _cubic(alogt, xla)
        cofeta[k] = _cubic(alogt, xeta)
    cofd = np.zeros((kk, kk, NO))
    for j in range(kk):
        for k in range(j + 1):
            sj, sk = lj[j], lj[k]
            sigm = 0.5 * (sj["sig"] + sk["sig"])
            epsm = np.sqrt(sj["eps"] * sk["eps"])
            if sj["dip"] < DIPMIN and sk["dip"] > DIPMIN:
                dst = 0.0   # k polar, j not
                xi = 1.0 + sj["pol"] * fcon * sk["dip"] ** 2 * np.sqrt(sk["eps"] / sj["eps"]) / \
                    (2.0 * sj["sig"] ** 3 * sk["eps"] * sk["sig"] ** 3)
            elif sj["dip"] > DIPMIN and sk["dip"] < DIPMIN:
                dst = 0.0   # j polar, k not
                xi = 1.0 + sk["pol"] * fcon * sj["dip"] ** 2 * np.sqrt(sj["eps"] / sk["eps"]) / \
                    (2.0 * sk["sig"] ** 3 * sj["eps"] * sj["sig"] ** 3)
            else:
                dst = fcon * sj["dip"] * sk["dip"] / (epsm * sigm ** 3)
                xi = 1.0
            sigm = sigm * xi ** _f32(-1.0 / 6.0)
            epsm = epsm * xi ** 2
            help1 = (wt[j] + wt[k]) / (2.0 * wt[j] * wt[k])
            xd = []
            for t in temps:
                tr = t / epsm
                help2 = _f32(0.002628) * np.sqrt(help1 * t ** 3)
                d = help2 * PATM / (omega(1, tr, dst) * sigm ** 2 * p)
                d = d * _d12_correction(wt[j], wt[k], t, sj["eps"], sk["eps"], sj["sig"],
                                        sk["sig"], dst)
                xd.append(np.log(d))
            cofd[j, k] = cofd[k, j] = _cubic(alogt, xd)
    kncf = np.asarray(mech.ncf, dtype=np.int64).T.copy()
    awt = np.array([_f32(cd._AWT[e]) for e in mech.elements]) / 1000.0
    return cd.TransportFits(list(mech.species), cofd, coflam, cofeta,
                            np.asarray(mech.wt, dtype=np.float64) / 1000.0, awt, kncf)


def write_tranfit(fits: "cd.TransportFits", path: str, nreac: int = 0) -> None:
    """The fits in the record order of the reference's tranfit.out
    (tranfit.f:262-318), readable by thd_cofd_read and read_tranfit."""
    kk = len(fits.names)
    nele = fits.awt.size
    L = [str(nreac), str(kk), str(NO), str(nele)]
    for j in range(kk):
        L += [str(j + 1), fits.names[j]]
        for k in range(j + 1):
            L.append(fits.names[k])
            L += [repr(float(x)) for x in fits.cofd[j, k]]
    L.append("-2")
    for k in range(kk):
        L.append(fits.names[k])
        L += [repr(float(x)) for x in fits.coflam[k]]
    L.append("-2")
    for k in range(kk):
        L.append(fits.names[k])
        L += [repr(float(x)) for x in fits.cofeta[k]]
    L.append("-2")
    for k in range(kk):
        L += [fits.names[k], repr(float(fits.W[k] * 1000.0))]
    L.append("-2")
    L += [repr(float(a * 1000.0)) for a in fits.awt]
    L.append("-2")
    for k in range(kk):
        L += [str(int(fits.kncf[n, k])) for n in range(nele)]
    L.append("-1")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")
