"""Transport-property fits from Lennard-Jones parameters, the input the
reference's solver reads as tranfit.out (devel-izem/src/tranfit/tranfit.f
makes that file; thd/thd_tools_perfectgas.f90:478-638 reads it).

For every species the viscosity and the thermal conductivity, and for every
pair the binary diffusion coefficient at 1 atm, are evaluated at 50
temperatures evenly spaced over the range that the thermodynamic data of all
species cover; ln(property) is then fitted by a least-squares cubic in ln T.
The methods are the published ones that CHEMKIN's transport package uses:

* dilute-gas kinetic theory (Chapman-Enskog; Hirschfelder, Curtiss and Bird,
  "Molecular Theory of Gases and Liquids", 1954): viscosity with the reduced
  collision integral Omega(2,2)*, self- and binary diffusion with
  Omega(1,1)*, both for the Stockmayer potential (Lennard-Jones plus
  point dipoles);
* Omega* from the tables of Monchick and Mason (J. Chem. Phys. 35, 1676,
  1961) over reduced temperature T* and reduced dipole moment delta*,
  interpolated through three neighbouring nodes, first across delta* and
  then across T*, and by cubics in T* above the tables;
* pure-species conductivity split into translational, rotational and
  vibrational parts (Warnatz, 1978; Kee et al., "Chemically Reacting
  Flow", 2003, sec. 12.4), the rotational relaxation number scaled from its
  298 K value by Parker's temperature dependence (Parker 1959; Brau and
  Jonkman 1970);
* binary pairs: the combination rules for sigma and epsilon, with the
  induction correction of a polar / non-polar pair, and the second
  Chapman-Enskog approximation to D_12 (Marrero and Mason, J. Phys. Chem.
  Ref. Data 1, 3, 1972).

The reference program writes its constants as single-precision literals; the
values here are rounded to single precision where it rounds them, so that
the fits match its tranfit.out to rounding of the double-precision
arithmetic.  The least-squares cubic is numpy's (the reference uses SLATEC's
orthogonal-polynomial fit; the two agree to the conditioning of a cubic in
ln T).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import chemdata as cd
from .chemdata import _f32

NT = 50     # fit temperatures
NO = 4      # coefficients of the cubic in ln T

# cgs constants, single precision as the reference program holds them
PI = _f32(3.1415926535)
R_CGS = _f32(8.314e7)                      # erg/(mol K)
P_ATM = _f32(1.01325e6)                    # dyn/cm^2
K_BOLTZ = R_CGS / _f32(6.02217e23)         # erg/K
DEBYE = _f32(1.0e-18)                      # esu cm
ANGSTROM = _f32(1.0e8)                     # 1/cm
POLAR_EPS = _f32(1.0e-20)                  # dipole [D] below which a species is non-polar
# Chapman-Enskog prefactors: eta = C_ETA sqrt(W T) / (sigma^2 Omega22*) [g/cm/s],
# D = C_D sqrt(T^3 / W_red) / (p[atm] sigma^2 Omega11*) [cm^2/s]
C_ETA = _f32(2.6693e-5)
C_D = _f32(2.6280e-3)
C_D12 = _f32(0.002628)
# Parker's rotational-relaxation function: 1 + a x^(1/2) + b x + c x^(3/2),
# x = eps / kT, with a = pi^(3/2) / 2, b = pi^2 / 4 + 2, c = pi^(3/2) to
# four decimals
PARKER = (_f32(2.7842), _f32(4.4674), _f32(5.5683))
# Marrero-Mason factors 6/5, 8/5, 12/5 and the ratios B*(T*), C*(T*) of
# collision integrals, cubic in T* up to 5 and constant above
MM_6_5, MM_8_5, MM_12_5 = _f32(1.2), _f32(1.6), _f32(2.4)
B_STAR = ((_f32(1.36), -_f32(0.223), _f32(0.0613), -_f32(0.00554)), _f32(1.095))
C_STAR = ((_f32(0.823), _f32(0.0128), _f32(0.0112), -_f32(0.00193)), _f32(0.9483))
# cubics in T* for non-polar pairs above the tables (T* > 75): Omega11*, Omega22*
HIGH_T = {1: (_f32(0.623), -_f32(0.136e-2), _f32(0.346e-5), -_f32(0.343e-8)),
          2: (_f32(0.703), -_f32(0.146e-2), _f32(0.357e-5), -_f32(0.343e-8))}

# Monchick-Mason reduced collision integrals: rows T*, columns delta*
T_STAR = np.array([_f32(x) for x in (
    .1, .2, .3, .4, .5, .6, .7, .8, .9, 1., 1.2, 1.4, 1.6, 1.8, 2., 2.5, 3., 3.5, 4., 5., 6., 7.,
    8., 9., 10., 12., 14., 16., 18., 20., 25., 30., 35., 40., 50., 75., 100.)])
DELTA_STAR = np.array([_f32(x) for x in (0., .25, .5, .75, 1., 1.5, 2., 2.5)])

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




def _parse_table(text: str) -> np.ndarray:
    return np.array([[_f32(float(x)) for x in row.split()] for row in text.strip().splitlines()])


TABLES = {1: _parse_table(_OMEGA11), 2: _parse_table(_OMEGA22)}


def _three_point(x, x3, y3):
    """Value at x of three samples (x3, y3), x3 increasing: the secant through
    the first two and the one through the last two, blended with a weight on
    the farther secant that falls off linearly from one half at the middle
    node -- a smooth curve through the three points that reduces to the
    nearer secant at the outer nodes."""
    x0, x1, x2 = x3
    s_lo = y3[0] + (x - x0) * (y3[1] - y3[0]) / (x1 - x0)
    s_hi = y3[1] + (x - x1) * (y3[2] - y3[1]) / (x2 - x1)
    if x >= x1:
        w = (x2 - x) / (x2 - x1) / 2.0
        return (s_lo * w + s_hi) / (1.0 + w)
    w = (x - x0) / (x1 - x0) / 2.0
    return (s_lo + s_hi * w) / (1.0 + w)


def _window(nodes: np.ndarray, x: float, lowest: float) -> int:
    """First index of the three table nodes around x: the triple whose middle
    node is the first node at or above x, clamped to the table."""
    if x <= lowest:
        return 0
    m = int(np.searchsorted(nodes, x, side="left"))
    return min(max(m - 1, 0), len(nodes) - 3)


def collision_integral(l: int, tstar: float, dstar: float) -> float:
    """Reduced collision integral Omega(l,l)* (l = 1 or 2) of the Stockmayer
    potential at reduced temperature tstar = kT/eps and reduced dipole
    moment dstar (Monchick-Mason tables; cubics in T* above T* = 75 for
    non-polar pairs)."""
    polar = abs(dstar) > 1.0e-5
    if dstar < -1.0e-5 or dstar > 2.5 or tstar < 0.09 or tstar > 500.0 or (polar and tstar > 75.0):
        raise ValueError(f"collision integral undefined at T* = {tstar}, delta* = {dstar}")
    if tstar > 75.0:
        c = HIGH_T[l]
        return c[0] + c[1] * tstar + c[2] * tstar * tstar + c[3] * tstar * tstar * tstar
    tab = TABLES[l]
    i = _window(T_STAR, tstar, 0.2)
    rows = range(i, i + 3)
    if polar:
        j = _window(DELTA_STAR, dstar, 0.25)
        col = [_three_point(dstar, DELTA_STAR[j:j + 3], tab[r, j:j + 3]) for r in rows]
    else:
        col = [tab[r, 0] for r in rows]
    return _three_point(tstar, T_STAR[i:i + 3], col)


def _ratio_fit(fit, tstar: float) -> float:
    (c0, c1, c2, c3), high = fit
    return c0 + c1 * tstar + c2 * tstar ** 2 + c3 * tstar ** 3 if tstar <= 5.0 else high


@dataclass
class LJSpecies:
    """Molecular parameters of a CHEMKIN transport database entry."""
    geometry: int      # 0 atom, 1 linear, 2 non-linear
    eps_k: float       # well depth over Boltzmann's constant [K]
    sigma: float       # collision diameter [A]
    dipole: float      # [Debye]
    polarizability: float  # [A^3]
    zrot298: float     # rotational relaxation collision number at 298 K

    def reduced_dipole_sq(self) -> float:
        """mu*^2 = mu^2 / (eps sigma^3) in cgs."""
        return (self.dipole * DEBYE) ** 2 / (self.eps_k * K_BOLTZ * (self.sigma / ANGSTROM) ** 3)


def read_tran_dat(species, path: str | None = None) -> dict:
    """Lennard-Jones parameters of the given species from a CHEMKIN transport
    database (name, geometry, eps/k, sigma, dipole, polarizability, Z_rot)."""
    path = path or os.path.join(cd.DATA_DIR, "tran.dat")
    want = {s.upper() for s in species}
    out = {}
    for raw in open(path).read().splitlines():
        ln = raw.split("!")[0].split()
        if len(ln) < 7 or ln[0].upper() not in want:
            continue
        g, e, s, d, a, z = (float(x) for x in ln[1:7])
        out[ln[0].upper()] = LJSpecies(int(g), e, s, d, a, z)
    missing = [s for s in species if s.upper() not in out]
    if missing:
        raise ValueError(f"transport data not given for {missing}")
    return out


def _zrot_scale(sp: LJSpecies, T: float) -> float:
    """Z_rot(T) / Z_rot(298 K) by Parker's function of eps / kT."""
    a, b, c = PARKER

    def F(x):
        return 1.0 + a * np.sqrt(x) + b * x + c * x ** 1.5

    return F(sp.eps_k / 298.0) / F(sp.eps_k / T)


def _cv_molar(mech, k: int, T: float) -> float:
    """c_v of species k [erg/(mol K)] from its NASA-7 c_p/R polynomial."""
    a = mech.a[k][1] if T > mech.tmid[k] else mech.a[k][0]
    cp = 0.0
    for n, tn in enumerate((1.0, T, T ** 2, T ** 3, T ** 4)):
        cp = cp + R_CGS * tn * a[n]
    return cp - R_CGS


def pure_properties(sp: LJSpecies, W: float, T: float, cv: float, p: float):
    """Viscosity [g/(cm s)] and thermal conductivity [erg/(cm s K)] of one
    species at T; cv is its molar c_v [erg/(mol K)]."""
    dstar = 0.5 * sp.reduced_dipole_sq()
    tstar = T / sp.eps_k
    eta = C_ETA * np.sqrt(W) / sp.sigma ** 2 * np.sqrt(T) / collision_integral(2, tstar, dstar)
    D = C_D * (P_ATM / p) / (np.sqrt(W) * sp.sigma ** 2) * np.sqrt(T ** 3) / \
        collision_integral(1, tstar, dstar)
    rhoD_eta = D * W * (p / (R_CGS * T)) / eta
    cv_tr = 1.5 * R_CGS
    if sp.geometry == 0:
        return eta, eta / W * (2.5 * cv_tr)
    # rotational part: c_v,rot = R (linear) or 3/2 R (non-linear)
    if sp.geometry == 1:
        cv_rot, rot_over_tr, five_thirds_cvrot = R_CGS, _f32(2.0 / 3.0), _f32(5.0 / 3.0)
    else:
        cv_rot, rot_over_tr, five_thirds_cvrot = 1.5 * R_CGS, 1.0, 2.5
    A = 2.5 - rhoD_eta
    B = sp.zrot298 * _zrot_scale(sp, T) + 2.0 * (five_thirds_cvrot + rhoD_eta) / PI
    f_tr = 2.5 * (1.0 - 2.0 * rot_over_tr * A / (B * PI))
    f_rot = rhoD_eta * (1.0 + 2.0 * A / (B * PI))
    f_vib = rhoD_eta
    cv_vib = cv - cv_tr - cv_rot
    if sp.geometry == 1:
        c = (f_tr * 1.5 + f_rot) * R_CGS + f_vib * cv_vib
    else:
        c = (f_tr + f_rot) * 1.5 * R_CGS + f_vib * cv_vib
    return eta, eta / W * c


def pair_potential(a: LJSpecies, b: LJSpecies):
    """eps/k [K], sigma [A] and delta* of the a-b interaction: arithmetic
    sigma and geometric eps, a polar / non-polar pair corrected by the
    induced dipole (xi = 1 + alpha*_n mu*_p^2 sqrt(eps_p / eps_n) / 4, sigma
    times xi^(-1/6), eps times xi^2; delta* = 0 for it), two polar species
    combined as delta* = mu_a mu_b / (2 eps sigma^3)."""
    sigma = 0.5 * (a.sigma + b.sigma)
    eps = np.sqrt(a.eps_k * b.eps_k)
    pa, pb = a.dipole > POLAR_EPS, b.dipole > POLAR_EPS
    if pa != pb:
        p, n = (a, b) if pa else (b, a)
        xi = 1.0 + 0.25 * (n.polarizability / n.sigma ** 3) * p.reduced_dipole_sq() * \
            np.sqrt(p.eps_k / n.eps_k)
        return eps * xi ** 2, sigma * xi ** _f32(-1.0 / 6.0), 0.0
    dstar = 0.5 * (a.dipole * DEBYE) * (b.dipole * DEBYE) / \
        (eps * K_BOLTZ * (sigma / ANGSTROM) ** 3)
    return eps, sigma, dstar


def second_order_factor(Wa, Wb, T, a: LJSpecies, b: LJSpecies, dstar) -> float:
    """Marrero-Mason's second Chapman-Enskog approximation to a binary
    diffusion coefficient, f = 1 + (6 C* - 5)^2 / 10 (P_a + P_b + P_ab) /
    (Q_a + Q_b + Q_ab), written with the pair's mass fractions m = W / (Wa +
    Wb) and the collision integrals of the like and unlike interactions."""
    ma, mb = Wa / (Wa + Wb), Wb / (Wa + Wb)
    sab = 0.5 * (a.sigma + b.sigma)
    t_ab = T / np.sqrt(a.eps_k * b.eps_k)
    o22a = collision_integral(2, T / a.eps_k, dstar)
    o22b = collision_integral(2, T / b.eps_k, dstar)
    o11ab = collision_integral(1, t_ab, dstar)
    A = collision_integral(2, t_ab, dstar) / o11ab
    Bs = _ratio_fit(B_STAR, t_ab)
    Cs = _ratio_fit(C_STAR, t_ab)
    ra, rb = o22a / o11ab, o22b / o11ab

    def P(mi, mj, si, sj, r):
        return (si / sj) ** 2 * np.sqrt(2.0 * mj) * 2.0 * mi * mi / mj * r

    def Q(mi, mj, si, r):
        return ((2.5 - MM_6_5 * Bs) * mi * mi + 3.0 * mj * mj + MM_8_5 * mi * mj * A) * \
            2.0 / mj * np.sqrt(2.0 * mj) * r * (si / sab) ** 2

    dm2 = (ma - mb) ** 2
    p_ab = 15.0 * dm2 + 8.0 * ma * mb * A
    q_ab = dm2 * (2.5 - MM_6_5 * Bs) * 15.0 + 4.0 * ma * mb * A * (11.0 - MM_12_5 * Bs) + \
        MM_8_5 * o22a * o22b / (np.sqrt(ma * mb) * o11ab ** 2) * a.sigma ** 2 * b.sigma ** 2 / sab ** 4
    num = P(ma, mb, a.sigma, b.sigma, ra) + P(mb, ma, b.sigma, a.sigma, rb) + p_ab
    den = Q(ma, mb, a.sigma, ra) + Q(mb, ma, b.sigma, rb) + q_ab
    return 1.0 + ((6.0 * Cs - 5.0) ** 2 / 10.0) * num / den


def binary_diffusion(a: LJSpecies, b: LJSpecies, Wa, Wb, T, p) -> float:
    """D_ab [cm^2/s] at T and p [dyn/cm^2]."""
    eps, sigma, dstar = pair_potential(a, b)
    d1 = C_D12 * np.sqrt((Wa + Wb) / (2.0 * Wa * Wb) * T ** 3) * P_ATM / \
        (collision_integral(1, T / eps, dstar) * sigma ** 2 * p)
    return d1 * second_order_factor(Wa, Wb, T, a, b, dstar)


def fit_temperatures(mech, therm_path=None) -> np.ndarray:
    """50 temperatures over the range every species' thermodynamic data cover."""
    nasa, defaults = cd.read_therm(therm_path)
    lo = max((nasa[s.upper()].tlo or defaults[0]) for s in mech.species)
    hi = min((nasa[s.upper()].thi or defaults[2]) for s in mech.species)
    return np.array([lo + n * ((hi - lo) / (NT - 1)) for n in range(NT)])


def _lsq_cubic(lnT, y) -> np.ndarray:
    return np.polynomial.polynomial.polyfit(lnT, np.asarray(y), NO - 1)


def fit_transport(mech: "cd.Mechanism", tran_path: str | None = None,
                  therm_path: str | None = None, p: float = P_ATM) -> "cd.TransportFits":
    """ln(viscosity), ln(conductivity) and ln(binary diffusion) of the
    mechanism's species as cubics in ln T: the content of tranfit.out."""
    lj = read_tran_dat(mech.species, tran_path)
    sp = [lj[s.upper()] for s in mech.species]
    W = [float(w) for w in mech.wt]
    T = fit_temperatures(mech, therm_path)
    lnT = np.log(T)
    kk = mech.kk
    cofeta, coflam = np.zeros((kk, NO)), np.zeros((kk, NO))
    for k in range(kk):
        props = [pure_properties(sp[k], W[k], t, _cv_molar(mech, k, t), p) for t in T]
        cofeta[k] = _lsq_cubic(lnT, [np.log(e) for e, _ in props])
        coflam[k] = _lsq_cubic(lnT, [np.log(l) for _, l in props])
    cofd = np.zeros((kk, kk, NO))
    for j in range(kk):
        for k in range(j + 1):
            d = [np.log(binary_diffusion(sp[j], sp[k], W[j], W[k], t, p)) for t in T]
            cofd[j, k] = cofd[k, j] = _lsq_cubic(lnT, d)
    kncf = np.asarray(mech.ncf, dtype=np.int64).T.copy()
    awt = np.array([_f32(cd._AWT[e]) for e in mech.elements]) / 1000.0
    return cd.TransportFits(list(mech.species), cofd, coflam, cofeta,
                            np.asarray(mech.wt, dtype=np.float64) / 1000.0, awt, kncf)


def write_tranfit(fits: "cd.TransportFits", path: str, nreac: int = 0) -> None:
    """The fits in the record layout of tranfit.out as the reference's
    reader (thd_tools_perfectgas.f90:478-638) and chemdata.read_tranfit take
    it: counts; the lower triangle of D fits per species; conductivity,
    viscosity, molar masses, atomic weights, element counts."""
    kk = len(fits.names)
    nele = fits.awt.size
    out = [str(nreac), str(kk), str(NO), str(nele)]
    num = lambda x: repr(float(x))
    for j in range(kk):
        out += [str(j + 1), fits.names[j]]
        for k in range(j + 1):
            out += [fits.names[k]] + [num(x) for x in fits.cofd[j, k]]
    for block in (fits.coflam, fits.cofeta):
        out.append("-2")
        for k in range(kk):
            out += [fits.names[k]] + [num(x) for x in block[k]]
    out.append("-2")
    for k in range(kk):
        out += [fits.names[k], num(fits.W[k] * 1000.0)]
    out.append("-2")
    out += [num(a * 1000.0) for a in fits.awt]
    out.append("-2")
    for k in range(kk):
        out += [str(int(fits.kncf[n, k])) for n in range(nele)]
    out.append("-1")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
