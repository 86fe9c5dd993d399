"""Resolved KORAL build configuration (the C struct ``koral_params``).

KORAL fixes its physics knobs at compile time (choices.h, the problem's
define.h).  The MI355X path takes the resolved values through one C struct,
declared in include/koral_hip.h; this module mirrors its layout with ctypes
and provides the presets for the benchmark workloads.

Unit macros follow ko.h:4-48 (GGG, CCC, K_BOLTZ, M_PROTON, SIGMA_RAD,
kappaCGS2GU) and are evaluated in the same left-to-right order so the
constants are bit-identical to the compiled macros.
"""
import ctypes

GGG = 6.674e-8 / 1.0        # ko.h:5,9  GGG0/gTILDA
CCC = 2.998e10 / 1.0        # ko.h:6,10
MSUNCM = 147700.0           # ko.h:13


class KoralParams(ctypes.Structure):
    _fields_ = [
        ("nv", ctypes.c_int), ("mag", ctypes.c_int), ("rad", ctypes.c_int), ("pad0", ctypes.c_int),
        ("gamma", ctypes.c_double),
        ("k_boltz", ctypes.c_double), ("m_proton", ctypes.c_double),
        ("sigma_rad", ctypes.c_double), ("mu_gas", ctypes.c_double),
        ("ggg", ctypes.c_double), ("ccc", ctypes.c_double), ("masscm", ctypes.c_double),
        ("kappa_model", ctypes.c_int), ("polar_nc", ctypes.c_int),
        ("kappa_a_coef", ctypes.c_double), ("kappa_es_coef", ctypes.c_double),
        ("rhofloor", ctypes.c_double), ("uurhoratiomin", ctypes.c_double),
        ("uurhoratiomax", ctypes.c_double), ("b2rhoratiomax", ctypes.c_double),
        ("gammamaxhd", ctypes.c_double), ("gammamaxrad", ctypes.c_double),
        ("eradfloor", ctypes.c_double),
        ("eerhoratiomin", ctypes.c_double), ("eerhoratiomax", ctypes.c_double),
        ("eeuuratiomin", ctypes.c_double), ("eeuuratiomax", ctypes.c_double),
        ("u2pconv", ctypes.c_double),
        ("radimpconv", ctypes.c_double), ("radimpentrconv", ctypes.c_double),
        ("radimpeps", ctypes.c_double),
        ("radimpmaxiter", ctypes.c_int), ("convrel_enabled", ctypes.c_int),
        ("dampingfactor", ctypes.c_double), ("maxenchangedown", ctypes.c_double),
        ("maxenchangeup", ctypes.c_double), ("maxradimpdamping", ctypes.c_double),
        ("radimplicitthreshold", ctypes.c_double),
        ("radimpconvrel", ctypes.c_double), ("radimpconvrelerr", ctypes.c_double),
        ("radimpconvrelentr", ctypes.c_double), ("radimpconvrelentrerr", ctypes.c_double),
        ("int_order", ctypes.c_int), ("fluxmethod", ctypes.c_int),
        ("tsteplim", ctypes.c_double), ("minmod_theta", ctypes.c_double),
        ("dofixups", ctypes.c_int), ("dou2pmhdfixups", ctypes.c_int),
        ("dou2pradfixups", ctypes.c_int), ("doradimpfixups", ctypes.c_int),
    ]

    @property
    def ee(self):
        return 6 + 3 * self.mag

    def copy(self):
        c = KoralParams()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(self))
        return c


def units(mass_msun):
    """ko.h:4-48 unit constants for a black-hole mass in solar masses."""
    masscm = mass_msun * MSUNCM
    k_boltz = 1.3806488e-16 * GGG / CCC / CCC / CCC / CCC / masscm
    m_proton = 1.67262158e-24 * GGG / CCC / CCC / masscm
    sigma_rad = 5.67e-5 * GGG / CCC / CCC / CCC / CCC / CCC * masscm * masscm
    return masscm, k_boltz, m_proton, sigma_rad


def kappa_cgs2gu(x, masscm):
    return x / GGG * CCC * CCC / masscm


def rho_cgs2gu(x, masscm):
    """rhoCGS2GU, ko.h:25"""
    return x * GGG / CCC / CCC * masscm * masscm


def base_params(mag, rad, gamma, mass_msun=10.0, masscm=None):
    """KORAL defaults (choices.h) for the given layout."""
    P = KoralParams()
    P.mag, P.rad = int(mag), int(rad)
    P.nv = 6 + 3 * P.mag + 4 * P.rad
    P.gamma = gamma
    mcm, kb, mp, sr = units(mass_msun)
    if masscm is not None:
        mcm = masscm
        kb = 1.3806488e-16 * GGG / CCC / CCC / CCC / CCC / mcm
        mp = 1.67262158e-24 * GGG / CCC / CCC / mcm
        sr = 5.67e-5 * GGG / CCC / CCC / CCC / CCC / CCC * mcm * mcm
    P.k_boltz, P.m_proton, P.sigma_rad, P.mu_gas = kb, mp, sr, 1.0
    P.ggg, P.ccc, P.masscm = GGG, CCC, mcm
    P.kappa_model = 0
    P.kappa_a_coef = 0.0
    P.kappa_es_coef = 0.0
    P.rhofloor = 1.e-50
    P.uurhoratiomin = 1.e-50
    P.uurhoratiomax = 1.e2
    P.b2rhoratiomax = 100.
    P.gammamaxhd = 100.
    P.gammamaxrad = 100.
    P.eradfloor = 10. * 1.e-50
    P.eerhoratiomin, P.eerhoratiomax = 1.e-30, 1.e30
    P.eeuuratiomin, P.eeuuratiomax = 1.e-30, 1.e30
    P.u2pconv = 1.e-10
    P.radimpconv = 1.e-10
    P.radimpentrconv = 1.e-5
    P.radimpeps = 1.e-6
    P.radimpmaxiter = 50
    P.convrel_enabled = 0
    P.dampingfactor = 3.
    P.maxenchangedown = 10.
    P.maxenchangeup = 10.
    P.maxradimpdamping = 1.e-5
    P.radimplicitthreshold = 1.e-2
    P.radimpconvrel = 1.e-10
    P.radimpconvrelerr = 1.e-8
    P.radimpconvrelentr = 1.e-6
    P.radimpconvrelentrerr = 1.e-6
    P.int_order = 4
    P.fluxmethod = 1
    P.tsteplim = 0.5
    P.minmod_theta = 1.5
    P.dofixups = 1
    P.dou2pmhdfixups = 1
    P.dou2pradfixups = 1
    P.doradimpfixups = 1
    return P


def T_from_urho(P, u, rho):
    """calc_PEQ_Tfromurho, physics.c:984-989."""
    p = u * (P.gamma - 1.)
    return p / (P.k_boltz * rho / P.mu_gas / P.m_proton)


def u_from_Trho(P, T, rho):
    """calc_PEQ_ufromTrho, physics.c:977-982."""
    p = P.k_boltz * rho * T / P.mu_gas / P.m_proton
    return p / (P.gamma - 1.)


def params_tube():
    """1-D flat-space radiation shock tube: Gamma=5/3, kappa_abs=0.4 rho,
    no scattering, PPM, HLL, Courant 0.5.  Length unit 1 cm
    (PROBLEMS/RADTUBE/define.h MASSCM 1); SIGMA_RAD chosen, as RADTUBE does,
    so that the left state (rho=1, u=1e-3, E=1e-2) is in LTE."""
    P = base_params(mag=0, rad=1, gamma=5. / 3., masscm=1.0)
    T_left = T_from_urho(P, 1.e-3, 1.0)
    P.sigma_rad = 1.e-2 / pow(T_left, 4.) / 4.
    P.kappa_model = 0
    P.kappa_a_coef = 0.4
    P.kappa_es_coef = 0.0
    P.int_order = 2
    P.fluxmethod = 1
    P.tsteplim = 0.5
    return P


def params_torus_mhd():
    """3-D GRMHD torus step: Gamma=13/9, MP5, HLL, flux-CT, no radiation."""
    P = base_params(mag=1, rad=0, gamma=13. / 9.)
    P.int_order = 4
    P.fluxmethod = 1
    P.tsteplim = 0.5
    return P


def params_implicit_batch(mass_msun=10.0):
    """Batched implicit solve: GRRMHD+M1 layout (NV=13), Gamma=5/3, Kramers
    free-free absorption (PROBLEMS/KATOTORUS/kappa.c) and Thomson scattering
    KAPPA_ES_COEFF (ko.h:48)."""
    P = base_params(mag=1, rad=1, gamma=5. / 3., mass_msun=mass_msun)
    P.kappa_model = 1
    P.kappa_a_coef = 6.4e22
    P.kappa_es_coef = kappa_cgs2gu(0.4, P.masscm)
    return P


def params_torus_rad(mass_msun=10.0):
    """3-D GRRMHD+M1 torus step: Gamma=13/9 torus plus the batched-solve
    opacities."""
    P = base_params(mag=1, rad=1, gamma=13. / 9., mass_msun=mass_msun)
    P.kappa_model = 1
    P.kappa_a_coef = 6.4e22
    P.kappa_es_coef = kappa_cgs2gu(0.4, P.masscm)
    P.int_order = 4
    P.fluxmethod = 1
    P.tsteplim = 0.5
    P.polar_nc = 2  # CORRECT_POLARAXIS, NCCORRECTPOLAR 2 (PROBLEMS/KATOTORUS/define.h:234-238)
    return P
