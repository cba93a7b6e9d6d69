"""Material types and the derivation of point parameters (material!).

Mirrors src/physics/material_parameters.jl:100-231, src/physics/fracture.jl:110-175,
src/physics/bond_based.jl:74-113, ordinary_state_based.jl:82-96, correspondence.jl:111-169."""
import math
from dataclasses import dataclass, field

from ._lib import (ArgumentError, PDX_BB, PDX_CPD, PDX_LINEAR_ELASTIC, PDX_NEO_HOOKE,
                   PDX_NHP, PDX_OSB, PDX_SVK)


def _arg_error(msg):
    return ArgumentError(-1, msg)


# constitutive models (constitutive_models.jl) --------------------------------------------
class SaintVenantKirchhoff:
    code = PDX_SVK

    def __repr__(self):
        return "SaintVenantKirchhoff()"


class NeoHookePenalty:
    code = PDX_NHP

    def __repr__(self):
        return "NeoHookePenalty()"


class LinearElastic:
    """constitutive_models.jl:81-97: S from the Hooke matrix in Voigt form."""
    code = PDX_LINEAR_ELASTIC

    def __repr__(self):
        return "LinearElastic()"


class NeoHooke:
    """constitutive_models.jl:167-175: S = mu (I - C^-1) + lambda log(J) C^-1."""
    code = PDX_NEO_HOOKE

    def __repr__(self):
        return "NeoHooke()"


@dataclass
class ZEMSilling:
    """ZEMSilling(; Cs=0.5): zem_stabilization.jl:16-21."""
    Cs: float = 0.5
    code = 0


class ZEMWan:
    """ZEMWan(): zem_stabilization.jl:30-42, the stabilisation stiffness C_1 = C : K^-1 of the
    point's Hooke tensor (correspondence.jl:253-285)."""
    code = 1

    def __repr__(self):
        return "ZEMWan()"


class CriticalStretch:
    """fracture.jl:10."""

    def __repr__(self):
        return "CriticalStretch()"


def const_one_kernel(delta, L):
    """kernels.jl:10: omega = 1."""
    return 1


def linear_kernel(delta, L):
    """kernels.jl:22: omega = delta / L."""
    return delta / L


def cubic_b_spline_kernel(delta, L):
    """kernels.jl:43-52: the cubic B-spline of xi = L / delta."""
    xi = L / delta
    if 0 < xi <= 0.5:
        return 2 / 3 - 4 * xi ** 2 + 4 * xi ** 3
    if 0.5 < xi <= 1:
        return 4 / 3 - 4 * xi + 4 * xi ** 2 - 4 / 3 * xi ** 3
    return 0


def cubic_b_spline_kernel_norm(delta, L):
    """kernels.jl:74-84: the cubic B-spline scaled to unit integral over the horizon."""
    xi = L / delta
    C = 8.0 / (math.pi * delta ** 3)
    if 0 < xi <= 0.5:
        return C * (1 - 6 * xi ** 2 + 6 * xi ** 3)
    if 0.5 < xi <= 1:
        return C * 2 * (1 - xi) ** 3
    return 0


# the influence functions the device evaluates (PDX_KERNEL_* of include/pdx.h)
KERNEL_IDS = {linear_kernel: 0, const_one_kernel: 1, cubic_b_spline_kernel: 2,
              cubic_b_spline_kernel_norm: 3}


def _kernel(kernel):
    if kernel not in KERNEL_IDS:
        raise _arg_error("the influence functions are linear_kernel, const_one_kernel, "
                         "cubic_b_spline_kernel and cubic_b_spline_kernel_norm")
    return kernel


class NoCorrection:
    """NoCorrection: bond_system_corrections.jl:10 (surface correction factor 1)."""
    code = 0


class EnergySurfaceCorrection:
    """EnergySurfaceCorrection: bond_system_corrections.jl:58-61 (Le and Bobaru 2018): per-bond
    factors from a one-time uniform-stretch energy comparison, computed on the device when
    the bonds are built."""
    code = 1


def _correction(correction):
    if correction in (None, NoCorrection):
        return NoCorrection
    if correction is EnergySurfaceCorrection:
        return EnergySurfaceCorrection
    raise _arg_error("correction must be NoCorrection or EnergySurfaceCorrection")


class BBMaterial:
    """BBMaterial() / BBMaterial{Correction}(): bond_based.jl:74-84.  The type parameter is
    the `correction` argument here: BBMaterial(correction=EnergySurfaceCorrection)."""
    kind = PDX_BB

    def __init__(self, dmgmodel=None, correction=None):
        self.dmgmodel = dmgmodel or CriticalStretch()
        self.correction = _correction(correction)

    def __repr__(self):
        return f"BBMaterial{{{self.correction.__name__}}}()"


class OSBMaterial:
    """OSBMaterial(; kernel=linear_kernel, dmgmodel=CriticalStretch()) and
    OSBMaterial{Correction}(...): ordinary_state_based.jl:82-94."""
    kind = PDX_OSB

    def __init__(self, kernel=linear_kernel, dmgmodel=None, correction=None):
        self.correction = _correction(correction)
        self.kernel = _kernel(kernel)
        self.dmgmodel = dmgmodel or CriticalStretch()

    def __repr__(self):
        return f"OSBMaterial{{{self.correction.__name__}}}()"


class CMaterial:
    """CMaterial(; kernel, model, zem, dmgmodel, maxdmg): correspondence.jl:111-128."""
    kind = PDX_CPD

    def __init__(self, kernel=linear_kernel, model=None, zem=None, dmgmodel=None, maxdmg=0.85):
        self.kernel = _kernel(kernel)
        self.constitutive_model = model or SaintVenantKirchhoff()
        if not isinstance(self.constitutive_model,
                          (SaintVenantKirchhoff, NeoHookePenalty, LinearElastic, NeoHooke)):
            raise _arg_error("the constitutive models are SaintVenantKirchhoff, LinearElastic, "
                             "NeoHooke and NeoHookePenalty")
        self.zem = zem or ZEMSilling()
        if not isinstance(self.zem, (ZEMSilling, ZEMWan)):
            raise _arg_error("the zero-energy-mode stabilisations are ZEMSilling and ZEMWan")
        self.dmgmodel = dmgmodel or CriticalStretch()
        self.maxdmg = float(maxdmg)

    def __repr__(self):
        return f"CMaterial(maxdmg={self.maxdmg})"


# parameter derivation -------------------------------------------------------------------
ALLOWED_KWARGS = ("horizon", "rho", "E", "nu", "G", "K", "lambda", "mu", "Gc", "epsilon_c")


def _isapprox(a, b, rtol=math.sqrt(2.220446049250313e-16), atol=0.0):
    return abs(a - b) <= max(atol, rtol * max(abs(a), abs(b)))


def get_given_elastic_params(p):
    """material_parameters.jl:131-172."""
    nan = float("nan")
    E = float(p["E"]) if "E" in p else nan
    if "E" in p and E <= 0:
        raise _arg_error("`E` should be larger than zero!")
    nu = float(p["nu"]) if "nu" in p else nan
    if "nu" in p:
        if nu <= 0:
            raise _arg_error("`nu` should be larger than zero!")
        if nu >= 1:
            raise _arg_error("too high value of `nu`! Condition: 0 < `nu` ≤ 1")
    G = float(p["G"]) if "G" in p else nan
    if "G" in p and G <= 0:
        raise _arg_error("`G` should be larger than zero!")
    K = float(p["K"]) if "K" in p else nan
    if "K" in p and K <= 0:
        raise _arg_error("`K` should be larger than zero!")
    lam = float(p["lambda"]) if "lambda" in p else nan
    mu = float(p["mu"]) if "mu" in p else nan
    if "mu" in p and mu <= 0:
        raise _arg_error("`μ` should be larger than zero!")
    return dict(E=E, nu=nu, G=G, K=K, lam=lam, mu=mu)


def check_elastic_params(par):
    """material_parameters.jl:174-188."""
    fin = [k for k, v in par.items() if math.isfinite(v)]
    if math.isfinite(par["G"]) and math.isfinite(par["mu"]):
        raise _arg_error("`G` and `μ` are defined! Please define either `G` or `μ`!")
    if len(fin) < 2:
        raise _arg_error("Not enough material parameters defined!")
    if len(fin) > 2:
        raise _arg_error("Too many material parameters defined!")


def get_E_and_nu(par):
    """material_parameters.jl:190-231 (same branch order)."""
    E, nu, G, K, lam, mu = (par[k] for k in ("E", "nu", "G", "K", "lam", "mu"))
    f = math.isfinite
    if f(E) and f(nu):
        pass
    elif f(E) and f(G):
        nu = E / (2 * G) - 1
    elif f(E) and f(K):
        nu = (3 * K - E) / (6 * K)
    elif f(E) and f(lam):
        nu = (-(E + lam) + math.sqrt((E + lam) ** 2 + 8 * (lam * lam))) / (4 * lam)
    elif f(E) and f(mu):
        nu = E / (2 * mu) - 1
    elif f(nu) and f(G):
        E = 2 * G * (1 + nu)
    elif f(nu) and f(K):
        E = 3 * K * (1 - 2 * nu)
    elif f(nu) and f(lam):
        E = (lam * (1 + nu) * (1 - 2 * nu)) / nu
    elif f(nu) and f(mu):
        E = 2 * mu * (1 + nu)
    elif f(G) and f(K):
        E = 9 * K * G / (3 * K + G)
        nu = (3 * K - 2 * G) / (2 * (3 * K + G))
    elif f(G) and f(lam):
        E = G * (3 * lam + 2 * G) / (lam + G)
        nu = lam / (2 * (lam + G))
    elif f(K) and f(lam):
        E = 9 * K * (K - lam) / (3 * K - lam)
        nu = lam / (3 * K - lam)
    elif f(K) and f(mu):
        E = 9 * K * mu / (3 * K + mu)
        nu = (3 * K - 2 * mu) / (2 * (3 * K + mu))
    elif f(lam) and f(mu):
        E = mu * (3 * lam + 2 * mu) / (lam + mu)
        nu = lam / (2 * (lam + mu))
    return E, nu


def get_elastic_params(p):
    """material_parameters.jl:118-129."""
    par = get_given_elastic_params(p)
    check_elastic_params(par)
    E, nu = get_E_and_nu(par)
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = G
    return dict(E=E, nu=nu, G=G, K=K, lam=lam, mu=mu)


def get_frac_params(p, delta, K):
    """get_frac_params(::CriticalStretch, ...): fracture.jl:110-130."""
    if "Gc" in p and "epsilon_c" not in p:
        Gc = float(p["Gc"])
        epsc = math.sqrt(5.0 * Gc / (9.0 * K * delta))
    elif "Gc" not in p and "epsilon_c" in p:
        epsc = float(p["epsilon_c"])
        Gc = 9.0 / 5.0 * K * delta * epsc ** 2
    elif "Gc" in p and "epsilon_c" in p:
        raise _arg_error("insufficient keywords for calculation of fracture parameters!\n"
                         "Define either Gc or epsilon_c, not both!")
    else:
        Gc, epsc = 0.0, 0.0
    return Gc, epsc


@dataclass
class PointParameters:
    """StandardPointParameters / CPointParameters (parameters.jl:107-126, correspondence.jl:148-169)."""
    delta: float
    rho: float
    E: float
    nu: float
    G: float
    K: float
    lam: float
    mu: float
    Gc: float
    epsc: float
    bc: float
    extra: dict = field(default_factory=dict)

    def has_fracture(self):
        """has_fracture(::CriticalStretch, params): fracture.jl:169-175."""
        eps = 2.220446049250313e-16
        return not (abs(self.Gc) <= eps or abs(self.epsc) <= eps)

    def key(self):
        return (self.delta, self.rho, self.E, self.nu, self.G, self.K, self.lam, self.mu,
                self.Gc, self.epsc, self.bc)


def get_point_params(mat, p):
    """StandardPointParameters(mat, p) (bond_based.jl:86-113 for BB, parameters.jl for OSB)
    and CPointParameters (correspondence.jl:162-169)."""
    for k in p:
        if k not in ALLOWED_KWARGS:
            raise _arg_error(f"keyword `{k}` is not allowed for {type(mat).__name__}")
    if "horizon" not in p:
        raise _arg_error("keyword `horizon` is required")
    delta = float(p["horizon"])
    if delta <= 0:
        raise _arg_error("`horizon` should be larger than zero!")
    if "rho" not in p:
        raise _arg_error("keyword `rho` is required")
    rho = float(p["rho"])
    if rho <= 0:
        raise _arg_error("`rho` should be larger than zero!")
    p = dict(p)
    if isinstance(mat, BBMaterial):
        par = get_given_elastic_params(p)
        nu = par["nu"]
        n_fin = sum(1 for v in par.values() if math.isfinite(v))
        if math.isfinite(nu) and not _isapprox(nu, 0.25):
            raise _arg_error("Bond-based peridynamics has a limitation on the Poisson's ratio!\n"
                             "With BBMaterial, no other values than nu=0.25 are allowed!")
        elif not math.isfinite(nu) and n_fin == 1:
            p["nu"] = 0.25
    el = get_elastic_params(p)
    if isinstance(mat, BBMaterial) and not _isapprox(el["nu"], 0.25):
        raise _arg_error("Bond-based peridynamics has a limitation on the Poisson's ratio!\n"
                         "The submitted parameter combination results in an illegal value for nu!")
    Gc, epsc = get_frac_params(p, delta, el["K"])
    bc = 18 * el["K"] / (math.pi * delta ** 4)
    return PointParameters(delta, rho, el["E"], el["nu"], el["G"], el["K"], el["lam"], el["mu"],
                           Gc, epsc, bc)
