"""Material objects with pSierra's interface, evaluated on the GPU.

* ``IsotropicElasticParameters`` -- the moduli container of
  src/psierra/material/parameters.py:6-138 (the setters the hot path uses).
* ``NeoHookean`` -- src/psierra/material/neohookean.py:9-122, one parameter
  set for every element.
* ``NeoHookeanField`` -- per-element (mu_e, lambda_e), with
  ``supports_element_indexing = True`` following the elem_index protocol of
  src/psierra/material/viscoelastic.py:80,315-343 and
  src/psierra/element/tet4_total_lagrangian.py:114-120.

``compute_stress_and_tangent`` runs the device material kernel
(``psx_nh_material_batch``) on a single F, as a drop-in for code that calls
the per-point API; the assembly path never goes through it.
"""
from __future__ import annotations

import numpy as np


class IsotropicElasticParameters:
    def __init__(self) -> None:
        self._youngs_modulus = None
        self._poissons_ratio = None
        self._shear_modulus = None
        self._lame_parameter_1 = None
        self._lame_parameter_2 = None
        self._bulk_modulus = None
        self._p_wave_modulus = None
        self._mass_density = None

    @staticmethod
    def _validate_density(rho: float) -> None:
        if rho <= 0.0:
            raise ValueError(f"Mass density must be positive. Got {rho}")

    def set_from_youngs_poisson(self, youngs_modulus, poissons_ratio, mass_density) -> None:
        E, nu, rho = float(youngs_modulus), float(poissons_ratio), float(mass_density)
        if E <= 0.0:
            raise ValueError(f"Young's modulus must be positive. Got {E}")
        if not (-1.0 < nu < 0.5):
            raise ValueError(f"Poisson's ratio must satisfy -1.0 < poissons_ratio < 0.5. Got {nu}")
        self._validate_density(rho)
        mu = E / (2.0 * (1.0 + nu))
        lam = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
        self._youngs_modulus, self._poissons_ratio, self._mass_density = E, nu, rho
        self._shear_modulus = self._lame_parameter_2 = mu
        self._lame_parameter_1 = lam
        self._bulk_modulus = E / (3.0 * (1.0 - 2.0 * nu))
        self._p_wave_modulus = lam + 2.0 * mu

    def set_from_bulk_shear(self, bulk_modulus, shear_modulus, mass_density) -> None:
        kappa, mu, rho = float(bulk_modulus), float(shear_modulus), float(mass_density)
        if kappa <= 0.0:
            raise ValueError(f"Bulk modulus must be positive. Got {kappa}")
        if mu <= 0.0:
            raise ValueError(f"Shear modulus must be positive. Got {mu}")
        self._validate_density(rho)
        self._bulk_modulus, self._shear_modulus, self._lame_parameter_2 = kappa, mu, mu
        self._mass_density = rho
        self._lame_parameter_1 = kappa - (2.0 / 3.0) * mu
        self._youngs_modulus = (9.0 * kappa * mu) / (3.0 * kappa + mu)
        self._poissons_ratio = (3.0 * kappa - 2.0 * mu) / (2.0 * (3.0 * kappa + mu))
        self._p_wave_modulus = kappa + (4.0 / 3.0) * mu

    def set_from_lame_parameters(self, lame_parameter_1, lame_parameter_2, mass_density) -> None:
        lam, mu, rho = float(lame_parameter_1), float(lame_parameter_2), float(mass_density)
        if mu <= 0.0:
            raise ValueError(f"Lamé parameter 2 (shear modulus) must be positive. Got {mu}")
        if (lam + mu) <= 0.0:
            raise ValueError(
                f"Constraint (lame_parameter_1 + lame_parameter_2 > 0) violated. Got lambda={lam}, mu={mu}"
            )
        self._validate_density(rho)
        self._lame_parameter_1, self._lame_parameter_2, self._shear_modulus = lam, mu, mu
        self._mass_density = rho
        self._bulk_modulus = lam + (2.0 / 3.0) * mu
        self._youngs_modulus = (mu * (3.0 * lam + 2.0 * mu)) / (lam + mu)
        self._poissons_ratio = lam / (2.0 * (lam + mu))
        self._p_wave_modulus = lam + 2.0 * mu

    youngs_modulus = property(lambda s: s._youngs_modulus)
    poissons_ratio = property(lambda s: s._poissons_ratio)
    shear_modulus = property(lambda s: s._shear_modulus)
    lame_parameter_1 = property(lambda s: s._lame_parameter_1)
    lame_parameter_2 = property(lambda s: s._lame_parameter_2)
    bulk_modulus = property(lambda s: s._bulk_modulus)
    p_wave_modulus = property(lambda s: s._p_wave_modulus)
    mass_density = property(lambda s: s._mass_density)


def lame_arrays(material, n_elements: int) -> tuple[np.ndarray, np.ndarray]:
    """Per-element (mu, lambda) float64 arrays for any supported material.

    Accepts ``NeoHookeanField``, ``NeoHookean`` here or pSierra's own
    ``NeoHookean`` (anything exposing ``parameters.shear_modulus`` and
    ``parameters.bulk_modulus``; lambda = K - 2 mu / 3 so that the kernel's
    K = lambda + 2 mu / 3 reproduces the stored bulk modulus).
    """
    if isinstance(material, NeoHookeanField):
        if material.mu.size != n_elements:
            raise ValueError(f"material has {material.mu.size} elements, mesh has {n_elements}")
        return material.mu, material.lam
    params = getattr(material, "parameters", None)
    if params is None or not hasattr(params, "shear_modulus"):
        raise TypeError(f"unsupported material {type(material).__name__}: the engine evaluates Neo-Hookean only")
    mu = float(params.shear_modulus)
    lam = float(params.bulk_modulus) - (2.0 / 3.0) * mu
    return np.full(n_elements, mu), np.full(n_elements, lam)


class NeoHookean:
    """Compressible Neo-Hookean (uniform parameters), GPU-evaluated."""

    def __init__(self, parameters: IsotropicElasticParameters) -> None:
        self._parameters = parameters

    @property
    def parameters(self) -> IsotropicElasticParameters:
        return self._parameters

    @property
    def mass_density(self) -> float:
        return self._parameters.mass_density

    def compute_constitutive_matrix(self) -> np.ndarray:
        """Small-strain isotropic C (src/psierra/material/neohookean.py:33-53)."""
        lam = self._parameters.lame_parameter_1
        mu = self._parameters.shear_modulus
        C = np.zeros((6, 6))
        C[0, 0] = C[1, 1] = C[2, 2] = lam + 2.0 * mu
        C[0, 1] = C[1, 0] = C[0, 2] = C[2, 0] = C[1, 2] = C[2, 1] = lam
        C[3, 3] = C[4, 4] = C[5, 5] = mu
        return C

    def compute_stress_and_tangent(self, deformation_gradient: np.ndarray, precision: str = "fp64"):
        from .engine import material_batch

        F = np.asarray(deformation_gradient, dtype=np.float64).reshape(1, 3, 3)
        C, S = material_batch(F, self._parameters.shear_modulus, self._parameters.bulk_modulus, precision)
        return C[0].cpu().numpy(), S[0].cpu().numpy()


class NeoHookeanField:
    """Compressible Neo-Hookean with per-element (mu_e, lambda_e)."""

    supports_element_indexing: bool = True

    def __init__(self, mu: np.ndarray, lam: np.ndarray, mass_density: float = 1.0) -> None:
        self.mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        if self.mu.shape != self.lam.shape:
            raise ValueError("mu and lam must have the same length")
        if np.any(self.mu <= 0.0) or np.any(self.lam + self.mu <= 0.0):
            raise ValueError("need mu > 0 and lambda + mu > 0 for every element")
        self._mass_density = float(mass_density)
        self.version = 0

    def set_parameters(self, mu: np.ndarray, lam: np.ndarray) -> None:
        self.mu = np.ascontiguousarray(mu, dtype=np.float64).reshape(-1)
        self.lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        self.version += 1

    @property
    def mass_density(self) -> float:
        return self._mass_density

    def compute_stress_and_tangent(self, deformation_gradient, elem_index: int | None = None, precision="fp64"):
        from .engine import material_batch

        e = 0 if elem_index is None else int(elem_index)
        mu = float(self.mu[e])
        bulk = float(self.lam[e]) + (2.0 / 3.0) * mu
        F = np.asarray(deformation_gradient, dtype=np.float64).reshape(1, 3, 3)
        C, S = material_batch(F, mu, bulk, precision)
        return C[0].cpu().numpy(), S[0].cpu().numpy()


class MooneyRivlinField:
    """Compressible Mooney-Rivlin (src/psierra/material/mooney_rivlin.py) with
    per-element (c10, c01, K); evaluated by the assembly kernels."""

    supports_element_indexing: bool = True

    def __init__(self, c10, c01, bulk_modulus, mass_density: float = 1.0) -> None:
        self.c10, self.c01, self.bulk = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                                         for a in (c10, c01, bulk_modulus))
        if not (self.c10.shape == self.c01.shape == self.bulk.shape):
            raise ValueError("c10, c01 and bulk_modulus must have the same length")
        if np.any(self.c10 + self.c01 <= 0.0) or np.any(self.bulk <= 0.0):
            raise ValueError("need c10 + c01 > 0 and K > 0 for every element")
        self._mass_density = float(mass_density)
        self.version = 0

    @property
    def mass_density(self) -> float:
        return self._mass_density


class YeohField:
    """Compressible Yeoh (src/psierra/material/yeoh.py) with per-element
    (c10, c20, c30, K); evaluated by the assembly kernels."""

    supports_element_indexing: bool = True

    def __init__(self, c10, c20, c30, bulk_modulus, mass_density: float = 1.0) -> None:
        self.c10, self.c20, self.c30, self.bulk = (np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                                                   for a in (c10, c20, c30, bulk_modulus))
        if not (self.c10.shape == self.c20.shape == self.c30.shape == self.bulk.shape):
            raise ValueError("c10, c20, c30 and bulk_modulus must have the same length")
        if np.any(self.c10 <= 0.0) or np.any(self.bulk <= 0.0):
            raise ValueError("need c10 > 0 and K > 0 for every element")
        self._mass_density = float(mass_density)
        self.version = 0

    @property
    def mass_density(self) -> float:
        return self._mass_density


def material_model(material, n_elements: int):
    """(model name, per-element parameter arrays) of a supported material:
    Neo-Hookean (mu, lambda), Mooney-Rivlin (c10, c01, K), Yeoh (c10, c20,
    c30, K).  Accepts the *Field classes here and pSierra's own materials
    (their ``parameters`` expose c10/c01/c20/c30 and bulk_modulus)."""

    def full(v):
        a = np.asarray(v, dtype=np.float64).reshape(-1)
        if a.size == 1:
            return np.full(n_elements, float(a[0]))
        if a.size != n_elements:
            raise ValueError(f"material has {a.size} elements, mesh has {n_elements}")
        return a

    if isinstance(material, MooneyRivlinField):
        return "mooney_rivlin", [full(material.c10), full(material.c01), full(material.bulk)]
    if isinstance(material, YeohField):
        return "yeoh", [full(material.c10), full(material.c20), full(material.c30), full(material.bulk)]
    params = getattr(material, "parameters", None)
    if params is not None and hasattr(params, "c01") and hasattr(params, "c10"):
        return "mooney_rivlin", [full(params.c10), full(params.c01), full(params.bulk_modulus)]
    if params is not None and hasattr(params, "c20") and hasattr(params, "c30"):
        return "yeoh", [full(params.c10), full(params.c20), full(params.c30), full(params.bulk_modulus)]
    mu, lam = lame_arrays(material, n_elements)
    return "neohookean", [mu, lam]
