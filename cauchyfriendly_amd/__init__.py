"""MI355X (gfx950) HIP implementation of the Multivariate Cauchy Estimator's
term-propagation step, behind the reference's pycauchy / CauchyEstimator API."""
from . import _native  # noqa: F401
