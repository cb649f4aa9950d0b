"""MI355X-native MolFlow+ test-particle Monte Carlo tracer (host package)."""
from .model import Model, Parameter, default_facets, default_settings, parse_user_moment  # noqa: F401
from . import geometry  # noqa: F401
