"""Assembly (python/bempp/assembly of the reference)."""
from . import boundary  # noqa: F401
