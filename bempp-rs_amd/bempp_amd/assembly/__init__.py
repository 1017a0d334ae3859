"""Assembly (python/bempp/assembly of the reference)."""
from . import batch, boundary, potential  # noqa: F401
