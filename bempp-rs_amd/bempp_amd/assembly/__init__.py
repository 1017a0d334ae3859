"""Assembly (python/bempp/assembly of the reference)."""
from . import batch, boundary  # noqa: F401
