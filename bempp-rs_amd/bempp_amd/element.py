"""Reference cells and Lagrange element families (the parts of ndelement the
reference's python layer uses: python/bempp/function_space.py:197-223,335-338)."""
from enum import Enum, IntEnum


class ReferenceCellType(IntEnum):
    """ndelement ReferenceCellType numbering."""

    Point = 0
    Interval = 1
    Triangle = 2
    Quadrilateral = 3
    Tetrahedron = 4
    Hexahedron = 5
    Prism = 6
    Pyramid = 7


class Family(Enum):
    Lagrange = 0


class Continuity(IntEnum):
    Standard = 0
    Discontinuous = 1


class LagrangeElementFamily:
    """Lagrange family of a given degree and continuity (degree 0 is always
    discontinuous; supported degrees: 0 and 1)."""

    def __init__(self, degree, continuity=Continuity.Standard, dtype=None):
        if degree not in (0, 1):
            raise ValueError("only Lagrange degree 0 and 1 are supported")
        self.family = Family.Lagrange
        self.degree = int(degree)
        self.continuity = Continuity(continuity)
        self.dtype = dtype

    def __repr__(self):
        return f"LagrangeElementFamily(degree={self.degree}, continuity={self.continuity.name})"


def create_family(family, degree, continuity=Continuity.Standard, dtype=None):
    """Mirror of ndelement.ciarlet.create_family for the Lagrange family."""
    if family != Family.Lagrange:
        raise ValueError(f"unsupported family: {family}")
    return LagrangeElementFamily(degree, continuity, dtype)
