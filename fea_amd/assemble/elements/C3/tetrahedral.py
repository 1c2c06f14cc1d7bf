"""C3D4 / C3D10 (reference: elements/C3/tetrahedral.py)."""
from .C3base import Element_3D


class C3D4(Element_3D):
    """Linear tet, 1 Gauss point (tetrahedral.py:6-64); faces 021, 013, 123, 032."""
    etype = 4
    ngp = 1
    face_nodes = {0: [0, 2, 1], 1: [0, 1, 3], 2: [1, 2, 3], 3: [0, 3, 2]}


class C3D10(Element_3D):
    """Quadratic tet, 4 Gauss points (tetrahedral.py:66-181).

    Mid nodes 4(01) 5(12) 6(02) 7(03) 8(13) 9(23); every face is a full-order
    T6 facet (the default surf_order of C3base.py:82 keeps all faces quadratic).
    """
    etype = 10
    ngp = 4
    face_nodes = {0: [0, 2, 1, 6, 5, 4], 1: [0, 1, 3, 4, 8, 7],
                  2: [1, 2, 3, 5, 9, 8], 3: [0, 3, 2, 7, 9, 6]}
    empty_surface_is_dropped = True
