"""C3D8 (reference: elements/C3/brick.py:6-201)."""
from .C3base import Element_3D


class C3D8(Element_3D):
    """Trilinear brick, 2x2x2 Gauss points; faces 0321, 4567, 0154, 1265, 2376, 0473."""
    etype = 8
    ngp = 8
    face_nodes = {0: [0, 3, 2, 1], 1: [4, 5, 6, 7], 2: [0, 1, 5, 4], 3: [1, 2, 6, 5],
                  4: [2, 3, 7, 6], 5: [0, 4, 7, 3]}
