"""Pin of the oracle's LevelSetGetVelocity (LevelSetMethod/Advect.c:145-192):
a known-answer test in which every fluid value encodes its own node and
component, so the expected window is written down directly: node (i, j, k)
of the window holds the codes of global node (p + (i, j, k)) and components
U_FACE..W_FACE (1..3) where that node is in the domain, and keeps its old
value elsewhere."""
import numpy as np

import oracle as O
from oracle import Win


def _code(x, y, z, c):
    return x + 1000.0 * y + 1e6 * z + 0.1 * c


def test_known_answer_3d():
    N = (7, 6, 5)  # (Nz, Ny, Nx)
    z, y, x, c = np.meshgrid(*(np.arange(s) for s in N + (4,)), indexing="ij")
    fluid = _code(x, y, z, c)
    for p, n in (((1, 2, 1), (3, 2, 4)), ((-2, -1, -3), (4, 5, 6)), ((3, 4, 5), (5, 5, 5)),
                 ((9, 0, 0), (2, 2, 2))):
        w = Win(p, n, (1, 1, 1))
        vel = np.full((n[2], n[1], n[0], 3), -7.0)
        got = O.get_velocity(fluid, w, vel)
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(n[0]):
                    X, Y, Z = p[0] + i, p[1] + j, p[2] + k
                    inside = 0 <= X < N[2] and 0 <= Y < N[1] and 0 <= Z < N[0]
                    want = [_code(X, Y, Z, cc) for cc in (1, 2, 3)] if inside else [-7.0] * 3
                    assert list(got[k, j, i]) == want, (p, n, i, j, k)


def test_known_answer_2d():
    N = (6, 8)  # (Ny, Nx)
    y, x, c = np.meshgrid(np.arange(N[0]), np.arange(N[1]), np.arange(3), indexing="ij")
    fluid = _code(x, y, 0, c)
    p, n = (-2, 3, 0), (5, 6, 1)
    w = Win(p, n, (1, 1, 1))
    vel = np.full((n[1], n[0], 2), -7.0)
    got = O.get_velocity(fluid, w, vel)
    for j in range(n[1]):
        for i in range(n[0]):
            X, Y = p[0] + i, p[1] + j
            inside = 0 <= X < N[1] and 0 <= Y < N[0]
            want = [_code(X, Y, 0, 1), _code(X, Y, 0, 2)] if inside else [-7.0, -7.0]
            assert list(got[j, i]) == want
