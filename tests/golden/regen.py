"""Regenerate the inputs of the reference's golden-vector scripts.

tests/test-demag/cmp-ref.py:24-51 and tests/test-exchange/cmp-ref.py:26-64 draw
Python's MT19937 stream after random.seed(0); Python 3's random.seed(int) and
random() reproduce Python 2's stream.  Loop order is k (z), j (y), i (x), and
the arrays are user-ordered [c][x][y][z]; they are returned here in internal
order (component 0 = user z, axes (z, y, x)).
"""
import math
import random

import numpy as np


def _to_internal_scalar(a):
    # user [x][y][z] -> internal [z][y][x]
    return np.ascontiguousarray(np.transpose(a, (2, 1, 0)))


def _to_internal_vector(a):
    # user [c][x][y][z] -> internal [2-c][z][y][x]
    return np.ascontiguousarray(np.transpose(a[::-1], (0, 3, 2, 1)))


def demag_inputs(n=64):
    """Msat0T0 mask and unit m of tests/test-demag/cmp-ref.py (seed 0)."""
    random.seed(0)
    Nx = Ny = Nz = n
    msk = np.zeros((Nx, Ny, Nz))
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                msk[i, j, k] = random.random()
    m = np.zeros((3, Nx, Ny, Nz))
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                mx = random.random()
                my = random.random()
                mz = random.random()
                l = math.sqrt(mx ** 2 + my ** 2 + mz ** 2)
                m[0, i, j, k] = mx / l
                m[1, i, j, k] = my / l
                m[2, i, j, k] = mz / l
    return _to_internal_scalar(msk), _to_internal_vector(m)


def exchange_inputs(n=64):
    """Msat0T0 mask, lex mask (1/msk) and unit m of tests/test-exchange/cmp-ref.py."""
    random.seed(0)
    Nx = Ny = Nz = n
    msk = np.zeros((Nx, Ny, Nz))
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                msk[i, j, k] = random.random()
    lex = np.zeros((Nx, Ny, Nz))
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                random.random()  # the script's unused rnd draw
                lex[i, j, k] = math.sqrt(1.0 / (msk[i, j, k] ** 2.0))
    m = np.zeros((3, Nx, Ny, Nz))
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                mx = float(random.random())
                my = float(random.random())
                mz = float(random.random())
                l = math.sqrt(mx ** 2.0 + my ** 2.0 + mz ** 2.0)
                m[0, i, j, k] = mx / l
                m[1, i, j, k] = my / l
                m[2, i, j, k] = mz / l
    return _to_internal_scalar(msk), _to_internal_scalar(lex), _to_internal_vector(m)


# physical set-up shared by both scripts (user x, y, z sizes / 64 cells)
CELL_USER = (320e-9 / 64, 160e-9 / 64, 64e-9 / 64)
CELL_INTERNAL = CELL_USER[::-1]
MU0_PY = 1.25663706e-6  # the Python client's mu0, src/python/hotspin.py.tail:4
