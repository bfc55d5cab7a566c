"""Two-sphere lubrication tables (lubdat/*.dat, distributed with the reference, read by no reference
source) and the library's interpolant of them.

Tables, each with the centre distance r (radius 1) in the first column:
  0  r2babc.dat  X11A X12A Y11A Y12A Y11B Y12B X11C X12C Y11C Y12C
  1  r2bgh.dat   X11G X12G Y11G Y12G Y11H Y12H
  2  r2bm.dat    XM YM ZM
r2bghd.dat holds the r-derivatives of r2bgh.dat; it is not needed by the interpolant (which takes
its own monotone slopes) and serves as a check of it.
"""
import ctypes as C
import os

import numpy as np

from .polystokes import lib

LUBDAT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lubdat")
TABLE_FILES = ("r2babc.dat", "r2bgh.dat", "r2bm.dat")
COLUMNS = (
    ("X11A", "X12A", "Y11A", "Y12A", "Y11B", "Y12B", "X11C", "X12C", "Y11C", "Y12C"),
    ("X11G", "X12G", "Y11G", "Y12G", "Y11H", "Y12H"),
    ("XM", "YM", "ZM"),
)


def load_table(name, directory=LUBDAT):
    """Rows of one table file (whitespace-separated, Fortran E notation) as float64."""
    return np.loadtxt(os.path.join(directory, name), ndmin=2)


def load_tables(directory=LUBDAT):
    return [load_table(f, directory) for f in TABLE_FILES]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def interp(rows, r):
    """The library's monotone cubic interpolant of every column of `rows` at distance r."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros(rows.shape[1] - 1)
    rc = lib().ps_lub_interp(_dp(rows), rows.shape[0], rows.shape[1], float(r), _dp(out))
    if rc != 0:
        raise ValueError("ps_lub_interp: invalid table")
    return out


def slopes(rows):
    """The interpolant's slopes at the nodes, shape (nrows, ncols - 1)."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    out = np.zeros((rows.shape[0], rows.shape[1] - 1))
    rc = lib().ps_lub_slopes(_dp(rows), rows.shape[0], rows.shape[1], _dp(out))
    if rc != 0:
        raise ValueError("ps_lub_slopes: invalid table")
    return out
