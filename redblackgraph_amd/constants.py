"""Red-black constants (redblackgraph/constants.py:71-164)."""
import numpy as np

BLACK_ONE = 1
RED_ONE = -1


def red_one_for_dtype(dtype):
    dt = np.dtype(dtype)
    if dt.kind == "u":
        return np.iinfo(dt).max
    return RED_ONE


def black_one_for_dtype(dtype):
    return BLACK_ONE


def is_red_one(value, dtype):
    dt = np.dtype(dtype)
    if dt.kind == "u":
        return value == np.iinfo(dt).max
    return value == RED_ONE


def is_black_one(value):
    return value == BLACK_ONE
