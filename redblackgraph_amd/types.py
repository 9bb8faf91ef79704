"""Result types (redblackgraph/types/transitive_closure.py:1-7)."""
from dataclasses import dataclass

import numpy as np


@dataclass
class TransitiveClosure:
    W: np.array
    diameter: int
