"""Result types (redblackgraph/types/transitive_closure.py:1-7, types/ordering.py:1-9)."""
from dataclasses import dataclass
from typing import Dict, Sequence

import numpy as np


@dataclass
class TransitiveClosure:
    W: np.array
    diameter: int


@dataclass
class Ordering:
    A: Sequence[Sequence[int]]
    label_permutation: Sequence[int]
    components: Dict[int, int]
