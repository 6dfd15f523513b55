"""Type aliases (fdtd/typing_.py)."""

from typing import List, Tuple, Union

import numpy as np
import torch

Number = Union[int, float]
Tensorlike = Union[np.ndarray, torch.Tensor]
ListOrSlice = Union[List[int], slice]
IntOrSlice = Union[int, slice]

__all__ = ["List", "Tuple", "Number", "Tensorlike", "ListOrSlice", "IntOrSlice"]
