"""Model constructors (tedeous/models.py)."""

from typing import Any

import torch


def mat_model(domain: Any, equation: Any, nn_model: torch.nn.Module = None) -> torch.Tensor:
    """Grid tensor ``[n_eq, n_1, ...]`` for 'mat' mode (tedeous/models.py:198-226):
    ones, or the values of ``nn_model`` on the grid points."""
    grid = domain.build('mat')
    shape = [len(equation.equation_lst)] + list(grid.shape)[1:]
    if nn_model is not None:
        pts = torch.vstack([grid[i].reshape(-1) for i in range(grid.shape[0])]).T.float()
        return nn_model(pts).detach().reshape(shape)
    return torch.ones(shape)
