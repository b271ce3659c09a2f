"""Host-side checks of the differentiable rollout's Python layer (no GPU):
a solid with friction is refused before anything reaches the device, since
the rollout's velocity constraint and its adjoint are those of a
frictionless solid (the engine's constraint with friction 0,
fluidsimulation.cpp:6391-6442)."""
import numpy as np
import pytest

from ffengine.rollout import Rollout


@pytest.mark.parametrize("friction", [0.25, 1.0])
def test_rollout_refuses_a_solid_with_friction(friction):
    n = 8
    pos = np.full((4, 3), 0.5, np.float32)
    solid = np.ones((n + 1, n + 1, n + 1), np.float32)
    with pytest.raises(ValueError, match="friction"):
        Rollout((n, n, n), 1.0 / n, pos, solid_nodes=solid, solid_friction=friction)
