"""ff-room on MI355X: the reference's scene/solver API over a gfx950 GPU core.

Exports the hot-path API of ddd3h/ff-room (python/ff_room/__init__.py):
configuration, Scene, SolverBridge, FieldSnapshot, SimResult. The rendering
helpers of the reference (Visualizer, ScenePlotter, plot_scene, Animator)
are not part of this build; asking for them raises ImportError.
"""
from .config import (FanConfig, ObstacleConfig, OpeningConfig, RoomConfig, SceneConfig,
                     SolverConfig)
from .scene import Scene
from .solver_bridge import FieldSnapshot, SimResult, SolverBridge

__all__ = [
    "RoomConfig", "FanConfig", "ObstacleConfig", "OpeningConfig", "SolverConfig",
    "SceneConfig", "Scene", "SolverBridge", "FieldSnapshot", "SimResult",
]

_NOT_BUILT = {"Visualizer", "ScenePlotter", "plot_scene", "Animator"}


def __getattr__(name):
    if name in _NOT_BUILT:
        raise ImportError(f"ff_room.{name} (rendering) is not part of the MI355X build")
    if name == "ResultStore":
        from .io import ResultStore
        return ResultStore
    raise AttributeError(name)
