from .grid_force import DXGrid  # noqa: F401
