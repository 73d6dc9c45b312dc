from .nam_simulator import NAMParameters  # noqa: F401
