from .model import Surrogate, build_surrogate, envelope  # noqa: F401
