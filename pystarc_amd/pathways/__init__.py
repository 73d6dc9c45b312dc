from .reaction_interface import ReactionInterface, PathwaySet  # noqa: F401
