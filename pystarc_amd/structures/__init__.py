from .molecules import Atom, Molecule, ContactPair, ReactionCriteria  # noqa: F401
