"""Exception types of the solver's public API, named as in the reference (torchtt/errors.py:4-29)."""


class ShapeMismatch(Exception):
    """The shapes of the operands do not match."""


class RankMismatch(Exception):
    """The ranks of consecutive TT cores do not match."""


class IncompatibleTypes(Exception):
    """The operands are of incompatible TT types (tensor vs operator)."""


class InvalidArguments(Exception):
    """An argument has an invalid value."""
