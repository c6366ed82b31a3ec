"""Exceptions raised by the grid API; names follow the Julia exceptions the
reference throws for the same conditions (see the C-ABI return codes)."""


class SpringsteelError(Exception):
    pass


class ArgumentError(SpringsteelError, ValueError):
    pass


class DomainError(SpringsteelError, ValueError):
    pass


class PosDefException(SpringsteelError, ArithmeticError):
    pass


class BoundsError(SpringsteelError, IndexError):
    pass


class NotSupportedError(SpringsteelError, NotImplementedError):
    """A feature of the reference this library does not run (C-ABI SSG_ENOTSUP)."""


class HIPError(SpringsteelError, RuntimeError):
    pass


class ExtensionMissingError(SpringsteelError, ImportError):
    """The HIP library has not been built or cannot be loaded."""
