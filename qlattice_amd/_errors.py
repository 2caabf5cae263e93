"""The error raised by failed library calls."""


class QlatDwfError(RuntimeError):
    """A call of the HIP library returned a non-zero status, or the library
    cannot be loaded (the reference raises RuntimeError from its C API,
    cqlat/lib.h:28-35)."""
