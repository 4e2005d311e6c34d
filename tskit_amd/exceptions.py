"""Exceptions, mirroring the reference's (python/tskit/exceptions.py and
_tskitmodule.c:231-304, 12152-12158): library errors carry the C error number."""


class TskitException(Exception):
    """Base class of the package's exceptions."""


class LibraryError(TskitException):
    """A negative error number returned by the C library.

    ``err`` holds the number; it is the same value the reference C library
    returns for the same input (c/tskit/core.h).
    """

    def __init__(self, err, message=None):
        self.err = err
        if message is None:
            from . import _lib

            message = _lib.lib().tsg_strerror(err).decode()
        super().__init__(message)


def check(err):
    if err != 0:
        raise LibraryError(err)
