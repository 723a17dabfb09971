"""Small stand-in for the twiggy logger used by the reference.

The reference exposes ``qmctorch.log`` built with twiggy
(qmctorch/__init__.py:9-14); call sites use ``log.info("{0}", x)`` brace
formatting and ``log.options(style="percent").info("%f" % x)``.  twiggy is not
installed, so this wraps :mod:`logging` with the same call shapes.
"""
import logging
import sys


class _Logger:
    def __init__(self, name: str, style: str = "braces"):
        self._log = logging.getLogger(name)
        if not self._log.handlers:
            h = logging.StreamHandler(sys.stdout)
            h.setFormatter(logging.Formatter("INFO:%(name)s|%(message)s"))
            self._log.addHandler(h)
            self._log.setLevel(logging.WARNING)
            self._log.propagate = False
        self._style = style

    def _fmt(self, msg, *args, **kwargs):
        if self._style == "percent" or (not args and not kwargs):
            return str(msg)
        try:
            return str(msg).format(*args, **kwargs)
        except (IndexError, KeyError, ValueError):
            return " ".join([str(msg)] + [str(a) for a in args])

    def options(self, style: str = "braces"):
        return _Logger(self._log.name, style)

    def set_level(self, level):
        self._log.setLevel(level)

    @property
    def min_level(self):
        return self._log.level

    @min_level.setter
    def min_level(self, level):
        self._log.setLevel(level)

    def debug(self, msg, *a, **k):
        if self._log.isEnabledFor(logging.DEBUG):
            self._log.debug(self._fmt(msg, *a, **k))

    def info(self, msg, *a, **k):
        if self._log.isEnabledFor(logging.INFO):
            self._log.info(self._fmt(msg, *a, **k))

    def warning(self, msg, *a, **k):
        self._log.warning(self._fmt(msg, *a, **k))

    def error(self, msg, *a, **k):
        self._log.error(self._fmt(msg, *a, **k))


log = _Logger("QMCTorch")
