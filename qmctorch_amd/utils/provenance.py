"""Version string with the checkout's git description (qmctorch/utils/provenance.py:6-26)."""
import os
import subprocess

from ..__version__ import __version__


def get_git_tag() -> str:
    """"<package version> - <git describe --always>", or a note when the
    package is not inside a git checkout."""
    try:
        tag = subprocess.check_output(["git", "describe", "--always"], cwd=os.path.dirname(os.path.abspath(__file__)),
                                      stderr=subprocess.DEVNULL).decode("utf-8").strip()
        return "%s - %s" % (__version__, tag)
    except Exception:
        return "%s - hash commit not found" % __version__
