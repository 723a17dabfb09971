"""Unit conversions, same values as qmctorch/utils/constants.py:1-2."""
ANGS2BOHR = 1.8897259886
BOHR2ANGS = 0.529177
