"""MI355X-native uplink LDPC decode path: Python host layer over the HIP C ABI.

Mirrors the reference's pyAerial LDPC surface (aerial.phy5g.ldpc) and its
C++ convenience layer; every call goes through lib/libcuphy_hip.so.
"""
from . import _lib  # noqa: F401
