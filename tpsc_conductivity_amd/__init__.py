"""MI355X-native TPSC hot path (chi0 -> Usp/Uch -> V -> Sigma -> mu -> Lambda_jj).

All computation runs in libtpsc_hip.so (hand-written HIP for gfx950) behind the C ABI of
include/tpsc_capi.h; this package is the thin host binding.
"""
import os
import subprocess

from .capi import LIB_PATH, TPSCError, exported_symbols  # noqa: F401

_CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")


def build(force=True, jobs=4):
    """Compile libtpsc_hip.so for gfx950 with hipcc (works without a GPU)."""
    args = ["make", "-C", _CSRC, f"-j{jobs}"]
    if force:
        subprocess.run(["make", "-C", _CSRC, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(args, check=True)
    return LIB_PATH


def __getattr__(name):
    if name == "CondOpt":
        from .tpsc import CondOpt

        return CondOpt
    raise AttributeError(name)
