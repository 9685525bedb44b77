"""MI355X-native KPolaris imaging kernel: Python host layer over the C ABI.

The computation lives in the HIP library kpolaris_amd/_lib/libkpolaris_amd.so
(built by ``python -m kpolaris_amd.build_lib``); ``kpolaris_amd.image`` binds
it.  Importing the package does not load the library.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
