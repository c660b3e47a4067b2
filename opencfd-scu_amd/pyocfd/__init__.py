"""Host-side Python mirror of the OpenCFD-SCU MI355X right-hand-side library.

``capi``  - ctypes bindings of libocfd_hip.so (include/ocfd.h).
``cases`` - seeded generators of the benchmark workloads and their input files.
"""
from . import capi, cases  # noqa: F401
