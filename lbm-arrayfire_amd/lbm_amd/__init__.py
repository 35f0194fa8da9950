"""lbm_amd — host-side Python mirror of the MI355X lattice Boltzmann library (liblbm_amd.so).

The compute path is the HIP library behind include/lbm_capi.h; this package only binds it
(ctypes), builds node-type masks and case descriptors, and drives batches and the job server.
"""
from . import capi, cases, geometry  # noqa: F401
from .capi import Batch, Lattice, LbmError, Server, ShanChen, load  # noqa: F401
