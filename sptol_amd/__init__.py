"""sptol_amd: ParTI!-compatible MTTKRP / CPD-ALS on MI355X.

The product is the C-ABI library libsptol_amd.so (include/ParTI.h,
include/sptol_hip.h), built from sptol_amd/csrc.  This package only loads it
(sptol_amd.parti) and provides seeded synthetic inputs (sptol_amd.gen).
"""
from .parti import load_library  # noqa: F401
