"""tessera_amd: an MI355X (gfx950) prover back end for Tessera's data-parallel core.

Hand-written HIP kernels behind a C ABI (include/tessera_mi355x.h, libtessera_mi355x.so): Goldilocks
Reed-Solomon row encoding, BLAKE3 column chains and Merkle commitments, and sumcheck / LogUp rounds.
This package is the Python mirror of the reference crates' API over that ABI.
"""
from .device import Device, DeviceBuffer  # noqa: F401
from .lib import BadShape, Declined, TsrError, declared_symbols  # noqa: F401
from .pcs import (ColumnAnswer, ColumnChains, CommitmentParamsPy, NttPlan, ReferenceCommitment,  # noqa: F401
                  RowEncoder, merkle_root)

__version__ = "0.1.0"
