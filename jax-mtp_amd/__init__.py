"""MI355X-native Moment Tensor Potential evaluator (drop-in for pair_style jax/mtp).

The compute path is libmtp_hip.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/mtp_hip.h).  This package holds the host-side mirror of
the reference's Python interface: the .mtp format (mtp_io), the padded
evaluation entry point (potential.MTPPotential), seeded workloads (configs),
neighbor lists (neighbors), the training path (training) and the distilled
head (head).

The directory name contains a hyphen, so import it with
``importlib.import_module("jax-mtp_amd")``.
"""
from . import capi, cfg_io, configs, distill, head, mtp_io, neighbors, training  # noqa: F401
from .mtp_io import MTPData, format_mtp, parse_mtp_text, read_mtp, untrained, write_mtp  # noqa: F401
from .potential import MTPPotential  # noqa: F401

__all__ = ["MTPData", "MTPPotential", "read_mtp", "write_mtp", "format_mtp", "parse_mtp_text", "untrained",
           "capi", "configs", "head", "mtp_io", "neighbors", "training"]
