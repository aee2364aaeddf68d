"""MI355X-native oxDNA2 molecular-dynamics hot path.

Modules:
  systems  synthetic workloads (B-helix duplexes, hairpins, bundles, jobs)
  io       oxDNA input / topology / configuration formats
  backend  HIPBatch: systems resident on the GPU through the C ABI
  build    compiles csrc/*.hip into liboxh.so for gfx950
"""
from . import io, systems  # noqa: F401

__all__ = ["io", "systems"]
