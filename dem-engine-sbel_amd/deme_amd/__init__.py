"""MI355X-native DEM hot path: Python plumbing over libdeme.so (C ABI in include/deme/deme.h).

Device memory, kernels and the step loop live in libdeme.so; this package only loads it through ctypes and builds
synthetic workloads for tests and benchmarks.
"""
from .abi import (DeviceSystem, DemeError, SimParams, build_sim_params, decode_positions, device_count,
                  encode_positions, load)

__all__ = ["DeviceSystem", "DemeError", "SimParams", "build_sim_params", "decode_positions", "device_count",
           "encode_positions", "load"]
