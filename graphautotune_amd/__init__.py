"""MI355X-native subgraph pattern-matching engine (GraphAutoTune `pm` path).

Host C++ and hand-written gfx950 HIP kernels behind the C ABI of
include/gat.h (libgat.so); this package is the Python side of that ABI.
"""
from .core import (CPU_DEVICE, GPU_DEVICE, DeviceContext, Executor, GlobalMemoryGraph, Graph, Pattern, Plan,
                   Schedule, count, default_config, device_count, schedule_candidates)

__all__ = ["Pattern", "Schedule", "GlobalMemoryGraph", "Graph", "DeviceContext", "Executor", "Plan", "count",
           "default_config", "device_count", "schedule_candidates", "CPU_DEVICE", "GPU_DEVICE"]
