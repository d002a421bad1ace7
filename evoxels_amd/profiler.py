"""Wall-time and memory report for the 'hip' backend (the interface of
evoxels/profiler.py:9-94: ``update_memory_stats`` / ``print_memory_stats``)."""
from __future__ import annotations

import os
import tracemalloc

import numpy as np
import psutil
import torch


class HipMemoryProfiler:
    def __init__(self, device):
        self.device = torch.device(device)
        self.max_used_cpu = 0.0
        tracemalloc.start()
        torch.cuda.reset_peak_memory_stats(device=self.device)

    def update_memory_stats(self):
        used_cpu = psutil.Process(os.getpid()).memory_info().rss / 1024 ** 2
        self.max_used_cpu = max(self.max_used_cpu, used_cpu)

    def print_memory_stats(self, start, end, iters):
        print(f"Wall time: {np.around(end - start, 4)} s after {iters} iterations "
              f"({np.around((end - start) / max(iters, 1), 4)} s/iter)")
        self.update_memory_stats()
        print(f"CPU-RAM (psutil)      current: {psutil.Process(os.getpid()).memory_info().rss / 1024 ** 2:.2f} MB "
              f"({self.max_used_cpu:.2f} MB max)")
        print(f"GPU-RAM (torch)       current: {torch.cuda.memory_allocated(self.device) / 1024 ** 2:.2f} MB "
              f"({torch.cuda.max_memory_allocated(self.device) / 1024 ** 2:.2f} MB max, "
              f"{torch.cuda.max_memory_reserved(self.device) / 1024 ** 2:.2f} MB reserved)")
        if tracemalloc.is_tracing():
            tracemalloc.stop()
