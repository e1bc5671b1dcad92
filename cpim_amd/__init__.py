"""cpim_amd: MI355X (gfx950) GAC/SAC propagation engine for cpim's GPU path.

The engine is libcpim_gpu.so (HIP, C-ABI in include/cpim_gpu.h); this package
is its Python mirror of the reference's GModel / GModelAdapter / GModelSolver /
CModel entry points, plus the seeded workload generators used by the
benchmarks and tests.
"""
from .model import ModelDesc, dense_from_tuples, domains_to_sets, from_tuples, words_for
from .gmodel import (Batch, CModel, GacStats, GModel, GModelAdapter, GModelOptions, GModelSolver,
                     GpuSearchStatistics, SacStats, SearchStatistics, Server)

__all__ = [
    "ModelDesc", "dense_from_tuples", "domains_to_sets", "from_tuples", "words_for",
    "Batch", "CModel", "GacStats", "GModel", "GModelAdapter", "GModelOptions", "GModelSolver",
    "GpuSearchStatistics", "SacStats", "SearchStatistics", "Server",
]
