"""gpugraphalgo_amd -- MI355X-native BFS/SSSP over CSR with node splitting.

The product is libgga.so: a C11 host layer plus hand-written gfx950 HIP
kernels behind the C-ABI of include/gga.h.  This package is its Python view
(ctypes); it holds no traversal logic of its own.
"""
from .serve import Server, latency_summary, poisson_trace
from .graph import (INF, REF_INF, CsrGraph, DeviceGraph, MultiSourceEngine, csr_from_arrays, device_count,
                    timeline_makespan, write_conv,
                    gen_grid, gen_rmat, read_gr, verify, write_degree_dist, write_dist, write_gr,
                    write_result_row)

__all__ = ["timeline_makespan", "write_conv", "Server", "latency_summary", "poisson_trace", "CsrGraph", "DeviceGraph", "MultiSourceEngine", "csr_from_arrays", "device_count", "gen_grid", "gen_rmat",
           "read_gr", "verify", "write_degree_dist", "write_dist", "write_gr", "write_result_row", "INF", "REF_INF"]
