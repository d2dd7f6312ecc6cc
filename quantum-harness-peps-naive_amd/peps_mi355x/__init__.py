"""peps_mi355x -- MI355X (gfx950) explicit-C PEPS row-boundary contraction.

Python mirror of the reference crate's library API (src/lib.rs), backed by
the C ABI of libpeps_mi355x.so (include/peps_mi355x.h).  Every contraction
runs on the GPU; without the built library or a gfx950 device the calls
raise ``PepsError`` -- there is no CPU fallback.
"""
from .api import (  # noqa: F401
    PepsError,
    PepsOptions,
    RequestOutcome,
    Server,
    contract_rows,
    contract_rows_d4_orbit_sort_reduce,
    contract_rows_sort_reduce,
    count_batched,
    count_prefix,
    crt_plan,
    crt_reconstruct,
    device_info,
    kernel_stats,
    known_count,
    lib_path,
    load,
    peak_rss_bytes,
    prefix_tasks,
    profile_kernels,
    version,
)
