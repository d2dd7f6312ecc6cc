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
    canonical_layer_bytes,
    compile_row_operator,
    contract_rows,
    contract_rows_d4_orbit_sort_reduce,
    contract_rows_hash_materialization,
    contract_rows_sitewise,
    contract_rows_sort_reduce,
    contract_rows_with_tensor,
    count_batched,
    count_prefix,
    count_prefix_no_alloc,
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
    set_progress,
    sha256,
    site_tensor_c,
    trace_layers,
    version,
)
