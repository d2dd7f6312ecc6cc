"""ctypes mirror of include/peps_mi355x.h (the library's C ABI structs).

The structs mirror the reference's result types: ``peps_layer_metric`` is
``LayerMetric`` (src/lib.rs:388-401) and ``peps_contraction_result`` is
``ContractionResult`` (src/lib.rs:403-415), each with a few device fields
appended.
"""
import ctypes

PEPS_MAX_N = 42
PEPS_MAX_DEVICE_N = 21
PEPS_MAX_LANES = 4

PEPS_OK = 0
PEPS_E_INVALID = 1
PEPS_E_OVERFLOW = 2
PEPS_E_DEVICE = 3
PEPS_E_MEMORY = 4
PEPS_E_UNKNOWN = 5
PEPS_E_BUSY = 6
PEPS_E_PENDING = 7
PEPS_E_INTERNAL = 8

PEPS_POLICY_SHORTEST_FIRST = 0
PEPS_POLICY_FIFO = 1
PEPS_POLICY_RESPONSE_RATIO = 2

PEPS_ROW_OPERATOR_COMPILED = 0
PEPS_ROW_OPERATOR_SITEWISE = 1

PEPS_SYMMETRY_NONE = 0
PEPS_SYMMETRY_TOP_ROW_ORBITS = 1

PEPS_RECORDS_AUTO = 0
PEPS_RECORDS_LANES = 1
PEPS_RECORDS_PACKED = 2
PEPS_FORCE_EXACT_LEVELS = 1
PEPS_FORCE_UPPER_RUNS = 2
PEPS_FORCE_LEVEL_OVERFLOW = 4
PEPS_FORCE_CAP_OVERFLOW = 8


# peps_progress_fn: (user, done, total)
PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64)


class U128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]

    @property
    def value(self) -> int:
        return (self.hi << 64) | self.lo

    @staticmethod
    def of(v: int) -> "U128":
        return U128(v & ((1 << 64) - 1), v >> 64)


class LayerMetric(ctypes.Structure):
    _fields_ = [
        ("row", ctypes.c_uint32),
        ("input_states", ctypes.c_uint64),
        ("tensor_entries_examined", U128),
        ("tensor_entries_matched", U128),
        ("row_operator_candidates", U128),
        ("row_operator_matched", U128),
        ("completed_row_terms", U128),
        ("output_states", ctypes.c_uint64),
        ("output_weight", U128),
        ("elapsed_s", ctypes.c_double),
        ("peak_rss_bytes", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
        ("partition_levels", ctypes.c_uint32),
        ("final_buckets", ctypes.c_uint32),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, typ in self._fields_:
            v = getattr(self, name)
            out[name] = v.value if isinstance(v, U128) else v
        return out


class ContractionResult(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint32),
        ("count", U128),
        ("elapsed_s", ctypes.c_double),
        ("peak_states", ctypes.c_uint64),
        ("tensor_entries_examined", U128),
        ("tensor_entries_matched", U128),
        ("row_operator_candidates", U128),
        ("row_operator_matched", U128),
        ("peak_rss_bytes", ctypes.c_uint64),
        ("layer_count", ctypes.c_uint32),
        ("layers", LayerMetric * PEPS_MAX_N),
        ("crt_lanes", ctypes.c_uint32),
        ("record_bytes", ctypes.c_uint32),
        ("primes", ctypes.c_uint64 * PEPS_MAX_LANES),
        ("residues", ctypes.c_uint64 * PEPS_MAX_LANES),
        ("device_peak_bytes", ctypes.c_uint64),
        ("kernel_s", ctypes.c_double),
        ("alloc_s", ctypes.c_double),
    ]

    def layer_list(self) -> list:
        return [self.layers[i].as_dict() for i in range(self.layer_count)]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            if name == "layers":
                out[name] = self.layer_list()
                continue
            v = getattr(self, name)
            if isinstance(v, U128):
                v = v.value
            elif name in ("primes", "residues"):
                v = list(v)[: self.crt_lanes]
            out[name] = v
        return out


class Options(ctypes.Structure):
    _fields_ = [
        ("crt_lanes", ctypes.c_uint32),
        ("radix_bits", ctypes.c_uint32),
        ("tile_states", ctypes.c_uint32),
        ("device", ctypes.c_uint32),
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("collect_layers", ctypes.c_uint32),
        ("expand_bits", ctypes.c_uint32),
        ("graph_mode", ctypes.c_uint32),
        ("slice_group", ctypes.c_uint32),
        ("merge_passes", ctypes.c_uint32),
        ("merge_waves", ctypes.c_uint32),
        ("site_counters", ctypes.c_uint32),
        ("record_format", ctypes.c_uint32),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32),
        ("launches", ctypes.c_uint64),
        ("seconds", ctypes.c_double),
        ("algorithmic_bytes", ctypes.c_double),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("ticket", ctypes.c_uint64),
        ("n", ctypes.c_uint32),
        ("prefix_len", ctypes.c_uint32),
        ("status", ctypes.c_int32),
        ("tasks", ctypes.c_uint32),
        ("launches", ctypes.c_uint32),
        ("max_batch", ctypes.c_uint32),
        ("count", U128),
        ("submit_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("finish_s", ctypes.c_double),
        ("error", ctypes.c_char * 160),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("failed", ctypes.c_uint64),
        ("pending", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("tasks", ctypes.c_uint64),
        ("retries", ctypes.c_uint64),
        ("busy_s", ctypes.c_double),
        ("uptime_s", ctypes.c_double),
    ]


LEG_NAMES = ("column_in", "column_out", "row_in", "row_out", "diag_dr_in", "diag_dr_out", "diag_dl_in",
             "diag_dl_out")


class TensorEntry(ctypes.Structure):
    """peps_tensor_entry: one entry of the explicit C tensor (CEntry, src/lib.rs:88-92)."""
    _fields_ = [(name, ctypes.c_uint8) for name in LEG_NAMES] + [("value", ctypes.c_uint64)]

    def legs(self):
        return tuple(getattr(self, k) for k in LEG_NAMES)

    def __repr__(self):
        return f"TensorEntry(legs={self.legs()}, value={self.value})"


class LayerDigest(ctypes.Structure):
    _fields_ = [
        ("row", ctypes.c_uint32),
        ("support", ctypes.c_uint64),
        ("weight_sum", U128),
        ("transitions", U128),
        ("sha256", ctypes.c_uint8 * 32),
    ]
