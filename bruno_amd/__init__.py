"""bruno_amd: MI355X-native de Bruijn graph construction and compaction.

The hot path (parse -> count -> filter -> chain extraction -> list ranking ->
unitigs) runs as HIP kernels in libbruno_hip.so behind the C ABI declared in
include/bruno_hip.h.  This package is a thin host binding: GraphJob mirrors
the reference driver's call sequence (compact_dbg_freq_construct_clean.cpp
:431-840) for tests, the benchmark and scripting.
"""
from ._lib import (BrunoError, lib, NODE_DTYPE, CHAIN_DTYPE, EXPORTS, LIB_PATH)  # noqa: F401
from .pipeline import (GraphBatch, GraphJob, SynthConfig, synth_reads_host, thresholds_L, DEFAULT_THRESH)  # noqa: F401
