"""cudadtw_amd -- MI355X (gfx950) z-normalised subsequence similarity search.

A drop-in for the gravitino/cudadtw search path: window statistics, LB_Kim /
LB_Keogh pruning, Sakoe-Chiba-banded cDTW, ED and greedy LDTW, best and
top-k matches.  Compute runs in hand-written HIP kernels behind the C ABI of
include/mdtw.h (libmdtw.so); this package is its Python face.
"""
from .api import (ALGO_CDTW, ALGO_ED, ALGO_ED_FFT, ALGO_LDTW, METRIC_L1, METRIC_L2, STATS_FP64,
                  STATS_NONE, STATS_REFCOMPAT, Batch, BatchResult, Engine, SearchResult, Stream, band_width,
                  envelope, znormalize)

__all__ = ["Engine", "SearchResult", "Batch", "BatchResult", "Stream", "znormalize", "envelope", "band_width", "ALGO_CDTW",
           "ALGO_ED", "ALGO_ED_FFT", "ALGO_LDTW", "METRIC_L1", "METRIC_L2", "STATS_FP64", "STATS_NONE",
           "STATS_REFCOMPAT"]
