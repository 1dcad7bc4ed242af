"""Readers and writers of the reference's data formats (host side).

- Raw little-endian FP32 / FP64 ``.bin`` series, as the drivers read them
  (``GPU-dtw.cu:59-64``: ``fread`` of M and N floats; ``GPU-ed.cu:51-56``:
  doubles) and the extraction scripts write them (``array('f')`` /
  ``array('d')``, ``data/metal/extract.py:45-55``).
- The metal data set: 1000 text files per series (``parent_0000.txt`` ..
  ``parent_0999.txt`` for the subject, ``child_*`` for the query source), one
  value per line, NaN lines dropped, concatenated in file order
  (``data/metal/extract.py:5-29``).
- The ECG data set: ``ECG_query.mat`` holds ``ecg_query``; ``ECG.mat`` (not
  shipped with the reference) holds ``ECG``, split into the first 100000
  samples (query source) and the rest (subject) (``data/ecg/extract_ecg.py:
  6-9``).
- Query extraction: windows of length L at random offsets in [0, n - L)
  (``extract.py:35-37``), here from a seeded generator so runs repeat.

``python -m cudadtw_amd.data metal <root> <outdir> [--length L --count K
--seed S]`` writes the files ``extract.py`` writes (``single_/double_subject
.bin``, ``single_/double_query<k>.bin``), without its plots.
"""
from __future__ import annotations

import argparse
import os

import numpy as np

__all__ = ["read_bin", "write_bin", "load_metal", "load_ecg_query", "split_ecg",
           "extract_queries", "write_extracted"]


def read_bin(path, dtype=np.float32, count: int = -1) -> np.ndarray:
    """The first ``count`` values (all with -1) of a raw little-endian file."""
    dt = np.dtype(dtype).newbyteorder("<")
    return np.fromfile(path, dtype=dt, count=count).astype(np.dtype(dtype), copy=False)


def write_bin(path, values, dtype=np.float32) -> None:
    """Write values as raw little-endian ``dtype`` (no header)."""
    np.ascontiguousarray(values, dtype=np.dtype(dtype).newbyteorder("<")).tofile(path)


def _read_text_values(path) -> np.ndarray:
    # One value per line; blank lines carry no value.
    with open(path, "r") as f:
        vals = [float(line) for line in f if line.strip()]
    return np.asarray(vals, dtype=np.float64)


def load_metal(root, part: str = "parent", files: int = 1000) -> np.ndarray:
    """Concatenate ``<root>/<part>/<part>_NNNN.txt`` for NNNN < files, in
    order, dropping NaN values (extract.py:5-13 subject, :20-28 query)."""
    if part not in ("parent", "child"):
        raise ValueError("part must be 'parent' or 'child'")
    chunks = []
    for i in range(files):
        x = _read_text_values(os.path.join(root, part, f"{part}_{i:04d}.txt"))
        chunks.append(x[~np.isnan(x)])
    return np.concatenate(chunks) if chunks else np.empty(0, np.float64)


def load_ecg_query(path) -> np.ndarray:
    """``ecg_query`` of ECG_query.mat as a flat FP64 array (a MATLAB data
    file; scipy's reader parses it, nothing in it is executed)."""
    import scipy.io
    return np.asarray(scipy.io.loadmat(path)["ecg_query"], dtype=np.float64).ravel()


def split_ecg(ecg) -> tuple[np.ndarray, np.ndarray]:
    """(query source, subject) = (first 100000 samples, the rest)
    (extract_ecg.py:9)."""
    x = np.asarray(ecg, dtype=np.float64).ravel()
    return x[:100000], x[100000:]


def extract_queries(series, length: int, count: int, seed: int = 0) -> list[np.ndarray]:
    """``count`` windows of ``length`` samples at offsets drawn uniformly from
    [0, n - length) (extract.py:35-37, seeded)."""
    x = np.asarray(series)
    if x.size <= length:
        raise ValueError("series shorter than the query length")
    rng = np.random.default_rng(seed)
    starts = rng.integers(0, x.size - length, size=count)
    return [x[s:s + length].copy() for s in starts]


def write_extracted(outdir, subject, queries) -> list[str]:
    """The ``.bin`` files extract.py writes: single/double subject and
    single/double query<k>.  Returns the paths written."""
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for prefix, dt in (("single", np.float32), ("double", np.float64)):
        p = os.path.join(outdir, f"{prefix}_subject.bin")
        write_bin(p, subject, dt)
        paths.append(p)
        for k, q in enumerate(queries):
            p = os.path.join(outdir, f"{prefix}_query{k}.bin")
            write_bin(p, q, dt)
            paths.append(p)
    return paths


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cudadtw_amd.data")
    ap.add_argument("dataset", choices=["metal"])
    ap.add_argument("root", help="the data set folder (holding parent/ and child/)")
    ap.add_argument("outdir")
    ap.add_argument("--length", type=int, default=8192)
    ap.add_argument("--count", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    subject = load_metal(a.root, "parent")
    source = load_metal(a.root, "child")
    queries = extract_queries(source, a.length, a.count, a.seed)
    write_extracted(a.outdir, subject, queries)
    print(len(source), len(subject))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
