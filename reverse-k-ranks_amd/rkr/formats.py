"""On-disk formats of the reference's US path.

* fvecs: per row an int32 dimension followed by dim float32 values
  (loadVector, src/include/util/VectorIO.hpp:14-37; dataset files are
  <dir>/<name>/<name>_{data_item,user,query_item}.fvecs, VectorIO.hpp:47-55).
* sampled score table: size_t n_sample, size_t n_item, size_t n_user,
  int32 ranks[n_sample], float table[n_user][n_sample]
  (SampleSearch::SaveIndex / LoadIndex,
  src/include/alg/RankBoundRefinement/SampleSearch.hpp:186-194, 229-241),
  stored as <index_dir>/memory_index/US-<dataset>-n_sample_<S>.index
  (SampleSearch.hpp:158-179 with the name "US" from NameTranslation.hpp).
* result CSVs: one line per query, comma-separated userID / rank / queryIP
  (WriteRankResult, src/include/util/FileIO.hpp:84-148).
"""
from __future__ import annotations

import os

import numpy as np


def read_fvecs(path: str) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.int32)
    if raw.size == 0:
        return np.zeros((0, 0), np.float32)
    d = int(raw[0])
    if d <= 0 or raw.size % (d + 1) != 0:
        raise ValueError(f"{path}: not an fvecs file (dim {d}, {raw.size} words)")
    rows = raw.reshape(-1, d + 1)
    if not np.all(rows[:, 0] == d):
        raise ValueError(f"{path}: inconsistent row dimensions")
    return rows[:, 1:].view(np.float32).copy()


def write_fvecs(path: str, x: np.ndarray) -> None:
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    out = np.empty((n, d + 1), np.int32)
    out[:, 0] = d
    out[:, 1:] = x.view(np.int32)
    out.tofile(path)


def read_dataset(basic_dir: str, name: str):
    """(users, items, queries) as readData loads them (users not yet normalised)."""
    base = os.path.join(basic_dir, name, name)
    items = read_fvecs(base + "_data_item.fvecs")
    users = read_fvecs(base + "_user.fvecs")
    queries = read_fvecs(base + "_query_item.fvecs")
    return users, items, queries


def table_path(index_dir: str, dataset: str, n_sample: int) -> str:
    return os.path.join(index_dir, "memory_index", f"US-{dataset}-n_sample_{n_sample}.index")


def write_table(path: str, ranks: np.ndarray, table: np.ndarray, n_item: int) -> None:
    ranks = np.ascontiguousarray(ranks, np.int32)
    table = np.ascontiguousarray(table, np.float32)
    n_user, S = table.shape
    with open(path, "wb") as f:
        np.array([S, n_item, n_user], dtype=np.uint64).tofile(f)
        ranks.tofile(f)
        table.tofile(f)


def read_table(path: str):
    """Returns (ranks, table, n_item)."""
    with open(path, "rb") as f:
        S, n_item, n_user = (int(v) for v in np.fromfile(f, dtype=np.uint64, count=3))
        ranks = np.fromfile(f, dtype=np.int32, count=S)
        table = np.fromfile(f, dtype=np.float32, count=S * n_user)
    if ranks.size != S or table.size != S * n_user:
        raise ValueError(f"{path}: truncated table file")
    return ranks, table.reshape(n_user, S), n_item


def _csv_rows(a) -> str:
    return "".join(",".join(str(v) for v in row) + "\n" for row in a)


def write_rank_result(out_dir: str, dataset: str, method: str, params: str, users, ranks, qips) -> list:
    """Writes <dataset>-<method>-<params>-{userID,rank,IP}.csv like WriteRankResult."""
    os.makedirs(out_dir, exist_ok=True)
    paths = []
    for suffix, arr in (("userID", users), ("rank", ranks), ("IP", qips)):
        p = os.path.join(out_dir, f"{dataset}-{method}-{params}-{suffix}.csv")
        with open(p, "w") as f:
            if suffix == "IP":
                f.write("".join(",".join(f"{float(v):g}" for v in row) + "\n" for row in arr))
            else:
                f.write(_csv_rows(arr))
        paths.append(p)
    return paths
