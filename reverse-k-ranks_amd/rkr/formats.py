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
* query distribution of QS training: <index_dir>/query_distribution/
  <dataset>-n_sample_item_<Q>-sample_topk_<K>/ holding sample_itemID_l.txt
  (one item ID per line, samples in k-th-rank order), sort_kth_rank_l.index
  (int32[Q]) and accu_n_user_rank_l.index (int32[Q][n_item + 1], same order)
  (WriteDistributionBelowTopk, src/app/build-index/FileIO.hpp:56-137); the
  trained ranks go to <index_dir>/qrs_to_sample_index/
  QS-<dataset>-n_sample_<S>-n_sample_query_<Q>-sample_topk_<K>.index in the
  table-file header layout without scores (SampleSearch.hpp:158-194).
* full score table: float[n_user][n_item], each user's scores against every
  item sorted in descending order, row-major with no header
  (ReadAll::BuildIndexLoop / ReadDisk, src/include/alg/DiskIndex/ReadAll.hpp:80-87, 100-106),
  stored as <index_dir>/<dataset>.index (src/app/BuildScoreTableBatch.cpp:168-169).
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


def query_distribution_dir(index_dir: str, dataset: str, n_sample_item: int, sample_topk: int) -> str:
    return os.path.join(index_dir, "query_distribution",
                        f"{dataset}-n_sample_item_{n_sample_item}-sample_topk_{sample_topk}")


def write_query_distribution(path: str, sample_item_ids, sort_kth_rank, accu_sorted) -> None:
    os.makedirs(path, exist_ok=True)
    with open(os.path.join(path, "sample_itemID_l.txt"), "w") as f:
        f.write("".join(f"{int(i)}\n" for i in sample_item_ids))
    np.ascontiguousarray(sort_kth_rank, np.int32).tofile(os.path.join(path, "sort_kth_rank_l.index"))
    np.ascontiguousarray(accu_sorted, np.int32).tofile(os.path.join(path, "accu_n_user_rank_l.index"))


def read_query_distribution(path: str, n_item: int):
    """Returns (sample_item_ids, sort_kth_rank, accu_sorted)."""
    ids = np.loadtxt(os.path.join(path, "sample_itemID_l.txt"), dtype=np.int64, ndmin=1).astype(np.int32)
    kth = np.fromfile(os.path.join(path, "sort_kth_rank_l.index"), dtype=np.int32)
    accu = np.fromfile(os.path.join(path, "accu_n_user_rank_l.index"), dtype=np.int32)
    if kth.size != ids.size or accu.size != kth.size * (n_item + 1):
        raise ValueError(f"{path}: query distribution files disagree in size")
    return ids, kth, accu.reshape(kth.size, n_item + 1)


def qs_ranks_path(index_dir: str, dataset: str, n_sample: int, n_sample_query: int, sample_topk: int) -> str:
    return os.path.join(index_dir, "qrs_to_sample_index",
                        f"QS-{dataset}-n_sample_{n_sample}-n_sample_query_{n_sample_query}-sample_topk_{sample_topk}.index")


def write_ranks(path: str, ranks, n_item: int, n_user: int) -> None:
    """The rank file fsr writes: size_t n_sample, n_item, n_user, then int32 ranks."""
    ranks = np.ascontiguousarray(ranks, np.int32)
    with open(path, "wb") as f:
        np.array([ranks.size, n_item, n_user], dtype=np.uint64).tofile(f)
        ranks.tofile(f)


def read_ranks(path: str):
    """Returns (ranks, n_item, n_user) of a rank file."""
    with open(path, "rb") as f:
        S, n_item, n_user = (int(v) for v in np.fromfile(f, dtype=np.uint64, count=3))
        ranks = np.fromfile(f, dtype=np.int32, count=S)
    if ranks.size != S:
        raise ValueError(f"{path}: truncated rank file")
    return ranks, n_item, n_user


def score_table_path(index_dir: str, dataset: str) -> str:
    return os.path.join(index_dir, f"{dataset}.index")


def write_score_table(path: str, rows: np.ndarray) -> None:
    np.ascontiguousarray(rows, np.float32).tofile(path)


def read_score_table(path: str, n_item: int, mmap: bool = True) -> np.ndarray:
    """The (n_user, n_item) score table; memory-mapped read-only unless mmap is False.

    The file has no header, so n_item comes from the dataset; a file whose size
    is not a whole number of rows is rejected.
    """
    size = os.path.getsize(path)
    row = 4 * int(n_item)
    if n_item <= 0 or size % row != 0:
        raise ValueError(f"{path}: {size} bytes is not a whole number of {n_item}-item rows")
    n_user = size // row
    if mmap:
        return np.memmap(path, dtype=np.float32, mode="r", shape=(n_user, n_item))
    return np.fromfile(path, dtype=np.float32).reshape(n_user, n_item)


def read_score_rows(path: str, n_item: int, user: int, start: int, count: int) -> np.ndarray:
    """Items [start, start + count) of one user's sorted row (ReadAll::ReadDisk, ReadAll.hpp:100-106)."""
    with open(path, "rb") as f:
        f.seek((int(user) * int(n_item) + int(start)) * 4)
        out = np.fromfile(f, dtype=np.float32, count=count)
    if out.size != count:
        raise ValueError(f"{path}: truncated score table")
    return out


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
