"""The snapshot Parquet writer (legendre_amd/csrc/lgd_parquet.cpp) read back by an
independent implementation, pyarrow: values bit for bit, the flat schema of required
DOUBLE / INT64 columns, row-group sizes, the snappy codec, dictionary pages where a
chunk has few distinct values (RLE and bit-packed runs at every index width up to
the 2^16-entry dictionary limit), min/max statistics under Parquet's float conventions, and
the error contract. Host code only: runs without a GPU."""
from __future__ import annotations

import numpy as np
import pytest

pq = pytest.importorskip("pyarrow.parquet")
lg = pytest.importorskip("legendre_amd")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def read(path):
    f = pq.ParquetFile(path)
    return f, f.read()


def test_special_values_and_row_groups_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n = 70_001
    x = rng.standard_normal(n)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 5e-324, -5e-324, np.finfo(float).max,
                        np.finfo(float).tiny])
    x[: special.size] = special
    x[100] = np.frombuffer(np.uint64(0x7FF8000000000123).tobytes(), dtype=np.float64)[0]  # NaN payload
    i = rng.integers(-(2**63), 2**63 - 1, n, dtype=np.int64, endpoint=True)
    i[:2] = [np.iinfo(np.int64).min, np.iinfo(np.int64).max]
    groups = [1, 30_000, 0, 40_000]
    lg.write_parquet(tmp_path / "a.parquet", {"x": x, "i": i, "step": 12, "t": 0.375}, groups)
    f, t = read(tmp_path / "a.parquet")
    assert f.schema_arrow.names == ["x", "i", "step", "t"]
    schema = f.metadata.schema
    assert [schema.column(c).physical_type for c in range(4)] == ["DOUBLE", "INT64", "INT64", "DOUBLE"]
    assert [schema.column(c).max_definition_level for c in range(4)] == [0, 0, 0, 0]  # required
    assert [f.metadata.row_group(g).num_rows for g in range(f.metadata.num_row_groups)] == [1, 30_000, 40_000]
    assert np.array_equal(bits(t["x"].to_numpy()), bits(x))
    assert np.array_equal(t["i"].to_numpy(), i)
    assert np.all(t["step"].to_numpy() == 12) and np.all(t["t"].to_numpy() == 0.375)
    for g in range(f.metadata.num_row_groups):
        for c in range(4):
            assert f.metadata.row_group(g).column(c).compression == "SNAPPY"


@pytest.mark.parametrize("ndistinct", [1, 2, 3, 255, 256, 257, 4097, 1 << 16])
def test_dictionary_chunks_at_every_index_width(tmp_path, ndistinct):
    rng = np.random.default_rng(ndistinct)
    n = max(300_000, 2 * ndistinct)
    values = rng.standard_normal(ndistinct)
    idx = rng.integers(0, ndistinct, n)
    idx[:ndistinct] = np.arange(ndistinct)
    # Runs long enough for RLE between bit-packed stretches, at odd offsets.
    for s in range(7, n - 100, 9973):
        idx[s:s + 37 + s % 11] = idx[s]
    x = values[idx]
    lg.write_parquet(tmp_path / "d.parquet", {"x": x})
    f, t = read(tmp_path / "d.parquet")
    assert np.array_equal(bits(t["x"].to_numpy()), bits(x))
    col = f.metadata.row_group(0).column(0)
    assert col.has_dictionary_page and "RLE_DICTIONARY" in col.encodings


def test_more_distinct_values_than_the_dictionary_holds_fall_back_to_plain(tmp_path):
    x = np.arange((1 << 16) + 1, dtype=np.float64)
    lg.write_parquet(tmp_path / "p.parquet", {"x": x})
    f, t = read(tmp_path / "p.parquet")
    col = f.metadata.row_group(0).column(0)
    assert not col.has_dictionary_page and col.encodings == ("PLAIN",)
    assert np.array_equal(t["x"].to_numpy(), x)


def test_statistics_follow_the_parquet_float_conventions(tmp_path):
    a = np.array([0.0, 3.0, np.nan, 2.0])          # min +0 is written as -0
    b = np.array([-5.0, -0.0, np.nan])            # max -0 is written as +0
    c = np.array([np.nan, np.nan])                # no min/max at all
    d = np.array([7, -3, 11], dtype=np.int64)
    for name, v in (("a", a), ("b", b), ("c", c), ("d", d)):
        lg.write_parquet(tmp_path / f"{name}.parquet", {name: v})
        st = pq.ParquetFile(tmp_path / f"{name}.parquet").metadata.row_group(0).column(0).statistics
        if name == "c":
            assert st is None or not st.has_min_max
            continue
        assert st.has_min_max and st.null_count == 0
        if name == "a":
            assert st.min == 0.0 and np.signbit(st.min) and st.max == 3.0
        elif name == "b":
            assert st.min == -5.0 and st.max == 0.0 and not np.signbit(st.max)
        else:
            assert (st.min, st.max) == (-3, 11)


def test_large_chunks_split_into_pages_and_compress(tmp_path):
    n = 3 * (1 << 20) + 5
    rng = np.random.default_rng(5)
    smooth = np.repeat(rng.standard_normal(n // 64 + 1), 64)[:n]  # repeated 8-byte runs
    lg.write_parquet(tmp_path / "l.parquet", {"s": smooth, "r": rng.standard_normal(n)}, [1 << 21, n - (1 << 21)])
    f, t = read(tmp_path / "l.parquet")
    assert np.array_equal(t["s"].to_numpy(), smooth)
    meta = f.metadata
    s_bytes = sum(meta.row_group(g).column(0).total_compressed_size for g in range(meta.num_row_groups))
    r_bytes = sum(meta.row_group(g).column(1).total_compressed_size for g in range(meta.num_row_groups))
    assert s_bytes < 0.2 * 8 * n           # long repeats compress
    assert r_bytes < 1.01 * 8 * n + 4096   # random doubles barely grow


def test_zero_rows(tmp_path):
    lg.write_parquet(tmp_path / "z.parquet", {"x": np.zeros(0)}, [])
    f, t = read(tmp_path / "z.parquet")
    assert t.num_rows == 0 and f.metadata.num_row_groups == 0


def test_errors(tmp_path):
    with pytest.raises(lg.LegendreError) as e:
        lg.write_parquet(tmp_path / "e.parquet", {"x": np.zeros(10)}, [4, 5])
    assert e.value.status == 1
    with pytest.raises(lg.LegendreError) as e:
        lg.write_parquet(tmp_path / "missing" / "dir" / "e.parquet", {"x": np.zeros(10)})
    assert e.value.status == 8
    assert not (tmp_path / "e.parquet").exists()


def test_the_reference_snapshot_layout_reads_back(tmp_path):
    """tests/parquet.rs:88-140 as host columns: a 4×4 grid of 2×2 blocks, row order
    block-major with dimension 0 fastest, phi = fma(100, x, y), theta0 = −fma(1000, y, x);
    the static file joins the snapshot by row order."""
    xs, ys = [], []
    for b in range(4):
        bx, by = b % 2, b // 2
        for j in range(2):
            for i in range(2):
                xs.append(((bx * 2 + i) + 0.5) * 0.5)
                ys.append(((by * 2 + j) + 0.5) * 0.5)
    xs, ys = np.array(xs), np.array(ys)
    phi = np.array([np.float64(100.0) * x + y for x, y in zip(xs, ys)])
    theta = -np.array([np.float64(1000.0) * y + x for x, y in zip(xs, ys)])
    lg.write_parquet(tmp_path / "static_0000000.parquet",
                     {"x": xs, "y": ys, "level": np.zeros(16), "patch": np.repeat(np.arange(4.0), 4),
                      "theta0": theta})
    lg.write_parquet(tmp_path / "snap_0000004.parquet", {"step": 4, "t": 1.0, "epoch": 0, "phi": phi})
    _, st = read(tmp_path / "static_0000000.parquet")
    _, sn = read(tmp_path / "snap_0000004.parquet")
    assert st.column_names == ["x", "y", "level", "patch", "theta0"]
    assert sn.column_names == ["step", "t", "epoch", "phi"]
    assert np.array_equal(st["x"].to_numpy(), xs) and np.array_equal(sn["phi"].to_numpy(), phi)
    assert list(sn["step"].to_numpy()) == [4] * 16 and list(sn["epoch"].to_numpy()) == [0] * 16
