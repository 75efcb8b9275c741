"""HDF5 output without h5py: the writer (ipie_amd.h5write) read back by the
reader (ipie_amd.h5read), and the EstimatorHandler drop-in's file in ipie's
layout (ipie/estimators/handler.py:148-242), read the way ipie's analysis
reads it (ipie/analysis/extraction.py:27-68).  CPU only: the estimates are
set on the host.  Compatibility with the HDF5 library itself is not checked
here (it is not installed); the writer uses the specification's oldest
layout (superblock 0, version-1 object headers, symbol-table groups)."""

import json

import numpy
import pytest


def test_writer_round_trip(tmp_path):
    from ipie_amd.h5read import H5File
    from ipie_amd.h5write import write_file

    rs = numpy.random.RandomState(0)
    many = {f"{i:09d}": rs.normal(size=(2, 3)) + 1j * rs.normal(size=(2, 3)) for i in range(150)}
    tree = {
        "metadata": json.dumps({"system": {"nup": 5}}),
        "g": {"f64": rs.normal(size=(4, 5)), "f32": rs.normal(size=7).astype(numpy.float32),
              "i64": numpy.arange(6, dtype=numpy.int64).reshape(2, 3), "i32": numpy.int32(-7),
              "scalar": 2.5, "flag": True, "names": "a b c", "hdr": numpy.array([b"x", b"yy"]),
              "empty": {}, "many": many},
    }
    path = str(tmp_path / "t.h5")
    write_file(path, tree)
    with H5File(path) as f:
        assert f.keys() == ["g", "metadata"]
        assert json.loads(f["metadata"].tobytes().rstrip(b"\0")) == {"system": {"nup": 5}}
        assert numpy.array_equal(f["g/f64"], tree["g"]["f64"])
        assert f["g/f32"].dtype == numpy.float32 and numpy.array_equal(f["g/f32"], tree["g"]["f32"])
        assert numpy.array_equal(f["g/i64"], tree["g"]["i64"]) and int(f["g/i32"]) == -7
        assert float(f["g/scalar"]) == 2.5 and int(f["g/flag"]) == 1
        assert f["g/names"].tobytes().rstrip(b"\0") == b"a b c"
        assert list(f["g/hdr"]) == [b"x", b"yy"]
        assert f.keys("g/empty") == []
        assert f.keys("g/many") == sorted(many)
        for k in ("000000000", "000000077", "000000149"):
            assert numpy.array_equal(f["g/many/" + k], many[k])


def _extract(path, block_idx=1):
    """ipie/analysis/extraction.py:27-47 over ipie_amd.h5read."""
    from ipie_amd.h5read import H5File

    with H5File(path) as f:
        base = f"block_size_{block_idx}"
        data = numpy.concatenate([f[f"{base}/data/{d}"].real for d in f.keys(f"{base}/data")])
        shapes = {}
        for k in f.keys(f"{base}/shape"):
            shapes[k] = {
                "names": f[f"{base}/names/{k}"][()],
                "shape": f[f"{base}/shape/{k}"][:],
                "offset": f[f"{base}/offset/{k}"][()],
                "size": f[f"{base}/size/{k}"][()],
                "scalar": bool(f[f"{base}/scalar/{k}"][()]),
                "num_walker_props": f[f"{base}/num_walker_props"][()],
                "walker_header": f[f"{base}/walker_prop_header"][()],
            }
        max_block = sum(f[f"{base}/max_block/{d}"][()] for d in f.keys(f"{base}/max_block"))
    return data[: max_block + 1], shapes


def test_estimator_handler_file_layout(tmp_path, capsys):
    from ipie_amd.estimators import EstimatorHandler
    from ipie_amd.walkers import WalkerAccumulator

    acc = WalkerAccumulator(["Weight", "WeightFactor", "HybridEnergy"], 10)
    h = EstimatorHandler(None, None, None, None, walker_state=acc, filename=str(tmp_path / "estimates.0.h5"))
    h.json_string = json.dumps({"qmc": {"dt": 0.005}})
    header = h.initialize(None)
    assert header.split() == ["Block", "Weight", "WeightFactor", "HybridEnergy", "ENumer", "EDenom", "ETotal",
                              "E1Body", "E2Body"]
    rows = []
    for block in range(3):
        acc._host[:] = [100.0 * (block + 1), 90.0, -550.0 * (block + 1)]
        h.local_estimates[3:] = [-530.0 - block, 100.0, 0.0, -1680.0, 1150.0 - block]
        rows.append(h.print_block(None, block, acc))
        assert acc.eshift == pytest.approx(-5.5)
    data, shapes = _extract(str(tmp_path / "estimates.0.h5"))
    assert data.shape == (3, 8)
    info = shapes["energy"]
    assert info["names"].tobytes().rstrip(b"\0").split() == [b"ENumer", b"EDenom", b"ETotal", b"E1Body", b"E2Body"]
    assert int(info["offset"]) == 3 and int(info["size"]) == 5 and info["scalar"]
    assert list(info["shape"]) == [5] and int(info["num_walker_props"]) == 3
    assert [bytes(x) for x in info["walker_header"]] == [b"Weight", b"WeightFactor", b"HybridEnergy"]
    # post-reduce hooks applied: ETotal = ENumer / EDenom, E1, E2 per unit weight; walker sums per step
    for block in range(3):
        factor = 1 if block == 0 else 10
        assert data[block, 0] == pytest.approx(100.0 * (block + 1) / factor)
        assert data[block, 2] == pytest.approx(-5.5)
        assert data[block, 3 + 2] == pytest.approx((-530.0 - block) / 100.0)
        assert data[block, 3 + 4] == pytest.approx((1150.0 - block) / 100.0)
        assert rows[block].split()[0] == str(block)
