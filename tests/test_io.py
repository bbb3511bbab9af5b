"""Tensor I/O in TAMM's HDF5 format (tmx_tensor_write_hdf5 / tmx_tensor_read_hdf5).

CPU: the ctypes HDF5 helper round-trips a dataset.  GPU: the engine reads files written by the
helper from the oracle's data (bit for bit), writes files the helper reads back as the dense
extent with the packed blocks at offset 0, and round-trips a spin + C2v tensor.
"""
import numpy as np
import pytest

from tamm_amd import configs
from tamm_amd import engine as E
from oracle import oracle as orc
from tests import h5util

pytestmark = pytest.mark.skipif(not h5util.available(), reason="libhdf5 not available")


def test_h5util_roundtrip(tmp_path):
    path = str(tmp_path / "x.h5")
    data = np.arange(10, dtype=np.float64)
    h5util.write_tensor_file(path, 16, data)
    n, vals = h5util.read_tensor_file(path)
    assert n == 16 and np.array_equal(vals[:10], data) and not vals[10:].any()


@pytest.mark.gpu
@pytest.mark.parametrize("irrep", [False, True])
def test_engine_hdf5_matches_tamm_layout(eng, tmp_path, irrep):
    wl = configs.tiny_grad(irrep)
    ow = orc.OracleWorkload(wl, configs.DEFAULT_SEED)
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    for name in ("V", "T", "L"):
        t = ew.tensors[name]
        dense_extent = int(np.prod([wl.spaces[m].n for m in wl.tensors[name].modes]))
        # engine -> file -> helper
        p1 = str(tmp_path / f"{name}_engine.h5")
        t.write_hdf5(p1)
        n, vals = h5util.read_tensor_file(p1)
        assert n == dense_extent
        assert np.array_equal(vals[:t.size], ow.data[name]), name
        # helper (oracle data) -> file -> engine
        p2 = str(tmp_path / f"{name}_oracle.h5")
        h5util.write_tensor_file(p2, dense_extent, ow.data[name])
        back = E.Tensor(t.spaces, wl.tensors[name].spin, wl.tensors[name].irrep_rule).allocate()
        back.read_hdf5(p2)
        assert np.array_equal(back.download(), ow.data[name]), name


@pytest.mark.gpu
def test_engine_hdf5_rejects_wrong_extent(eng, tmp_path):
    wl = configs.tiny_grad(False)
    ew = E.EngineWorkload(wl, configs.DEFAULT_SEED)
    p = str(tmp_path / "bad.h5")
    h5util.write_tensor_file(p, 7, np.zeros(7))
    with pytest.raises(E.EngineError):
        ew.tensors["V"].read_hdf5(p)
