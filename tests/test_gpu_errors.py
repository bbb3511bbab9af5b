"""Error paths of the C ABI on the GPU: pool exhaustion, the TAMM_GPU_POOL setting, bad
handles and calls out of order.

TAMM's pool is sized once from TAMM_GPU_POOL percent of free device memory and never grows
(src/tamm/rmm_memory_pool.hpp:230-237); a request it cannot serve terminates the run
(mr/stream_ordered_memory_resource.hpp:141-149).  The engine returns TMX_ERR_OOM instead and
stays usable: memory freed afterwards serves the same request.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

ERR_INVALID, ERR_OOM, ERR_STATE = 1, 3, 5
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_malloc_beyond_pool_is_oom(eng):
    cap, used, _ = eng.pool_info()
    with pytest.raises(eng.EngineError) as ei:
        eng.DeviceBuffer(cap + 1)
    assert ei.value.code == ERR_OOM
    assert "exhausted" in str(ei.value)
    assert eng.pool_info()[1] == used  # nothing leaked


def test_plan_creation_oom_then_recovers(eng):
    wl = configs.config0()
    ew = eng.EngineWorkload(wl, configs.DEFAULT_SEED)
    term = wl.terms[0]
    ref = ew.contraction(term, eng.EXEC_SECTOR_F64)
    ws = ref.workspace
    assert ws > (1 << 20)
    R = ew.tensors["R"]
    R.fill(0.0)
    ref.execute()
    expect = R.download()
    del ref
    hog = _fill_pool(eng, leave=1 << 20)  # leaves ~1 MiB, less than the plan needs
    full = eng.pool_info()[1]
    for mode in (eng.EXEC_SECTOR_F64, eng.EXEC_SPLIT_I8):
        with pytest.raises(eng.EngineError) as ei:
            ew.contraction(term, mode)
        assert ei.value.code == ERR_OOM, eng.EXEC_NAMES[mode]
        assert eng.pool_info()[1] == full  # the failed plan released what it had taken
    del hog
    c = ew.contraction(term, eng.EXEC_SECTOR_F64)
    R.fill(0.0)
    c.execute()
    assert np.array_equal(R.download(), expect)


def _fill_pool(eng, leave):
    """Takes pool memory until at most `leave` bytes (possibly in fragments) stay free."""
    bufs = []
    while True:
        cap, used, _ = eng.pool_info()
        want = cap - used - leave
        if want <= 0:
            return bufs
        while want >= 4096:
            try:
                bufs.append(eng.DeviceBuffer(want))
                break
            except eng.EngineError as e:
                assert e.code == ERR_OOM
                want //= 2
        else:
            return bufs


def test_bad_handles_and_order(eng):
    lib = eng.lib()
    assert lib.tmx_free(12345) == ERR_INVALID  # not a pool pointer
    assert lib.tmx_init(0, 0.5) == ERR_STATE  # already initialised
    assert lib.tmx_contraction_execute(None, None, None, None, None) == ERR_INVALID
    assert lib.tmx_set_split_bits(60) == ERR_INVALID
    assert lib.tmx_graph_launch(None, None) == ERR_INVALID


_POOL_PROBE = (
    "import sys; sys.path.insert(0, %r)\n"
    "from tamm_amd import engine as E\n"
    "E.init(0, 0.0)\n"
    "print(E.pool_info()[0])\n"
)


def _pool_capacity(percent):
    env = dict(os.environ, TAMM_GPU_POOL=str(percent))
    r = subprocess.run([sys.executable, "-c", _POOL_PROBE % ROOT], env=env, capture_output=True,
                       text=True, timeout=120)
    return r


def test_tamm_gpu_pool_is_honoured():
    """tmx_init(dev, 0) sizes the pool from TAMM_GPU_POOL percent of free memory (fresh
    processes, one at a time)."""
    small, big = _pool_capacity(10), _pool_capacity(30)
    assert small.returncode == 0 and big.returncode == 0, small.stderr + big.stderr
    c10, c30 = int(small.stdout.split()[-1]), int(big.stdout.split()[-1])
    assert abs(c30 / c10 - 3.0) < 0.05, (c10, c30)
    bad = _pool_capacity(99)  # more than the 95% the engine allows
    assert bad.returncode != 0 and "pool fraction" in bad.stderr
