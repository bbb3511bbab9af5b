"""Split-K parts of sector plans on the GPU.

Each sector GEMM can be cut along K into parts with their own partial results, reduced in the
scatter rounds.  The part count comes from the launch model unless fixed with
tmx_set_parallel_split; here it is fixed to 1, 2, 3 and 5 on cfg0 (dense O=20/t10, V=80/t20)
and every result is checked against the oracle -- FP64 to 1e-12 relative, split precision to
1e-10 -- with the plan's schedule reporting the expected number of rounds: parts are at least
1024 columns, and K blocks are padded to 32 columns (FP64) or 64 (split), so the ladder's K
range (16 blocks of 400) is 6656 / 7168 columns (up to 6 / 7 parts) and the ring's (8 blocks
of 200) 1792 / 2048 (1 / 2 parts).
"""
import numpy as np
import pytest

from tamm_amd import configs

pytestmark = pytest.mark.gpu

SEED = 20261202


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="module")
def cfg0_case(orc):
    wl = configs.config0()
    ow = orc.OracleWorkload(wl, SEED)
    refs = {}
    for term in wl.terms:
        base = np.random.default_rng(7).standard_normal(ow.tensors[term.c[0]].size)
        ref = base.copy()
        ow.run_term(term, out=ref)
        refs[term.name] = (base, ref)
    return wl, ow, refs


@pytest.mark.parametrize("parts", [1, 2, 3, 5])
@pytest.mark.parametrize("mode", ["sector_f64", "split_i8"])
def test_forced_parts_match_oracle(eng, cfg0_case, mode, parts):
    wl, ow, refs = cfg0_case
    m = {"sector_f64": eng.EXEC_SECTOR_F64, "split_i8": eng.EXEC_SPLIT_I8}[mode]
    tol = 1e-12 if mode == "sector_f64" else 1e-10
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    eng.check(eng.lib().tmx_set_parallel_split(parts))
    try:
        for term in wl.terms:
            base, ref = refs[term.name]
            cn = term.c[0]
            c = ew.contraction(term, m)
            rounds, launches, tasks = c.schedule_info()
            kmax = {"ladder": (6, 7), "ring": (1, 2)}[term.name][mode == "split_i8"]
            want = min(parts, kmax)
            assert rounds == want, (term.name, rounds, want)
            assert launches >= 1 and tasks >= rounds
            for rep in range(2):  # the partial results are rewritten on every execution
                ew.tensors[cn].upload(base)
                c.execute()
                got = ew.tensors[cn].download()
                err = _rel(got - base, ref - base)
                assert err <= tol, (term.name, mode, parts, rep, err)
            del c
    finally:
        eng.check(eng.lib().tmx_set_parallel_split(0))


def test_model_choice_matches_oracle(eng, cfg0_case):
    """The launch model's own part counts (whatever it picks) on both paths."""
    wl, ow, refs = cfg0_case
    ew = eng.EngineWorkload(wl, SEED, allocate=False)
    for name in wl.tensors:
        ew.tensors[name].upload(ow.data[name])
    for m, tol in ((eng.EXEC_SECTOR_F64, 1e-12), (eng.EXEC_SPLIT_I8, 1e-10)):
        for term in wl.terms:
            base, ref = refs[term.name]
            cn = term.c[0]
            ew.tensors[cn].upload(base)
            c = ew.contraction(term, m)
            rounds, _, _ = c.schedule_info()
            assert 1 <= rounds <= 16
            c.execute()
            err = _rel(ew.tensors[cn].download() - base, ref - base)
            assert err <= tol, (term.name, eng.EXEC_NAMES[m], rounds, err)
            del c


def test_parallel_split_rejects_out_of_range(eng):
    assert eng.lib().tmx_set_parallel_split(17) != 0
    assert eng.lib().tmx_set_parallel_split(-1) != 0
    eng.check(eng.lib().tmx_set_parallel_split(0))


@pytest.mark.parametrize("mode", ["sector_f64", "split_i8"])
def test_autotune_split_k_keeps_results(eng, mode):
    """tmx_contraction_autotune also times split-K counts 1..4 (for plans over 0.5 ms) and may
    replace the plan; the replacement reassociates FP64 sums, so results are compared to the
    plain FP64 sector result at the mode's tolerance, after tuning and on a later execution."""
    wl = configs.scaled_spin(30, 180, 16, 36)
    ew = eng.EngineWorkload(wl, SEED)
    R = ew.tensors["R"]
    term = wl.terms[0]  # ladder: ~300 GFLOP, several ms
    c0 = ew.contraction(term, eng.EXEC_SECTOR_F64)
    R.fill(0.0)
    c0.execute()
    ref = R.download()
    del c0
    m = {"sector_f64": eng.EXEC_SECTOR_F64, "split_i8": eng.EXEC_SPLIT_I8}[mode]
    tol = 1e-12 if mode == "sector_f64" else 1e-10
    c = ew.contraction(term, m)
    R.fill(0.0)
    c.autotune()
    print(f"{mode}: schedule after autotune {c.schedule_info()}")
    assert _rel(R.download(), ref) <= tol
    R.fill(0.0)
    c.execute()
    assert _rel(R.download(), ref) <= tol
