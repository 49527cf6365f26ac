"""GPU checks of the autotuner: every candidate launch shape gives identical
results (the tuner raises otherwise), the best one is applied and cached."""
import numpy as np
import pytest

from gf2_amd import _lib, autotune, codes
from gf2_amd.device import DeviceBuffer
from gf2_amd.ldpc import Context, LdpcCode
from gf2_amd.mixed import CodeSet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    autotune.apply(c, {"path": _lib.PATH_AUTO, "wg_per_cu": 3, "xcd_map": 1, "graph_iters": 4, "chunk": -1,
                       "lds_groups": 0})


def test_tune_single_code(ctx, tmp_path):
    code = LdpcCode(codes.n2304_r12().qc(), ctx)
    cache = autotune.TuningCache(str(tmp_path / "t.json"))
    r = autotune.tune_code(code, 256, 2.0, max_it=20, repeats=1,
                           knobs=("path", "wg_per_cu", "xcd_map"), cache=cache)
    C = autotune.CANDIDATES  # every path can run this code
    assert len(r["results"]) == len(C["path"]) * len(C["wg_per_cu"]) * len(C["xcd_map"])
    assert r["config"] in [cfg for cfg, _ in r["results"]]
    again = autotune.TuningCache(str(tmp_path / "t.json"))
    assert autotune.apply_cached(ctx, [code], 200, cache=again) == r["config"]  # same power-of-two bucket


def test_tune_code_set(ctx, tmp_path):
    specs = codes.mixed_twelve()[:3]
    gcodes = [LdpcCode(s.qc(), ctx) for s in specs]
    cs = CodeSet(gcodes)
    ids = np.repeat(np.arange(3, dtype=np.uint16), 40)
    n = np.array([s.n for s in specs])
    off = np.concatenate([[0], np.cumsum(n[ids])])
    d_llr = DeviceBuffer(ctx, int(off[-1]) * 4)
    lib = _lib.load()
    for c in range(3):
        f0 = int(np.nonzero(ids == c)[0][0])
        _lib.check(ctx.handle, lib.gf2l_frames_generate(ctx.handle, gcodes[c].handle, 9 + c, 0, 40, 2.5, 2.5,
                                                        d_llr.ptr + int(off[f0]) * 4, None))
    ctx.synchronize()
    cache = autotune.TuningCache(str(tmp_path / "m.json"))
    r = autotune.tune_codeset(cs, ids, d_llr, max_it=20, repeats=1,
                              knobs=("wg_per_cu", "graph_iters"), cache=cache)
    assert len(r["results"]) == len(autotune.CANDIDATES["wg_per_cu"]) * len(autotune.CANDIDATES["graph_iters"])
    assert r["key"].startswith("mixed:") and cache.get(r["key"]) == r["config"]


def test_tune_dvb_code(ctx, tmp_path):
    """A DVB-T2 code has no device encoder (all-zero frames) and runs on the
    state kernels only: the forced STREAM/LDS shapes are skipped, the rest
    agree bit for bit."""
    from gf2_amd.ldpc import CodeRate
    code = LdpcCode.dvb_t2_short(CodeRate.RATE_2_3, ctx)
    cache = autotune.TuningCache(str(tmp_path / "d.json"))
    r = autotune.tune_code(code, 256, 2.3, max_it=20, repeats=1, knobs=("path", "chunk"), cache=cache)
    paths = {cfg["path"] for cfg, _ in r["results"]}
    assert paths == {_lib.PATH_AUTO, _lib.PATH_STATE}
    assert len(r["results"]) == 2 * len(autotune.CANDIDATES["chunk"])
    assert r["key"].startswith("single:sparse-n16200")


def test_tune_lds_groups(ctx, tmp_path):
    """The N=2304 code runs on the LDS path, so the default knobs are the
    LDS path's thread groups (its block-row tiling); every group count gives
    the same results and the best one is applied."""
    code = LdpcCode(codes.n2304_r12().qc(), ctx)
    assert autotune.default_knobs(code) == ("lds_groups",)
    cache = autotune.TuningCache(str(tmp_path / "g.json"))
    r = autotune.tune_code(code, 512, 2.0, max_it=20, repeats=1, cache=cache)
    assert [cfg["lds_groups"] for cfg, _ in r["results"]] == autotune.CANDIDATES["lds_groups"]
    assert cache.get(r["key"]) == r["config"]
    ctx.set_tuning(_lib.TUNE_LDS_GROUPS, 0)
    big = LdpcCode(codes.n64800_r12().qc(), ctx)
    assert autotune.default_knobs(big) == ("path", "wg_per_cu", "chunk")
    assert ctx.code_path(big) == _lib.PATH_STATE
