"""GPU: the H*V launch-shape autotuner (ftlm_hip_autotune / _set_shape / _get_shape).

A launch shape changes how the work is cut into launches and workgroups, never what is summed
per output element, so y = H x must not change beyond rounding (checked within 1e-13 relative,
max-norm, against the default shape on the same input), and an FTLM run with a tuned operator
must give the same lnZ within 1e-11 relative (the dot partials are summed in another order; with
complex64 vectors and M=60 the last-bit differences of alpha reach 2e-12 of lnZ at beta=8).
"""

import json
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

api = pytest.importorskip("ftlm_extended_hubbard_amd.api")
from ftlm_extended_hubbard_amd import _lib as L  # noqa: E402

P = dict(t=1.0, tp=-0.25, U=8.0, V=1.0, phi_x=0.05, phi_y=0.0)


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _device():
    if api.device_count() < 1:
        pytest.skip("no HIP device")


def _apply_block(op, xb, R, dtype, torch):
    y = torch.empty_like(xb)
    op.apply_device(xb, y, R=R, dtype=dtype)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _ops():
    p = api.HubbardParams(4, 3, **P)
    yield "sector 4x3 (3,3) two-pass", api.HubbardOperator.from_params(p, 3, 3)
    yield "sector 4x2 (2,2) flat", api.HubbardOperator.from_params(api.HubbardParams(4, 2, **P), 2, 2)
    p0 = api.HubbardParams(4, 3, 1.0, -0.25, 8.0, 1.0, 0.0, 0.0)
    yield "kblock 4x3 (3,3) K=(1,0)", api.KBlockOperator.from_params(p0, 3, 3, api.MomentumSector(1, 0, 4, 3))


@pytest.mark.parametrize("dtype", [L.FTLM_C128, L.FTLM_C64])
@pytest.mark.parametrize("R", [1, 12, 16])
def test_tuned_apply_equals_default(dtype, R):
    import torch

    for name, op in _ops():
        say("autotune", name, "R", R, "dtype", dtype)
        dim = op.dim()
        rng = np.random.default_rng(R + 7 * dtype)
        x = (rng.standard_normal((dim, R)) + 1j * rng.standard_normal((dim, R))).astype(
            np.complex64 if dtype == L.FTLM_C64 else np.complex128)
        xb = torch.from_numpy(x.reshape(-1)).to("cuda")
        assert op.get_shape(R, dtype) == {"rb": 0, "bpx": 0, "wlog": -1, "grid": 0}
        y0 = _apply_block(op, xb, R, dtype, torch)
        res = op.autotune(R, dtype, reps=2)
        say(name, res)
        assert res["n_tried"] >= 1 and res["ms"] > 0 and res["ms"] <= res["ms_default"]
        if "two-pass" in name:
            assert res["n_tried"] >= 4  # bpx and stripe-width candidates at least
        assert op.get_shape(R, dtype) == {k: res[k] for k in ("rb", "bpx", "wlog", "grid")}
        y1 = _apply_block(op, xb, R, dtype, torch)
        tol = 1e-5 if dtype == L.FTLM_C64 else 1e-13
        assert np.max(np.abs(y1 - y0)) <= tol * np.max(np.abs(y0)), name
        # every shape the tuner may pick: force a few non-default ones and check them too
        for sh in ({"rb": 1}, {"bpx": 64, "wlog": 0}, {"grid": 256}, {"rb": 2, "bpx": 512, "wlog": 3}):
            op.set_shape(R, dtype, **sh)
            y2 = _apply_block(op, xb, R, dtype, torch)
            assert np.max(np.abs(y2 - y0)) <= tol * np.max(np.abs(y0)), (name, sh)
        op.set_shape(R, dtype)  # back to the defaults


def test_tuned_ftlm_matches_default(tmp_path):
    say("autotune ftlm 4x3 (3,3) R=16 c64, cache")
    p = api.HubbardParams(4, 3, **P)
    op = api.HubbardOperator.from_params(p, 3, 3)
    par = api.FtlmParams(n_random=16, lanczos_steps=60, seed=4, start=L.FTLM_START_PHASE, dtype=L.FTLM_C64,
                         quad=L.FTLM_QUAD_CONVERGED, path=L.FTLM_PATH_PER_OP)
    betas = np.array([0.5, 2.0, 8.0])
    before = api.ftlm_log_partition_complex(op, betas, par)
    cache = str(tmp_path / "shapes.json")
    res = op.autotune(16, L.FTLM_C64, reps=2, cache=cache)
    assert not res["cached"]
    after = api.ftlm_log_partition_complex(op, betas, par)
    assert np.all(np.abs(after - before) <= 1e-11 * np.maximum(1.0, np.abs(before))), (before, after)
    # a fresh operator of the same sector restores the shape from the cache without timing
    op2 = api.HubbardOperator.from_params(p, 3, 3)
    res2 = op2.autotune(16, L.FTLM_C64, cache=cache)
    assert res2["cached"] and op2.get_shape(16, L.FTLM_C64) == op.get_shape(16, L.FTLM_C64)
    with open(cache) as f:
        assert len(json.load(f)) == 1


def test_bad_shapes_rejected():
    say("autotune refusals")
    op = api.HubbardOperator.from_params(api.HubbardParams(4, 2, **P), 2, 2)
    for bad in ({"rb": 3}, {"rb": 32}, {"bpx": 1024}, {"wlog": 9}, {"grid": 9000}):
        with pytest.raises(ValueError):
            op.set_shape(4, L.FTLM_C64, **bad)
    with pytest.raises(ValueError):
        op.set_shape(4, L.FTLM_C128, rb=16)  # complex128 launches carry at most 8 vectors
