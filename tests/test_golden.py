"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py and cross-checked there against numpy's FFT,
numpy filter steps, direct AGC window sums and direct lag sums) against the
oracle on the CPU, and against the GPU path (marked gpu)."""
import os

import numpy as np
import pytest

import processseg_amd as P
from processseg_amd import workloads as W
from tests.common import rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))  # plain arrays: allow_pickle stays False


FLOWS = {"plain": W.plain_flow, "panel": W.big_flow}


def test_oracle_fft_matches_golden(oracle):
    g = load("fft.npz")
    for n in g["sizes"]:
        re, im = oracle.rfft(g[f"x{n}"])
        want = g[f"re{n}"] + 1j * g[f"im{n}"]
        assert np.max(np.abs(re + 1j * im - want)) <= 1e-13 * np.max(np.abs(want)), n
        back = oracle.irfft(re, im, int(n))
        assert np.max(np.abs(back - g[f"x{n}"])) <= 1e-13 * np.max(np.abs(g[f"x{n}"]))


@pytest.mark.parametrize("name", ["plain", "panel"])
def test_oracle_flows_match_golden(oracle, name):
    g = load(f"{name}.npz")
    ops = FLOWS[name]().ops
    got = oracle.pipeline_apply(ops, g["x"].astype(np.float64), g["live"], float(g["dt"]))
    assert rel_l2(got, g["out"]) <= 1e-13
    mirror = oracle.pipeline_apply(ops, g["x"], g["live"], float(g["dt"]))
    assert rel_l2(mirror, g["out"]) <= 1e-5


def test_oracle_correlation_matches_golden(oracle):
    g = load("vibroseis.npz")
    x = g["x"]
    rl = np.full(x.shape[0], x.shape[1], np.int32)
    got = oracle.correlate(x.astype(np.float64), rl, g["pilot"], 2.0, x.shape[1])
    assert rel_l2(got, g["out"]) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,tol", [(np.float32, 1e-5), (np.float64, 1e-12)])
@pytest.mark.parametrize("name", ["plain", "panel"])
def test_gpu_flows_match_golden(gpu, name, dtype, tol):
    import torch
    g = load(f"{name}.npz")
    x = g["x"].astype(dtype)
    plan = FLOWS[name]().compile(dtype, float(g["dt"]), x.shape[1])
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    plan.apply(src, dst, row_len=torch.from_numpy(g["live"]).cuda())
    torch.cuda.synchronize()
    e = rel_l2(dst.cpu().numpy(), g["out"])
    print(name, dtype.__name__, "rel L2 vs golden", e)
    assert e <= tol


@pytest.mark.gpu
def test_gpu_correlation_matches_golden(gpu):
    import torch
    g = load("vibroseis.npz")
    x = g["x"]
    pilot = P.Pilot(g["pilot"], 2.0)
    plan = P.initCorrelationPlan(np.float32, pilot, 2.0, x.shape[1])
    src = torch.from_numpy(x).cuda()
    dst = torch.empty(x.shape[0], plan.outSamples, device="cuda")
    rl = torch.full((x.shape[0],), x.shape[1], dtype=torch.int32, device="cuda")
    plan.correlateInto(src, dst, row_len=rl)
    torch.cuda.synchronize()
    e = rel_l2(dst.cpu().numpy(), g["out"])
    print("vibroseis rel L2 vs golden", e)
    assert e <= 1e-5
