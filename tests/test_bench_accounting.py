"""bench.py's accounting on the CPU: the whole-step roofline is the sum of
its stages' rooflines, each the larger of its byte and FLOP times at the
published peaks, and the host description is self-consistent."""
import math

import bench


def test_step_roofline_is_the_sum_of_stage_rooflines():
    n, m, l, p = 1 << 20, 800_000, 10 ** 8, 9 * 10 ** 7
    r = bench.step_roofline(n, m, l, p, {"traverse": 1.0, "p2p": 0.25, "sort": 0.1}, 2.0)
    hbm, fp = bench.HBM_PEAK_GBS * 1e9, bench.FP32_PEAK_TFLOPS * 1e12
    want = {"prepare": 16 * n / hbm, "sort": 48 * n / hbm, "build": (4 * n + 32 * m) / hbm,
            "upward": (16 * n + 32 * m) / hbm, "m2l": 41 * l / fp, "l2l": 24 * m / hbm,
            "p2p_l2p": max(16 * n / hbm, 20 * p / fp), "frame": 68 * n / hbm}
    for k, t in want.items():
        assert math.isclose(r["stages"][k]["roof_ms"], 1e3 * t, rel_tol=1e-3, abs_tol=1e-5), k
    total = 1e3 * sum(want.values())
    assert math.isclose(r["t_roof_ms"], total, rel_tol=1e-3)
    assert math.isclose(r["frac"], total / 2.0, rel_tol=1e-3)
    # stage fractions against the measured stage times
    assert math.isclose(r["stages"]["m2l"]["frac"], 1e3 * want["m2l"] / 1.0, abs_tol=1e-4)
    assert r["stages"]["m2l"]["bound"] == "valu" and r["stages"]["sort"]["bound"] == "hbm"
    assert "frac" not in r["stages"]["prepare"]  # no stage time measured for it


def test_host_cpus_is_consistent():
    info = bench.host_cpus()
    assert 1 <= info["usable"] <= info["affinity"] <= (info["nproc"] or info["affinity"])
    counts = bench.cpu_thread_counts(info)
    assert counts[0] == 1 and counts[-1] == info["usable"] and counts == sorted(set(counts))
