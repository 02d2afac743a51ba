"""The benchmark's bookkeeping, checked on the CPU: its table of register-kernel
schedules matches the library's, its kernel signatures name real kernels, and
the committed result line carries the fields of the benchmark contract."""
import json
import os
import re

from conftest import ROOT

import bench


def library_schedules():
    text = open(os.path.join(ROOT, "se2dr_fork_amd", "csrc", "hip", "se2w_kernels_reg.hpp")).read()
    body = re.search(r"kRegSchedules\[\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    return [tuple(int(x) for x in m.split(",")) for m in re.findall(r"\{([^{}]*)\}", body)]


def test_schedule_table_matches_the_library():
    assert [tuple(s) for s in bench.REG_SCHEDULES] == library_schedules()
    text = open(os.path.join(ROOT, "se2dr_fork_amd", "csrc", "hip", "se2w_kernels_reg64.hpp")).read()
    body = re.search(r"kReg64Schedules\[\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    assert [tuple(s) for s in bench.REG64_SCHEDULES] == [tuple(int(x) for x in m.split(","))
                                                          for m in re.findall(r"\{([^{}]*)\}", body)]
    header = open(os.path.join(ROOT, "include", "se2w.h")).read()
    assert f"#define SE2W_KERNEL_REG64 {bench.KERNEL_REG64} " in header


def test_kernel_signatures():
    # as rocprofv3 names the kernels: the LDS tile kernel with its thread count
    assert bench.kernel_signature(0, 4, 2) == "tile_kernel<4, 2, true, 256>"
    assert bench.kernel_signature(11, 8, 1) == "tile_kernel<8, 1, true, 512>"
    ex, ey, w, pf, lpe = bench.REG64_SCHEDULES[2]
    assert bench.kernel_signature(bench.KERNEL_REG64 + 2, 4, 0) == f"reg64_kernel<4, {ex}, {ey}, {w}, {pf}, {lpe}>"
    ex, ey, w, pf, lpe = bench.REG_SCHEDULES[6]
    assert bench.kernel_signature(7, 4, 2) == f"reg_kernel<4, 2, true, {ex}, {ey}, {w}, {pf}, {lpe}>"
    assert bench.ens_kernel_signature(4, 6, 1).startswith("ens_reg_kernel<6, 1, ")
    ex, ey, w, pf, lpe = bench.REG64_SCHEDULES[1]
    assert bench.ens_kernel_signature(bench.KERNEL_REG64 + 1, 6, 0) == f"ens_reg64_kernel<6, {ex}, {ey}, {w}, {pf}, {lpe}>"


def test_committed_result_line():
    d = json.load(open(os.path.join(ROOT, "profiles", "bench_512x512_p4.json")))
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    r = d["roofline"]
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9
    assert r["traffic"] is None or r["traffic"] > 0
    c = d["cpu_baseline"]
    assert c["kind"] in ("port", "reference") and c["cores"] >= 1 and c["value"] > 0
    assert d["n_gpus"] == 1 and d["higher_is_better"] is True


def test_split_schedule_bytes():
    # The two kernels of the split schedule together move the two-pass
    # scheme's bytes per node less the stored acceleration (its write and
    # its read: 2 x 2 components), i.e. 60 B in FP32.
    for mode, word in ((0, 8), (1, 4), (2, 4)):
        split = bench.SPLIT_ELEMENT_BYTES[mode] + bench.SPLIT_UPDATE_BYTES[mode]
        assert split == bench.CANONICAL_B_NODE[mode] - 2 * 2 * word
    assert bench.SPLIT_ELEMENT_BYTES[1] + bench.SPLIT_UPDATE_BYTES[1] == 60
    # the fused kernel drops the f round trip and the second read of u
    for mode, word in ((0, 8), (1, 4), (2, 4)):
        split = bench.SPLIT_ELEMENT_BYTES[mode] + bench.SPLIT_UPDATE_BYTES[mode]
        assert split - bench.BYTES_PER_NODE[mode] == 3 * 2 * word
