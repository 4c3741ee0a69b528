"""Host layer (reference operator interface + runExperiment driver).

CPU: the C++ host_tests binary, YAML->CSV runs with CPU executors checked
against the oracle.  GPU: the same driver with GPU executors must report the
same matches/checksums as the CPU executors on the same generated streams.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG

BIN = os.path.join(PKG, "host", "bin")


def run_driver(tmp_path, yaml_text, name="out.csv", env=None):
    cfg = tmp_path / "cfg.yaml"
    cfg.write_text(yaml_text)
    out = tmp_path / name
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([os.path.join(BIN, "runExperiment"), str(cfg), str(out)], capture_output=True,
                       text=True, timeout=600, env=e, cwd=str(tmp_path))
    assert p.returncode == 0, p.stderr[-2000:]
    with open(out) as f:
        return list(csv.DictReader(f))


def yaml_cfg(alg, dev, prog, method, window_mb=1, batch_kb=128, keys=1000, theta=0.0, data=1 << 17,
             par="INTRA", pk="No", source_threads=1):
    return f"""algTypes: [ "{alg}" ]
deviceTypes: [ "{dev}" ]
progressivenessModes: [ "{prog}" ]
distinctKeys: [ {keys} ]
zipfThetas: [ {theta} ]
resultMaterializationMethod: [ "{method}" ]
windowSizesMB: [ {window_mb} ]
batchSizeKB: [ {batch_kb} ]
measuredMetric: "Throughput"
measurePowerUsage: "No"
repeat: 1
parStrats: [ "{par}" ]
dataSizeTuples: {data}
rateSearch: "No"
extendedCsv: "Yes"
sourceThreads: {source_threads}
usePK: "{pk}"
"""


def test_host_unit_binary():
    from conftest import _gpu_available
    env = dict(os.environ)
    if _gpu_available():
        env["SJB_HAVE_GPU"] = "1"  # the no-device error path only exists without a GPU
    p = subprocess.run([os.path.join(BIN, "host_tests")], capture_output=True, text=True, timeout=300,
                       env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "0 failures" in p.stdout


@pytest.mark.parametrize("san", ["tsan", "asan"])
def test_host_sanitizers(san):
    """host_tests built with ThreadSanitizer (32 source threads into the
    windowing, pooled executors) and with AddressSanitizer + UBSan."""
    p = subprocess.run(["make", "-s", "-C", os.path.join(PKG, "host"), f"{BIN}/host_tests_{san}"],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66",
               ASAN_OPTIONS="detect_leaks=1 exitcode=66", UBSAN_OPTIONS="halt_on_error=1")
    p = subprocess.run([os.path.join(BIN, f"host_tests_{san}")], capture_output=True, text=True, timeout=600,
                       env=env)
    assert p.returncode == 0, (p.stdout + p.stderr)[-4000:]
    assert "0 failures" in p.stdout and "Sanitizer" not in p.stderr


def test_driver_rejects_bad_yaml(tmp_path):
    cfg = tmp_path / "bad.yaml"
    cfg.write_text(yaml_cfg("HJ", "CPU", "LAZY", "NoOutput").replace('"HJ"', '"XJ"'))
    p = subprocess.run([os.path.join(BIN, "runExperiment"), str(cfg), str(tmp_path / "o.csv")],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "Unknown algorithm type: XJ" in p.stderr
    p = subprocess.run([os.path.join(BIN, "runExperiment")], capture_output=True, text=True, timeout=60)
    assert p.returncode == 1


def expected_lazy(sjbmod, orc, data, window, keys, theta=0.0):
    dist = sjbmod.DIST_ZIPF if theta > 0 else sjbmod.DIST_UNIFORM
    L = sjbmod.generate(data, dist, keys, theta, seed=123)
    R = sjbmod.generate(data, dist, keys, theta, seed=456)
    cnt = c1 = 0
    for o in range(0, data - window + 1, window):
        r, _ = orc.lazy_smj(L[o:o + window], R[o:o + window])
        cnt += r.count
        c1 = (c1 + r.c1) % 2**64
    return cnt, c1


@pytest.mark.parametrize("alg", ["HJ", "SMJ", "NLJ"])
def test_driver_cpu_lazy(tmp_path, sjbmod, orc, alg):
    data, window_mb = 1 << 17, 1
    rows = run_driver(tmp_path, yaml_cfg(alg, "CPU", "LAZY", "NoOutput", window_mb=window_mb, data=data,
                                         keys=1000 if alg != "NLJ" else 5000))
    assert len(rows) == 1
    r = rows[0]
    assert list(r.keys())[:11] == ["algorithm_name", "write_result", "ingestion_rate", "batch_size",
                                   "zipf_theta", "distinct_key", "window_size", "probe_threads",
                                   "joiner_threads", "selectivity", "time"]
    assert r["algorithm_name"] == {"HJ": "lazy_cpu_hj_nm", "SMJ": "lazy_cpu_smj", "NLJ": "lazy_cpu_nlj_nm"}[alg]
    cnt, c1 = expected_lazy(sjbmod, orc, data, window_mb * 65536, 1000 if alg != "NLJ" else 5000)
    assert int(r["match_count"]) == cnt
    assert int(r["checksum1"]) == c1
    sel = 100.0 * cnt / ((data // (window_mb * 65536)) * (window_mb * 65536) ** 2)
    assert abs(float(r["selectivity"]) - sel) < 1e-6 * max(sel, 1)


@pytest.mark.parametrize("prog", ["LAZY", "EAGER"])
def test_driver_cpu_concurrent_sources(tmp_path, sjbmod, orc, prog):
    # 32 source threads (the reference driver's default): windows are
    # placed by stream position, so the result is the canonical one
    data = 1 << 18
    for par in ("INTRA", "INTER"):
        rows = run_driver(tmp_path, yaml_cfg("HJ", "CPU", prog, "NoOutput", data=data, par=par,
                                             source_threads=32), name=f"{par}.csv")
        cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 1000)
        assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1, par


def test_driver_power_without_gpu(tmp_path, sjbmod, orc):
    # measurePowerUsage: "Yes" with no GPU / no AMD SMI reading: the run
    # goes on and reports the same results, without power samples
    from conftest import _gpu_available
    if _gpu_available():
        pytest.skip("a GPU is present: test_driver_gpu_power covers the polling")
    y = yaml_cfg("HJ", "CPU", "LAZY", "NoOutput").replace('measurePowerUsage: "No"', 'measurePowerUsage: "Yes"')
    rows = run_driver(tmp_path, y)
    cnt, c1 = expected_lazy(sjbmod, orc, 1 << 17, 65536, 1000)
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1


def test_driver_cpu_eager(tmp_path, sjbmod, orc):
    data = 1 << 17
    rows = run_driver(tmp_path, yaml_cfg("HJ", "CPU", "EAGER", "NoOutput", data=data))
    # canonical eager windows equal the lazy windows
    cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 1000)
    assert rows[0]["algorithm_name"] == "eager_cpu_hj_nm"
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1


def test_driver_rate_search(tmp_path):
    # the reference rate search doubles dataSize while the median stays < 1 s;
    # a tiny precision bound keeps it short on the CPU
    y = yaml_cfg("HJ", "CPU", "LAZY", "NoOutput", data=1 << 17).replace('rateSearch: "No"', 'rateSearch: "Yes"')
    y += "precisionTuples: 2000000000\n"
    rows = run_driver(tmp_path, y)
    assert len(rows) == 1 and int(rows[0]["ingestion_rate"]) >= 1 << 17


GPU_CASES = [("HJ", "LAZY", "NoOutput", "lazy_gpu_hj_nm"), ("HJ", "LAZY", "CountKernel", "lazy_gpu_hj_ck"),
             ("HJ", "LAZY", "EstimatedSelectivity", "lazy_gpu_hj_es"), ("HJ", "LAZY", "Atomic", "lazy_gpu_hj_at"),
             ("SMJ", "LAZY", "NoOutput", "lazy_gpu_smj_nm"), ("SMJ", "LAZY", "CountKernel", "lazy_gpu_smj_ck"),
             ("SMJ", "LAZY", "EstimatedSelectivity", "lazy_gpu_smj_es"), ("SMJ", "LAZY", "Atomic", "lazy_gpu_smj_at"),
             ("NLJ", "LAZY", "NoOutput", "lazy_gpu_nlj_nm"), ("NLJ", "LAZY", "CountKernel", "lazy_gpu_nlj_ck"),
             ("NLJ", "LAZY", "EstimatedSelectivity", "lazy_gpu_nlj_es"), ("NLJ", "LAZY", "Bitmap", "lazy_gpu_nlj_bm"),
             ("HJ", "EAGER", "NoOutput", "eager_gpu_hj_nm"), ("HJ", "EAGER", "Atomic", "eager_gpu_hj_at"),
             ("NLJ", "EAGER", "NoOutput", "eager_gpu_nlj_nm"), ("NLJ", "EAGER", "Bitmap", "eager_gpu_nlj_bm")]


@pytest.mark.gpu
@pytest.mark.parametrize("alg,prog,method,name", GPU_CASES)
def test_driver_gpu(tmp_path, sjbmod, orc, alg, prog, method, name):
    print(f"case driver {name}", flush=True)
    data = 1 << 17
    rows = run_driver(tmp_path, yaml_cfg(alg, "GPU", prog, method, data=data, keys=5000),
                      env={"SJB_HAVE_GPU": "1"})
    assert rows[0]["algorithm_name"] == name
    cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 5000)
    assert int(rows[0]["match_count"]) == cnt
    assert int(rows[0]["checksum1"]) == c1


@pytest.mark.gpu
def test_driver_gpu_pk(tmp_path, sjbmod, orc):
    data = 1 << 17
    rows = run_driver(tmp_path, yaml_cfg("HJ", "GPU", "EAGER", "NoOutput", data=data, keys=5000, pk="Yes"),
                      env={"SJB_HAVE_GPU": "1"})
    assert rows[0]["algorithm_name"] == "eager_gpu_hj_nm_pk"
    cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 5000)
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1


@pytest.mark.gpu
@pytest.mark.parametrize("prog,method", [("LAZY", "NoOutput"), ("LAZY", "Atomic"), ("EAGER", "NoOutput")])
def test_driver_gpu_concurrent_sources(tmp_path, sjbmod, orc, prog, method):
    print(f"case driver sources=32 {prog} {method}", flush=True)
    data = 1 << 19
    rows = run_driver(tmp_path, yaml_cfg("HJ", "GPU", prog, method, data=data, keys=5000, par="INTER",
                                         source_threads=32), env={"SJB_HAVE_GPU": "1"})
    cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 5000)
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1


@pytest.mark.gpu
@pytest.mark.parametrize("prog,window_mb,data", [("LAZY", 1, 1 << 17), ("EAGER", 5, 2 * 5 * 65536)])
def test_driver_gpu_bitmap_dense(tmp_path, sjbmod, orc, prog, window_mb, data):
    # 1000 distinct keys (the reference configs' value): a 1 MB lazy window
    # has ~4.2M (left tuple, 64-key word) masks with a match, a 5 MB eager
    # window ~1.6M per batch pass; the mask buffer grows to what is needed.
    # Streams hold whole windows: eager executors also join a partial last
    # window, which expected_lazy does not count.
    print(f"case driver bitmap D=1000 {prog} {window_mb}MB", flush=True)
    rows = run_driver(tmp_path, yaml_cfg("NLJ", "GPU", prog, "Bitmap", window_mb=window_mb, data=data, keys=1000),
                      env={"SJB_HAVE_GPU": "1"})
    cnt, c1 = expected_lazy(sjbmod, orc, data, window_mb * 65536, 1000)
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1


@pytest.mark.gpu
def test_driver_gpu_power(tmp_path, sjbmod, orc):
    # AMD SMI polling while the source runs: power_data.txt in the working
    # directory, "<ms since epoch>,<watts>" per line (the reference format)
    print("case driver power", flush=True)
    data = 1 << 21
    y = yaml_cfg("HJ", "GPU", "LAZY", "NoOutput", data=data, keys=5000).replace(
        'measurePowerUsage: "No"', 'measurePowerUsage: "Yes"')
    rows = run_driver(tmp_path, y, env={"SJB_HAVE_GPU": "1"})
    cnt, c1 = expected_lazy(sjbmod, orc, data, 65536, 5000)
    assert int(rows[0]["match_count"]) == cnt and int(rows[0]["checksum1"]) == c1
    lines = (tmp_path / "power_data.txt").read_text().split()
    assert len(lines) >= 1
    ts, w = zip(*[(int(a), float(b)) for a, b in (ln.split(",") for ln in lines)])
    assert all(10.0 <= x <= 3000.0 for x in w), w[:10]
    assert list(ts) == sorted(ts)


def test_driver_rate_search_capped(tmp_path):
    # the rate search doubles the stream size while runs stay under 1 s; at
    # maxDataSizeTuples it stops and reports that size (the CPU hash join
    # runs 2^16-tuple streams far below 1 s)
    cap = 1 << 16
    y = yaml_cfg("HJ", "CPU", "LAZY", "NoOutput", data=1 << 14).replace('rateSearch: "No"', 'rateSearch: "Yes"')
    y += f"precisionTuples: 1024\nmaxDataSizeTuples: {cap}\n"
    rows = run_driver(tmp_path, y)
    assert len(rows) == 1 and int(rows[0]["ingestion_rate"]) == cap


def test_driver_refuses_stream_above_cap(tmp_path):
    y = yaml_cfg("HJ", "CPU", "LAZY", "NoOutput", data=1 << 17) + "maxDataSizeTuples: 65536\n"
    cfg = tmp_path / "cfg.yaml"
    cfg.write_text(y)
    p = subprocess.run([os.path.join(BIN, "runExperiment"), str(cfg), str(tmp_path / "o.csv")],
                       capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert p.returncode != 0 and "exceeds maxDataSize" in p.stderr
