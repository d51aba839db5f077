"""Host code under sanitizers (CPU only; built by gst_amd/build.py):

* the server's scheduler, admission, release and failure paths driven from several host
  threads in dry-run mode, with ThreadSanitizer and with AddressSanitizer;
* the oracle decoding the reference-encoded fixtures to DXT1 and RGB8 on one and on
  several threads, with AddressSanitizer + UndefinedBehaviorSanitizer and with
  ThreadSanitizer, including a garbled and a truncated file;
* the encoder's C++ API test with AddressSanitizer + UndefinedBehaviorSanitizer, its
  output byte-identical to the reference encoder's fixture.
"""
import os
import subprocess

import pytest

from conftest import ROOT, load_golden

CPP = os.path.join(ROOT, "tests", "cpp")


def _exe(name):
    path = os.path.join(CPP, name)
    if not os.path.exists(path):
        pytest.skip("%s is not built (gst_amd/build.py build_sanitizer_tests)" % name)
    return path


def _run(args, timeout=300):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               TSAN_OPTIONS="halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    proc = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=env)
    assert proc.returncode == 0, proc.stdout[-2000:] + proc.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in proc.stderr
    assert "ERROR: AddressSanitizer" not in proc.stderr
    assert "runtime error" not in proc.stderr
    return proc


@pytest.mark.parametrize("name", ["serve_dryrun_tsan", "serve_dryrun_asan"])
def test_server_dry_run_under_sanitizer(name):
    assert _run([_exe(name)]).stdout.strip().endswith("ok")


@pytest.mark.parametrize("name", ["oracle_sanitize_asan", "oracle_sanitize_tsan"])
@pytest.mark.parametrize("fixture", ["ref_small", "ref_c0"])
def test_oracle_under_sanitizer(name, fixture, tmp_path):
    g = load_golden(fixture)
    (tmp_path / "f.gtc").write_bytes(g["gtc"].tobytes())
    (tmp_path / "f.dxt").write_bytes(g["dxt"].tobytes())
    rgb = "-"
    if fixture == "ref_small":
        (tmp_path / "f.rgb").write_bytes(load_golden("ref_rgb")["small_rgb"].tobytes())
        rgb = str(tmp_path / "f.rgb")
    _run([_exe(name), str(tmp_path / "f.gtc"), str(tmp_path / "f.dxt"), rgb, "4"])


def test_encoder_under_address_sanitizer(tmp_path):
    outs = [str(tmp_path / n) for n in ("a.gtc", "a.dxt", "b.gtc", "b.dxt")]
    _run([_exe("encoder_asan")] + outs)
    g = load_golden("ref_c0")
    assert open(outs[0], "rb").read() == g["gtc"].tobytes()
    assert open(outs[1], "rb").read() == g["dxt"].tobytes()
