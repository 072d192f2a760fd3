"""CPU tests of the kernel selector's fitting step: the committed table is
exactly what `selector.py fit` produces from the committed corpus, the
leave-one-matrix-out report agrees with it, and the nearest-neighbour vote
prefers a configuration that is fast on the neighbours."""
import json
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "da-spmmul-kernel_amd")
sys.path.insert(0, PKG)
import selector as S  # noqa: E402

CORPUS = os.path.join(PKG, "selector_data", "corpus_mi355x.csv")
REPORT = os.path.join(PKG, "selector_data", "report_mi355x.json")
TABLE = os.path.join(PKG, "csrc", "selector_table.inc")


def test_committed_table_is_generated_from_committed_corpus(tmp_path):
    table = tmp_path / "t.inc"
    report = tmp_path / "r.json"
    subprocess.run([sys.executable, os.path.join(PKG, "selector.py"), "fit", "--csv", CORPUS,
                    "--table", str(table), "--report", str(report)], check=True,
                   capture_output=True)
    assert table.read_text() == open(TABLE).read()
    assert json.load(open(report)) == json.load(open(REPORT))


def test_report_shows_selector_beats_static_default():
    r = json.load(open(REPORT))["summary"]
    assert r["selector_covered"] == r["entries"] == r["heuristic_covered"]
    assert r["selector_geomean_slowdown_vs_best"] < r["heuristic_geomean_slowdown_vs_best"]
    assert r["selector_geomean_slowdown_vs_best"] < 1.2


def test_table_compiles():
    src = f'#include <cstdint>\n#include "{TABLE}"\nint main() {{ return kSelEntries > 0 ? 0 : 1; }}\n'
    out = subprocess.run(["g++", "-std=c++17", "-x", "c++", "-", "-o", "/dev/null"], input=src,
                         text=True, capture_output=True)
    assert out.returncode == 0, out.stderr


def test_vote_prefers_fast_configuration_of_neighbours():
    f = {k: 0.0 for k in S.FEATURES}
    a, b = (1, 4, 8, 32, 8, 0), (2, 0, 0, 0, 0, 0)
    near = [("m1", 32, dict(f), [(a, 0.0), (b, 0.5)]),
            ("m2", 32, dict(f), [(a, 0.1), (b, 0.0)]),
            ("m3", 32, dict(f), [(a, 0.0), (b, 0.9)])]
    far = dict(f, log_rows=50.0)
    entries = near + [("far", 32, far, [(b, 0.0)])]
    sd = {k: 1.0 for k in S.FEATURES}
    assert S.predict(dict(f), 32, entries, sd) == a
    # the nearest measured width is used when the query width was not measured
    assert S.predict(dict(f), 24, entries, sd) == a


def test_heuristic_config_matches_library_defaults():
    table = {(1, 4, 8, 32, 8, 0): 1.0, (1, 2, 16, 256, 8, 0): 2.0, (1, 4, 8, 32, 4, 0): 1.0}
    assert S.heuristic_cfg(1, table) == (6, 0, 0, 1024, 0, 0)
    # nv = 32 with vec 4 -> 8 lanes -> tile 16 * 8 = 128, unroll 8
    assert S.heuristic_cfg(32, table) == (1, 4, 8, 128, 8, 0)
