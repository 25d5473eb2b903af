"""Result API statistics against known answers: VectorValue getSum /
getVariance / getSumSTD / getMean / getMeanSTD and CountValue
(quantity.cpp:27-116), in the C++ header and in the Python mirror, on fixed
event lists whose statistics numpy's two-pass formulas give; plus the
reference's edge cases (no value, one value) and a constant non-dyadic value,
where moments can round below zero."""
import math
import os
import subprocess

import numpy as np
import pytest

import midsx_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

LISTS = {
    "small": [1.0, 2.0, 3.0, 4.0],
    "deposits": list(np.random.default_rng(3).exponential(25e3, 1000)),
    "constant": [0.1] * 1000,            # non-dyadic: sum and sum of squares round
    "constant_big": [60000.3] * 4097,    # a monoenergetic beam's IncidentEnergy
    "one": [42.0],
    "none": [],
}

PROGRAM = r"""
#include <MIDSX/Core.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
int main(int argc, char** argv) {
  VectorValue v, a, b;
  std::vector<double> xs;
  for (int i = 1; i < argc; ++i) xs.push_back(std::strtod(argv[i], nullptr));
  for (size_t i = 0; i < xs.size(); ++i) {
    v.addValue(xs[i]);
    if (i % 2) a.addValue(xs[i]); else b.addValue(xs[i]);
  }
  VectorValue ab = a + b;
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", v.getSum(), v.getVariance(), v.getSumSTD(), v.getMean(),
              v.getMeanSTD(), v.getCount());
  std::printf("%.17g %.17g\n", ab.getSum(), ab.getVariance());
  VectorValue w;
  w.addValues(xs);
  std::printf("%.17g %.17g\n", w.getSum(), w.getVariance());
  CountValue c, d;
  for (size_t i = 0; i < xs.size(); ++i) c.addCount();
  d.addCounts(5);
  CountValue cd = c + d;
  std::printf("%d %.17g %d\n", c.getCount(), c.getCountSTD(), cd.getCount());
  return 0;
}
"""


@pytest.fixture(scope="module")
def stats_exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("stats")
    src = d / "stats.cpp"
    src.write_text(PROGRAM)
    exe = str(d / "stats")
    lib = os.path.join(ROOT, "midsx_amd", "_lib")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "midsx_amd", "cpp"),
                        str(src), "-o", exe, "-L", lib, "-lmidsx_amd", "-Wl,-rpath," + lib],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    return exe


def expected(xs):
    """The reference's two-pass statistics (quantity.cpp:27-94)."""
    x = np.asarray(xs, dtype=np.float64)
    n = len(x)
    s = float(x.sum()) if n else 0.0
    if n == 0:
        return dict(sum=0.0, var=0.0, sum_std=0.0, mean=math.nan, mean_std=math.nan)
    mean = s / n
    var = float(((x - mean) ** 2).sum()) / (n - 1) if n > 1 else math.nan
    return dict(sum=s, var=var, sum_std=math.sqrt(n * var) if n > 1 else math.nan, mean=mean,
                mean_std=math.sqrt(var / n) if n > 1 else math.nan)


def close(a, b, rel=1e-12, abs_=0.0):
    if math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= max(rel * abs(b), abs_)


@pytest.mark.parametrize("name", list(LISTS))
def test_cpp_vector_and_count_values(stats_exe, name):
    xs = LISTS[name]
    r = subprocess.run([stats_exe] + ["%.17g" % v for v in xs], stdout=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0
    l1, l2, l3, l4 = [list(map(float, ln.split())) for ln in r.stdout.strip().splitlines()]
    e = expected(xs)
    s, var, sstd, mean, mstd, cnt = l1
    assert cnt == len(xs)
    assert close(s, e["sum"])
    const = len(set(xs)) == 1 and len(xs) > 1
    if const:  # the two-pass variance of a constant is (almost) 0; from moments it must not go below
        scale = xs[0] * xs[0]
        assert 0.0 <= var <= 1e-12 * scale and math.isfinite(sstd) and sstd >= 0, (var, sstd)
    else:
        assert close(var, e["var"], 1e-9), (var, e["var"])
        assert close(sstd, e["sum_std"], 1e-9) and close(mstd, e["mean_std"], 1e-9)
    assert close(mean, e["mean"])
    # a + b of the two halves and addValues give the same statistics
    assert close(l2[0], e["sum"]) and close(l3[0], e["sum"])
    if not const:
        assert close(l2[1], e["var"], 1e-9) and close(l3[1], e["var"], 1e-9)
    c, cstd, cd = l4
    assert c == len(xs) and close(cstd, math.sqrt(len(xs))) and cd == len(xs) + 5


@pytest.mark.parametrize("name", list(LISTS))
def test_python_vector_stats(name):
    xs = np.asarray(LISTS[name], dtype=np.float64)
    n = len(xs)
    r = M.vector_stats((n, float(xs.sum()) if n else 0.0, float((xs * xs).sum()) if n else 0.0))
    e = expected(xs)
    assert r["count"] == n and close(r["sum"], e["sum"]) and close(r["mean"], e["mean"])
    if len(set(LISTS[name])) == 1 and n > 1:
        assert 0.0 <= r["variance"] <= 1e-12 * xs[0] ** 2
        assert isinstance(r["sum_std"], float) and r["sum_std"] >= 0
    else:
        assert close(r["variance"], e["var"], 1e-9)
        assert close(r["sum_std"], e["sum_std"], 1e-9) and close(r["mean_std"], e["mean_std"], 1e-9)


def test_moment_form_would_go_negative():
    """The clamp matters: for this constant list the raw-moment variance is
    negative in FP64, which made getSumSTD NaN before."""
    xs = np.full(3, 0.1)
    n, s, s2 = 3, float(xs.sum()), float((xs * xs).sum())
    raw = (s2 - s * s / n) / (n - 1)
    assert raw < 0
    assert M.vector_stats((n, s, s2))["variance"] == 0.0
