"""Host-side logic without a device: the C-ABI library loads and exports
every symbol include/dfrocm.h declares; the generator's numpy twin matches
SplitMix64's published output; the predicate compiler's truth tables agree
with the oracle's independent Kleene evaluation."""
import os
import re

import numpy as np
import pytest

from dataframe_mojo_amd import Column, DataFrame, Series, _lib, col, datagen, provider
from dataframe_mojo_amd.plan import Unsupported, compile_plan, compile_predicate
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "dfrocm.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dfr_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert names, "no declarations found"
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(_lib.EXPORTS) | set(names)
    assert lib.dfr_abi_version() == 1


def test_only_dfr_symbols_exported():
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.library_path()], capture_output=True, text=True).stdout
    syms = [line.split()[-1] for line in out.splitlines() if line.strip()]
    assert syms and all(s.startswith("dfr_") for s in syms), [s for s in syms if not s.startswith("dfr_")][:10]


def test_splitmix64_known_output():
    # SplitMix64 seeded with 0 first returns 0xE220A8397B1DCDAF
    assert int(datagen.mix64(np.array([0x9E3779B97F4A7C15], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF


def test_uniform_range_and_validity_rate():
    u = datagen.uniform(3, 1, 200000, -5, 5)
    assert u.min() == -5 and u.max() == 5
    v = datagen.validity_bits(3, 2, 200000, 0.05)
    assert abs((~v).mean() - 0.05) < 0.005
    cdf = datagen.zipf_cdf(1.1, 1000)
    z = datagen.zipf(3, 3, 100000, cdf)
    assert z.min() >= 1 and z.max() <= 1000 and (z == 1).mean() > 0.1


def _frame(n=4000, seed=1):
    rng = np.random.default_rng(seed)
    return DataFrame([
        Series("a", Column.int64(rng.integers(-20, 20, n), rng.random(n) > 0.1)),
        Series("b", Column.int32(rng.integers(-9, 9, n))),
        Series("x", Column.float64(rng.normal(size=n), rng.random(n) > 0.05)),
        Series("s", Column.string([["p", "q", None, "r"][i % 4] for i in range(n)])),
    ])


PREDICATES = [
    col("a") > 0,
    ~(col("a") > 0),
    (col("a") > 0) & (col("b") <= 3),
    (col("a") > 0) | ~(col("b") == 2),
    ~((col("a") % 3 != 0) | (col("x") < 0.5)),
    col("x").is_null() | (col("s") == "q"),
    (col("s") != "r") & col("a").is_not_null(),
    (col("a") % -4 == -1) & (col("x") >= -1.0) & (col("b") != 0) & ~(col("s") == "zzz"),
    0 < col("a"),
    col("a").is_in([1, -3, None, 2.5, 7.0]),
    ~col("a").is_in([0, 5]) | col("s").is_in(["q"]),
    ~col("a").is_in([None]),
    col("x").is_between(-0.5, 1.0, closed="right") & col("b").is_between(-2, 4),
    ~col("b").is_between(None, 3),
    col("a") > col("b"),
    ~(col("b") <= col("a")) | (col("x") == col("x")),
    (col("a") % 3 == 1) & (col("b") != col("a")),
]


@pytest.mark.parametrize("i", range(len(PREDICATES)))
def test_truth_tables_match_kleene_evaluation(i):
    df = _frame()
    prog = compile_predicate(df, [PREDICATES[i]])
    n = df.height()
    bits = np.zeros(n, dtype=np.int64)
    for j, leaf in enumerate(prog.leaves):
        c = df.column(leaf.column).column
        v, m = c.values, c.is_valid()
        if leaf.op in (_lib.IS_NULL, _lib.IS_NOT_NULL):
            t = ~m if leaf.op == _lib.IS_NULL else m
        else:
            lit = leaf.literal
            if c.dtype == "float64":
                lit = np.int64(lit).view(np.float64)
            if leaf.flags & _lib.LEAF_COLUMN:
                other = df.column(leaf.column2).column
                lit = other.values if c.dtype == "float64" else other.values.astype(np.int64)
            x = v.astype(np.int64) if c.dtype != "float64" else v
            ok = m.copy()
            if leaf.flags & _lib.LEAF_COLUMN:
                ok &= df.column(leaf.column2).column.is_valid()
            if leaf.flags & _lib.LEAF_MOD:
                ok &= leaf.modulus != 0
                x = np.remainder(x, leaf.modulus if leaf.modulus else 1)
            with np.errstate(invalid="ignore"):
                t = [np.equal, np.not_equal, np.less, np.less_equal, np.greater, np.greater_equal][leaf.op](x, lit)
            if leaf.flags & _lib.LEAF_NOT:
                t = ~t
            t = t & ok
        bits |= t.astype(np.int64) << j
    keep = np.array([(prog.truth[m >> 5] >> (m & 31)) & 1 for m in bits], dtype=bool)
    assert (keep == oracle.kept_rows(df, [PREDICATES[i]])).all()


def test_unsupported_plans_raise_before_submission():
    df = _frame(10)
    with pytest.raises(Unsupported):
        compile_plan(df.lazy().group_by("x", "a").agg(col("a").sum()))  # Float64 key beside another
    with pytest.raises(Unsupported):
        compile_plan(df.filter(col("x") > col("a")).group_by("a").agg(col("b").sum()))  # Float64 with Int64
    with pytest.raises(Unsupported):
        compile_plan(df.filter(col("s") == col("s")).group_by("a").agg(col("b").sum()))
    compile_plan(df.filter(col("a") > col("b")).group_by("a").agg(col("b").sum()))  # two integer columns
    with pytest.raises(Unsupported):
        compile_plan(df.lazy().group_by("a").agg(col("s").sum()))
    compile_plan(df.lazy().group_by("x").agg(col("a").sum()))  # a Float64 key alone is on the path
    assert provider.pipeline_supports(df.lazy().group_by("a").agg(col("b").sum()), 0)
    assert not provider.pipeline_supports(df.lazy().group_by("x", "a").agg(col("b").sum()), 0)
    assert provider.describe(df.lazy().group_by("x", "a").agg(col("b").sum())).startswith("ENGINE accel: not on")


def test_provider_boundary_without_device():
    assert provider.handoff_supports(_frame(4).lazy()) is False
    assert provider.handoff_estimate(_frame(4).lazy(), 10) == ("", -1.0)
    tm = provider.transfer_model(0)
    assert isinstance(tm, provider.TransferModel)
    with pytest.raises(RuntimeError):
        provider.handoff_execute(_frame(4).lazy())


def test_memory_limit_env_rejects_bad_values(monkeypatch):
    from dataframe_mojo_amd import runtime

    monkeypatch.setenv("DATAFRAME_ACCEL_MEMORY_LIMIT", "lots")
    with pytest.raises(ValueError):
        runtime.memory_limit()
    monkeypatch.setenv("DATAFRAME_ACCEL_MEMORY_LIMIT", "1048576")
    assert runtime.memory_limit() == 1048576


def _jitc():
    exe = os.path.join(ROOT, "dataframe_mojo_amd", "dfr_jitc")
    if not os.path.exists(exe):
        pytest.skip("dfr_jitc is not built")
    return exe


JITC_OPTS = ["--offload-arch=gfx950", "-O3", "-std=c++17"]


def test_runtime_compiler_program_writes_a_code_object(tmp_path):
    """dfr_jitc, which the library runs as a child process for each runtime
    compile, turns a kernel source into a code object (exit 0), and reports
    a source that does not compile with exit 2 and the compiler's log."""
    import subprocess

    exe = _jitc()
    src = tmp_path / "k.hip"
    src.write_text('extern "C" __global__ void k(int *p) { p[threadIdx.x] = threadIdx.x; }\n')
    out = tmp_path / "k.co"
    r = subprocess.run([exe, str(src), str(out), *JITC_OPTS], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert out.stat().st_size > 0 and out.read_bytes()[:4] == b"\x7fELF"
    bad = tmp_path / "bad.hip"
    bad.write_text('extern "C" __global__ void k(int *p) { p[threadIdx.x] = undefined_name; }\n')
    r = subprocess.run([exe, str(bad), str(tmp_path / "bad.co"), *JITC_OPTS], capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "undefined_name" in r.stderr


def test_mojo_provider_binds_only_declared_symbols_and_constants():
    """The Mojo provider package (uncompiled here: no Mojo toolchain) names
    only entry points include/dfrocm.h declares, and its status, type,
    leaf, aggregate and buffer constants equal the header's."""
    pkg = os.path.join(ROOT, "providers", "rocm", "dataframe_provider")
    text = ""
    for name in sorted(os.listdir(pkg)):
        if name.endswith(".mojo"):
            with open(os.path.join(pkg, name)) as f:
                text += f.read()
    declared = set(declared_symbols())
    named = set(re.findall(r'"(dfr_[a-z_0-9]+)"', text))
    assert named and named <= declared, sorted(named - declared)
    with open(os.path.join(ROOT, "include", "dfrocm.h")) as f:
        header = f.read()
    defines = dict((k, int(v)) for k, v in re.findall(r"#define (DFR_[A-Z_0-9]+) (-?\d+)", header))
    consts = dict((k, int(v)) for k, v in re.findall(r"comptime (DFR_[A-Z_0-9]+) = (-?\d+)", text))
    assert len(consts) >= 30
    wrong = {k: (v, defines.get(k)) for k, v in consts.items() if defines.get(k) != v}
    assert not wrong, wrong
    # the 13 provider functions of providers/cpu/dataframe_provider/__init__.mojo
    for fn in ["installed", "execute", "describe", "execute_profiled", "handoff_supports", "handoff_execute",
               "handoff_estimate", "devices", "transfer_model", "pipeline_supports", "pipeline_estimate",
               "pipeline_costs", "run_pipeline"]:
        assert re.search(rf"^def {fn}\(", text, flags=re.M), fn
