"""The tuned parameter table (host only): the file beside the library that
the engine reads at the start of every run, and the autotuner's candidate
lists it is written from (graspan_g/tune.py)."""
import os

from graspan_g import PKG_DIR
from graspan_g import tune


def read_table(path):
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if not line:
                continue
            name, value = line.split()
            out[name] = value
    return out


def test_table_names_are_the_tuners():
    t = read_table(os.path.join(PKG_DIR, "tuned_gfx950.cfg"))
    assert set(t) == set(tune.CANDIDATES)


def test_table_values_are_candidates():
    t = read_table(os.path.join(PKG_DIR, "tuned_gfx950.cfg"))
    for name, value in t.items():
        cands = tune.CANDIDATES[name]
        kind = type(cands[0])
        assert kind(value) in cands, (name, value, cands)


def test_start_setting_is_complete():
    assert set(tune.START) == set(tune.CANDIDATES)
    for name, v in tune.START.items():
        assert v in tune.CANDIDATES[name], (name, v)


def test_engine_parses_every_name():
    """Every tuned name has a parser entry in the engine and is in its list
    of GG_TUNE_<NAME> overrides (a name it did not know would be skipped
    without a word, and the run would use the default)."""
    src = open(os.path.join(PKG_DIR, "csrc", "closure.hip")).read()
    names = src[src.index("kTuningNames[] = {"):]
    names = names[:names.index("};")]
    for name in tune.CANDIDATES:
        assert f'name == "{name}"' in src, name
        assert f'"{name}"' in names, name
