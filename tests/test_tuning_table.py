"""The tuning table as the library reads it (CPU only: fmh_get_tuning needs
no device).  Batched launches of a class without a batched entry ("b L"
line) use the class's single-product geometry, not the built-in defaults:
the table tunes batches only up to 2^17, and groups of several long products
(the request loop's, or split products' pieces) must not fall back to an
untuned geometry."""

import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(os.path.dirname(HERE), "fast_mul_ntt_amd", "tuning_gfx950.txt")


def _table():
    single, batched = {}, {}
    with open(TABLE) as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] == "b":
                batched[int(parts[1])] = parts[2]
            else:
                single[int(parts[0])] = parts[1]
    return single, batched


def _as_dict(spec):
    return dict(kv.split("=") for kv in spec.split(","))


def test_table_entries_are_what_the_library_reads(fm):
    single, batched = _table()
    assert single and batched
    for L, spec in single.items():
        got = _as_dict(fm.get_tuning(L))
        assert all(got[k] == v for k, v in _as_dict(spec).items()), (L, spec, got)
    for L, spec in batched.items():
        got = _as_dict(fm.get_tuning(L, batched=True))
        assert all(got[k] == v for k, v in _as_dict(spec).items()), (L, spec, got)


def test_batched_launches_without_an_entry_use_the_single_geometry(fm):
    single, batched = _table()
    missing = [L for L in single if L not in batched]
    assert missing, "every class has a batched entry"
    for L in missing:
        assert fm.get_tuning(L, batched=True) == fm.get_tuning(L), L
