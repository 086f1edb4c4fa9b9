"""The tile autotuner's candidate list and cache keys (host logic, no GPU)."""
from sim_juncs_amd import autotune


def test_candidates_cover_forms_rows_chunks_and_z_passes():
    c = autotune.candidates((1024, 768, 512))
    forms = {f for f, _, _, _ in c}
    assert forms == {"register_rows", "thread_rows"}
    # 512-cell lines: 128- and 256-cell passes and one pass over the line, one row per thread
    zsegs = {z for f, r, _, z in c if f == "register_rows" and r == 1}
    assert zsegs == {0, 128, 256, 512}
    assert all(r == 1 for f, r, _, z in c if z)
    # a row-of-threads tile holds (rows + 1) z lines of 512 cells within 1024 threads: one row
    assert {r for f, r, _, _ in c if f == "thread_rows"} == {1}
    assert {xc for _, _, xc, _ in c} == set(autotune.XCHUNKS)


def test_candidates_respect_the_grid():
    c = autotune.candidates((8, 3, 64))
    # no x-chunk beyond nx except the smallest, no register rows beyond ny
    assert all(xc <= 8 or xc == autotune.XCHUNKS[0] for _, _, xc, _ in c)
    assert all(r <= 3 for f, r, _, _ in c if f == "register_rows")
    # short lines take no z passes; thread rows up to min(1024 // 64 - 1, ny) with the largest
    assert all(z == 0 for *_, z in c)
    rows = {r for f, r, _, _ in c if f == "thread_rows"}
    assert rows == {1, 2, 3}
    # lines above 1024 cells or below 512 get no pass candidates of their own
    assert all(z == 0 for *_, z in autotune.candidates((64, 64, 256)))


def test_cache_key_and_lookup():
    assert autotune.key((96, 96, 96), "fp64") == "96x96x96:fp64:"
    t = autotune.lookup((1024, 768, 512), "fp32")
    assert t is not None and t["form"] == "register_rows" and t["xchunk"] >= 1
