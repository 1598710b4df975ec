"""Pass wisdom files without a GPU: bfpsh_wisdom_import / export /
forget (the counterparts of fftwf_import_wisdom_from_filename and
fftwf_export_wisdom_to_filename, bfps/cpp/full_code/main_code.hpp:101-104,
145-149) touch only the host-side wisdom table, so they run here."""
import os

import pytest


@pytest.fixture()
def B():
    import bfps_amd
    bfps_amd.capi.load()
    bfps_amd.wisdom_forget()
    yield bfps_amd
    bfps_amd.wisdom_forget()


LINES = ["gfx950/f32/1024/y_inv 16 16 0 2.460000 1",
         "gfx950/f32/1024/z_fwd 32 8 1 1.950000 2",
         "gfx950/f64/512/z_inv 8 4 0 0.812500 0"]


def read_entries(path):
    return sorted(l.strip() for l in open(path) if l.strip() and not l.startswith("#"))


def test_round_trip(B, tmp_path):
    src = tmp_path / "in.txt"
    src.write_text("# comment line\n" + "\n".join(LINES) + "\n")
    B.wisdom_import(str(src))
    out = tmp_path / "out.txt"
    B.wisdom_export(str(out))
    assert read_entries(out) == sorted(LINES)
    # a later import replaces entries of the same key and keeps the others
    upd = tmp_path / "upd.txt"
    upd.write_text("gfx950/f32/1024/y_inv 8 8 1 2.000000 2\n")
    B.wisdom_import(str(upd))
    B.wisdom_export(str(out))
    want = sorted(["gfx950/f32/1024/y_inv 8 8 1 2.000000 2"] + LINES[1:])
    assert read_entries(out) == want


def test_forget_and_missing_file(B, tmp_path):
    src = tmp_path / "in.txt"
    src.write_text(LINES[0] + "\n")
    B.wisdom_import(str(src))
    B.wisdom_forget()
    out = tmp_path / "out.txt"
    B.wisdom_export(str(out))
    assert read_entries(out) == []
    # a missing file is the start of a new run, not an error
    B.wisdom_import(str(tmp_path / "none.txt"))


@pytest.mark.parametrize("bad", ["gfx950/f32/1024/y_inv 16 16 0 1.0\n",     # too few fields
                                 "gfx950/f32/1024/y_inv 0 16 0 1.0 1\n",    # no shape
                                 "gfx950/f32/1024/y_inv 16 16 2 1.0 1\n",   # persistent neither 0 nor 1
                                 "gfx950/f32/1024/y_inv 16 16 0 1.0 7\n"])  # unknown rigor
def test_bad_lines_rejected(B, tmp_path, bad):
    src = tmp_path / "bad.txt"
    src.write_text(LINES[1] + "\n" + bad)
    with pytest.raises(B.capi.BfpshError):
        B.wisdom_import(str(src))
    # nothing of a rejected file is kept
    out = tmp_path / "out.txt"
    B.wisdom_export(str(out))
    assert read_entries(out) == []
