"""Text embedding I/O of the host layer (gvizEmbeddedGraphSaveEmbedding /
LoadEmbedding) against files and results the reference library produced
(src/embedders/gvizEmbeddedGraph.c:428-485; fixtures written by
tests/golden/make_golden.py from the compiled reference).

Save must write the reference's bytes exactly ("%f " per coordinate, so
rounding at the sixth decimal, tiny values, negative zero and large values
print the same). Load must return what the reference returns on the
reference's own file and on edited ones (scientific notation, tabs and
blank lines, a name line longer than the reference's 256-byte buffer, a
wrong vertex count or dimension, a truncated list, a missing file), with the
same positions afterwards (a failed load keeps the coordinates it parsed
before failing, as the reference does). CPU only."""
import os

import numpy as np
import pytest

from conftest import ROOT, golden
import refio

DRIVER = os.path.join(ROOT, "tests", "host", "api_driver")


@pytest.mark.parametrize("name", ["embedding_io_d2", "embedding_io_d3"])
def test_embedding_text_io_matches_reference(tmp_path, name):
    f = golden(name)
    n, dim = int(f["n"]), int(f["dim"])
    names = [str(s) for s in f["names"]]
    paths = []
    for k, nm in enumerate(names):
        p = tmp_path / f"{nm}.txt"
        if nm != "missing":
            p.write_bytes(f[f"t{k}"].tobytes())
        paths.append(str(p))
    rc, saved, loads = refio.embedding_io(n, f["edges"], dim, f["pos"], "layout",
                                          str(tmp_path / "mine.txt"), paths, exe=DRIVER)
    assert rc == int(f["save_rc"]) == 0
    assert saved == f["saved"].tobytes()
    for k, nm in enumerate(names):
        lrc, lpos = loads[k]
        assert lrc == int(f[f"l{k}_rc"]), nm
        np.testing.assert_array_equal(lpos, f[f"l{k}_pos"], err_msg=nm)
